"""Host side of the serving loop: the clients' capped moves keep every tet (hence every
tet-split hex) positively oriented; the node averaging reproduces the mesh generator."""
import numpy as np

from femocs_amd import mesh as M
from femocs_amd import serving


def _tet_volumes(x, tets):
    P = x[tets]
    return np.einsum("ij,ij->i", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]))


def _center_dets(x, cells_ucd):
    lex = cells_ucd[:, M.UCD_TO_LEX]
    p = x[lex]
    J = np.stack([0.25 * ((p[:, 1] - p[:, 0]) + (p[:, 3] - p[:, 2]) + (p[:, 5] - p[:, 4]) + (p[:, 7] - p[:, 6])),
                  0.25 * ((p[:, 2] - p[:, 0]) + (p[:, 3] - p[:, 1]) + (p[:, 6] - p[:, 4]) + (p[:, 7] - p[:, 5])),
                  0.25 * ((p[:, 4] - p[:, 0]) + (p[:, 5] - p[:, 1]) + (p[:, 6] - p[:, 2]) + (p[:, 7] - p[:, 3]))], 2)
    return np.linalg.det(J)


def test_hex_node_weights_rebuild_nodes():
    m = M.make_tip_mesh(M.small_geometry())
    W = M.hex_node_weights(m.n_tetnodes, m.tets)
    X = W @ m.nodes[: m.n_tetnodes]
    assert np.abs(X - m.nodes).max() <= 1e-12 * np.abs(m.nodes).max()


def test_altitudes_of_a_regular_tet():
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    h = serving.vertex_altitudes(x, np.array([[0, 1, 2, 3]]))
    assert np.allclose(h, [1 / np.sqrt(3), 1, 1, 1])


def test_capped_moves_keep_orientation():
    m = M.make_tip_mesh(M.small_geometry())
    c = serving.Client(0, m, sigma=0.5, seed=3)          # 5x the MD noise
    v0 = _tet_volumes(m.nodes, m.tets)
    d0 = _center_dets(m.nodes, m.hexes)
    assert np.all(d0 * np.sign(d0) > 0)
    for _ in range(5):
        X = c.next_positions()
        v = _tet_volumes(X, m.tets)
        assert np.all(v * np.sign(v0) > 0)
        d = _center_dets(X, m.hexes)
        assert np.all(np.sign(d) == np.sign(d0))
    # the uncapped noise does invert some cells of this mesh
    c.limit[:] = np.inf
    flips = 0
    for _ in range(5):
        X = c.next_positions()
        flips += int(np.sum(np.sign(_center_dets(X, m.hexes)) != np.sign(d0)))
    assert flips > 0

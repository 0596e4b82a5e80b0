"""Synthetic mesh generator and the tet -> hex split (src/Tethex.cpp:965-1102 layout)."""
import numpy as np
import pytest

from femocs_amd import mesh as M


@pytest.fixture(scope="module")
def tip():
    return M.make_tip_mesh(M.small_geometry())


def test_split_structure(tip):
    T = tip.n_tets
    assert tip.n_hexes == 4 * T
    kinds = tip.node_kind[tip.hexes]                      # (4T, 8)
    # each hex: its tet vertex, 3 edge midpoints, 3 face centroids and the tet centroid
    for k, cnt in ((M.KIND_TETNODE, 1), (M.KIND_EDGE, 3), (M.KIND_FACE, 3), (M.KIND_TET, 1)):
        assert np.all((kinds == k).sum(axis=1) == cnt)
    # hex 4t+k surrounds vertex k of tet t and carries +-(1 + that vertex)
    assert np.array_equal(tip.hexes[:, 0], tip.tets.reshape(-1))
    v = tip.tets.reshape(-1)
    assert np.array_equal(np.abs(tip.hex_marker), 1 + v)
    assert np.array_equal(tip.hex_marker > 0, np.repeat(tip.tet_marker == M.TYPES_VACUUM, 4))


def test_new_nodes_are_averages(tip):
    x = tip.nodes
    t = tip.tets[0]
    cen = x[t].mean(0)
    assert np.allclose(x[tip.hexes[0, 6]], cen, rtol=0, atol=1e-12)


def test_orientation_and_neighbours(tip):
    x = tip.nodes[tip.hexes][:, M.UCD_TO_LEX]
    det = np.einsum("ij,ij->i", x[:, 1] - x[:, 0], np.cross(x[:, 2] - x[:, 0], x[:, 4] - x[:, 0]))
    assert np.all(det > 0)
    nbr = tip.tet_nbr
    for t in range(0, tip.n_tets, 97):
        for k in range(4):
            o = nbr[t, k]
            if o >= 0:
                assert t in nbr[o]
                shared = set(tip.tets[t]) & set(tip.tets[o])
                assert len(shared) == 3 and tip.tets[t][k] not in shared


def test_box_boundary_is_exact(tip):
    lo, hi = tip.geometry.box()
    assert np.allclose(tip.nodes.min(0), lo) and np.allclose(tip.nodes.max(0), hi)


def test_surface_atoms(tip):
    a = tip.atoms
    g = tip.geometry
    sd = M.signed_distance(g, a)
    assert len(a) > 100 and np.all(sd < 0.1) and np.all(sd > -g.lattice)

"""The oracle's restatement of the reference's tet -> 4-hex split (Tethex Mesh::convert_3D,
src/Tethex.cpp:965-1169) pinned by its own definitions: the same hexes as the generator's
split up to the numbering, edges numbered by (larger, smaller) vertex, faces by first
encounter, every hex positively oriented, quadrangles of the boundary triangles."""
import numpy as np
import pytest

from femocs_amd import mesh as M
from oracle import tethex

LEX = [0, 1, 4, 5, 3, 2, 7, 6]


@pytest.fixture(scope="module")
def split():
    tip = M.make_tip_mesh(M.small_geometry(), with_atoms=False)
    nodes = tip.nodes[:tip.n_tetnodes]
    tris, _ = tip.surface_triangles()
    return tip, nodes, tethex.split(nodes, tip.tets, tris), tris


def test_same_nodes_and_hexes_as_the_generator(split):
    tip, nodes, (out, kind, hexes, quads, E, F), _ = split
    assert len(out) == len(tip.nodes) and len(hexes) == len(tip.hexes)
    # identical coordinates (same sums), so nodes map one to one by value
    a = {tuple(p): i for i, p in enumerate(tip.nodes)}
    assert len(a) == len(tip.nodes)
    perm = np.array([a[tuple(p)] for p in out])
    assert np.array_equal(np.sort(perm), np.arange(len(out)))
    assert np.array_equal(np.sort(perm[hexes], axis=1), np.sort(tip.hexes, axis=1))
    assert np.array_equal(np.bincount(kind), [0, len(nodes), E, F, len(tip.tets)])


def test_edge_and_face_numbering(split):
    tip, nodes, (out, kind, hexes, quads, E, F), _ = split
    nv = len(nodes)
    pairs = set()
    for t in tip.tets:
        for i in range(4):
            for j in range(i + 1, 4):
                pairs.add((max(t[i], t[j]), min(t[i], t[j])))
    edges = sorted(pairs)
    assert len(edges) == E
    mid = np.array([(nodes[lo] + nodes[hi]) / 2 for hi, lo in edges])
    assert np.array_equal(out[nv:nv + E], mid)
    # faces: first appearance over (tet, local edge with v_i > v_j, opposite vertex)
    seen, order = set(), []
    for t in tip.tets:
        for i in range(4):
            for j in range(4):
                if t[i] <= t[j]:
                    continue
                for w in t:
                    if w in (t[i], t[j]):
                        continue
                    key = tuple(sorted((t[i], t[j], w)))
                    if key not in seen:
                        seen.add(key)
                        order.append(key)
    assert len(order) == F
    cen = np.array([(nodes[a] + nodes[b] + nodes[c]) / 3 for a, b, c in order])
    assert np.allclose(out[nv + E:nv + E + F], cen, rtol=0, atol=1e-12)


def test_hexes_oriented_and_laid_out_around_their_vertex(split):
    tip, nodes, (out, kind, hexes, quads, E, F), _ = split
    x = out[hexes[:, LEX]]
    d0 = 0.25 * ((x[:, 1] - x[:, 0]) + (x[:, 3] - x[:, 2]) + (x[:, 5] - x[:, 4]) + (x[:, 7] - x[:, 6]))
    d1 = 0.25 * ((x[:, 2] - x[:, 0]) + (x[:, 3] - x[:, 1]) + (x[:, 6] - x[:, 4]) + (x[:, 7] - x[:, 5]))
    d2 = 0.25 * ((x[:, 4] - x[:, 0]) + (x[:, 5] - x[:, 1]) + (x[:, 6] - x[:, 2]) + (x[:, 7] - x[:, 3]))
    assert np.all(np.einsum("ij,ij->i", d0, np.cross(d1, d2)) > 0)
    # hex 4t + k holds tet vertex k, and the front/back swap happens for part of them
    v = tip.tets.reshape(-1)
    front = hexes[:, 0] == v
    back = hexes[:, 4] == v
    assert np.all(front | back) and front.any() and back.any()


def test_boundary_quadrangles(split):
    tip, nodes, (out, kind, hexes, quads, E, F), tris = split
    assert len(quads) == 3 * len(tris)
    for n, tr in enumerate(tris[:200]):
        for k in range(3):
            q = quads[3 * n + k]
            assert q[0] == tr[k] and kind[q[2]] == 3 and kind[q[1]] == 2 and kind[q[3]] == 2
            assert np.allclose(out[q[2]], nodes[tr].mean(0), atol=1e-12)
            xy = out[[q[0], q[1], q[3], q[2]], :2]
            assert tethex._measure_2d(xy) >= 0 or np.isclose(tethex._measure_2d(xy), 0.0, atol=1e-9)


def _volume_lex(x):
    """Exact volume of a trilinear hex (lexicographic 8 x 3): det J is of degree <= 2 in each
    reference coordinate, so QGauss(2) integrates it exactly.  This is the quantity the
    reference's cell_measure_3D (src/Tethex.cpp:1603-1630) evaluates in closed form."""
    g = (0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0))
    vol = 0.0
    for z in g:
        for y in g:
            for xx in g:
                J = np.zeros((3, 3))
                for n in range(8):
                    a, b, c = n & 1, (n >> 1) & 1, n >> 2
                    N = (xx if a else 1 - xx, y if b else 1 - y, z if c else 1 - z)
                    d = (1 if a else -1, 1 if b else -1, 1 if c else -1)
                    J += np.outer(x[n], [d[0] * N[1] * N[2], N[0] * d[1] * N[2], N[0] * N[1] * d[2]])
                vol += np.linalg.det(J) / 8.0
    return vol


def test_orientation_predicate_agrees_with_cell_measure():
    """The split orients each hex by the sign of its centre Jacobian; the reference swaps the
    front and back faces when cell_measure_3D < 0.  A sub-hex of the split is the image of one
    reference sub-hex under the affine map of its tet, so both signs are sign(det A) for every
    tet, slivers included: checked on random tets, flat slivers (height 1e-7 of the edge) and
    needles, before (both orders) and after the orientation fix."""
    rng = np.random.default_rng(17)
    tets_x = [rng.normal(size=(4, 3)) for _ in range(40)]
    for h in (1e-3, 1e-5, 1e-7):
        base = rng.normal(size=(3, 3))
        base[:, 2] = 0.0
        apex = base.mean(0) + [0.1, -0.2, h]
        tets_x.append(np.vstack([base, apex]))                  # sliver: apex near the base plane
        tets_x.append(np.array([[0, 0, 0], [1, h, 0], [1, 0, h], [1, h, h]], float))   # needle
    for P in tets_x:
        nodes, kind, hexes, _, _, _ = tethex.split(P, np.array([[0, 1, 2, 3]]))
        for hx in hexes:
            x = nodes[hx[LEX]]
            c, v = tethex._center_det_lex(x), _volume_lex(x)
            assert c > 0 and v > 0
            xf = x[[4, 5, 6, 7, 0, 1, 2, 3]]                     # front and back swapped
            assert tethex._center_det_lex(xf) < 0 and _volume_lex(xf) < 0

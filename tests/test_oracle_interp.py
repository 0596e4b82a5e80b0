"""The oracle's transfer-to-atoms routines pinned by what the mathematics fixes: quadratic tets
whose edge nodes are the edge midpoints, exact reproduction of linear (ranks 1, 3) and quadratic
(rank 2) fields, nodal fields of linear potentials, smoothing that keeps constants and stays in
the neighbours' hull, and the located cells of the chain."""
import numpy as np
import pytest

from femocs_amd import mesh as M
from oracle import oracle as O

A = np.array([0.3, -0.7, 1.1])


@pytest.fixture(scope="module")
def tip():
    m = M.make_tip_mesh(M.small_geometry(), with_atoms=False)
    L = O.LinTet(m.nodes, m.tets, m.tet_nbr, m.tet_marker, 3)
    q10 = O.quadtet_cells(m.tets, m.hexes, m.node_kind)
    return m, L, q10


def _inside_points(m, n, seed, region=3):
    rng = np.random.default_rng(seed)
    ts = rng.choice(np.nonzero(m.tet_marker == region)[0], n)
    b = rng.dirichlet(np.ones(4) * 2.0, n)
    return np.einsum("nk,nkd->nd", b, m.nodes[m.tets[ts]]), ts


def test_quadtet_edge_nodes_are_midpoints(tip):
    m, _, q10 = tip
    x = m.nodes
    pairs = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    assert np.array_equal(q10[:, :4], m.tets)
    for k, (i, j) in enumerate(pairs):
        mid = 0.5 * (x[m.tets[:, i]] + x[m.tets[:, j]])
        assert np.array_equal(x[q10[:, 4 + k]], mid)
        assert np.all(m.node_kind[q10[:, 4 + k]] == M.KIND_EDGE)


def _solution(x, f):
    v = f(x)
    return np.column_stack([v, 2 * v, -v, 0.5 * v, v])


@pytest.mark.parametrize("rank", [1, 2, 3])
def test_interpolation_reproduces_polynomials(tip, rank):
    """Linear fields are exact for every rank; rank 2 is exact for quadratics too."""
    m, L, q10 = tip
    pts, _ = _inside_points(m, 300, 11 + rank)
    lin = lambda x: x @ A + 0.25
    quad = lambda x: 1e-3 * (x[..., 0] ** 2 - 2 * x[..., 1] * x[..., 2]) + x @ A
    for f, exact in ((lin, True), (quad, rank == 2)):
        sol = _solution(m.nodes, f)
        cells, out = O.interp_points(L, m.nodes, m.tets, q10, m.hexes, sol, pts, rank)
        ref = _solution(pts, f)
        err = np.abs(out - ref).max() / np.abs(ref).max()
        if exact:
            assert err < 1e-11, (rank, err)
        else:
            assert err > 1e-9          # the quadratic is not in the linear spaces
        assert np.all(cells > 0)
        # the cached path interpolates in the same cells to the same bits
        assert np.array_equal(O.interp_cells(L, m.nodes, m.tets, q10, m.hexes, sol, pts, rank, cells), out)


def test_interpolation_cells_follow_the_locate_chain(tip):
    m, L, q10 = tip
    rng = np.random.default_rng(5)
    lo, hi = m.nodes.min(0), m.nodes.max(0)
    pts = np.concatenate([_inside_points(m, 200, 3)[0], rng.uniform(lo, hi, (100, 3))])
    sol = _solution(m.nodes, lambda x: x @ A)
    for rank in (1, 2, 3):
        cells, _ = O.interp_points(L, m.nodes, m.tets, q10, m.hexes, sol, pts, rank)
        assert np.array_equal(cells, L.locate_chain(pts, rank=3 if rank == 3 else 1))


def test_extract_nodal_linear_potential(tip):
    m, _, _ = tip
    used = np.zeros(len(m.nodes), np.uint8)
    used[m.hexes[m.hex_marker > 0]] = 1
    phi = m.nodes @ A
    sol = O.extract_nodal(m.nodes, m.hexes, m.hex_marker, 1, used, phi, empty=-7.0)
    u = used > 0
    assert np.abs(sol[u, :3] + A).max() < 1e-10
    assert np.array_equal(sol[u, 4], phi[u]) and np.all(sol[u, 3] == 0)
    assert np.all(sol[~u] == -7.0)


def test_smoothing_keeps_constants_and_hull(tip):
    m, _, _ = tip
    used = np.zeros(len(m.nodes), np.uint8)
    used[m.hexes[m.hex_marker > 0]] = 1
    decay = O.decay_factor(m.nodes, m.tets)
    e = m.nodes[m.tets][:, [0, 0, 0, 1, 1, 2]] - m.nodes[m.tets][:, [1, 2, 3, 2, 3, 3]]
    assert decay == -1.0 / np.sqrt((e ** 2).sum(-1).max())
    sol = np.zeros((len(m.nodes), 5))
    sol[:, :3] = [1.5, -2.0, 0.25]
    out = O.smooth_nodal(sol, m.nodes, m.n_tetnodes, m.node_kind, m.hexes, m.hex_marker, True, decay)
    assert np.abs(out[:, :3] - sol[:, :3]).max() < 1e-14
    rng = np.random.default_rng(1)
    sol[:, :3] = rng.normal(size=(len(m.nodes), 3))
    out = O.smooth_nodal(sol, m.nodes, m.n_tetnodes, m.node_kind, m.hexes, m.hex_marker, True, decay)
    # only tet nodes with a region hex change; each is a convex combination of hex nodes
    # of kind >= 2 tied to it
    changed = np.nonzero(np.any(out != sol, axis=1))[0]
    assert len(changed) > 0 and changed.max() < m.n_tetnodes
    vac = m.hex_marker > 0
    owner = m.hex_marker[vac] - 1
    for i in changed[:200]:
        nb = np.unique(m.hexes[vac][owner == i].ravel())
        nb = nb[(nb != i) & (m.node_kind[nb] >= 2)]
        lo, hi = sol[nb, :3].min(0), sol[nb, :3].max(0)
        assert np.all(out[i, :3] >= lo - 1e-12) and np.all(out[i, :3] <= hi + 1e-12)
    assert np.array_equal(out[:, 3:], sol[:, 3:])

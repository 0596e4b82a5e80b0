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


@pytest.fixture(scope="module")
def surface(tip):
    m, L, q10 = tip
    tris, t2t = m.surface_triangles()
    return tris, t2t, O.LinTri(m.nodes, tris, t2t)


def test_surface_triangles_separate_the_regions(tip, surface):
    m, _, _ = tip
    tris, t2t, T = surface
    assert len(tris) > 100
    assert np.all(m.tet_marker[t2t[:, 0]] == M.TYPES_VACUUM) and np.all(m.tet_marker[t2t[:, 1]] != M.TYPES_VACUUM)
    for k in range(3):      # every triangle node is a vertex of both tets
        assert np.all((m.tets[t2t[:, 0]] == tris[:, k, None]).any(1))
        assert np.all((m.tets[t2t[:, 1]] == tris[:, k, None]).any(1))


def test_triangle_weights_and_distance(tip, surface):
    """On the triangle: barycentric weights (+1e-15 each) that reproduce the point, zero
    distance, containment; off the plane: the distance along the normal times |e2|^2 ... sign."""
    m, _, _ = tip
    tris, _, T = surface
    rng = np.random.default_rng(3)
    for f in rng.choice(len(tris), 50, replace=False):
        b = rng.dirichlet(np.ones(3))
        x = m.nodes[tris[f]]
        p = b @ x
        w = T.shape_functions(p, f)
        assert abs(w.sum() - 1.0) < 1e-14 and np.allclose(w - 1e-15, b, atol=1e-10)
        assert abs(T.fast_distance(p, f)) < 1e-9 * np.abs(x).max() ** 2
        assert T.point_in_cell(p, f)
        n = np.cross(x[1] - x[0], x[2] - x[0])
        far = p + 3.0 * np.linalg.norm(x[2] - x[0]) * n / np.linalg.norm(n)
        assert not T.point_in_cell(far, f)


def test_triangle_locate_against_brute_force(tip, surface):
    m, _, _ = tip
    tris, _, T = surface
    rng = np.random.default_rng(4)
    f = rng.choice(len(tris), 60)
    pts = np.einsum("nk,nkd->nd", rng.dirichlet(np.ones(3), 60), m.nodes[tris[f]])
    pts = np.concatenate([pts, pts[:20] + rng.normal(0, 5.0, (20, 3))])
    for p in pts:
        c = T.locate(p, -1)
        inside = [t for t in range(T.n) if T.point_in_cell(p, t)]
        if inside:
            assert c == inside[0]
        else:
            d = ((T.centroid - p) ** 2).sum(1)
            assert c == -int(np.argmin(d))


def test_surface_interpolation_of_linear_fields(tip, surface):
    """Atoms on the surface get the linear field exactly (rank 1 in the tet chosen)."""
    m, L, q10 = tip
    tris, t2t, T = surface
    rng = np.random.default_rng(6)
    f = rng.choice(len(tris), 200)
    pts = np.einsum("nk,nkd->nd", rng.dirichlet(np.ones(3), 200), m.nodes[tris[f]])
    pts = pts + rng.normal(0, 0.05, pts.shape)
    sol = _solution(m.nodes, lambda x: x @ A + 0.25)
    cells, out = T.interp(L, m.tets, sol, pts)
    ref = _solution(pts, lambda x: x @ A + 0.25)
    assert np.abs(out - ref).max() <= 1e-10 * np.abs(ref).max()
    assert np.array_equal(T.interp(L, m.tets, sol, pts, cells)[1], out)


@pytest.fixture(scope="module")
def surface_quads(tip):
    m, L, q10 = tip
    tris, t2t = m.surface_triangles()
    quads = m.surface_quads(tris)
    T = O.LinTri(m.nodes, tris, t2t, quads, m.node_kind, m.tets)
    return tris, t2t, quads, T


def test_surface_quads_and_quadratic_faces(tip, surface_quads):
    """Quads: vertex, edge midpoint, face centroid, edge midpoint, all in both hexes mapped to
    them; the 6-node faces carry the edge midpoints (0,1), (1,2), (2,0)."""
    m, _, _ = tip
    tris, t2t, quads, T = surface_quads
    x = m.nodes
    assert np.array_equal(m.node_kind[quads], np.tile([M.KIND_TETNODE, M.KIND_EDGE, M.KIND_FACE, M.KIND_EDGE],
                                                      (len(quads), 1)))
    assert np.array_equal(quads[:, 0], tris.ravel())
    assert np.allclose(x[quads[:, 2]], x[np.repeat(tris, 3, 0)].mean(1), atol=1e-12)
    for r in range(2):
        hx = m.hexes[T.quad2hex[:, r]]
        assert np.all((hx[:, :, None] == quads[:, None, :]).any(1))
        assert np.array_equal(T.quad2hex[:, r] // 4, np.repeat(t2t[:, r], 3))
    for k, (a, b) in enumerate([(0, 1), (1, 2), (2, 0)]):
        assert np.array_equal(x[T.tri6[:, 3 + k]], 0.5 * (x[tris[:, a]] + x[tris[:, b]]))
    assert np.array_equal(T.quad_centroid, x[quads].sum(1) * 0.25)


@pytest.mark.parametrize("rank", [2, 3])
def test_surface_interpolation_higher_ranks(tip, surface_quads, rank):
    """Rank 2 reproduces quadratics and rank 3 linear fields at atoms near the surface; the
    located quads are the faces' sub-quads of the largest barycentric weight."""
    m, L, q10 = tip
    tris, t2t, quads, T = surface_quads
    rng = np.random.default_rng(20 + rank)
    f = rng.choice(len(tris), 200)
    pts = np.einsum("nk,nkd->nd", rng.dirichlet(np.ones(3), 200), m.nodes[tris[f]])
    pts = pts + rng.normal(0, 0.05, pts.shape)
    fn = (lambda x: 1e-3 * (x[..., 0] ** 2 - 2 * x[..., 1] * x[..., 2]) + x @ A) if rank == 2 else \
        (lambda x: x @ A + 0.25)
    sol = _solution(m.nodes, fn)
    kw = dict(rank=rank, xyz=m.nodes, cells10=q10, hexes=m.hexes)
    cells, out = T.interp(L, m.tets, sol, pts, **kw)
    ref = _solution(pts, fn)
    assert np.abs(out - ref).max() <= 1e-10 * np.abs(ref).max()
    assert np.array_equal(T.interp(L, m.tets, sol, pts, cells, **kw)[1], out)
    if rank == 3:
        prev = -1
        for p, c in zip(pts, cells):
            t = T.locate(p, abs(prev) // 3)
            w = T.shape_functions(p, abs(t))
            assert c == np.sign(t if t else 1) * (3 * abs(t) + int(np.argmax(w)))
            prev = c


@pytest.mark.parametrize("rank", [1, 2, 3])
def test_surface_centroid_mode(tip, surface_quads, rank):
    """Atoms on (within 1e-9 of) a face or quad centroid get the mean of its nodes; others
    get -1 and keep the previous cell."""
    m, L, q10 = tip
    tris, t2t, quads, T = surface_quads
    rng = np.random.default_rng(30 + rank)
    cen = T.quad_centroid if rank == 3 else T.centroid
    nodes = quads if rank == 3 else (T.tri6 if rank == 2 else tris)
    idx = rng.choice(len(cen), 100, replace=False)
    pts = cen[idx] + rng.uniform(-1e-9, 1e-9, (100, 3))
    pts[::7] += 1e-6                                     # off by more than sqrt(1e-15)
    sol = rng.normal(size=(len(m.nodes), 5))
    kw = dict(rank=rank, centroids=True, xyz=m.nodes, cells10=q10, hexes=m.hexes)
    cells, out = T.interp(L, m.tets, sol, pts, **kw)
    miss = np.zeros(100, bool)
    miss[::7] = True
    assert np.all(out[miss] == -1.0)
    assert np.array_equal(cells[~miss], idx[~miss])
    assert np.abs(out[~miss] - sol[nodes[idx[~miss]]].mean(1)).max() < 1e-14 * np.abs(sol).max() * 10
    # a miss keeps the previous cell
    prev = np.maximum.accumulate(np.where(miss, -1, np.arange(100)))
    expect = np.where(prev >= 0, idx[np.maximum(prev, 0)], -1)
    assert np.array_equal(cells, expect)
    c2 = np.where(cells < 0, 0, cells)
    assert np.array_equal(T.interp(L, m.tets, sol, pts, c2, **kw)[1][~miss], out[~miss])


# ------------------------------------------------------------------------------------------
# space charge deposit (PIC field mode)
# ------------------------------------------------------------------------------------------
def _particles_in_cells(m, n, seed):
    """Points at random natural coordinates of random solver cells (trilinear map of the
    cell's lexicographic vertices)."""
    rng = np.random.default_rng(seed)
    cells = rng.integers(0, m.n_cells, n)
    a = rng.uniform(0.02, 0.98, (n, 3))
    lex = np.array([[k & 1, (k >> 1) & 1, k >> 2] for k in range(8)])
    w = np.prod(np.where(lex[None] == 1, a[:, None, :], 1 - a[:, None, :]), axis=2)      # (n, 8)
    pts = np.einsum("nk,nkd->nd", w, m.x[m.cells[cells]])
    return pts, cells.astype(np.int32)


def test_space_charge_deposit_conserves_charge_and_first_moment(tip):
    """Shape functions sum to one and reproduce linear functions: the deposit carries the total
    charge and its first moment (sum_i b_i x_i = factor * sum_p x_p) -- a wrong order between
    the femocs hex and the solver cell breaks the second."""
    tip = tip[0]
    m = O.Mesh(tip.nodes, tip.vacuum_cells(), 8, 2, 4, 2)
    d2f = np.nonzero(tip.hex_marker > 0)[0].astype(np.int32)
    pts, cells = _particles_in_cells(m, 500, 3)
    f = 0.37
    b = O.space_charge(m, tip.nodes, tip.hexes, d2f, pts, cells, f)
    assert abs(b.sum() - f * len(pts)) < 1e-12 * len(pts)
    assert np.allclose(m.x.T @ b, f * pts.sum(0), rtol=1e-10, atol=1e-9)
    assert np.count_nonzero(b) <= 8 * len(pts)


def test_space_charge_at_a_vertex_lands_on_its_dof(tip):
    tip = tip[0]
    m = O.Mesh(tip.nodes, tip.vacuum_cells(), 8, 2, 4, 2)
    d2f = np.nonzero(tip.hex_marker > 0)[0].astype(np.int32)
    rng = np.random.default_rng(4)
    for c in rng.integers(0, m.n_cells, 5):
        for k in (0, 3, 6):
            dof = m.cells[c][k]
            b = O.space_charge(m, tip.nodes, tip.hexes, d2f, m.x[dof][None], np.array([c], np.int32), 2.0)
            e = np.zeros(m.n_dofs)
            e[dof] = 2.0
            assert np.abs(b - e).max() < 1e-12


_UVW = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])


def _points_in_hexes(m, n, seed):
    """Points at random natural coordinates of random vacuum hexes (femocs vertex order)."""
    rng = np.random.default_rng(seed)
    hx = rng.choice(np.nonzero(m.hex_marker > 0)[0], n).astype(np.int32)
    uvw = rng.uniform(-0.9, 0.9, (n, 3))
    w = np.prod(1 + _UVW[None] * uvw[:, None, :], axis=2) / 8.0
    return np.einsum("nk,nkd->nd", w, m.nodes[m.hexes[hx]]), hx


def test_hex_gradient_of_linear_field_is_exact(tip):
    m = tip[0]
    sol = np.zeros((len(m.nodes), 5))
    sol[:, 4] = m.nodes @ A + 0.25
    pts, hx = _points_in_hexes(m, 300, 21)
    g = O.linhex_gradient(m.nodes, m.hexes, sol, pts, hx)
    assert np.abs(g + A).max() < 1e-11


def test_hex_gradient_matches_differentiated_shape_functions(tip):
    """-grad of the trilinear interpolant by central differences of the shape functions (which
    are pinned on their own), for a field that is not linear."""
    m = tip[0]
    rng = np.random.default_rng(22)
    sol = np.zeros((len(m.nodes), 5))
    sol[:, 4] = rng.standard_normal(len(m.nodes))
    pts, hx = _points_in_hexes(m, 60, 23)
    g = O.linhex_gradient(m.nodes, m.hexes, sol, pts, hx)
    h = 1e-6
    for p, c, gi in zip(pts, hx, g):
        xe = m.nodes[m.hexes[c]]
        phi = sol[m.hexes[c], 4]
        fd = [(O.linhex_shape_functions(xe, p + h * e) @ phi - O.linhex_shape_functions(xe, p - h * e) @ phi) / (2 * h)
              for e in np.eye(3)]
        assert np.allclose(gi, -np.array(fd), rtol=1e-5, atol=1e-6 * np.abs(phi).max())


def test_particle_cell_update_against_brute_force(tip):
    """Pic::update_point_cell: the returned solver cell is the region hex whose tet holds the
    point and whose vertex has the largest barycentric coordinate; points outside every
    region tet get -1; a point still inside its own cell keeps it."""
    m, L, _ = tip
    d2f = np.nonzero(m.hex_marker > 0)[0].astype(np.int32)
    f2d = -np.ones(len(m.hexes), np.int32)
    f2d[d2f] = np.arange(len(d2f))
    rng = np.random.default_rng(41)
    pts, ts = _inside_points(m, 300, 42)
    guess = rng.integers(0, len(d2f), len(pts)).astype(np.int32)
    out = np.array([[0.0, 0.0, -1e4]])                  # far below the slab: outside
    cells = O.update_cells(L, d2f, f2d, np.vstack([pts, out]), np.concatenate([guess, [0]]))
    assert cells[-1] == -1
    for p, c in zip(pts, cells[:-1]):
        assert c >= 0
        h = d2f[c]
        t, k = h // 4, h % 4
        b = L.bcc(p, t)
        assert min(b) > -1e-12                           # the tet holds the point
        assert b[k] == max(b)
    again = O.update_cells(L, d2f, f2d, pts, cells[:-1])
    assert np.array_equal(again, cells[:-1])


def test_smoothing_weights_on_one_regular_tet():
    """One regular tet (edge a) split around its vertices: vertex 0's pseudo-Voronoi neighbours
    are its three edge midpoints (distance a/2), three face centroids (a/sqrt 3) and the tet
    centroid (a sqrt 6 / 4), so average_nodal_fields (src/Interpolator.cpp:142-170) with
    weights exp(decay * distance) has a closed form from the geometry alone.  Checks the
    neighbour set, the weight law and its sign, the normalisation, decay = 0 (plain mean) and
    that edge/face/centre nodes and the scalar slots are left alone."""
    a = 2.0
    V = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float) * (a / np.sqrt(8))
    edges = [(i, j) for i in range(4) for j in range(i + 1, 4)]
    faces = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
    nodes = np.concatenate([V, [(V[i] + V[j]) / 2 for i, j in edges], [V[list(f)].mean(0) for f in faces],
                            [V.mean(0)]])
    kind = np.array([1] * 4 + [2] * 6 + [3] * 4 + [4], np.int32)
    hexes, marker = [], []
    for v in range(4):   # node order inside a hex does not enter the average
        hexes.append([v] + [4 + e for e, (i, j) in enumerate(edges) if v in (i, j)]
                     + [10 + f for f, fv in enumerate(faces) if v in fv] + [14])
        marker.append(v + 1)
    hexes, marker = np.array(hexes, np.int32), np.array(marker, np.int32)
    assert hexes.shape == (4, 8)
    value = np.array([0.0] * 4 + [1.0] * 6 + [2.0] * 4 + [5.0])
    sol = np.zeros((15, 5))
    sol[:, 0] = value
    sol[:, 1] = 7.0
    sol[:, 3] = np.arange(15)
    sol[:, 4] = -1.0
    for decay in (-1.0 / a, 0.0):
        out = O.smooth_nodal(sol, nodes, 4, kind, hexes, marker, True, decay)
        we, wf, wc = np.exp(decay * a / 2), np.exp(decay * a / np.sqrt(3)), np.exp(decay * a * np.sqrt(6) / 4)
        expect = (3 * we * 1.0 + 3 * wf * 2.0 + wc * 5.0) / (3 * we + 3 * wf + wc)
        assert out[0, 0] == pytest.approx(expect, rel=1e-14)
        assert np.allclose(out[:4, 0], expect, rtol=1e-14)       # every vertex sees the same
        assert np.allclose(out[:4, 1], 7.0, rtol=1e-15)          # constants are kept
        assert np.array_equal(out[4:], sol[4:])                  # non-vertex nodes untouched
        assert np.array_equal(out[:, 3:], sol[:, 3:])            # scalars untouched
    # the nearer classes weigh more with the negative decay: the mean moves towards 1
    plain = (3 * 1.0 + 3 * 2.0 + 5.0) / 7
    near = O.smooth_nodal(sol, nodes, 4, kind, hexes, marker, True, -1.0 / a)[0, 0]
    assert near < plain

"""Q2 Laplace solve on the GPU (fgpu_q2_*) against the Q2 oracle (oracle/q2.py, a restatement of
the reference recompiled with shape_degree = 2): support points and connectivity, the
matrix-free apply against the oracle's assembled matrix, the right-hand side and the solution
(compared at the support points), Jacobi and Chebyshev-Jacobi."""
import numpy as np
import pytest
import scipy.sparse as sp

from femocs_amd import fgpu
from femocs_amd import mesh as M
from oracle import oracle as O
from oracle import q2

pytestmark = pytest.mark.gpu
MARKS = (8, 2, 4, 2)


def _box(n, distort, seed):
    """n^3 hexes of a 10 x 10 x 20 box (lexicographic), interior vertices jittered, as UCD cells."""
    g = np.linspace(0.0, 1.0, n + 1)
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    idx = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k
    lex = np.array([[idx(i + a, j + b, k + c) for c in (0, 1) for b in (0, 1) for a in (0, 1)]
                    for i in range(n) for j in range(n) for k in range(n)])
    inner = np.all((X > 1e-9) & (X < 1 - 1e-9), axis=1)
    X[inner] += np.random.default_rng(seed).uniform(-distort, distort, (inner.sum(), 3)) / n
    X = X * [10.0, 10.0, 20.0]
    ucd = np.empty_like(lex)
    ucd[:, M.UCD_TO_LEX] = lex
    return X, ucd.astype(np.int32)


@pytest.fixture(scope="module")
def box():
    return _box(5, 0.3, 11)


def _match(Pa, Pb):
    """index into Pb of every point of Pa (equal up to rounding)."""
    from scipy.spatial import cKDTree
    d, i = cKDTree(Pb).query(Pa)
    assert d.max() <= 1e-9
    assert len(np.unique(i)) == len(Pa)
    return i


def test_q2_dofs_and_apply_match_oracle(ctx, box):
    X, ucd = box
    gm = fgpu.Mesh(ctx, X, ucd, *MARKS)
    s = fgpu.Q2System(gm)
    om = O.Mesh(X, ucd, *MARKS)
    c27o, no = q2.cell_dofs(om.cells, om.n_dofs)
    assert s.n_dofs == no == 11 ** 3 and s.n_cells == om.n_cells
    P, cells = s.dofs()
    Po = q2.support_points(om.x, om.cells, c27o, no)
    _match(P, Po)
    # vertices keep the Q1 numbering; every cell's 27 dofs sit at its 27 support points
    xv, _ = gm.dofs()
    assert np.array_equal(P[: gm.n_dofs], xv)
    assert np.array_equal(cells[:, q2.VERTEX_NODE], om.cells)
    Pc = q2.support_points(xv, om.cells, c27o, no)
    assert np.abs(P[cells] - Pc[c27o]).max() <= 1e-12
    # the apply against the oracle's element matrices scattered with the GPU connectivity
    rows, cols, vals = [], [], []
    for c in range(om.n_cells):
        K = q2.element_matrix(xv[om.cells[c]])
        rows.append(np.repeat(cells[c], 27))
        cols.append(np.tile(cells[c], 27))
        vals.append(K.reshape(-1))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(no, no))
    rng = np.random.default_rng(2)
    for _ in range(3):
        u = rng.standard_normal(no)
        ref = A @ u
        assert np.linalg.norm(s.apply(u) - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("precond", [fgpu.FGPU_PC_JACOBI, fgpu.FGPU_PC_CHEBYSHEV])
def test_q2_laplace_matches_oracle(ctx, box, precond):
    X, ucd = box
    gm = fgpu.Mesh(ctx, X, ucd, *MARKS)
    s = fgpu.Q2System(gm)
    s.reset()
    s.add_dirichlet(2, 0.0)
    s.add_neumann(8, -1.0)
    s.apply_dirichlet()
    s.fill_solution(0.0)
    st = s.solve(fgpu.opts(tol=1e-13, relative=True, precond=precond))
    assert st.iterations > 0
    om = O.Mesh(X, ucd, *MARKS)
    _, Po, xo = q2.laplace(om, -1.0)
    P, _ = s.dofs()
    i = _match(P, Po)
    x = s.solution()
    assert np.linalg.norm(x - xo[i]) <= 1e-10 * np.linalg.norm(xo)


def test_q2_on_tip_mesh(ctx, small_mesh):
    """A tet-split tip (affine sub-hexes): the Q2 solution's vertex values approach the Q1
    solution's (same problem, finer space), both match their own references."""
    gm = fgpu.Mesh(ctx, small_mesh.nodes, small_mesh.vacuum_cells(), *MARKS)
    s = fgpu.Q2System(gm)
    s.reset()
    s.add_dirichlet(2, 0.0)
    s.add_neumann(8, -1.0)
    s.apply_dirichlet()
    s.fill_solution(0.0)
    st = s.solve(fgpu.opts(tol=1e-12, relative=True, precond=fgpu.FGPU_PC_CHEBYSHEV))
    assert st.iterations > 0
    x2 = s.solution()[: gm.n_dofs]
    _, x1, it, _, _ = O.poisson_laplace(small_mesh.nodes, small_mesh.vacuum_cells(), -1.0, tol=1e-12, relative=True)
    assert it > 0
    # the two discretisations differ by the discretisation error, not by more
    assert np.linalg.norm(x2 - x1) <= 0.05 * np.linalg.norm(x1)


def test_q2_tet_split_form_matches_general_form(ctx, small_mesh):
    """The tet-split form of the Q2 apply (a per-cell metric and per-point constants instead of
    the trilinear Jacobian at each of the 27 Gauss points) gives the general form's result to
    rounding, on the generator's hexes and on hexes in the reference's vertex order (about half
    of them eta-flipped, computed with the node grid's middle index reversed); the box mesh
    (not a tet split) keeps the general form."""
    from oracle import tethex
    nodes, _, hexes, _, _, _ = tethex.split(small_mesh.nodes[:small_mesh.n_tetnodes], small_mesh.tets)
    rng = np.random.default_rng(21)
    for X, cells in ((small_mesh.nodes, small_mesh.vacuum_cells()), (nodes, hexes[small_mesh.hex_marker > 0])):
        gm = fgpu.Mesh(ctx, X, cells, *MARKS)
        s = fgpu.Q2System(gm)
        assert s.uses_tet_split()
        for _ in range(2):
            u = rng.standard_normal(s.n_dofs)
            assert gm.use_affine(True) and s.uses_tet_split()
            ya = s.apply(u)
            assert not gm.use_affine(False) and not s.uses_tet_split()
            yg = s.apply(u)
            assert np.linalg.norm(ya - yg) <= 1e-13 * np.linalg.norm(yg)
        assert gm.use_affine(True)
        # the solve on the tet-split form equals the general form's solution
        sols = []
        for allow in (True, False):
            gm.use_affine(allow)
            s.reset()
            s.add_dirichlet(2, 0.0)
            s.add_neumann(8, -1.0)
            s.apply_dirichlet()
            s.fill_solution(0.0)
            st = s.solve(fgpu.opts(tol=1e-12, relative=True, precond=fgpu.FGPU_PC_JACOBI))
            assert st.iterations > 0
            sols.append(s.solution())
        assert np.linalg.norm(sols[0] - sols[1]) <= 1e-10 * np.linalg.norm(sols[1])


def test_q2_box_keeps_general_form(ctx, box):
    X, ucd = box
    s = fgpu.Q2System(fgpu.Mesh(ctx, X, ucd, *MARKS))
    assert not s.uses_tet_split()

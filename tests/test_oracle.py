"""Pins the CPU oracle with checks that do not restate its own formulas: textbook element
matrices, invariants, closed-form solutions, convergence order, brute force on tiny inputs
and the reference's own data tables and mesh file."""
import os

import numpy as np
import pytest

from femocs_amd import mesh as M
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
UNIT_LEX = np.array([[i & 1, (i >> 1) & 1, (i >> 2) & 1] for i in range(8)], dtype=np.float64)


def box_tets(n=(3, 3, 3), L=(1.0, 1.0, 1.0), jitter=0.0, seed=0):
    """Structured box split into 6 Kuhn tets per cube (jitter moves interior nodes)."""
    nx, ny, nz = n
    xs = [np.linspace(0, L[d], n[d] + 1) for d in range(3)]
    P = np.stack(np.meshgrid(*xs, indexing="ij"), -1).reshape(-1, 3)
    idx = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for p in perms:
                    c = [0, 0, 0]
                    path = [idx(i, j, k)]
                    for d in p:
                        c[d] += 1
                        path.append(idx(i + c[0], j + c[1], k + c[2]))
                    tets.append(path)
    tets = np.array(tets)
    rng = np.random.default_rng(seed)
    inner = np.all((P > 1e-9) & (P < np.array(L) - 1e-9), axis=1)
    P[inner] += rng.uniform(-jitter, jitter, (inner.sum(), 3))
    vol = M._tet_volumes(P, tets)
    tets[vol < 0] = tets[vol < 0][:, [0, 2, 1, 3]]
    return P, tets


def box_hexes(n=(3, 3, 4), L=(1.0, 1.0, 2.0), jitter=0.05, seed=0):
    P, T = box_tets(n, L, jitter, seed)
    return M.tet_to_hex(P, T)


# ------------------------------------------------------------------------------------------
# element matrices
# ------------------------------------------------------------------------------------------
def test_unit_cube_stiffness_is_textbook():
    """Q1 Laplace on the unit cube: 1/3 on the diagonal, 0 for edge neighbours, -1/12 for
    face- and body-diagonal pairs (exact with 2-point Gauss)."""
    K = O.element_matrix(UNIT_LEX)
    for i in range(8):
        for j in range(8):
            d = bin(i ^ j).count("1")
            expect = {0: 1 / 3, 1: 0.0, 2: -1 / 12, 3: -1 / 12}[d]
            assert abs(K[i, j] - expect) < 1e-15


def test_unit_cube_mass_matrix_is_textbook():
    """Q1 mass on the unit cube: (1/27) * 2^(3-d) / 8 ... i.e. 1/27, 1/54, 1/108, 1/216 by
    the number d of differing coordinates (exact with 2-point Gauss)."""
    K = O.element_matrix(UNIT_LEX, kappa_q=np.zeros(8), gamma_q=np.ones(8))
    for i in range(8):
        for j in range(8):
            d = bin(i ^ j).count("1")
            assert abs(K[i, j] - {0: 1 / 27, 1: 1 / 54, 2: 1 / 108, 3: 1 / 216}[d]) < 1e-16


def test_element_matrix_scales_and_is_invariant():
    rng = np.random.default_rng(3)
    base = UNIT_LEX + rng.uniform(-0.15, 0.15, (8, 3))
    K = O.element_matrix(base)
    assert np.allclose(K, K.T, atol=1e-15)
    assert np.abs(K.sum(axis=1)).max() < 1e-14          # constants are in the kernel
    assert np.linalg.eigvalsh(K)[1] > 0                  # only one zero eigenvalue
    # stiffness scales linearly with size, is invariant under rotation and translation
    assert np.allclose(O.element_matrix(2.0 * base), 2.0 * K, rtol=1e-13, atol=1e-14)
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    assert np.allclose(O.element_matrix(base @ R.T + 5.0), K, rtol=1e-12, atol=1e-13)


def test_stiffness_quadratic_form_matches_exact_energy_for_linear_field():
    """For u = a.x the energy u^T K u equals |a|^2 * volume on any trilinear hex."""
    rng = np.random.default_rng(4)
    xe = UNIT_LEX * np.array([1.0, 2.0, 0.5]) + rng.uniform(-0.1, 0.1, (8, 3))
    a = np.array([0.3, -1.2, 0.7])
    u = xe @ a
    K = O.element_matrix(xe)
    M1 = O.element_matrix(xe, kappa_q=np.zeros(8), gamma_q=np.ones(8))
    vol = M1.sum()
    assert abs(u @ K @ u - (a @ a) * vol) < 1e-12 * (a @ a) * vol


# ------------------------------------------------------------------------------------------
# mesh import and boundary marking
# ------------------------------------------------------------------------------------------
def test_import_compacts_vertices_and_marks_box_faces():
    nodes, _, hexes = box_hexes()
    extra = np.vstack([nodes, [[9.0, 9.0, 9.0]]])        # an unused vertex is dropped
    m = O.Mesh(extra, hexes, 8, 2, 4, 2)
    assert m.n_dofs == len(nodes) and np.array_equal(m.vertex_of_dof, np.arange(len(nodes)))
    # every boundary face lies on the box; ids by position
    FV = [[0, 2, 4, 6], [1, 3, 5, 7], [0, 1, 4, 5], [2, 3, 6, 7], [0, 1, 2, 3], [4, 5, 6, 7]]
    for c, f, b in zip(m.bface_cell, m.bface_face, m.bface_id):
        ctr = m.x[m.cells[c][FV[f]]].mean(0)
        if np.isclose(ctr[0], 0) or np.isclose(ctr[0], 1) or np.isclose(ctr[1], 0) or np.isclose(ctr[1], 1):
            assert b == 4
        elif np.isclose(ctr[2], 2.0):
            assert b == 8
        else:
            assert np.isclose(ctr[2], 0.0) and b == 2
    # lexicographic cells are positively oriented
    for c in m.cells[:50]:
        x = m.x[c]
        assert np.linalg.det(np.array([x[1] - x[0], x[2] - x[0], x[4] - x[0]])) > 0


def test_neumann_load_sums_to_flux():
    nodes, _, hexes = box_hexes()
    m = O.Mesh(nodes, hexes, 8, 2, 4, 2)
    b = m.face_rhs(8, constant=2.5)
    assert abs(b.sum() - 2.5 * 1.0) < 1e-13                 # top face area 1
    top = m.x[:, 2] > 2.0 - 1e-12
    assert np.all(b[~top] == 0)


def grid_hexes(n=(3, 4, 2), h=0.5, shear=0.0):
    """Structured grid of cubes of side h in UCD vertex order; shear maps (x, y, z) to
    (x + shear z, y, z), so the cells are congruent parallelepipeds."""
    nx, ny, nz = n
    xs = [np.arange(n[d] + 1) * h for d in range(3)]
    P = np.stack(np.meshgrid(*xs, indexing="ij"), -1).reshape(-1, 3)
    P[:, 0] += shear * P[:, 2]
    idx = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
    lex_to_ucd = [0, 1, 4, 5, 3, 2, 7, 6]
    cells = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                ucd = [0] * 8
                for l in range(8):
                    ucd[lex_to_ucd[l]] = idx(i + (l & 1), j + ((l >> 1) & 1), k + (l >> 2))
                cells.append(ucd)
    return P, np.array(cells)


def test_dof_volumes_count_the_cells_of_each_dof():
    """Congruent cells of volume h^3: a dof's volume is h^3 times the number of cells it
    belongs to (1 at a corner, 2 on an edge, 4 on a face, 8 inside)."""
    h = 0.5
    P, cells = grid_hexes(h=h, shear=0.3)
    m = O.Mesh(P, cells, 8, 2, 4, 2)
    count = np.bincount(m.cells.reshape(-1), minlength=m.n_dofs)
    assert set(count.tolist()) == {1, 2, 4, 8}
    assert np.allclose(m.dof_volumes(), h ** 3 * count, rtol=1e-14, atol=0)


def test_dof_volumes_sum_to_eight_mesh_volumes():
    nodes, _, hexes = box_hexes()                        # jittered 1 x 1 x 2 box
    m = O.Mesh(nodes, hexes, 8, 2, 4, 2)
    v = m.dof_volumes()
    assert np.all(v > 0) and abs(v.sum() - 8 * 2.0) < 1e-12


def test_charge_density_of_a_uniform_source():
    """Each cell gives every one of its dofs the load c*V/8 of a constant source c on a
    parallelepiped (the integral of a Q1 shape function) and the volume V: rho = c/8 at every
    dof.  On a curved mesh, rho_i vol_i still sums to c times the mesh volume."""
    c = 3.0
    P, cells = grid_hexes(shear=0.3)
    m = O.Mesh(P, cells, 8, 2, 4, 2)
    rho = m.charge_density(m.cell_rhs(np.full(8 * m.n_cells, c)))
    assert np.allclose(rho, c / 8, rtol=1e-13, atol=0)
    nodes, _, hexes = box_hexes()
    m = O.Mesh(nodes, hexes, 8, 2, 4, 2)
    rho = m.charge_density(m.cell_rhs(np.full(8 * m.n_cells, c)))
    assert abs((rho * m.dof_volumes()).sum() - c * 2.0) < 1e-12
    assert not np.allclose(rho, c / 8)


def test_charge_density_of_a_neumann_load():
    """The top-face Neumann load of the field solve shows up as charge on the top dofs only:
    on the grid a top dof of k top faces of area h^2 and flux q gets k q h^2/4 over the k h^3
    of the k cells under those faces."""
    h, q = 0.5, 2.5
    P, cells = grid_hexes(h=h)
    m = O.Mesh(P, cells, 8, 2, 4, 2)
    rho = m.charge_density(m.face_rhs(8, constant=q))
    top = m.x[:, 2] > m.x[:, 2].max() - 1e-12
    assert np.all(rho[~top] == 0)
    assert np.allclose(rho[top], q / (4 * h), rtol=1e-13, atol=0)


# ------------------------------------------------------------------------------------------
# solutions
# ------------------------------------------------------------------------------------------
def test_patch_test_linear_potential_exact():
    """phi = 0 at the bottom, dphi/dn = -1 on the top: the exact solution -z is linear and
    reproduced to round-off on distorted trilinear hexes."""
    nodes, _, hexes = box_hexes(jitter=0.07)
    om, x, it, res, _ = O.poisson_laplace(nodes, hexes, -1.0, tol=1e-13, relative=True)
    assert it > 0
    assert np.abs(x + om.x[:, 2]).max() < 1e-12


def _dirichlet_solve(nodes, hexes, g):
    m = O.Mesh(nodes, hexes, 2, 2, 2, 2).assemble()
    b = np.zeros(m.n_dofs)
    x = np.zeros(m.n_dofs)
    mask = m.dirichlet_mask(2)
    m.apply_dirichlet(b, x, mask, g(m.x))
    it, _ = m.pcg(b, x, tol=1e-14, relative=False, precond="ssor", max_it=5000)
    assert it >= 0
    return m, x


def test_harmonic_solution_converges_at_second_order():
    """Dirichlet data of the harmonic phi = x^2 + y^2 - 2 z^2 on the whole boundary: the
    nodal error falls by ~4x when the mesh is refined by 2."""
    g = lambda x: x[:, 0] ** 2 + x[:, 1] ** 2 - 2 * x[:, 2] ** 2
    errs = []
    for n in (3, 6):
        nodes, _, hexes = box_hexes(n=(n, n, n), L=(1.0, 1.0, 1.0), jitter=0.0)
        m, x = _dirichlet_solve(nodes, hexes, g)
        errs.append(np.abs(x - g(m.x)).max())
    assert 3.0 < errs[0] / errs[1] < 6.0


def test_lifted_system_equals_dense_constrained_solve():
    """deal.II-style lifting (row and column elimination) gives the same solution as solving
    the reduced system for the free dofs directly (numpy dense solve)."""
    nodes, _, hexes = box_hexes(n=(2, 2, 2), L=(1, 1, 1), jitter=0.05)
    m = O.Mesh(nodes, hexes, 8, 2, 4, 2).assemble()
    rp, ci, vv = m.csr()
    A = np.zeros((m.n_dofs, m.n_dofs))
    for i in range(m.n_dofs):
        A[i, ci[rp[i]:rp[i + 1]]] = vv[rp[i]:rp[i + 1]]
    b0 = m.face_rhs(8, constant=-0.7)
    mask = m.dirichlet_mask(2)
    g = np.where(mask > 0, 0.3 * m.x[:, 0], 0.0)
    free = mask == 0
    xd = np.where(mask > 0, g, 0.0)
    xf = np.linalg.solve(A[np.ix_(free, free)], b0[free] - A[np.ix_(free, ~free)] @ xd[~free])
    b = b0.copy()
    x = np.zeros(m.n_dofs)
    m.apply_dirichlet(b, x, mask, g)
    it, _ = m.pcg(b, x, tol=1e-15, relative=True, precond="ssor")
    assert np.allclose(x[free], xf, rtol=1e-12, atol=1e-12)
    assert np.array_equal(x[~free], g[~free])


@pytest.mark.parametrize("precond", ["none", "jacobi", "ssor"])
def test_pcg_matches_dense_solve(precond):
    nodes, _, hexes = box_hexes(n=(2, 2, 3), L=(1, 1, 1.5), jitter=0.05)
    om, x, it, res, b = O.poisson_laplace(nodes, hexes, -1.0, tol=1e-14, relative=True, precond=precond)
    rp, ci, vv = om.csr()
    A = np.zeros((om.n_dofs, om.n_dofs))
    for i in range(om.n_dofs):
        A[i, ci[rp[i]:rp[i + 1]]] = vv[rp[i]:rp[i + 1]]
    xd = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xd) <= 1e-11 * np.linalg.norm(xd)


def test_absolute_tolerance_and_failure_sign():
    """SolverControl semantics: absolute |r| <= tol, negative iteration count on failure
    (src/DealSolver.cpp:443-458)."""
    nodes, _, hexes = box_hexes()
    om, x, it, res, b = O.poisson_laplace(nodes, hexes, -1.0, tol=1e-9, relative=False, max_it=3)
    assert it == -3 and res > 1e-9
    om, x, it, res, b = O.poisson_laplace(nodes, hexes, -1.0, tol=1e-9, relative=False)
    assert it > 0 and res <= 1e-9


def test_gradient_export_exact_for_linear_field():
    nodes, _, hexes = box_hexes(jitter=0.06)
    m = O.Mesh(nodes, hexes, 8, 2, 4, 2)
    a = np.array([0.4, -0.2, 1.1])
    g = m.export_solution_grad(m.x @ a)
    assert np.abs(g + a).max() < 1e-12


# ------------------------------------------------------------------------------------------
# interpolation
# ------------------------------------------------------------------------------------------
def test_nodal_field_exact_for_linear_potential():
    """Interpolator::store_elfield averages exact hex-corner gradients: -grad of a linear
    potential at every node of the region."""
    m = M.make_tip_mesh(M.small_geometry(), with_atoms=False)
    a = np.array([0.2, 0.5, -1.0])
    E = O.nodal_field(m.nodes, m.hexes, m.hex_marker, 1, m.nodes @ a)
    used = np.zeros(len(m.nodes), bool)
    used[m.hexes[m.hex_marker > 0]] = True
    assert np.abs(E[used] + a).max() < 1e-10
    assert np.all(E[~used] == 0)


def test_linhex_shape_functions_invert_trilinear_map():
    rng = np.random.default_rng(5)
    ucd = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
    xe = 2.0 * ucd + rng.uniform(-0.2, 0.2, (8, 3))
    for _ in range(20):
        uvw = rng.uniform(-0.9, 0.9, 3)
        N = np.prod(1 + ucd * uvw, axis=1) / 8
        p = N @ xe
        sf = O.linhex_shape_functions(xe, p)
        assert np.abs(sf - N).max() < 1e-12
    for k in range(8):
        sf = O.linhex_shape_functions(xe, xe[k])
        assert abs(sf[k] - 1) < 1e-12


@pytest.fixture(scope="module")
def lintet_small():
    m = M.make_tip_mesh(M.small_geometry(), with_atoms=True)
    return m, O.LinTet(m.nodes, m.tets, m.tet_nbr, m.tet_marker, 3)


def test_barycentric_coordinates(lintet_small):
    m, L = lintet_small
    rng = np.random.default_rng(6)
    for t in rng.integers(0, m.n_tets, 50):
        v = m.nodes[m.tets[t]]
        for k in range(4):
            b = L.bcc(v[k], t)
            e = np.eye(4)[k] + 1e-15
            assert np.abs(b - e).max() < 1e-9
        w = rng.dirichlet(np.ones(4))
        b = L.bcc(w @ v, t)
        assert abs(b.sum() - (1 + 4e-15)) < 1e-10 and np.abs(b - w - 1e-15).max() < 1e-9


def _brute_containing(L, p):
    return [c for c in range(L.det0.size) if np.all(L.det0[c] * (L.det[c] @ np.append(p, 1.0)) >= -1e-15)]


def test_locate_cell_semantics_against_brute_force(lintet_small):
    """locate_cell: a containing searchable cell when the guess chain misses, else minus the
    cell with the nearest centroid (src/InterpolatorCells.cpp:268-307)."""
    m, L = lintet_small
    rng = np.random.default_rng(7)
    lo, hi = m.nodes.min(0), m.nodes.max(0)
    pts = rng.uniform(lo, hi, (60, 3))
    cells = L.locate_chain(pts, rank=1)
    for p, c in zip(pts, cells):
        inside = _brute_containing(L, p)
        if c > 0 or (c == 0 and 0 in inside):
            assert c in inside
        else:
            searchable = [k for k in inside if L.marker[k] == 0]
            assert not searchable
            d2 = ((L.centroid - p) ** 2).sum(1)
            assert -c == int(np.argmin(d2))


def test_linhex_locate_picks_largest_barycentric(lintet_small):
    m, L = lintet_small
    rng = np.random.default_rng(8)
    vac = np.nonzero(m.tet_marker == 3)[0]
    for t in rng.choice(vac, 30):
        w = rng.dirichlet(np.ones(4))
        p = w @ m.nodes[m.tets[t]]
        h = L.locate_chain(p[None], rank=3)[0]
        assert h // 4 == t and h % 4 == int(np.argmax(w))


# ------------------------------------------------------------------------------------------
# physical quantities: the reference's own resistivity table
# ------------------------------------------------------------------------------------------
def test_resistivity_table_and_laws():
    # table nodes (T, rho [Ohm nm]) of bulk Cu, src/PhysicalQuantityData.cpp:47-76
    assert O.resistivity(300.0) == 10.0 * 1.72303253767393e1
    assert O.resistivity(1400.0) == 10.0 * 1.01349915510292e2
    assert O.resistivity(100.0) == O.resistivity(200.0)          # clamped below
    assert O.resistivity(5000.0) == O.resistivity(1400.0)        # clamped above
    mid = 0.5 * (1.72303253767393e1 + 2.05822342557969e1) * 10.0
    assert abs(O.resistivity(325.0) - mid) < 1e-12 * mid          # linear between nodes
    for T in (250.0, 600.0, 1234.5):
        assert O.sigma(T) == 1.0 / O.resistivity(T)
        assert abs(O.kappa(T) - 2.44e-8 * T * O.sigma(T)) <= 1e-18


# ------------------------------------------------------------------------------------------
# the reference's own mesh
# ------------------------------------------------------------------------------------------
def test_reference_hemicone_mesh_solves():
    from femocs_amd import gmsh
    m = gmsh.tip_mesh_from_msh(os.path.join(HERE, "golden", "hemicone.msh"))
    assert m.n_tets == 7338 and m.n_hexes == 4 * 7338 and len(m.nodes) == 32885
    om, x, it, res, _ = O.poisson_laplace(m.nodes, m.vacuum_cells(), -1.0, tol=1e-12, relative=True)
    assert it > 0
    # grounded copper, field -1 V/A on the top at z = 500: 0 >= phi >= -500
    assert x.max() <= 1e-12 and x.min() > -500.0
    top = om.x[:, 2] > 500 - 1e-9
    assert np.all(x[top] < -490.0)


def test_qp_fields_of_linear_function():
    """Values and gradients at the Gauss points: exact for a linear function, points at
    (1 +- 1/sqrt 3) / 2 of a unit cube, x fastest."""
    g = (1 - 1 / np.sqrt(3)) / 2
    x = np.array([[i, j, k] for k in (0, 1) for j in (0, 1) for i in (0, 1)], float)
    ucd = np.array([[0, 1, 5, 4, 2, 3, 7, 6]], np.int32)       # UCD order of the lexicographic cube
    m = O.Mesh(x, ucd, 8, 2, 4, 2)
    a = np.array([0.3, -0.7, 1.1])
    v, gr = m.qp_fields(m.x @ a + 2.0)
    t = lambda bit: 1 - g if bit else g
    pts = np.array([[t(q & 1), t((q >> 1) & 1), t((q >> 2) & 1)] for q in range(8)])
    assert np.allclose(gr[0], a, atol=1e-14)
    assert np.allclose(v[0], pts @ a + 2.0, atol=1e-14)


# ------------------------------------------------------------------------------------------
# heat step: implicit Euler and the theta scheme on a discrete eigenmode
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_heat_step_amplifies_an_eigenmode_by_the_theta_factor(theta):
    """Below 200 K the conductivity table is clamped, so kappa is one constant and the step is
    linear.  Starting from T = eps v, v a generalized eigenvector of the free block
    (K v = lambda gamma M v, from scipy, with T = 0 on the Dirichlet bottom), one step gives
    g(lambda) eps v with g = (1 - (1 - theta) lambda) / (1 + theta lambda): 1/(1 + lambda) for
    implicit Euler, the Crank-Nicolson factor for theta = 1/2."""
    import scipy.linalg
    nodes, _, hexes = box_hexes(n=(2, 2, 2), L=(1.0, 1.0, 1.0), jitter=0.05)
    m = O.Mesh(nodes, hexes, 2, 7, 4, 2)
    dt = 4e-15
    gamma = O.CU_RHO_CP / dt
    k0 = O.kappa(0.0)
    nq = 8 * m.n_cells

    def dense(kq, gq):
        m.assemble(kappa_q=np.full(nq, kq), gamma_q=np.full(nq, gq))
        rp, col, val = m.csr()
        A = np.zeros((m.n_dofs, m.n_dofs))
        for i in range(m.n_dofs):
            A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
        return A

    K, Mm = dense(k0, 0.0), dense(0.0, 1.0)
    free = m.dirichlet_mask(7) == 0
    lam, V = scipy.linalg.eigh(K[np.ix_(free, free)], gamma * Mm[np.ix_(free, free)])
    n_face = int(np.sum(np.asarray(m.bface_id) == 2))
    for j in (0, len(lam) // 3, len(lam) - 1):
        v = np.zeros(m.n_dofs)
        v[free] = V[:, j] / np.abs(V[:, j]).max()
        eps = 50.0
        T, it = O.heat_step(m, eps * v, np.zeros(m.n_dofs), dt, np.zeros(n_face), 0.0, tol=1e-14, theta=theta)
        assert it > 0
        g = (1 - (1 - theta) * lam[j]) / (1 + theta * lam[j])
        assert np.abs(T - g * eps * v).max() <= 1e-9 * eps, (j, lam[j])
    assert lam[-1] > 10                       # stiff modes: the two schemes differ in sign there


def test_face_centroids_of_a_box_grid():
    """Mesh.face_centroids (DealSolver::export_surface_centroids): on a regular grid the top
    faces' centres are (x_c, y_c, z_top) of their cells, listed in cell order; a sheared grid
    moves them with the shear."""
    n, h = (3, 2, 4), 0.5
    ijk = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3)
    vid = lambda i, j, k: (i * (n[1] + 1) + j) * (n[2] + 1) + k
    lex = np.stack([vid(ijk[:, 0] + (l & 1), ijk[:, 1] + ((l >> 1) & 1), ijk[:, 2] + (l >> 2)) for l in range(8)], 1)
    ucd = np.empty_like(lex)
    ucd[:, M.UCD_TO_LEX] = lex
    P = np.stack(np.meshgrid(*[np.arange(k + 1) * h for k in n], indexing="ij"), -1).reshape(-1, 3)
    for shear in (0.0, 0.3):
        X = P.copy()
        X[:, 0] += shear * X[:, 2]
        m = O.Mesh(X, ucd, 8, 2, 4, 2)
        top = ijk[:, 2] == n[2] - 1
        zt = n[2] * h
        ref = np.stack([(ijk[top, 0] + 0.5) * h + shear * zt, (ijk[top, 1] + 0.5) * h, np.full(top.sum(), zt)], 1)
        c = m.face_centroids(8)
        assert c.shape == ref.shape
        assert np.abs(c - ref).max() <= 1e-14

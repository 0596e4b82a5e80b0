"""GPU current and heat solves on the bulk mesh against the oracle: the sigma = 1 current
potential with per-face current densities, its gradient export, implicit-Euler heat steps with
temperature-dependent kappa and Joule heat, and the nodal (J, phi_c, T) transfer."""
import numpy as np
import pytest

from femocs_amd import fgpu, interp, solvers
from femocs_amd import mesh as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MARKS = (2, 7, 4, 2)
DT = 4e-15


@pytest.fixture(scope="module")
def tip():
    return M.make_tip_mesh(M.small_geometry())


def _face_values(gm, bid, scale, seed):
    c = gm.face_centroids(bid)
    rng = np.random.default_rng(seed)
    return scale * (1.0 + 0.5 * np.sin(c[:, 0] / 7.0) * np.cos(c[:, 2] / 5.0) + 0.1 * rng.uniform(size=len(c)))


def _gpu_current(ctx, gm, J):
    s = fgpu.System(gm)
    s.reset()
    s.add_neumann(2, J, 0.0)
    s.add_dirichlet(7, 0.0)
    s.apply_dirichlet()
    s.fill_solution(0.0)
    st = s.solve(fgpu.opts(tol=1e-13, relative=True))
    assert st.iterations > 0
    return s


def test_current_potential_and_gradient_export(ctx, tip):
    gm = fgpu.Mesh(ctx, tip.nodes, tip.bulk_cells(), *MARKS)
    J = _face_values(gm, 2, 1e-3, 1)
    s = _gpu_current(ctx, gm, J)
    om, xo, it = O.current_solve(tip.nodes, tip.bulk_cells(), J, tol=1e-14)
    assert it > 0
    x = s.solution()
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
    g = s.solution_grad()
    go = om.export_solution_grad(xo)
    assert np.linalg.norm(g - go) <= 1e-9 * np.linalg.norm(go)
    # the gradient of the exact same vector agrees to rounding
    assert np.abs(g - om.export_solution_grad(x)).max() <= 1e-12 * np.abs(go).max()


@pytest.mark.parametrize("precision", [fgpu.FGPU_FP64, fgpu.FGPU_MIXED])
def test_heat_steps_match_oracle(ctx, tip, precision):
    gm = fgpu.Mesh(ctx, tip.nodes, tip.bulk_cells(), *MARKS)
    J = _face_values(gm, 2, 5e-3, 2)
    N = _face_values(gm, 2, 1e-9, 3)
    cur = _gpu_current(ctx, gm, J)
    phic = cur.solution()
    om, _, _ = O.current_solve(tip.nodes, tip.bulk_cells(), J, tol=1e-14)
    xyz, _ = gm.dofs()
    T0 = 300.0 + 400.0 * np.exp(-((xyz[:, 2] - xyz[:, 2].max()) / 20.0) ** 2)
    heat = fgpu.System(gm)
    heat.set_solution(T0)
    To = T0.copy()
    o = fgpu.opts(tol=1e-13, relative=True, precision=precision)
    for step in range(3):
        heat.reset()
        heat.heat_terms(cur, O.CU_RHO_CP / DT, O.LORENTZ)
        heat.add_neumann(2, N, 0.0)
        heat.add_dirichlet(7, 300.0)
        heat.apply_dirichlet()
        st = heat.solve(o)
        assert st.iterations > 0, step
        To, it = O.heat_step(om, To, phic, DT, N, 300.0, tol=1e-14)
        assert it > 0
        T = heat.solution()
        assert np.linalg.norm(T - To) <= 1e-10 * np.linalg.norm(To), step
    assert T.max() > 300.0 + 1e-6                    # the source terms did heat the tip


def test_current_heat_solver_and_transfer(ctx, tip):
    chs = solvers.CurrentHeatSolver(ctx=ctx)
    assert chs.import_mesh(tip.nodes, tip.bulk_cells())
    chs.setup(300.0)
    chs.current.bc_values = _face_values(chs.mesh, 2, 1e-3, 4)
    chs.heat.bc_values = _face_values(chs.mesh, 2, 1e-9, 5)
    chs.current.relative = chs.heat.relative = True
    chs.current.conf.cg_tolerance = 1e-12
    chs.current.assemble()
    assert chs.current.solve() > 0
    chs.heat.assemble(DT)
    assert chs.heat.solve() > 0
    T, rho = chs.export_temp_rho()
    assert np.all(np.isfinite(T)) and T.min() >= 300.0 - 1e-6
    ip = interp.Interpolator(ctx, tip, interp.TYPES_BULK, empty_value=300.0)
    ip.extract_heat(chs.current.system, chs.heat.system)
    sol = ip.nodal()
    xyz, vod = chs.mesh.dofs()
    assert np.array_equal(sol[vod, :3], rho)
    assert np.array_equal(sol[vod, 3], chs.current.export_solution())
    assert np.array_equal(sol[vod, 4], T)
    outside = np.ones(len(tip.nodes), bool)
    outside[vod] = False
    assert np.all(sol[outside, :4] == 0.0) and np.all(sol[outside, 4] == 300.0)


def test_heat_diagonal_matches_operator(ctx, tip):
    """The one-pass element diagonal (stiffness kappa(T) + mass gamma) equals e_i^T A e_i of
    the operator, seen through the lifting b_i = A_ii g_i on constrained rows."""
    gm = fgpu.Mesh(ctx, tip.nodes, tip.bulk_cells(), *MARKS)
    cur = _gpu_current(ctx, gm, _face_values(gm, 2, 5e-3, 6))
    xyz, _ = gm.dofs()
    heat = fgpu.System(gm)
    heat.set_solution(300.0 + 100.0 * np.exp(-((xyz[:, 2] - xyz[:, 2].max()) / 20.0) ** 2))
    heat.reset()
    heat.heat_terms(cur, O.CU_RHO_CP / DT, O.LORENTZ)
    heat.add_dirichlet(7, 1.0)
    heat.apply_dirichlet()
    b = heat.rhs()
    fixed = np.nonzero(heat.solution() == 1.0)[0]
    assert len(fixed) > 10
    for i in fixed[:: max(1, len(fixed) // 6)][:6]:
        e = np.zeros(gm.n_dofs)
        e[i] = 1.0
        aii = heat.apply(e)[i]
        assert abs(b[i] - aii) <= 1e-13 * abs(aii), (i, b[i], aii)


@pytest.mark.parametrize("precision", [fgpu.FGPU_FP64, fgpu.FGPU_MIXED])
def test_coupled_crank_nicolson_with_sigma_T(ctx, tip, precision):
    """The coupled extension: each step a current solve with sigma(T) = sigma0 / (1 + alpha
    (T - 300 K)) from the previous temperature, then a Crank-Nicolson heat step (theta = 1/2)
    with its Joule source, dt = 4 fs, 20 steps; the temperature of every step within 1e-9 of
    the FP64 oracle's."""
    sigma0, alpha = 1.0, 3.9e-3
    gm = fgpu.Mesh(ctx, tip.nodes, tip.bulk_cells(), *MARKS)
    J = _face_values(gm, 2, 3e-5, 4)         # larger currents run away as sigma(T) falls
    N = _face_values(gm, 2, 1e-9, 5)
    om = O.Mesh(tip.nodes, tip.bulk_cells(), *MARKS)
    xyz, _ = gm.dofs()
    T0 = 300.0 + 400.0 * np.exp(-((xyz[:, 2] - xyz[:, 2].max()) / 20.0) ** 2)
    cur, heat = fgpu.System(gm), fgpu.System(gm)
    heat.set_solution(T0)
    cur.fill_solution(0.0)
    To = T0.copy()
    o = fgpu.opts(tol=1e-13, relative=True, precision=precision)
    worst = 0.0
    for step in range(20):
        cur.reset()
        cur.sigma_terms(heat, sigma0, alpha)
        cur.add_neumann(2, J, 0.0)
        cur.add_dirichlet(7, 0.0)
        cur.apply_dirichlet()
        assert cur.solve(o).iterations >= 0, step
        heat.reset()
        heat.heat_terms(cur, O.CU_RHO_CP / DT, O.LORENTZ, theta=0.5)
        heat.add_neumann(2, N, 0.0)
        heat.add_dirichlet(7, 300.0)
        heat.apply_dirichlet()
        assert heat.solve(o).iterations > 0, step
        Tq, _ = om.qp_fields(To)
        _, phic, _ = O.current_solve(tip.nodes, tip.bulk_cells(), J, tol=1e-14, sigma_q=O.sigma_T(Tq, sigma0, alpha))
        To, it = O.heat_step(om, To, phic, DT, N, 300.0, tol=1e-14, theta=0.5)
        T = heat.solution()
        err = np.linalg.norm(T - To) / np.linalg.norm(To)
        worst = max(worst, err)
        assert err <= 1e-9, (step, err)
    assert np.abs(T - T0).max() > 10.0              # the steps did change the temperature
    with pytest.raises(fgpu.FgpuError):
        heat.reassemble()                            # the explicit term is not replayable


def test_solver_mirror_crank_nicolson_sigma_T(ctx, tip):
    """CurrentHeatSolver with the extensions switched on in its config (theta = 1/2,
    sigma(T)) gives the oracle's coupled steps."""
    conf = solvers.HeatingConfig(theta=0.5, sigma_T=(1.0, 3.9e-3), cg_tolerance=1e-13)
    chs = solvers.CurrentHeatSolver(conf=conf, ctx=ctx)
    assert chs.import_mesh(tip.nodes, tip.bulk_cells())
    chs.setup(300.0)
    J = _face_values(chs.mesh, 2, 3e-5, 7)
    N = _face_values(chs.mesh, 2, 1e-9, 8)
    chs.current.bc_values, chs.heat.bc_values = J, N
    chs.current.relative = chs.heat.relative = True
    om = O.Mesh(tip.nodes, tip.bulk_cells(), *MARKS)
    xyz, _ = chs.mesh.dofs()
    T0 = 300.0 + 200.0 * np.exp(-((xyz[:, 2] - xyz[:, 2].max()) / 20.0) ** 2)
    chs.heat.import_solution(T0)
    To = T0.copy()
    for step in range(3):
        chs.current.assemble()
        assert chs.current.solve() >= 0
        chs.heat.assemble(DT)
        assert chs.heat.solve() > 0
        Tq, _ = om.qp_fields(To)
        _, phic, _ = O.current_solve(tip.nodes, tip.bulk_cells(), J, tol=1e-14, sigma_q=O.sigma_T(Tq, 1.0, 3.9e-3))
        To, _ = O.heat_step(om, To, phic, DT, N, 300.0, tol=1e-14, theta=0.5)
        T = chs.heat.export_solution()
        assert np.linalg.norm(T - To) <= 1e-9 * np.linalg.norm(To), step

"""GPU parity of the Poisson path: matrix-free apply, right-hand side, lifting and the CG
solution against the CPU oracle on the same seeded meshes."""
import numpy as np
import pytest

from femocs_amd import fgpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MARKS = (8, 2, 4, 2)


def _pair(ctx, mesh):
    om = O.Mesh(mesh.nodes, mesh.vacuum_cells(), *MARKS)
    gm = fgpu.Mesh(ctx, mesh.nodes, mesh.vacuum_cells(), *MARKS)
    return om, gm


def test_mesh_import_matches_oracle(ctx, small_mesh):
    om, gm = _pair(ctx, small_mesh)
    assert gm.n_dofs == om.n_dofs and gm.n_cells == om.n_cells
    x, vod = gm.dofs()
    assert np.array_equal(vod, om.vertex_of_dof)
    assert np.array_equal(x, om.x)
    c, f, b = gm.bfaces()
    assert np.array_equal(c, om.bface_cell) and np.array_equal(f, om.bface_face)
    assert np.array_equal(b, om.bface_id)


def test_apply_matches_csr(ctx, small_mesh):
    om, gm = _pair(ctx, small_mesh)
    om.assemble()
    s = fgpu.System(gm)
    rng = np.random.default_rng(1)
    for _ in range(3):
        u = rng.standard_normal(om.n_dofs)
        y_ref = om.spmv(u)
        y = s.apply(u)
        assert np.linalg.norm(y - y_ref) <= 1e-13 * np.linalg.norm(y_ref)


def test_neumann_and_lifting_match_oracle(ctx, small_mesh):
    om, gm = _pair(ctx, small_mesh)
    om.assemble()
    b = om.face_rhs(8, constant=1.0)
    x = np.zeros(om.n_dofs)
    mask = om.dirichlet_mask(2)
    vals = np.where(mask > 0, 0.5, 0.0)
    om.apply_dirichlet(b, x, mask, vals)
    s = fgpu.System(gm)
    s.reset()
    s.add_dirichlet(2, 0.5)
    s.add_neumann(8, None, 1.0)
    s.apply_dirichlet()
    bg = s.rhs()
    assert np.linalg.norm(bg - b) <= 1e-13 * np.linalg.norm(b)


def test_charge_density_matches_oracle(ctx, small_mesh):
    """Write-step charge density: the right-hand side before the Dirichlet conditions (Neumann
    load and a cell source) over the dof volumes; zeros until saved and after a reset; refused
    once the conditions are applied."""
    om, gm = _pair(ctx, small_mesh)
    rng = np.random.default_rng(4)
    f_q = rng.uniform(-1.0, 1.0, 8 * om.n_cells)
    b = om.face_rhs(8, constant=-1.0) + om.cell_rhs(f_q)
    ref = om.charge_density(b)
    s = fgpu.System(gm)
    s.reset()
    s.add_dirichlet(2, 0.0)
    s.add_neumann(8, None, -1.0)
    s.add_cell_rhs(f_q)
    assert np.array_equal(s.charge_density(), np.zeros(om.n_dofs))
    s.save_charge_density()
    rho = s.charge_density()
    assert np.abs(rho - ref).max() <= 1e-12 * np.abs(ref).max()
    s.apply_dirichlet()
    assert np.array_equal(s.charge_density(), rho)
    with pytest.raises(fgpu.FgpuError):
        s.save_charge_density()
    s.reset()
    assert np.array_equal(s.charge_density(), np.zeros(om.n_dofs))


def test_solver_mirror_write_step(ctx, small_mesh):
    """PoissonSolver.assemble keeps the charge density on a write step only, as
    PoissonSolver::assemble does (src/PoissonSolver.cpp:196-207)."""
    from femocs_amd import solvers
    p = solvers.PoissonSolver(ctx=ctx)
    assert p.import_mesh(small_mesh.nodes, small_mesh.vacuum_cells())
    p.setup(-1.0, 0.0)
    p.assemble()
    assert not p.export_charge_dens().any()
    p.write_step = True
    p.assemble()
    om = O.Mesh(small_mesh.nodes, small_mesh.vacuum_cells(), *MARKS)
    ref = om.charge_density(om.face_rhs(8, constant=-1.0))
    rho = p.export_charge_dens()
    assert np.abs(rho - ref).max() <= 1e-12 * np.abs(ref).max()
    assert p.solve() > 0


def test_charge_density_follows_reassembly(ctx, small_mesh):
    """After a vertex move, fgpu_sys_reassemble recomputes a saved charge density for the moved
    mesh: equal to the oracle's on a fresh import of the moved vertices."""
    from femocs_amd import serving
    c = serving.Client(0, small_mesh, sigma=0.1, seed=7)
    gm = fgpu.Mesh(ctx, small_mesh.nodes, small_mesh.vacuum_cells(), *MARKS)
    s = fgpu.System(gm)
    s.reset()
    s.add_dirichlet(2, 0.0)
    s.add_neumann(8, None, -1.0)
    f_q = np.random.default_rng(8).uniform(-1.0, 1.0, 8 * gm.n_cells)
    s.add_cell_rhs(f_q)                     # charge on the dofs around the moved vertices too
    s.save_charge_density()
    s.apply_dirichlet()
    before = s.charge_density()
    ids, xyz = c.next_move()
    gm.move_vertices(ids, xyz)
    s.reassemble()
    moved = c.positions_after((ids, xyz))
    om = O.Mesh(moved, small_mesh.vacuum_cells(), *MARKS)
    ref = om.charge_density(om.face_rhs(8, constant=-1.0) + om.cell_rhs(f_q))
    rho = s.charge_density()
    assert not np.array_equal(rho, before)
    assert np.abs(rho - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("precision", [fgpu.FGPU_FP64, fgpu.FGPU_MIXED])
def test_poisson_solution_matches_oracle(ctx, small_mesh, precision):
    om, xo, it, res, _ = O.poisson_laplace(small_mesh.nodes, small_mesh.vacuum_cells(), -1.0,
                                           tol=1e-13, relative=True, precond="ssor")
    assert it > 0
    gm = fgpu.Mesh(ctx, small_mesh.nodes, small_mesh.vacuum_cells(), *MARKS)
    s = fgpu.System(gm)
    s.reset()
    s.add_dirichlet(2, 0.0)
    s.add_neumann(8, None, -1.0)
    s.apply_dirichlet()
    st = s.solve(fgpu.opts(tol=1e-12, relative=True, precision=precision))
    assert st.iterations > 0
    x = s.solution()
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    assert err <= 1e-10, err


def _poisson_system(ctx, mesh, field=-1.0):
    gm = fgpu.Mesh(ctx, mesh.nodes, mesh.vacuum_cells(), *MARKS)
    s = fgpu.System(gm)
    s.reset()
    s.add_dirichlet(2, 0.0)
    s.add_neumann(8, None, field)
    s.apply_dirichlet()
    s.fill_solution(0.0)
    return s


@pytest.fixture(scope="module")
def three_meshes(small_mesh):
    from femocs_amd import mesh as M
    m2 = M.make_tip_mesh(M.small_geometry(radius=10.0, height=24.0, seed=7))
    m3 = M.make_tip_mesh(M.small_geometry(radius=14.0, height=40.0, h_min=7.0, seed=99))
    return [small_mesh, m2, m3]


@pytest.mark.parametrize("precision", [fgpu.FGPU_FP64, fgpu.FGPU_MIXED])
def test_batch_results_equal_alone_and_in_mix(ctx, three_meshes, precision):
    """A system's solution is bit-identical whether it is solved alone or batched with others
    of different sizes (no floating-point atomics, per-system fixed reduction order)."""
    o = fgpu.opts(tol=1e-10, relative=True, precision=precision)
    alone = []
    for i, m in enumerate(three_meshes):
        s = _poisson_system(ctx, m, field=-1.0 - 0.1 * i)
        st = s.solve(o)
        assert st.iterations > 0
        alone.append((s.solution(), st.iterations))
    systems = [_poisson_system(ctx, m, field=-1.0 - 0.1 * i) for i, m in enumerate(three_meshes)]
    stats = ctx.solve_batch(systems, o)
    for (xa, ita), s, st in zip(alone, systems, stats):
        assert st.iterations == ita
        assert np.array_equal(s.solution(), xa)


def test_context_closed_before_children(small_mesh):
    c = fgpu.Context(0)
    s = _poisson_system(c, small_mesh)
    c.close()                      # deferred: the system still holds the context
    st = s.solve(fgpu.opts(tol=1e-8, relative=True))
    assert st.iterations > 0
    del s                          # the last child frees the context


@pytest.mark.parametrize("precision", [fgpu.FGPU_FP64, fgpu.FGPU_MIXED])
def test_continuous_batching_joins_and_leaves(ctx, three_meshes, precision):
    """Systems that join a running batch at different steps finish with the result and the
    iteration count they have alone."""
    o = fgpu.opts(tol=1e-10, relative=True, precision=precision, use_graph=False)
    alone = []
    for i, m in enumerate(three_meshes):
        s = _poisson_system(ctx, m, field=-1.0 - 0.1 * i)
        st = s.solve(o)
        alone.append((s.solution(), st.iterations))
    b = fgpu.Batch(ctx, o)
    systems = [_poisson_system(ctx, m, field=-1.0 - 0.1 * i) for i, m in enumerate(three_meshes)]
    done = {}
    b.add(systems[0])
    for step in range(400):
        if step == 3:
            b.add(systems[1])
        if step == 7:
            b.add(systems[2])
        b.step(16)
        for s, st in b.retire():
            done[systems.index(s)] = (s.solution(), st.iterations)
        if len(done) == 3:
            break
    assert len(done) == 3 and len(b) == 0
    for i in range(3):
        assert done[i][1] == alone[i][1]
        assert np.array_equal(done[i][0], alone[i][0])


def test_moved_vertices_match_fresh_import(ctx, small_mesh):
    """After a vertex update (same topology) the operator, the Jacobi diagonal and the solution
    follow the new coordinates: the result matches the oracle on the moved mesh."""
    s = _poisson_system(ctx, small_mesh)
    o = fgpu.opts(tol=1e-12, relative=True)
    assert s.solve(o).iterations > 0
    X = small_mesh.nodes.copy()
    lo, hi = X.min(axis=0), X.max(axis=0)
    # the box faces stay planes, so the boundary ids read at import still hold
    X[:, 0] += 0.3 * np.sin(np.pi * (X[:, 0] - lo[0]) / (hi[0] - lo[0])) * np.cos(X[:, 2] / 17.0)
    X[:, 2] *= 1.02
    s.mesh.update_vertices(X)
    s.reset()
    s.add_dirichlet(2, 0.0)
    s.add_neumann(8, None, -1.0)
    s.apply_dirichlet()
    s.fill_solution(0.0)
    st = s.solve(o)
    assert st.iterations > 0
    _, xo, it, _, _ = O.poisson_laplace(X, small_mesh.vacuum_cells(), -1.0, tol=1e-13,
                                        relative=True, precond="ssor")
    assert it > 0
    x = s.solution()
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)


def test_inverting_move_is_rejected(ctx, small_mesh):
    """A vertex update that turns cells inside out (here: a mirror image) is refused with
    FGPU_EMESH and leaves the mesh as it was."""
    s = _poisson_system(ctx, small_mesh)
    o = fgpu.opts(tol=1e-10, relative=True)
    s.solve(o)
    x0 = s.solution()
    X = small_mesh.nodes.copy()
    X[:, 0] = -X[:, 0]
    with pytest.raises(fgpu.MeshInvertedError):
        s.mesh.update_vertices(X)
    s2 = _poisson_system(ctx, small_mesh)
    s.reset()
    s.add_dirichlet(2, 0.0)
    s.add_neumann(8, None, -1.0)
    s.apply_dirichlet()
    s.fill_solution(0.0)
    s.solve(o)
    assert np.array_equal(s.solution(), x0)
    del s2


def test_move_and_reassemble_equal_fresh_assembly(ctx, small_mesh):
    """Moving a vertex subset and rebuilding the system on the device gives the right-hand
    side, lifting and solution of a fresh import of the moved mesh assembled from scratch."""
    from femocs_amd import serving
    c = serving.Client(0, small_mesh, sigma=0.1, seed=5)
    s = _poisson_system(ctx, small_mesh)
    ids, xyz = c.next_move()
    s.mesh.move_vertices(ids, xyz)
    s.reassemble()
    s.fill_solution(0.0)
    gm = fgpu.Mesh(ctx, small_mesh.nodes, small_mesh.vacuum_cells(), *MARKS)
    gm.update_vertices(c.positions_after((ids, xyz)))
    f = fgpu.System(gm)
    f.reset()
    f.add_dirichlet(2, 0.0)
    f.add_neumann(8, None, -1.0)
    f.apply_dirichlet()
    f.fill_solution(0.0)
    assert np.array_equal(s.rhs(), f.rhs())
    o = fgpu.opts(tol=1e-10, relative=True, precision=fgpu.FGPU_MIXED)
    st1, st2 = s.solve(o), f.solve(o)
    assert st1.iterations == st2.iterations > 0
    assert np.array_equal(s.solution(), f.solution())


def test_device_motion_map(ctx, small_mesh):
    """Control-point displacements applied through the motion map on the device place the
    vertices where the host-side averaging puts them, and the system built on them equals a
    fresh import of those positions; an inverting displacement is refused and reverted."""
    from femocs_amd import serving
    c = serving.Client(0, small_mesh, sigma=0.1, seed=9)
    s = _poisson_system(ctx, small_mesh)
    s.mesh.set_motion(c.moved_ids, c.W_moved)
    d = c.displacements()
    s.mesh.displace(d)
    s.reassemble()
    s.fill_solution(0.0)
    xyz, vod = s.mesh.dofs()
    X = small_mesh.nodes.copy()
    X[vod] = xyz
    Xh = c.positions_after((c.moved_ids, small_mesh.nodes[c.moved_ids] + c.W_moved @ d))
    # vertices of the other (bulk) hexes are not part of this mesh and stay put
    assert np.abs(X[vod] - Xh[vod]).max() <= 1e-12 * np.abs(Xh).max()
    gm = fgpu.Mesh(ctx, small_mesh.nodes, small_mesh.vacuum_cells(), *MARKS)
    gm.update_vertices(X)
    f = fgpu.System(gm)
    f.reset()
    f.add_dirichlet(2, 0.0)
    f.add_neumann(8, None, -1.0)
    f.apply_dirichlet()
    f.fill_solution(0.0)
    assert np.array_equal(s.rhs(), f.rhs())
    o = fgpu.opts(tol=1e-10, relative=True, precision=fgpu.FGPU_MIXED)
    assert s.solve(o).iterations == f.solve(o).iterations > 0
    assert np.array_equal(s.solution(), f.solution())
    with pytest.raises(fgpu.MeshInvertedError):
        s.mesh.displace(d * 1e4)
    xyz2, _ = s.mesh.dofs()
    assert np.array_equal(xyz2, xyz)


def test_tet_split_fp32_operator(ctx, small_mesh):
    """The FP32 operator of tet-split meshes (per-cell metric) is the FP64 operator to FP32
    precision, before and after a device-side move; mixed solves reach the same solution."""
    from femocs_amd import serving
    s = _poisson_system(ctx, small_mesh)
    gm = s.mesh
    assert gm.use_affine(True)
    x = np.random.default_rng(4).normal(size=gm.n_dofs)
    y64 = s.apply(x)
    ya = s.apply32(x)
    assert not gm.use_affine(False)
    yg = s.apply32(x)
    ea = np.linalg.norm(ya - y64) / np.linalg.norm(y64)
    eg = np.linalg.norm(yg - y64) / np.linalg.norm(y64)
    # FP32 rounding of element sums that cancel: both kernels alike
    assert ea <= 5e-5 and eg <= 5e-5 and ea <= 3 * eg, (ea, eg)
    assert gm.use_affine(True)
    c = serving.Client(0, small_mesh, sigma=0.1, seed=3)
    gm.set_motion(c.moved_ids, c.W_moved)
    gm.displace(c.displacements())
    s.reassemble()
    assert gm.use_affine(True)
    assert np.linalg.norm(s.apply32(x) - s.apply(x)) <= 5e-5 * np.linalg.norm(s.apply(x))
    o = fgpu.opts(tol=1e-10, relative=True, precision=fgpu.FGPU_MIXED)
    s.fill_solution(0.0)
    its_a = s.solve(o).iterations
    xa = s.solution()
    gm.use_affine(False)
    s.fill_solution(0.0)
    its_g = s.solve(o).iterations
    xg = s.solution()
    assert its_a > 0 and its_g > 0 and abs(its_a - its_g) <= 0.1 * its_g
    assert np.linalg.norm(xa - xg) <= 1e-9 * np.linalg.norm(xg)


def test_general_hex_mesh_keeps_trilinear_fp32_operator(ctx):
    """A mesh that is not a tet split (a distorted box of hexes) keeps the general kernel."""
    n = 6
    g = np.stack(np.meshgrid(*[np.arange(n + 1.0)] * 3, indexing="ij"), -1).reshape(-1, 3)
    g += 0.15 * np.sin(g[:, [1, 2, 0]])
    idx = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k
    cells = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                lex = [idx(i + a, j + b, k + c) for c in (0, 1) for b in (0, 1) for a in (0, 1)]
                cells.append([lex[u] for u in (0, 1, 5, 4, 2, 3, 7, 6)])
    gm = fgpu.Mesh(ctx, g, np.array(cells, np.int32), *MARKS)
    assert not gm.use_affine(True)
    s = fgpu.System(gm)
    x = np.random.default_rng(2).normal(size=gm.n_dofs)
    assert np.linalg.norm(s.apply32(x) - s.apply(x)) <= 5e-5 * np.linalg.norm(s.apply(x))


def test_reference_ordered_hexes_keep_the_tet_split_fp32_operator(ctx, small_mesh):
    """Hexes in the reference's vertex order (Tethex swaps the front and back faces of about
    half of them, which puts the sub-hex eta-flipped in lexicographic order) still run on the
    tet-split FP32 operator, with the same accuracy as the generator's order, and the mixed
    solution matches the oracle's."""
    from oracle import tethex
    nodes, _, hexes, _, _, _ = tethex.split(small_mesh.nodes[:small_mesh.n_tetnodes], small_mesh.tets)
    cells = hexes[small_mesh.hex_marker > 0]
    gm = fgpu.Mesh(ctx, nodes, cells, *MARKS)
    assert gm.use_affine(True)
    om = O.Mesh(nodes, cells, *MARKS)
    om.assemble()
    s = fgpu.System(gm)
    rng = np.random.default_rng(12)
    u = rng.standard_normal(om.n_dofs)
    y64 = om.spmv(u)
    ya = s.apply32(u)
    assert not gm.use_affine(False)
    yg = s.apply32(u)
    ea = np.linalg.norm(ya - y64) / np.linalg.norm(y64)
    eg = np.linalg.norm(yg - y64) / np.linalg.norm(y64)
    assert ea <= 5e-5 and eg <= 5e-5 and ea <= 3 * eg, (ea, eg)
    assert gm.use_affine(True)
    _, xo, it, _, _ = O.poisson_laplace(nodes, cells, -1.0, tol=1e-13, relative=True)
    s.reset()
    s.add_dirichlet(2, 0.0)
    s.add_neumann(8, None, -1.0)
    s.apply_dirichlet()
    s.fill_solution(0.0)
    st = s.solve(fgpu.opts(tol=1e-11, relative=True, precision=fgpu.FGPU_MIXED))
    assert st.iterations > 0
    assert np.linalg.norm(s.solution() - xo) <= 1e-9 * np.linalg.norm(xo)


def test_minmax_and_check_limits(ctx, small_mesh):
    """DealSolver::check_limits (src/DealSolver.cpp:158-167) through fgpu_sys_minmax: the
    extremes of the solution, and out of limits on either side."""
    from femocs_amd import solvers
    p = solvers.PoissonSolver(ctx=ctx)
    assert p.import_mesh(small_mesh.nodes, small_mesh.vacuum_cells())
    p.setup(-1.0, 0.0)
    p.assemble()
    assert p.solve() > 0
    x = p.export_solution()
    mn, mx = p.system.minmax()
    assert mn == x.min() and mx == x.max()
    assert not p.check_limits(mn, mx)
    assert p.stat.sol_min == mn and p.stat.sol_max == mx
    assert p.check_limits(mn + 1e-9 * abs(mx - mn), mx)          # minimum below the low limit
    assert p.check_limits(mn, mx - 1e-9 * abs(mx - mn))          # maximum above the high limit


@pytest.mark.parametrize("bid", [2, 8])
def test_face_centroids_match_oracle(ctx, small_mesh, bid):
    """DealSolver::export_surface_centroids (src/DealSolver.cpp:229-245): the centres of the
    faces with an id, in the cell x face order of the Neumann load."""
    om, gm = _pair(ctx, small_mesh)
    c = gm.face_centroids(bid)
    ref = om.face_centroids(bid)
    assert c.shape == ref.shape and len(c) > 0
    assert np.abs(c - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("precision", [fgpu.FGPU_FP64, fgpu.FGPU_MIXED])
def test_patches_per_workgroup_give_the_same_bits(ctx, small_mesh, precision):
    """Workgroups that run several patches in turn compute every patch as one-patch workgroups
    do: the applies and the solves are bit-identical for 1, 2 and 3 patches per workgroup (3: a
    partial last group)."""
    gm = fgpu.Mesh(ctx, small_mesh.nodes, small_mesh.vacuum_cells(), *MARKS)
    gm.set_patch_cells(64)
    s = fgpu.System(gm)
    rng = np.random.default_rng(31)
    u = rng.standard_normal(gm.n_dofs)
    outs = []
    for ppb in (1, 2, 3):
        s.set_patches_per_block(ppb)
        y32, y64 = s.apply32(u), s.apply(u)
        s.reset()
        s.add_dirichlet(2, 0.0)
        s.add_neumann(8, None, -1.0)
        s.apply_dirichlet()
        s.fill_solution(0.0)
        st = s.solve(fgpu.opts(tol=1e-10, relative=True, precision=precision))
        assert st.iterations > 0
        outs.append((y32, y64, st.iterations, s.solution()))
    for y32, y64, it, x in outs[1:]:
        assert np.array_equal(y32, outs[0][0]) and np.array_equal(y64, outs[0][1])
        assert it == outs[0][2] and np.array_equal(x, outs[0][3])

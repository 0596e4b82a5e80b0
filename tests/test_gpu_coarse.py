"""Two-level preconditioner of the mixed solver: the coarse operator E = Z^T A Z built on the
device (coloured applies) against the oracle's assembled matrix, and mixed solves with it
against FP64 solves (same tolerance, fewer iterations, bit-identical alone and batched)."""
import numpy as np
import pytest

from femocs_amd import fgpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MARKS = (8, 2, 4, 2)


def _system(ctx, mesh, field=-1.0):
    gm = fgpu.Mesh(ctx, mesh.nodes, mesh.vacuum_cells(), *MARKS)
    s = fgpu.System(gm)
    s.reset()
    s.add_dirichlet(2, 0.0)
    s.add_neumann(8, None, field)
    s.apply_dirichlet()
    s.fill_solution(0.0)
    return s


@pytest.fixture(scope="module")
def meshes(small_mesh):
    from femocs_amd import mesh as M
    m2 = M.make_tip_mesh(M.small_geometry(radius=10.0, height=24.0, seed=7))
    m3 = M.make_tip_mesh(M.small_geometry(radius=14.0, height=40.0, h_min=7.0, seed=99))
    return [small_mesh, m2, m3]


@pytest.mark.parametrize("agg", [256, 1024])
def test_coarse_operator_matches_assembled(ctx, meshes, agg):
    """E^-1 from the device equals the inverse of Z^T A Z of the oracle's CSR matrix, to its FP16
    storage."""
    import scipy.sparse as sp
    m = meshes[2]
    s = _system(ctx, m)
    ncol = s.set_coarse(agg)
    assert ncol > 1
    block, einv = s.coarse()
    om = O.Mesh(m.nodes, m.vacuum_cells(), *MARKS)
    om.assemble()
    b = om.face_rhs(8, constant=-1.0)
    x = np.zeros(om.n_dofs)
    mask = om.dirichlet_mask(2)
    om.apply_dirichlet(b, x, mask, None)
    rp, ci, va = om.csr()
    A = sp.csr_matrix((va, ci, rp), shape=(om.n_dofs, om.n_dofs))
    nc = einv.shape[0]
    assert nc == (om.n_dofs + agg - 1) // agg and block.max() == nc - 1
    free = mask == 0
    Z = sp.csr_matrix((free.astype(float), (np.arange(om.n_dofs), block)), shape=(om.n_dofs, nc))
    E = (Z.T @ (A @ Z)).toarray()
    empty = np.asarray(Z.sum(0)).ravel() == 0
    E[empty, :] = 0
    E[:, empty] = 0
    E[empty, empty] = 1
    ref = np.linalg.inv(E)
    # stored in FP16 with a power-of-two scale that puts the largest (diagonal) entry in
    # [2^13, 2^14): every entry is the FP16 rounding of the scaled inverse
    scale = 2.0 ** (np.frexp(np.diag(ref).max())[1] - 14)
    q = einv.astype(np.float64) / scale
    assert np.array_equal(q, q.astype(np.float16).astype(np.float64))
    assert 2 ** 13 <= np.abs(q).max() < 2 ** 14
    assert np.all(np.abs(einv - ref) <= 2.0 ** -11 * np.abs(ref) + 1e-12 * np.abs(ref).max())
    assert s.set_coarse(False) == 0 and s.coarse() is None


def test_two_level_mixed_solve(ctx, meshes):
    """Same answer as the FP64 solver to the stop tolerance, in fewer iterations than Jacobi."""
    m = meshes[2]
    o64 = fgpu.opts(tol=1e-12, relative=True)
    s = _system(ctx, m)
    assert s.solve(o64).iterations > 0
    x64 = s.solution()
    om = fgpu.opts(tol=1e-9, relative=True, precision=fgpu.FGPU_MIXED)
    s = _system(ctx, m)
    it_j = s.solve(om).iterations
    xj = s.solution()
    s = _system(ctx, m)
    s.set_coarse(True)
    st = s.solve(om)
    assert st.iterations > 0
    assert st.iterations < 0.8 * it_j, (st.iterations, it_j)
    x2 = s.solution()
    ej = np.linalg.norm(xj - x64) / np.linalg.norm(x64)
    e2 = np.linalg.norm(x2 - x64) / np.linalg.norm(x64)
    assert e2 <= 1e-8 and ej <= 1e-8, (e2, ej)
    assert st.residual <= 1e-9 * st.rhs_norm * 1.0001


def test_two_level_alone_equals_batched(ctx, meshes):
    """With the coarse level the solve stays deterministic: alone, in a batch and joining a
    running batch give the same bits and iterations; a batch may mix coarse and plain systems."""
    o = fgpu.opts(tol=1e-9, relative=True, precision=fgpu.FGPU_MIXED, use_graph=False)
    alone = []
    for i, m in enumerate(meshes):
        s = _system(ctx, m, field=-1.0 - 0.1 * i)
        if i != 1:
            s.set_coarse(True)
        st = s.solve(o)
        alone.append((s.solution(), st.iterations))
    systems = []
    for i, m in enumerate(meshes):
        s = _system(ctx, m, field=-1.0 - 0.1 * i)
        if i != 1:
            s.set_coarse(True)
        systems.append(s)
    stats = ctx.solve_batch(systems, o)
    for (xa, ita), s, st in zip(alone, systems, stats):
        assert st.iterations == ita
        assert np.array_equal(s.solution(), xa)
    b = fgpu.Batch(ctx, o)
    for s in systems:
        s.fill_solution(0.0)
    done = {}
    b.add(systems[0])
    for step in range(400):
        if step == 2:
            b.add(systems[1])
        if step == 5:
            b.add(systems[2])
        b.step(16)
        for s, st in b.retire():
            done[systems.index(s)] = (s.solution(), st.iterations)
        if len(done) == 3:
            break
    for i in range(3):
        assert done[i][1] == alone[i][1]
        assert np.array_equal(done[i][0], alone[i][0])


def test_stale_coarse_level_after_move(ctx, meshes):
    """The coarse operator of the unmoved mesh still preconditions the moved one: the solve
    converges to the tolerance with close to the fresh iteration count."""
    m = meshes[2]
    s = _system(ctx, m)
    s.set_coarse(True)
    X = m.nodes.copy()
    lo, hi = X.min(axis=0), X.max(axis=0)
    X[:, 0] += 0.3 * np.sin(np.pi * (X[:, 0] - lo[0]) / (hi[0] - lo[0])) * np.cos(X[:, 2] / 17.0)
    s.mesh.update_vertices(X)
    s.reassemble()
    s.fill_solution(0.0)
    o = fgpu.opts(tol=1e-9, relative=True, precision=fgpu.FGPU_MIXED)
    st_stale = s.solve(o)
    s.set_coarse(True)
    s.fill_solution(0.0)
    st_fresh = s.solve(o)
    assert st_stale.iterations > 0 and st_fresh.iterations > 0
    assert st_stale.iterations <= 1.2 * st_fresh.iterations + 5

"""Readers on the GPU path: the analytic check of a hemisphere on a plane in a uniform field
(FieldReader, src/SolutionReader.cpp:465-589) on a field solved by the library, the export of
what a reader interpolated, and the heat reader's transfer to the heat system's vertices
(HeatReader::interpolate_dofs, :670-683)."""
import numpy as np
import pytest

from femocs_amd import fgpu, interp, solvers
from femocs_amd import mesh as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu
R, E0 = 12.0, 1.0


@pytest.fixture(scope="module")
def hemisphere(ctx):
    g = M.TipGeometry(radius=R, height=0.0, h_min=3.0, growth=0.3, slab_width=160.0, slab_thickness=8.0,
                      top_factor=12.0)
    tip = M.make_tip_mesh(g)
    conf = solvers.FieldConfig(E0=E0, cg_tolerance=1e-10)
    p = solvers.PoissonSolver(conf=conf, ctx=ctx)
    p.relative = True
    assert p.import_mesh(tip.nodes, tip.vacuum_cells())
    p.setup(-conf.E0, conf.V0)
    p.assemble()
    assert p.solve() > 0
    ip = interp.Interpolator(ctx, tip, interp.TYPES_VACUUM)
    ip.extract_solution(p.system, smoothen=True)
    return g, tip, conf, p, ip


def _shell_points(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((4 * n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p = d * rng.uniform(lo, hi, (4 * n, 1))
    return p[p[:, 2] > 0.5 * R][:n]


def test_hemisphere_potential_and_enhancement(hemisphere):
    """The solved potential follows phi = -E0 z (1 - (R/r)^3) to within the faceting of the
    atomistic surface (10 % between 2R and 3R), the field too (median 10 %), and the apex
    enhancement measured on the cap atoms passes the reference's default limits (0.1 to 5 times
    the hemisphere's 3; the smoothed nodal field on this coarse surface measures about 0.4)."""
    g, tip, conf, p, ip = hemisphere
    fr = interp.FieldReader(ip, rank=1)
    fr.set_check_params(conf, radius=R, tip_height=R, box_height=g.z_top)
    assert fr.get_analyt_enhancement() == 3.0
    pts = _shell_points(400, 2 * R, 3 * R, 1)
    sol = fr.interpolate(pts)
    ana = np.array([fr.get_analyt_potential(i) for i in range(fr.size())])
    assert np.all(np.abs(sol[:, 4] - ana) <= 0.1 * np.abs(ana))
    Ea = np.array([fr.get_analyt_field(i) for i in range(fr.size())])
    assert np.median(np.linalg.norm(sol[:, :3] - Ea, axis=1) / np.linalg.norm(Ea, axis=1)) < 0.1
    cap = tip.atoms[tip.atoms[:, 2] > 0.5 * R]
    fr.interpolate(cap)
    assert not fr.check_limits()
    assert 0.3 < fr.get_beta() < 1.2
    assert fr.E_max == max(fr.get_elfield_norm(i) for i in range(fr.size()))


def test_interpolate_results_exports_what_was_interpolated(hemisphere):
    g, tip, conf, p, ip = hemisphere
    fr = interp.FieldReader(ip, rank=2)
    pts = _shell_points(300, 1.5 * R, 4 * R, 2)
    sol = fr.interpolate(pts)
    n = len(pts)
    phi = np.full(n, 5.0)
    assert fr.interpolate_results(n, "POTENTIAL", pts[:, 0], pts[:, 1], pts[:, 2], phi) == 0
    assert np.array_equal(phi, sol[:, 4])
    E = np.zeros(3 * n)
    assert fr.export_results(n, "Elfield", E) == 0
    assert np.array_equal(E.reshape(-1, 3), sol[:, :3])
    rho = np.zeros(n)
    fr.export_results(n, "charge_density", rho)
    assert not rho.any()                     # not a write step: zeros


def test_heat_reader_hands_back_vertex_values(ctx, hemisphere):
    """interpolate_dofs installs the rank-3 (linear hex) interpolation at the system's own
    vertices as its solution, bit-exact with the oracle; at the vertices located inside the
    search region it reproduces the nodal scalar2 (vertices on edges shared with the other
    region can test outside every tet of the region and are extrapolated from the nearest,
    as the reference does).  Atoms other than SURFACE and BULK are skipped."""
    g, tip, conf, p, ip = hemisphere
    hr = interp.HeatReader(ip, rank=3)
    target = fgpu.System(p.mesh)
    phi_dof = p.system.solution()
    t = hr.interpolate_dofs(target)
    assert np.array_equal(target.solution(), t)
    xyz, _ = p.mesh.dofs()
    L = O.LinTet(tip.nodes, tip.tets, tip.tet_nbr, tip.tet_marker, 3)
    q10 = O.quadtet_cells(tip.tets, tip.hexes, tip.node_kind)
    cells, ref = O.interp_points(L, tip.nodes, tip.tets, q10, tip.hexes, ip.nodal(), xyz, 3)
    assert np.array_equal(t, ref[:, 4])
    inside = cells >= 0
    assert inside.mean() > 0.9
    assert np.abs(t - phi_dof)[inside].max() <= 1e-9 * np.abs(phi_dof).max()
    markers = np.resize(np.array([interp.TYPES_SURFACE, interp.TYPES_BULK, interp.TYPES_VACUUM]), len(xyz))
    sol = hr.interpolate_atoms(xyz, markers)
    keep = markers != interp.TYPES_VACUUM
    assert hr.size() == keep.sum() and np.array_equal(hr.ids, np.nonzero(keep)[0])
    assert np.abs(sol[:, 4] - t[keep])[inside[keep]].max() <= 1e-9 * np.abs(phi_dof).max()

"""Host side of the readers: label lookup and the export of interpolated values into caller
arrays (SolutionReader::contains / export_results, src/SolutionReader.cpp:269-421), and the
analytic hemisphere check of FieldReader (src/SolutionReader.cpp:498-589), pinned against
finite differences, boundary values and known limits rather than its own formulas."""
import numpy as np
import pytest

from femocs_amd import interp
from femocs_amd.solvers import FieldConfig


def _reader(sol, ids):
    r = interp.FieldReader(None)
    r.interpolation = np.asarray(sol, np.float64)
    r.ids = np.asarray(ids, np.int64)
    r.points = np.zeros((len(ids), 3))
    return r


def test_labels_own_spelling_appends_upper_case_overwrites():
    r = interp.FieldReader(None)
    assert [r.contains(s) for s in ("elfield", "elfield_norm", "charge_density", "potential")] == [1, 2, 3, 4]
    assert [r.contains(s) for s in ("Elfield", "ELFIELD_NORM", "Charge_Density", "POTENTIAL")] == [-1, -2, -3, -4]
    assert r.contains("temperature") == 0
    h = interp.HeatReader(None)
    assert [h.contains(s) for s in ("rho", "rho_norm", "potential", "Temperature")] == [1, 2, 3, -4]


def test_export_layout_ids_and_modes():
    rng = np.random.default_rng(3)
    sol = rng.standard_normal((4, 5))
    ids = [2, 0, 7, -1]                        # ids 7 and -1 fall outside 5 points
    r = _reader(sol, ids)
    n = 5
    # vector, overwrite: x1 y1 z1 x2 ..., untouched points cleared
    data = np.full(3 * n, 9.0)
    assert r.export_results(n, "Elfield", data) == 0
    expect = np.zeros(3 * n)
    expect[6:9], expect[0:3] = sol[0, :3], sol[1, :3]
    assert np.array_equal(data, expect)
    # vector, append
    data = np.ones(3 * n)
    r.export_results(n, "elfield", data)
    assert np.array_equal(data, expect + 1.0)
    # norm and both scalars
    for label, col in (("elfield_norm", None), ("charge_density", 3), ("potential", 4)):
        vals = np.linalg.norm(sol[:, :3], axis=1) if col is None else sol[:, col]
        data = np.full(n, 2.0)
        r.export_results(n, label.upper(), data)
        e = np.zeros(n)
        e[2], e[0] = vals[0], vals[1]
        assert np.array_equal(data, e)
        data = np.full(n, 2.0)
        r.export_results(n, label, data)
        assert np.array_equal(data, np.full(n, 2.0) + e)
    with pytest.raises(ValueError):
        r.export_results(n, "temperature", np.zeros(n))
    assert interp.FieldReader(None).export_results(n, "elfield", np.zeros(3 * n)) == 1


def test_export_appends_repeated_ids():
    r = _reader(np.arange(10.0).reshape(2, 5), [1, 1])
    data = np.zeros(3)
    r.export_results(3, "potential", data)
    assert np.array_equal(data, [0.0, 4.0 + 9.0, 0.0])


def _hemisphere_reader(points, E0=0.7, R=12.0):
    r = interp.FieldReader(None)
    r.E0, r.radius1, r.radius2 = E0, R, R
    r.points = np.asarray(points, np.float64)
    r.interpolation = np.zeros((len(points), 5))
    r.ids = np.arange(len(points))
    return r


def test_analytic_potential_vanishes_on_the_metal_and_tends_to_the_uniform_field():
    R, E0 = 12.0, 0.7
    rng = np.random.default_rng(5)
    d = rng.standard_normal((20, 3))
    d[:, 2] = np.abs(d[:, 2])
    on_sphere = R * d / np.linalg.norm(d, axis=1, keepdims=True)
    on_plane = np.c_[rng.uniform(20, 80, (20, 2)), np.zeros(20)]
    far = np.array([[0.0, 0.0, 1e5], [3e4, 0.0, 2e4]])
    r = _hemisphere_reader(np.vstack([on_sphere, on_plane, far]), E0, R)
    phi = np.array([r.get_analyt_potential(i) for i in range(r.size())])
    assert np.abs(phi[:40]).max() < 1e-12
    assert np.allclose(phi[40:], -E0 * r.points[40:, 2], rtol=1e-9)
    # the apex field is three times the applied one (the hemisphere's enhancement)
    r2 = _hemisphere_reader([[0.0, 0.0, R]], E0, R)
    assert np.allclose(r2.get_analyt_field(0), [0.0, 0.0, 3 * E0], atol=1e-15)
    assert r2.get_analyt_enhancement() == 3.0


def test_analytic_field_is_minus_the_potential_gradient():
    rng = np.random.default_rng(6)
    p = rng.uniform(-40, 40, (30, 3))
    p[:, 2] = np.abs(p[:, 2]) + 13.0
    h = 1e-5
    origin = np.array([1.0, -2.0, 0.5])
    r = _hemisphere_reader(np.vstack([p + origin] + [p + origin + h * e for e in np.eye(3)]
                                     + [p + origin - h * e for e in np.eye(3)]))
    n = len(p)
    for i in range(n):
        E = r.get_analyt_field(i, origin)
        g = [(r.get_analyt_potential(n * (1 + k) + i, origin) - r.get_analyt_potential(n * (4 + k) + i, origin))
             / (2 * h) for k in range(3)]
        assert np.allclose(E, -np.array(g), rtol=1e-6, atol=1e-9)


def test_hemi_ellipsoid_enhancement_limits():
    """gamma -> 3 as the ellipsoid becomes a sphere; gamma ~ nu^2 / (ln(2 nu) - 1) for a
    slender one (the textbook asymptote of a prolate conductor)."""
    r = interp.FieldReader(None)
    r.radius1, r.radius2 = 10.0, 10.0 * (1 + 1e-4)
    assert abs(r.get_analyt_enhancement() - 3.0) < 1e-3
    r.radius2 = 10.0 * 2000
    nu = 2000.0
    assert abs(r.get_analyt_enhancement() / (nu * nu / (np.log(2 * nu) - 1)) - 1) < 1e-3
    r.radius1 = 0.0
    with pytest.raises(ValueError):
        r.get_analyt_enhancement()


def test_check_params_and_limits():
    r = interp.FieldReader(None)
    r.set_check_params(FieldConfig(E0=0.5, anode_BC="neumann"), radius=10.0, tip_height=40.0, box_height=400.0)
    assert (r.E0, r.radius1, r.radius2) == (0.5, 20.0, 40.0)
    r.set_check_params(FieldConfig(V0=100.0, anode_BC="dirichlet"), radius=10.0, tip_height=5.0, box_height=400.0)
    assert (r.E0, r.radius1, r.radius2) == (0.25, 10.0, 10.0)
    assert (r.limit_min, r.limit_max) == (0.1, 5.0)
    with pytest.raises(ValueError):
        r.set_check_params(FieldConfig(anode_BC="robin"), 1.0, 1.0, 1.0)
    sol = np.zeros((3, 5))
    sol[1, :3] = [0.0, 0.3, 0.4]               # |E| = 0.5 -> gamma1 = 2, beta = 2/3
    assert not r.check_limits(sol)
    assert abs(r.get_beta() - 2.0 / 3.0) < 1e-15
    sol[1, :3] *= 0.1                           # beta 1/15 < 0.1
    assert r.check_limits(sol)
    r.limit_min = r.limit_max = 1.0             # equal limits switch the check off
    assert not r.check_limits(sol)

"""Interpolator tables built on the device against the host build (LinearTetrahedra::precompute
and its neighbour lists, src/InterpolatorCells.cpp:523-742; QuadraticTetrahedra::calc_cell,
:1151-1174; the node -> hex and pseudo-Voronoi lists of Interpolator::initialize and
TetgenMesh::calc_pseudo_3D_vorocells; the search tree): every table, including the host copy
of the tet tables the chained guess reads, hashes to the same bits; transfers to atoms give the
same cells and values on either build."""
import numpy as np
import pytest

from femocs_amd import fgpu, interp
from femocs_amd import mesh as M

pytestmark = pytest.mark.gpu
MARKS = (8, 2, 4, 2)


def _same(ctx, tip, region):
    print(f"interpolator tables: {len(tip.tets)} tets, region {region}", flush=True)
    dev = interp.Interpolator(ctx, tip, region)
    host = interp.Interpolator(ctx, tip, region, host_build=True)
    a, b = dev.digest(), host.digest()
    assert [i for i in range(len(a)) if a[i] != b[i]] == []
    return dev, host


def test_small_tip_both_regions(ctx, small_mesh):
    _same(ctx, small_mesh, interp.TYPES_VACUUM)
    _same(ctx, small_mesh, interp.TYPES_BULK)


def test_transfer_is_identical_on_either_build(ctx, small_mesh):
    dev, host = _same(ctx, small_mesh, interp.TYPES_VACUUM)
    gm = fgpu.Mesh(ctx, small_mesh.nodes, small_mesh.vacuum_cells(), *MARKS)
    s = fgpu.System(gm)
    s.reset()
    s.add_dirichlet(2, 0.0)
    s.add_neumann(8, None, -1.0)
    s.apply_dirichlet()
    s.fill_solution(0.0)
    assert s.solve(fgpu.opts(tol=1e-9, relative=True)).iterations > 0
    rng = np.random.default_rng(3)
    lo, hi = small_mesh.nodes.min(0), small_mesh.nodes.max(0)
    pts = np.concatenate([small_mesh.atoms[:2000], rng.uniform(lo, hi, (2000, 3))])
    for ip in (dev, host):
        ip.extract_solution(s, smoothen=True)
    for rank in (1, 2, 3):
        cd, vd = dev.points(pts, rank)
        ch, vh = host.points(pts, rank)
        assert np.array_equal(cd, ch) and np.array_equal(vd, vh)
    assert np.array_equal(dev.nodal(), host.nodal())


def test_large_tip(ctx):
    tip = M.make_tip_mesh(M.tip_r10nm_geometry(), with_atoms=False)
    _same(ctx, tip, interp.TYPES_VACUUM)

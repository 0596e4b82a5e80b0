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


def _create(ctx, tip, nbr, n_hexes, flags):
    import ctypes as C
    h = fgpu._P()
    hexes = np.ascontiguousarray(tip.hexes if n_hexes else np.zeros((0, 8)), np.int32).reshape(-1)
    hm = np.ascontiguousarray(tip.hex_marker if n_hexes else np.zeros(0), np.int32)
    rc = ctx.lib.fgpu_interp_create_ex(
        ctx.h, len(tip.nodes), np.ascontiguousarray(tip.nodes, np.float64).reshape(-1),
        np.ascontiguousarray(tip.node_kind, np.int32), len(tip.tets), np.ascontiguousarray(tip.tets, np.int32).reshape(-1),
        np.ascontiguousarray(nbr, np.int32).reshape(-1), np.ascontiguousarray(tip.tet_marker, np.int32),
        n_hexes, hexes, hm, interp.TYPES_VACUUM, 0.0, flags, C.byref(h))
    return rc, h


def _digest(ctx, h):
    import ctypes as C
    out = (C.c_ulonglong * fgpu.FGPU_INTERP_DIGEST_TABLES)()
    ctx.check(ctx.lib.fgpu_interp_digest(h, fgpu.FGPU_INTERP_DIGEST_TABLES, out), "fgpu_interp_digest")
    return list(out)


def test_tets_only_and_invalid_neighbours(ctx, small_mesh):
    """Without hexes (no 10-node tets, no node lists) both builds agree; a neighbour index that
    is neither -1 nor a tet is refused before anything is built."""
    hs = []
    for flags in (0, fgpu.FGPU_INTERP_HOST):
        rc, h = _create(ctx, small_mesh, small_mesh.tet_nbr, 0, flags)
        assert rc == fgpu.FGPU_OK
        hs.append(h)
    try:
        assert _digest(ctx, hs[0]) == _digest(ctx, hs[1])
    finally:
        for h in hs:
            ctx.lib.fgpu_interp_destroy(h)
    bad = np.array(small_mesh.tet_nbr, np.int32)
    bad[5, 2] = len(small_mesh.tets)
    for flags in (0, fgpu.FGPU_INTERP_HOST):
        rc, h = _create(ctx, small_mesh, bad, len(small_mesh.hexes), flags)
        assert rc == fgpu.FGPU_EINVAL and not h

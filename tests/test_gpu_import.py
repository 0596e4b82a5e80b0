"""Mesh import on the device against the host import (the deal.II import_mesh + mark_boundary
steps, src/DealSolver.cpp:192-209,461-518, and the internal order of the element kernels):
every table, host mirror and device array, hashes to the same bits on generated tips (patches
of 64 and 256 cells, the node cap binding), on the reference's own mesh file, and on a shuffled
grid with unused vertices and inverted cells; invalid meshes fail with the same message; a
solve on either import gives the same bits."""
import os

import numpy as np
import pytest

from femocs_amd import fgpu, gmsh
from femocs_amd import mesh as M

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
MARKS = (8, 2, 4, 2)
LEX_TO_UCD = [0, 1, 4, 5, 3, 2, 7, 6]


def _same(ctx, nodes, cells):
    print(f"import: {len(nodes)} vertices, {len(cells)} cells", flush=True)
    dev = fgpu.Mesh(ctx, nodes, cells, *MARKS)
    host = fgpu.Mesh(ctx, nodes, cells, *MARKS, host_import=True)
    assert (dev.n_dofs, dev.n_cells, dev.n_bfaces) == (host.n_dofs, host.n_cells, host.n_bfaces)
    a, b = dev.digest(), host.digest()
    assert [i for i in range(len(a)) if a[i] != b[i]] == []
    return dev, host


def _grid(n, h=0.5, seed=0):
    """Structured grid in UCD order with shuffled vertex numbers, unused vertices between the
    used ones and every third cell given inside out (its UCD faces swapped)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    ijk = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    vid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
    cells = np.empty((len(ijk), 8), np.int64)
    for l in range(8):
        cells[:, LEX_TO_UCD[l]] = vid(ijk[:, 0] + (l & 1), ijk[:, 1] + ((l >> 1) & 1), ijk[:, 2] + (l >> 2))
    P = np.stack(np.meshgrid(*[np.arange(k + 1) * h for k in n], indexing="ij"), -1).reshape(-1, 3)
    P += rng.uniform(-0.05 * h, 0.05 * h, P.shape)
    inv = np.arange(len(cells)) % 3 == 0
    cells[inv] = cells[inv][:, [4, 5, 6, 7, 0, 1, 2, 3]]
    nv = len(P)
    perm = rng.permutation(2 * nv)[:nv]           # new number of each vertex, half the slots unused
    X = rng.uniform(-1, 1, (2 * nv, 3))
    X[perm] = P
    return X, perm[cells]


def test_small_tip_vacuum_and_bulk(ctx, small_mesh):
    _same(ctx, small_mesh.nodes, small_mesh.vacuum_cells())
    _same(ctx, small_mesh.nodes, small_mesh.bulk_cells())


def test_shuffled_grid_with_unused_vertices_and_inverted_cells(ctx):
    X, cells = _grid((9, 7, 5))
    dev, _ = _same(ctx, X, cells)
    assert dev.n_dofs == 10 * 8 * 6
    X, cells = _grid((72, 64, 60), seed=1)        # 276k cells: patches of 256 cells
    _same(ctx, X, cells)


def test_reference_mesh_file(ctx):
    nodes, tets, _, tris, _ = gmsh.read_msh(os.path.join(HERE, "golden", "hemicone.msh"))
    used = np.unique(tets)
    remap = -np.ones(len(nodes), np.int64)
    remap[used] = np.arange(len(used))
    nodes, tets = nodes[used], remap[tets]
    xyz, _, hexes, _, _, _ = fgpu.tethex_split(ctx, nodes, tets)
    _same(ctx, xyz, hexes)


def test_large_tip_where_the_patch_cap_binds(ctx):
    tip = M.make_tip_mesh(M.tip_r10nm_geometry(), with_atoms=False)
    dev, _ = _same(ctx, tip.nodes, tip.vacuum_cells())
    assert dev.n_cells > 256 * 1024


def test_invalid_meshes_fail_alike(ctx):
    X, cells = _grid((2, 2, 2))
    for host in (False, True):
        bad = cells.copy()
        bad[3, 5] = len(X)
        with pytest.raises(fgpu.FgpuError, match="out of range"):
            fgpu.Mesh(ctx, X, bad, *MARKS, host_import=host)
        with pytest.raises(fgpu.FgpuError, match="more than two cells"):
            fgpu.Mesh(ctx, X, np.concatenate([cells, cells[:1], cells[:1]]), *MARKS, host_import=host)


def test_solve_on_either_import_is_identical(ctx, small_mesh):
    out = []
    for host in (False, True):
        gm = fgpu.Mesh(ctx, small_mesh.nodes, small_mesh.vacuum_cells(), *MARKS, host_import=host)
        s = fgpu.System(gm)
        s.reset()
        s.add_dirichlet(2, 0.0)
        s.add_neumann(8, None, -1.0)
        s.apply_dirichlet()
        s.fill_solution(0.0)
        st = s.solve(fgpu.opts(tol=1e-9, relative=True))
        out.append((st.iterations, s.solution()))
    assert out[0][0] == out[1][0] > 0
    assert np.array_equal(out[0][1], out[1][1])


def test_patch_rebuild_after_either_import(ctx, small_mesh):
    """A new patch size rebuilds the patches from the internal cells, which a device import keeps
    on the device only: both imports still agree table by table and solve to the same bits."""
    dev, host = _same(ctx, small_mesh.nodes, small_mesh.vacuum_cells())
    out = []
    for m in (dev, host):
        m.set_patch_cells(128)
    assert dev.digest() == host.digest()
    for m in (dev, host):
        s = fgpu.System(m)
        s.reset()
        s.add_dirichlet(2, 0.0)
        s.add_neumann(8, None, -1.0)
        s.apply_dirichlet()
        s.fill_solution(0.0)
        st = s.solve(fgpu.opts(tol=1e-9, relative=True))
        out.append((st.iterations, s.solution()))
    assert out[0][0] == out[1][0] > 0
    assert np.array_equal(out[0][1], out[1][1])


def test_collapsed_cells_where_the_patch_cap_binds(ctx):
    """Cells with a repeated vertex (one edge collapsed) count each distinct vertex once in the
    patch node cap on both imports: the same tables where the cap binds."""
    tip = M.make_tip_mesh(M.tip_r10nm_geometry(), with_atoms=False)
    cells = np.array(tip.vacuum_cells(), copy=True)
    cells[::997, 1] = cells[::997, 0]              # collapse the edge 0-1 of every 997th cell
    _same(ctx, tip.nodes, cells)

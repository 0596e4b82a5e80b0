"""Autotuner of the element kernel's work decomposition and of the CG's launch unrolling.

The matrix-free apply processes patches of consecutive cells per workgroup (their nodes
staged in LDS); two knobs trade LDS reuse against parallelism and latency hiding:
cells per patch (the elements-per-block of the element kernel) and patches handled in turn
by one workgroup.  The best pair depends on the mesh size and on the precision, so it is
measured on the system's own data: each candidate times the element apply plus the dof
gather that consumes its output (the patch size changes the gather's work too).

The third knob is the unroll of a batched solve: how many CG iterations are captured in one
hipGraph between two host convergence checks (more: fewer host round trips; fewer: less
work after a system has converged).  It is timed on whole solves of the batch itself.
"""
from __future__ import annotations

import time

from . import fgpu

UNROLLS = (8, 16, 32, 64)

CELLS = (64, 128, 256)
PATCHES_PER_BLOCK = (1, 2, 4, 8)


def autotune(system: fgpu.System, precision: int = fgpu.FGPU_MIXED, cells=CELLS, ppbs=PATCHES_PER_BLOCK,
             reps: int = 20, apply: bool = True):
    """Returns (best (cells, ppb), table of {(cells, ppb): ms per apply + gather}) and, with
    apply=True, leaves the system's mesh and the system set to the best pair."""
    mesh = system.mesh
    table = {}
    for c in cells:
        mesh.set_patch_cells(c)
        gather = system.time_kernel(1, precision, reps)
        for ppb in ppbs:
            system.set_patches_per_block(ppb)
            table[(c, ppb)] = system.time_kernel(0, precision, reps) + gather
    best = min(table, key=table.get)
    if apply:
        mesh.set_patch_cells(best[0])
        system.set_patches_per_block(best[1])
    return best, table


def autotune_unroll(ctx: fgpu.Context, systems, opts: fgpu.SolveOpts, unrolls=UNROLLS):
    """Times a solve of the batch from zero for every unroll (CG iterations per captured graph);
    returns (best unroll, {unroll: seconds}).  The systems are left at their last solution."""
    table = {}
    for u in unrolls:
        opts.check_every = int(u)
        opts.use_graph = 1
        for s in systems:
            s.fill_solution(0.0)
        ctx.sync()
        t = time.perf_counter()
        st = ctx.solve_batch(systems, opts)
        table[u] = time.perf_counter() - t
        if min(x.iterations for x in st) <= 0:
            raise RuntimeError("a batched solve did not converge while tuning the unroll")
    return min(table, key=table.get), table

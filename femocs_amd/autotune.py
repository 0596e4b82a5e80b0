"""Autotuner of the element kernel's work decomposition.

The matrix-free apply processes patches of consecutive cells per workgroup (their nodes
staged in LDS); two knobs trade LDS reuse against parallelism and latency hiding:
cells per patch (the elements-per-block of the element kernel) and patches handled in turn
by one workgroup.  The best pair depends on the mesh size and on the precision, so it is
measured on the system's own data: each candidate times the element apply plus the dof
gather that consumes its output (the patch size changes the gather's work too).
"""
from __future__ import annotations

from . import fgpu

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

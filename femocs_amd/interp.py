"""Field and temperature transfer to atoms on the GPU, mirroring the reference's Interpolator
(src/Interpolator.cpp) and SolutionReader (src/SolutionReader.cpp) for dim 3 and, on the surface faces, dim 2 rank 1.

    ip = Interpolator(ctx, tip, search_region=TYPES_VACUUM)
    ip.extract_solution(poisson_system)            # nodal (E, rho, phi), smoothed
    reader = SolutionReader(ip, rank=1)
    sol = reader.calc_interpolation(atom_xyz)       # (n, 5): Ex Ey Ez rho phi

The C library does the work; this module only holds the handles and the atom-to-cell map.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import fgpu

TYPES_BULK = 1
TYPES_VACUUM = 3


class Interpolator:
    """Interpolator::initialize (src/Interpolator.cpp:28-77) on a tet mesh and its tet-split
    hexes (hex 4t+k around vertex k of tet t)."""

    def __init__(self, ctx: fgpu.Context, tip, search_region: int = TYPES_VACUUM, empty_value: float = 0.0):
        self.ctx = ctx
        self.n_nodes = len(tip.nodes)
        self.search_region = search_region
        h = fgpu._P()
        ctx.check(ctx.lib.fgpu_interp_create(
            ctx.h, self.n_nodes, np.ascontiguousarray(tip.nodes, np.float64).reshape(-1),
            np.ascontiguousarray(tip.node_kind, np.int32), len(tip.tets),
            np.ascontiguousarray(tip.tets, np.int32).reshape(-1),
            np.ascontiguousarray(tip.tet_nbr, np.int32).reshape(-1),
            np.ascontiguousarray(tip.tet_marker, np.int32), len(tip.hexes),
            np.ascontiguousarray(tip.hexes, np.int32).reshape(-1),
            np.ascontiguousarray(tip.hex_marker, np.int32), int(search_region), float(empty_value), C.byref(h)),
            "fgpu_interp_create")
        self.h = h

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.fgpu_interp_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def extract_solution(self, system: fgpu.System, smoothen: bool = True):
        """Interpolator::extract_solution(PoissonSolver&, smoothen) (src/Interpolator.cpp:172-190)."""
        self.ctx.check(self.ctx.lib.fgpu_interp_extract(self.h, system.h, int(bool(smoothen))), "fgpu_interp_extract")

    def extract_heat(self, current: fgpu.System, heat: fgpu.System):
        """Interpolator::extract_solution(CurrentHeatSolver&) (src/Interpolator.cpp:207-216)."""
        self.ctx.check(self.ctx.lib.fgpu_interp_extract_heat(self.h, current.h, heat.h), "fgpu_interp_extract_heat")

    def nodal(self) -> np.ndarray:
        out = np.zeros(5 * self.n_nodes)
        self.ctx.check(self.ctx.lib.fgpu_interp_get_nodal(self.h, out), "fgpu_interp_get_nodal")
        return out.reshape(-1, 5)

    def set_nodal(self, sol5):
        s = np.ascontiguousarray(sol5, np.float64).reshape(-1)
        if s.size != 5 * self.n_nodes:
            raise ValueError("5 values per node")
        self.ctx.check(self.ctx.lib.fgpu_interp_set_nodal(self.h, s), "fgpu_interp_set_nodal")

    def points(self, xyz, rank: int):
        pts = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        n = pts.size // 3
        cells = np.zeros(n, np.int32)
        out = np.zeros(5 * n)
        self.ctx.check(self.ctx.lib.fgpu_interp_points(self.h, n, pts, int(rank), cells, out), "fgpu_interp_points")
        return cells, out.reshape(-1, 5)

    def in_cells(self, xyz, rank: int, cells):
        pts = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        n = pts.size // 3
        c = np.ascontiguousarray(cells, np.int32)
        out = np.zeros(5 * n)
        self.ctx.check(self.ctx.lib.fgpu_interp_cells(self.h, n, pts, int(rank), c, out), "fgpu_interp_cells")
        return out.reshape(-1, 5)


    def set_triangles(self, tris, tri2tet, quads=None):
        """LinearTriangles on the surface faces (3 nodes and 2 tets per face); quads (3 per
        face, quad 3f+k from vertex k) for rank 3 and the rank 2 centroids."""
        t = np.ascontiguousarray(tris, np.int32).reshape(-1)
        t2 = np.ascontiguousarray(tri2tet, np.int32).reshape(-1)
        if 2 * t.size != 3 * t2.size:
            raise ValueError("3 nodes and 2 tets per face")
        q = None
        if quads is not None:
            q = np.ascontiguousarray(quads, np.int32).reshape(-1)
            if q.size != 4 * t.size:
                raise ValueError("3 quadrangles of 4 nodes per face")
        self._quads = q
        self.ctx.check(self.ctx.lib.fgpu_interp_set_triangles(
            self.h, t.size // 3, t, t2, None if q is None else q.ctypes.data), "fgpu_interp_set_triangles")

    def surface_points(self, xyz, rank: int = 1, centroids: bool = False):
        pts = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        n = pts.size // 3
        cells = np.zeros(n, np.int32)
        out = np.zeros(5 * n)
        self.ctx.check(self.ctx.lib.fgpu_interp_surface_points(self.h, n, pts, int(rank), int(bool(centroids)), cells,
                                                               out), "fgpu_interp_surface_points")
        return cells, out.reshape(-1, 5)

    def surface_cells(self, xyz, cells, rank: int = 1, centroids: bool = False):
        pts = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        n = pts.size // 3
        c = np.ascontiguousarray(cells, np.int32)
        out = np.zeros(5 * n)
        self.ctx.check(self.ctx.lib.fgpu_interp_surface_cells(self.h, n, pts, int(rank), int(bool(centroids)), c, out),
                       "fgpu_interp_surface_cells")
        return out.reshape(-1, 5)


class SolutionReader:
    """SolutionReader for dim 3 without atom sorting (src/SolutionReader.cpp:43-190): the first
    call maps the atoms to cells (markers), later calls interpolate in the mapped cells."""

    def __init__(self, interpolator: Interpolator, rank: int = 1, dim: int = 3, interp_centroids: bool = False):
        self.interpolator = interpolator
        self.markers = None
        self.set_preferences(False, dim, rank, interp_centroids)

    def set_preferences(self, sort_atoms: bool, dim: int, rank: int, interp_centroids: bool = False):
        """SolutionReader::set_preferences (include/SolutionReader.h:60-67).  Atom sorting
        (sort_spatial) is not offered: the caller orders the atoms as the guess chain should run."""
        if sort_atoms:
            raise ValueError("atom sorting is not supported")
        if rank not in (1, 2, 3) or dim not in (2, 3):
            raise ValueError("dim 2 or 3, rank 1, 2 or 3")
        if interp_centroids and dim != 2:
            raise ValueError("centroid interpolation is offered on the surface (dim 2)")
        self.dim, self.rank, self.interp_centroids = dim, rank, bool(interp_centroids)

    def calc_full_interpolation(self, xyz) -> np.ndarray:
        if self.dim == 2:
            self.markers, sol = self.interpolator.surface_points(xyz, self.rank, self.interp_centroids)
        else:
            self.markers, sol = self.interpolator.points(xyz, self.rank)
        return sol

    def calc_interpolation(self, xyz) -> np.ndarray:
        n = np.asarray(xyz).reshape(-1, 3).shape[0]
        if self.markers is None or len(self.markers) != n:
            return self.calc_full_interpolation(xyz)
        if self.dim == 2:
            return self.interpolator.surface_cells(xyz, self.markers, self.rank, self.interp_centroids)
        return self.interpolator.in_cells(xyz, self.rank, self.markers)

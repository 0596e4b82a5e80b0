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
TYPES_SURFACE = 2
TYPES_VACUUM = 3


class Interpolator:
    """Interpolator::initialize (src/Interpolator.cpp:28-77) on a tet mesh and its tet-split
    hexes (hex 4t+k around vertex k of tet t)."""

    def __init__(self, ctx: fgpu.Context, tip, search_region: int = TYPES_VACUUM, empty_value: float = 0.0,
                 host_build: bool = False, host_chain: bool = False):
        """The tables are built on the device; host_build=True builds the same bits on the host.
        host_chain=True resolves the located transfers' chained guess on the host, point after
        point (the device resolves it as a prefix scan by default; same cells)."""
        self.ctx = ctx
        self.n_nodes = len(tip.nodes)
        self.search_region = search_region
        h = fgpu._P()
        ctx.check(ctx.lib.fgpu_interp_create_ex(
            ctx.h, self.n_nodes, np.ascontiguousarray(tip.nodes, np.float64).reshape(-1),
            np.ascontiguousarray(tip.node_kind, np.int32), len(tip.tets),
            np.ascontiguousarray(tip.tets, np.int32).reshape(-1),
            np.ascontiguousarray(tip.tet_nbr, np.int32).reshape(-1),
            np.ascontiguousarray(tip.tet_marker, np.int32), len(tip.hexes),
            np.ascontiguousarray(tip.hexes, np.int32).reshape(-1),
            np.ascontiguousarray(tip.hex_marker, np.int32), int(search_region), float(empty_value),
            (fgpu.FGPU_INTERP_HOST if host_build else 0) | (fgpu.FGPU_INTERP_HOST_CHAIN if host_chain else 0),
            C.byref(h)),
            "fgpu_interp_create_ex")
        self.h = h

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.fgpu_interp_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def digest(self):
        """FNV-1a hash of every table (host copy of the tet tables, device tables, decay)."""
        out = (C.c_ulonglong * fgpu.FGPU_INTERP_DIGEST_TABLES)()
        self.ctx.check(self.ctx.lib.fgpu_interp_digest(self.h, fgpu.FGPU_INTERP_DIGEST_TABLES, out),
                       "fgpu_interp_digest")
        return list(out)

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

    def hex_gradient(self, xyz, hexes) -> np.ndarray:
        """LinearHexahedra::interp_gradient(point, hex) (src/InterpolatorCells.cpp:407-421):
        minus the gradient of the nodal scalar2 at each point in its hex, (n, 3)."""
        x = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        h = np.ascontiguousarray(hexes, np.int32).reshape(-1)
        if x.size != 3 * h.size:
            raise ValueError("one hex per point")
        out = np.zeros(3 * h.size)
        self.ctx.check(self.ctx.lib.fgpu_interp_hex_gradient(self.h, h.size, x, h, out), "fgpu_interp_hex_gradient")
        return out.reshape(-1, 3)

    def update_cells(self, xyz, cells) -> np.ndarray:
        """Pic::update_point_cell (src/Pic.cpp:185-195) for every particle: the solver cell
        holding each point, located from the particle's previous cell; -1 outside the region."""
        x = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        c = np.array(cells, np.int32).reshape(-1)
        if x.size != 3 * c.size:
            raise ValueError("one cell per particle")
        self.ctx.check(self.ctx.lib.fgpu_interp_update_cells(self.h, c.size, x, c), "fgpu_interp_update_cells")
        return c

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


# labels of the interpolated data (Labels, include/Globals.h:134-165)
ELFIELD, POTENTIAL, CHARGE_DENSITY = "elfield", "potential", "charge_density"
RHO, TEMPERATURE = "rho", "temperature"


class SolutionReader:
    """SolutionReader without atom sorting (src/SolutionReader.cpp:43-190, 269-436): the first
    call maps the atoms to cells (markers), later calls interpolate in the mapped cells.  The
    last result is kept as ``interpolation`` (n, 5): vector x y z, scalar1, scalar2, with the
    atom ``ids`` that export_results writes them to."""

    def __init__(self, interpolator: Interpolator, rank: int = 1, dim: int = 3, interp_centroids: bool = False,
                 vec_label: str = "vec", scalar1_label: str = "scalar1", scalar2_label: str = "scalar2"):
        self.interpolator = interpolator
        self.markers = None
        self.vec_label, self.vec_norm_label = vec_label, vec_label + "_norm"
        self.scalar1_label, self.scalar2_label = scalar1_label, scalar2_label
        self.interpolation = np.zeros((0, 5))
        self.ids = np.zeros(0, np.int64)
        self.points = np.zeros((0, 3))
        self.set_preferences(False, dim, rank, interp_centroids)

    def size(self) -> int:
        return len(self.interpolation)

    def interpolate(self, xyz, ids=None) -> np.ndarray:
        """SolutionReader::interpolate(n_points, x, y, z) (src/SolutionReader.cpp:428-436):
        atoms with ids 0..n-1 unless given."""
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        self.points = xyz
        self.ids = np.arange(len(xyz)) if ids is None else np.asarray(ids, np.int64).reshape(-1)
        if len(self.ids) != len(xyz):
            raise ValueError("one id per point")
        self.interpolation = self.calc_interpolation(xyz)
        return self.interpolation

    def contains(self, data_label: str) -> int:
        """SolutionReader::contains (src/SolutionReader.cpp:269-285): 1..4 for the vector, its
        norm, scalar1, scalar2 under their own (lower-case) label, negated when the label was
        given with upper-case letters, 0 if absent."""
        labels = (self.vec_label, self.vec_norm_label, self.scalar1_label, self.scalar2_label)
        for k, lab in enumerate(labels):
            if data_label == lab:
                return k + 1
        low = data_label.lower()
        for k, lab in enumerate(labels):
            if low == lab:
                return -(k + 1)
        return 0

    def export_results(self, n_points: int, data_type: str, data: np.ndarray) -> int:
        """SolutionReader::export_results (src/SolutionReader.cpp:303-421) into ``data`` in place:
        the label's own spelling adds to what ``data`` holds, an upper-case spelling clears
        ``data`` first (3*n_points values for the vector, x1 y1 z1 x2 ...); atoms whose id is
        outside 0..n_points-1 are skipped.  Returns 0; 1 when nothing was interpolated."""
        if self.size() == 0:
            return 1
        dtype = self.contains(data_type)
        if dtype == 0:
            raise ValueError(f"{type(self).__name__} does not contain {data_type}")
        append = dtype > 0
        width = 3 if abs(dtype) == 1 else 1
        out = np.asarray(data).reshape(-1)
        if out.size < width * n_points:
            raise ValueError("data holds fewer values than the points need")
        if not append:
            out[: width * n_points] = 0.0
        ok = (self.ids >= 0) & (self.ids < n_points)
        ids, sol = self.ids[ok], self.interpolation[ok]
        if abs(dtype) == 1:
            vals, idx = sol[:, :3].reshape(-1), (3 * ids[:, None] + np.arange(3)).reshape(-1)
        elif abs(dtype) == 2:
            vals, idx = np.linalg.norm(sol[:, :3], axis=1), ids
        else:
            vals, idx = sol[:, abs(dtype)], ids
        if append:
            np.add.at(out, idx, vals)
        else:
            out[idx] = vals                  # a repeated id keeps its last atom's value
        return 0

    def interpolate_results(self, n_points: int, data_type: str, x, y, z, data: np.ndarray) -> int:
        """SolutionReader::interpolate_results (src/SolutionReader.cpp:396-414): a fresh reader
        with these preferences interpolates at the points and exports one quantity."""
        if self.size() == 0:
            return 1
        sr = SolutionReader(self.interpolator, self.rank, self.dim, self.interp_centroids, self.vec_label,
                            self.scalar1_label, self.scalar2_label)
        sr.interpolate(np.stack([np.asarray(x, np.float64)[:n_points], np.asarray(y, np.float64)[:n_points],
                                 np.asarray(z, np.float64)[:n_points]], axis=1))
        return sr.export_results(n_points, data_type, data)

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


class FieldReader(SolutionReader):
    """FieldReader (src/SolutionReader.cpp:465-597): (E, charge density, potential) at the
    atoms, the field norms and their maximum, and the analytic check against a hemisphere
    (or hemi-ellipsoid) on a plane in the uniform field E0."""

    def __init__(self, interpolator: Interpolator, rank: int = 1, dim: int = 3, interp_centroids: bool = False):
        super().__init__(interpolator, rank, dim, interp_centroids, ELFIELD, CHARGE_DENSITY, POTENTIAL)
        self.E0 = self.radius1 = self.radius2 = self.beta = self.E_max = 0.0
        self.limit_min = self.limit_max = 0.0
        self.field_norm = np.zeros(0)

    def interpolate(self, xyz, ids=None) -> np.ndarray:
        """FieldReader::interpolate(Surface) (:469-484) and calc_interpolation (:486-496)."""
        sol = super().interpolate(xyz, ids)
        self.field_norm = np.linalg.norm(sol[:, :3], axis=1)
        self.E_max = float(self.field_norm.max()) if len(sol) else -1e100
        return sol

    def get_elfield(self, i: int) -> np.ndarray:
        return self.interpolation[i, :3]

    def get_elfield_norm(self, i: int) -> float:
        return float(self.field_norm[i])

    def get_charge_dens(self, i: int) -> float:
        return float(self.interpolation[i, 3])

    def get_potential(self, i: int) -> float:
        return float(self.interpolation[i, 4])

    def get_analyt_potential(self, i: int, origin=(0.0, 0.0, 0.0)) -> float:
        """phi = -E0 z (1 - (R/r)^3) around the hemisphere centre ``origin`` (:498-505)."""
        if not 0 <= i < self.size():
            raise IndexError(f"Invalid index: {i}")
        p = self.points[i] - np.asarray(origin, np.float64)
        r = float(np.linalg.norm(p))
        return -self.E0 * p[2] * (1.0 - (self.radius1 / r) ** 3)

    def get_analyt_field(self, i: int, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
        """E = -grad phi of the same solution (:507-521)."""
        if not 0 <= i < self.size():
            raise IndexError(f"Invalid index: {i}")
        p = self.points[i] - np.asarray(origin, np.float64)
        r5 = float(p @ p) ** 2.5
        r3 = self.radius1 ** 3
        f = p[0] * p[0] + p[1] * p[1] - 2.0 * p[2] * p[2]
        return np.array([3 * self.E0 * r3 * p[0] * p[2] / r5, 3 * self.E0 * r3 * p[1] * p[2] / r5,
                         self.E0 * (1.0 - r3 * f / r5)])

    def get_analyt_enhancement(self) -> float:
        """Apex enhancement of a hemisphere (3) or of a hemi-ellipsoid with semi-axes
        radius1 < radius2 (:523-534)."""
        if self.radius1 <= 0:
            raise ValueError(f"Invalid nanotip minor semi-axis: {self.radius1}")
        if self.radius2 <= self.radius1:
            return 3.0
        nu = self.radius2 / self.radius1
        zeta = np.sqrt(nu * nu - 1)
        return float(zeta ** 3 / (nu * np.log(zeta + nu) - zeta))

    def max_field(self, solutions=None) -> float:
        """:555-565: the largest |E| of the given (n, 5) rows, else of the last interpolation."""
        if solutions is not None:
            sol = np.asarray(solutions, np.float64).reshape(-1, 5)
            return float(np.sqrt((sol[:, :3] ** 2).sum(1).max())) if len(sol) else float("nan")
        return self.E_max

    def check_limits(self, solutions=None, verbose: bool = False) -> bool:
        """:536-553: beta = |E_max / E0| / analytic enhancement; True when outside the limits
        (never when the limits are equal)."""
        if self.limit_min == self.limit_max:
            return False
        gamma1 = abs(self.max_field(solutions) / self.E0)
        gamma2 = self.get_analyt_enhancement()
        self.beta = abs(gamma1 / gamma2)
        if verbose:
            print(f"field enhancements:  (M)easured:{gamma1:.3f}  (A)nalyt:{gamma2:.3f}  M/A:{gamma1 / gamma2:.3f}")
        return self.beta < self.limit_min or self.beta > self.limit_max

    def get_beta(self) -> float:
        return self.beta

    def set_check_params(self, conf, radius: float, tip_height: float, box_height: float,
                         field_min: float = 0.1, field_max: float = 5.0):
        """:568-589: E0 from the Neumann field or V0 / box height; the ellipsoid semi-axes from
        the tip height and radius.  conf: anode_BC, E0, V0 (FieldConfig); the tolerances
        default to the reference's (src/Config.cpp:60-61)."""
        if conf.anode_BC not in ("neumann", "dirichlet"):
            raise ValueError("Invalid anode boundary condition: " + conf.anode_BC)
        self.E0 = conf.E0 if conf.anode_BC == "neumann" else conf.V0 / box_height
        self.limit_min, self.limit_max = field_min, field_max
        if tip_height > radius:
            self.radius2, self.radius1 = tip_height, float(np.sqrt(tip_height * radius))
        else:
            self.radius2 = self.radius1 = radius


class HeatReader(SolutionReader):
    """HeatReader (src/SolutionReader.cpp:591-683): (current density, potential, temperature)
    at the atoms; interpolate_dofs hands the temperatures at the heat solver's vertices back
    to the solver."""

    def __init__(self, interpolator: Interpolator, rank: int = 1, dim: int = 3):
        super().__init__(interpolator, rank, dim, False, RHO, POTENTIAL, TEMPERATURE)

    def interpolate_atoms(self, xyz, markers, ids=None) -> np.ndarray:
        """HeatReader::interpolate(AtomReader) (:655-668): the atoms marked SURFACE or BULK."""
        xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
        markers = np.asarray(markers).reshape(-1)
        ids = np.arange(len(xyz)) if ids is None else np.asarray(ids, np.int64).reshape(-1)
        keep = (markers == TYPES_SURFACE) | (markers == TYPES_BULK)
        return self.interpolate(xyz[keep], ids[keep])

    def interpolate_dofs(self, heat_system: fgpu.System) -> np.ndarray:
        """HeatReader::interpolate_dofs (:670-683) without sorting: temperatures interpolated at
        the heat system's vertices become its solution."""
        xyz, _ = heat_system.mesh.dofs()
        sol = self.interpolate(xyz)
        temperatures = np.ascontiguousarray(sol[:, 4])
        heat_system.set_solution(temperatures)
        return temperatures

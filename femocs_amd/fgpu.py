"""Python front end of libfemocs_gpu.so (the C ABI in include/femocs_gpu.h).

Thin ctypes wrappers: every call goes to the HIP library; there is no CPU fallback.  If the
library is missing or no GPU is visible, the constructors raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfemocs_gpu.so")

FGPU_FP64, FGPU_MIXED = 0, 1
FGPU_PC_NONE, FGPU_PC_JACOBI, FGPU_PC_CHEBYSHEV = 0, 1, 2
FGPU_IMPORT_HOST = 1
FGPU_MESH_DIGEST_TABLES = 32
FGPU_INTERP_HOST = 1
FGPU_INTERP_HOST_CHAIN = 2
FGPU_INTERP_DIGEST_TABLES = 23
BID_COPPER_SURFACE, BID_VACUUM_TOP, BID_COPPER_BOTTOM, BID_SIDES = 2, 8, 7, 4

_lib = None


FGPU_OK, FGPU_EINVAL = 0, -1
FGPU_EMESH = -5


class FgpuError(RuntimeError):
    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


class MeshInvertedError(FgpuError):
    """A vertex update would turn a cell inside out; the mesh is left as it was."""


class SolveOpts(C.Structure):
    _fields_ = [("max_it", C.c_int), ("tol", C.c_double), ("relative", C.c_int), ("precond", C.c_int),
                ("precision", C.c_int), ("inner_tol", C.c_double), ("check_every", C.c_int),
                ("use_graph", C.c_int), ("cheb_degree", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer", C.c_int), ("residual", C.c_double), ("rhs_norm", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_P = C.c_void_p

# (name, restype, argtypes): the exported C ABI
API = [
    ("fgpu_ctx_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("fgpu_ctx_destroy", None, [_P]),
    ("fgpu_ctx_error", C.c_char_p, [_P]),
    ("fgpu_ctx_sync", C.c_int, [_P]),
    ("fgpu_ctx_alloc", C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    ("fgpu_ctx_free", C.c_int, [_P, _P]),
    ("fgpu_mesh_import", C.c_int, [_P, C.c_int, _D, C.c_int, _I, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_P)]),
    ("fgpu_mesh_import_ex", C.c_int, [_P, C.c_int, _D, C.c_int, _I, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(_P)]),
    ("fgpu_mesh_digest", C.c_int, [_P, C.c_int, C.POINTER(C.c_ulonglong)]),
    ("fgpu_mesh_destroy", None, [_P]),
    ("fgpu_mesh_sizes", C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("fgpu_mesh_export_dofs", C.c_int, [_P, _P, _P]),
    ("fgpu_mesh_export_bfaces", C.c_int, [_P, _P, _P, _P]),
    ("fgpu_mesh_export_face_centroids", C.c_int, [_P, C.c_int, _P, C.POINTER(C.c_int)]),
    ("fgpu_mesh_update_vertices", C.c_int, [_P, _D]),
    ("fgpu_sys_create", C.c_int, [_P, C.POINTER(_P)]),
    ("fgpu_sys_destroy", None, [_P]),
    ("fgpu_sys_set_coefficients", C.c_int, [_P, _P, _P]),
    ("fgpu_sys_reset", C.c_int, [_P]),
    ("fgpu_sys_add_neumann", C.c_int, [_P, C.c_int, _P, C.c_double]),
    ("fgpu_sys_add_cell_rhs", C.c_int, [_P, _D]),
    ("fgpu_sys_add_dirichlet", C.c_int, [_P, C.c_int, C.c_double]),
    ("fgpu_sys_apply_dirichlet", C.c_int, [_P]),
    ("fgpu_sys_fill_solution", C.c_int, [_P, C.c_double]),
    ("fgpu_sys_set_solution", C.c_int, [_P, _D]),
    ("fgpu_sys_get_solution", C.c_int, [_P, _D]),
    ("fgpu_sys_solution_to_device", C.c_int, [_P, C.c_void_p]),
    ("fgpu_sys_solution_from_device", C.c_int, [_P, C.c_void_p]),
    ("fgpu_sys_get_rhs", C.c_int, [_P, _D]),
    ("fgpu_sys_apply", C.c_int, [_P, _D, _D]),
    ("fgpu_sys_solve", C.c_int, [_P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    ("fgpu_solve_batch", C.c_int, [_P, C.c_int, C.POINTER(_P), C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    ("fgpu_batch_time_kernel", C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    ("fgpu_interp_create", C.c_int, [_P, C.c_int, _D, _I, C.c_int, _I, _I, _I, C.c_int, _I, _I, C.c_int,
                                     C.c_double, C.POINTER(_P)]),
    ("fgpu_interp_create_ex", C.c_int, [_P, C.c_int, _D, _I, C.c_int, _I, _I, _I, C.c_int, _I, _I, C.c_int,
                                        C.c_double, C.c_int, C.POINTER(_P)]),
    ("fgpu_interp_digest", C.c_int, [_P, C.c_int, C.POINTER(C.c_ulonglong)]),
    ("fgpu_interp_destroy", None, [_P]),
    ("fgpu_interp_extract", C.c_int, [_P, _P, C.c_int]),
    ("fgpu_interp_extract_heat", C.c_int, [_P, _P, _P]),
    ("fgpu_interp_get_nodal", C.c_int, [_P, _D]),
    ("fgpu_interp_hex_gradient", C.c_int, [_P, C.c_int, _D, _I, _D]),
    ("fgpu_interp_update_cells", C.c_int, [_P, C.c_int, _D, _I]),
    ("fgpu_tethex_split", C.c_int, [_P, C.c_int, _D, C.c_int, _I, C.c_int, C.c_void_p, C.POINTER(_P)]),
    ("fgpu_tethex_sizes", C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("fgpu_tethex_export", C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("fgpu_tethex_destroy", None, [_P]),
    ("fgpu_interp_set_nodal", C.c_int, [_P, _D]),
    ("fgpu_interp_points", C.c_int, [_P, C.c_int, _D, C.c_int, _I, _D]),
    ("fgpu_interp_cells", C.c_int, [_P, C.c_int, _D, C.c_int, _I, _D]),
    ("fgpu_interp_set_triangles", C.c_int, [_P, C.c_int, _I, _I, C.c_void_p]),
    ("fgpu_interp_surface_points", C.c_int, [_P, C.c_int, _D, C.c_int, C.c_int, _I, _D]),
    ("fgpu_interp_surface_cells", C.c_int, [_P, C.c_int, _D, C.c_int, C.c_int, _I, _D]),
    ("fgpu_sys_heat_terms", C.c_int, [_P, _P, C.c_double, C.c_double]),
    ("fgpu_sys_heat_terms_theta", C.c_int, [_P, _P, C.c_double, C.c_double, C.c_double]),
    ("fgpu_sys_sigma_terms", C.c_int, [_P, _P, C.c_double, C.c_double]),
    ("fgpu_sys_apply32", C.c_int, [_P, _D, _D]),
    ("fgpu_sys_set_coarse", C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    ("fgpu_sys_get_coarse", C.c_int, [_P, C.POINTER(C.c_int), C.c_void_p, C.c_void_p]),
    ("fgpu_mesh_use_affine", C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    ("fgpu_sys_minmax", C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("fgpu_sys_solution_grad", C.c_int, [_P, _D]),
    ("fgpu_sys_time_kernel", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    ("fgpu_mesh_set_patch_cells", C.c_int, [_P, C.c_int]),
    ("fgpu_mesh_patch_info", C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]),
    ("fgpu_sys_set_patches_per_block", C.c_int, [_P, C.c_int]),
    ("fgpu_mesh_move_vertices", C.c_int, [_P, C.c_int, _I, _D]),
    ("fgpu_mesh_set_motion", C.c_int, [_P, C.c_int, _I, _I, _I, _D, C.c_int]),
    ("fgpu_mesh_displace", C.c_int, [_P, _D]),
    ("fgpu_mesh_displace_noalloc", C.c_int, [_P, _D]),
    ("fgpu_sys_reassemble", C.c_int, [_P]),
    ("fgpu_sys_save_charge_density", C.c_int, [_P]),
    ("fgpu_sys_add_space_charge", C.c_int, [_P, _P, C.c_int, _D, _I, C.c_double]),
    ("fgpu_sys_export_charge_density", C.c_int, [_P, _D]),
    ("fgpu_mesh_device_bytes", C.c_int, [_P, C.POINTER(C.c_size_t)]),
    ("fgpu_sys_device_bytes", C.c_int, [_P, C.POINTER(C.c_size_t)]),
    ("fgpu_batch_create", C.c_int, [_P, C.POINTER(SolveOpts), C.POINTER(_P)]),
    ("fgpu_batch_destroy", None, [_P]),
    ("fgpu_batch_size", C.c_int, [_P]),
    ("fgpu_batch_add", C.c_int, [_P, _P]),
    ("fgpu_batch_step", C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    ("fgpu_batch_retire", C.c_int, [_P, C.c_int, C.POINTER(_P), C.POINTER(SolveStats), C.POINTER(C.c_int)]),
    ("fgpu_batch_step_async", C.c_int, [_P, C.c_int]),
    ("fgpu_batch_wait", C.c_int, [_P, C.POINTER(C.c_int)]),
    ("fgpu_q2_create", C.c_int, [_P, C.POINTER(_P)]),
    ("fgpu_q2_destroy", None, [_P]),
    ("fgpu_q2_sizes", C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("fgpu_q2_export_dofs", C.c_int, [_P, C.c_void_p, C.c_void_p]),
    ("fgpu_q2_reset", C.c_int, [_P]),
    ("fgpu_q2_add_dirichlet", C.c_int, [_P, C.c_int, C.c_double]),
    ("fgpu_q2_add_neumann", C.c_int, [_P, C.c_int, C.c_double]),
    ("fgpu_q2_apply_dirichlet", C.c_int, [_P]),
    ("fgpu_q2_fill_solution", C.c_int, [_P, C.c_double]),
    ("fgpu_q2_solve", C.c_int, [_P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    ("fgpu_q2_get_solution", C.c_int, [_P, _D]),
    ("fgpu_q2_get_rhs", C.c_int, [_P, _D]),
    ("fgpu_q2_apply", C.c_int, [_P, _D, _D]),
    ("fgpu_q2_time_apply", C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    ("fgpu_q2_uses_tet_split", C.c_int, [_P, C.POINTER(C.c_int)]),
]


def load(path: str = LIB_PATH):
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise FgpuError(f"{path} is missing: build it with `python -m femocs_amd.build`")
        lib = C.CDLL(path)
        for name, res, args in API:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


class Context:
    def __init__(self, device: int = 0):
        self.lib = load()
        h = _P()
        rc = self.lib.fgpu_ctx_create(device, C.byref(h))
        if rc != 0:
            raise FgpuError(f"fgpu_ctx_create failed ({rc}): no usable GPU")
        self.h = h

    def check(self, rc, what):
        if rc != 0:
            msg = self.lib.fgpu_ctx_error(self.h)
            cls = MeshInvertedError if rc == FGPU_EMESH else FgpuError
            raise cls(f"{what} failed ({rc}): {msg.decode() if msg else ''}", rc)

    def sync(self):
        self.check(self.lib.fgpu_ctx_sync(self.h), "sync")

    def alloc(self, nbytes: int) -> "DeviceBuffer":
        """Device memory on the context's device, freed when the buffer is dropped."""
        return DeviceBuffer(self, nbytes)

    def close(self):
        if getattr(self, "h", None):
            self.lib.fgpu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve_batch(self, systems, opts: SolveOpts):
        n = len(systems)
        arr = (_P * n)(*[s.h for s in systems])
        stats = (SolveStats * n)()
        self.check(self.lib.fgpu_solve_batch(self.h, n, arr, C.byref(opts), stats), "fgpu_solve_batch")
        return [stats[i] for i in range(n)]


class DeviceBuffer:
    """A block of device memory from fgpu_ctx_alloc (ptr: its address)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        p = _P()
        ctx.check(ctx.lib.fgpu_ctx_alloc(ctx.h, C.c_size_t(int(nbytes)), C.byref(p)), "fgpu_ctx_alloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def __del__(self):
        try:
            if self.ptr:
                self.ctx.lib.fgpu_ctx_free(self.ctx.h, _P(self.ptr))
                self.ptr = None
        except Exception:
            pass


class Mesh:
    """Device mesh: DealSolver::import_mesh + mark_boundary."""

    def __init__(self, ctx: Context, vertices, ucd_cells, top, bottom, sides, other, host_import=False):
        """The tables are built on the device; host_import=True builds the same bits on the host."""
        self.ctx = ctx
        v = np.ascontiguousarray(vertices, np.float64).reshape(-1)
        c = np.ascontiguousarray(ucd_cells, np.int32).reshape(-1)
        h = _P()
        flags = FGPU_IMPORT_HOST if host_import else 0
        ctx.check(ctx.lib.fgpu_mesh_import_ex(ctx.h, len(v) // 3, v, len(c) // 8, c, top, bottom, sides, other,
                                              flags, C.byref(h)), "fgpu_mesh_import_ex")
        self.h = h
        nd, nc, nb = C.c_int(), C.c_int(), C.c_int()
        ctx.lib.fgpu_mesh_sizes(h, C.byref(nd), C.byref(nc), C.byref(nb))
        self.n_dofs, self.n_cells, self.n_bfaces = nd.value, nc.value, nb.value

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.fgpu_mesh_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def dofs(self):
        x = np.zeros(3 * self.n_dofs)
        v = np.zeros(self.n_dofs, np.int32)
        self.ctx.lib.fgpu_mesh_export_dofs(self.h, x.ctypes.data, v.ctypes.data)
        return x.reshape(-1, 3), v

    def bfaces(self):
        c, f, b = (np.zeros(self.n_bfaces, np.int32) for _ in range(3))
        self.ctx.lib.fgpu_mesh_export_bfaces(self.h, c.ctypes.data, f.ctypes.data, b.ctypes.data)
        return c, f, b

    def face_centroids(self, bid):
        n = C.c_int()
        self.ctx.lib.fgpu_mesh_export_face_centroids(self.h, bid, None, C.byref(n))
        out = np.zeros(3 * n.value)
        self.ctx.lib.fgpu_mesh_export_face_centroids(self.h, bid, out.ctypes.data, C.byref(n))
        return out.reshape(-1, 3)

    def device_bytes(self) -> int:
        b = C.c_size_t()
        self.ctx.lib.fgpu_mesh_device_bytes(self.h, C.byref(b))
        return b.value

    def digest(self):
        """FNV-1a hash of every import table (host mirrors, then device arrays)."""
        out = (C.c_ulonglong * FGPU_MESH_DIGEST_TABLES)()
        self.ctx.check(self.ctx.lib.fgpu_mesh_digest(self.h, FGPU_MESH_DIGEST_TABLES, out), "fgpu_mesh_digest")
        return list(out)

    def set_motion(self, ids, W):
        """Rows ids follow the control points through the sparse weights W (len(ids) x n_ctrl)."""
        from scipy.sparse import csr_matrix
        W = csr_matrix(W)
        ids = np.ascontiguousarray(ids, np.int32)
        if W.shape[0] != ids.size:
            raise ValueError("one weight row per vertex id")
        self.n_ctrl = W.shape[1]
        self.ctx.check(self.ctx.lib.fgpu_mesh_set_motion(
            self.h, int(ids.size), ids, np.ascontiguousarray(W.indptr, np.int32),
            np.ascontiguousarray(W.indices, np.int32), np.ascontiguousarray(W.data, np.float64), int(W.shape[1])),
            "fgpu_mesh_set_motion")

    def displace(self, disp):
        d = np.ascontiguousarray(disp, np.float64).reshape(-1)
        if d.size != 3 * self.n_ctrl:
            raise ValueError("3 displacement components per control point")
        self.ctx.check(self.ctx.lib.fgpu_mesh_displace(self.h, d), "fgpu_mesh_displace")

    def displace_noalloc(self, disp):
        """displace() through the entry point that allocates nothing (the library's fixed
        scratch array holds the validity check's partial minima)."""
        d = np.ascontiguousarray(disp, np.float64).reshape(-1)
        if d.size != 3 * self.n_ctrl:
            raise ValueError("3 displacement components per control point")
        self.ctx.check(self.ctx.lib.fgpu_mesh_displace_noalloc(self.h, d), "fgpu_mesh_displace_noalloc")

    def use_affine(self, allow: bool = True) -> bool:
        """Tet-split FP32 operator on or off; returns whether it is in use."""
        a = C.c_int()
        self.ctx.check(self.ctx.lib.fgpu_mesh_use_affine(self.h, int(bool(allow)), C.byref(a)), "fgpu_mesh_use_affine")
        return bool(a.value)

    def move_vertices(self, ids, xyz):
        ids = np.ascontiguousarray(ids, np.int32)
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        if xyz.size != 3 * ids.size:
            raise ValueError("xyz must hold 3 coordinates per vertex id")
        self.ctx.check(self.ctx.lib.fgpu_mesh_move_vertices(self.h, int(ids.size), ids, xyz), "fgpu_mesh_move_vertices")

    def set_patch_cells(self, cells: int):
        self.ctx.check(self.ctx.lib.fgpu_mesh_set_patch_cells(self.h, int(cells)), "fgpu_mesh_set_patch_cells")

    def patch_info(self):
        v = [C.c_int() for _ in range(4)]
        self.ctx.lib.fgpu_mesh_patch_info(self.h, *[C.byref(x) for x in v])
        return dict(zip(("cells", "n_patches", "node_cap", "n_slots"), (x.value for x in v)))

    def update_vertices(self, vertices):
        v = np.ascontiguousarray(vertices, np.float64).reshape(-1)
        self.ctx.check(self.ctx.lib.fgpu_mesh_update_vertices(self.h, v), "fgpu_mesh_update_vertices")


class System:
    def __init__(self, mesh: Mesh):
        self.mesh = mesh
        self.ctx = mesh.ctx
        h = _P()
        self.ctx.check(self.ctx.lib.fgpu_sys_create(mesh.h, C.byref(h)), "fgpu_sys_create")
        self.h = h

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.fgpu_sys_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _c(self, rc, what):
        self.ctx.check(rc, what)

    def set_coefficients(self, kappa_q=None, gamma_q=None):
        k = None if kappa_q is None else np.ascontiguousarray(kappa_q, np.float64)
        g = None if gamma_q is None else np.ascontiguousarray(gamma_q, np.float64)
        self._c(self.ctx.lib.fgpu_sys_set_coefficients(self.h, _ptr(k), _ptr(g)), "set_coefficients")

    def reset(self):
        self._c(self.ctx.lib.fgpu_sys_reset(self.h), "reset")

    def add_neumann(self, bid, values=None, constant=0.0):
        v = None if values is None else np.ascontiguousarray(values, np.float64)
        self._c(self.ctx.lib.fgpu_sys_add_neumann(self.h, bid, _ptr(v), float(constant)), "add_neumann")

    def add_cell_rhs(self, f_q):
        self._c(self.ctx.lib.fgpu_sys_add_cell_rhs(self.h, np.ascontiguousarray(f_q, np.float64)), "add_cell_rhs")

    def add_dirichlet(self, bid, value):
        self._c(self.ctx.lib.fgpu_sys_add_dirichlet(self.h, bid, float(value)), "add_dirichlet")

    def apply_dirichlet(self):
        self._c(self.ctx.lib.fgpu_sys_apply_dirichlet(self.h), "apply_dirichlet")

    def heat_terms(self, current: "System", gamma: float, lorentz: float = 2.44e-8, theta: float = 1.0):
        """Heat-step terms from T_old (this system's solution) and the current potential; theta
        1 is implicit Euler, 1/2 Crank-Nicolson."""
        if theta == 1.0:
            self._c(self.ctx.lib.fgpu_sys_heat_terms(self.h, current.h, float(gamma), float(lorentz)), "heat_terms")
        else:
            self._c(self.ctx.lib.fgpu_sys_heat_terms_theta(self.h, current.h, float(gamma), float(lorentz),
                                                           float(theta)), "heat_terms_theta")

    def sigma_terms(self, heat: "System", sigma0: float, alpha: float):
        """sigma(T) = sigma0 / (1 + alpha (T - 300)) as this current system's conductivity."""
        self._c(self.ctx.lib.fgpu_sys_sigma_terms(self.h, heat.h, float(sigma0), float(alpha)), "sigma_terms")

    def add_space_charge(self, interpolator, xyz, cells, charge_factor: float):
        """b += the particles' charge (PoissonSolver::assemble_space_charge_fast); cells are the
        solver cells of the particles, interpolator the one over this mesh's region hexes."""
        x = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        c = np.ascontiguousarray(cells, np.int32).reshape(-1)
        if x.size != 3 * c.size:
            raise ValueError("one cell per particle")
        self._c(self.ctx.lib.fgpu_sys_add_space_charge(self.h, interpolator.h, c.size, x, c, float(charge_factor)),
                "add_space_charge")

    def save_charge_density(self):
        self._c(self.ctx.lib.fgpu_sys_save_charge_density(self.h), "save_charge_density")

    def charge_density(self):
        rho = np.zeros(self.mesh.n_dofs)
        self._c(self.ctx.lib.fgpu_sys_export_charge_density(self.h, rho), "export_charge_density")
        return rho

    def reassemble(self):
        self._c(self.ctx.lib.fgpu_sys_reassemble(self.h), "reassemble")

    def fill_solution(self, v):
        self._c(self.ctx.lib.fgpu_sys_fill_solution(self.h, float(v)), "fill_solution")

    def set_solution(self, x):
        self._c(self.ctx.lib.fgpu_sys_set_solution(self.h, np.ascontiguousarray(x, np.float64)), "set_solution")

    def solution(self):
        x = np.zeros(self.mesh.n_dofs)
        self._c(self.ctx.lib.fgpu_sys_get_solution(self.h, x), "get_solution")
        return x

    def solution_to_device(self, dev_ptr: int):
        """Copies the solution (internal dof order) to device memory at dev_ptr (n_dofs
        doubles), ordered on the context's stream."""
        self._c(self.ctx.lib.fgpu_sys_solution_to_device(self.h, C.c_void_p(int(dev_ptr))), "solution_to_device")

    def solution_from_device(self, dev_ptr: int):
        """Sets the solution from device memory written by solution_to_device on this mesh."""
        self._c(self.ctx.lib.fgpu_sys_solution_from_device(self.h, C.c_void_p(int(dev_ptr))), "solution_from_device")

    def rhs(self):
        b = np.zeros(self.mesh.n_dofs)
        self._c(self.ctx.lib.fgpu_sys_get_rhs(self.h, b), "get_rhs")
        return b

    def apply(self, x):
        y = np.zeros(self.mesh.n_dofs)
        self._c(self.ctx.lib.fgpu_sys_apply(self.h, np.ascontiguousarray(x, np.float64), y), "apply")
        return y

    def set_coarse(self, enable=True) -> int:
        """Two-level preconditioner of the mixed solver: False removes it, True picks the
        aggregate size by mesh size, 256 / 512 / 1024 sets it.  Returns the number of operator
        applications the build took."""
        mode = int(enable) if not isinstance(enable, bool) else int(enable)
        n = C.c_int(0)
        self._c(self.ctx.lib.fgpu_sys_set_coarse(self.h, mode, C.byref(n)), "set_coarse")
        return n.value

    def coarse(self):
        """(block of every dof, E^-1) of the two-level preconditioner, or None without one."""
        nc = C.c_int(0)
        self._c(self.ctx.lib.fgpu_sys_get_coarse(self.h, C.byref(nc), None, None), "get_coarse")
        if nc.value == 0:
            return None
        block = np.zeros(self.mesh.n_dofs, np.int32)
        einv = np.zeros((nc.value, nc.value), np.float32)
        self._c(self.ctx.lib.fgpu_sys_get_coarse(self.h, C.byref(nc), block.ctypes.data, einv.ctypes.data),
                "get_coarse")
        return block, einv

    def apply32(self, x):
        y = np.zeros(self.mesh.n_dofs)
        self._c(self.ctx.lib.fgpu_sys_apply32(self.h, np.ascontiguousarray(x, np.float64), y), "apply32")
        return y

    def solve(self, opts: SolveOpts) -> SolveStats:
        st = SolveStats()
        self._c(self.ctx.lib.fgpu_sys_solve(self.h, C.byref(opts), C.byref(st)), "solve")
        return st

    def minmax(self):
        a, b = C.c_double(), C.c_double()
        self._c(self.ctx.lib.fgpu_sys_minmax(self.h, C.byref(a), C.byref(b)), "minmax")
        return a.value, b.value

    def device_bytes(self) -> int:
        b = C.c_size_t()
        self.ctx.lib.fgpu_sys_device_bytes(self.h, C.byref(b))
        return b.value

    def set_patches_per_block(self, ppb: int):
        self._c(self.ctx.lib.fgpu_sys_set_patches_per_block(self.h, int(ppb)), "set_patches_per_block")

    def time_kernel(self, which=0, precision=FGPU_FP64, reps=50):
        ms = C.c_double()
        self._c(self.ctx.lib.fgpu_sys_time_kernel(self.h, which, precision, reps, C.byref(ms)), "time_kernel")
        return ms.value

    def solution_grad(self):
        g = np.zeros(3 * self.mesh.n_dofs)
        self._c(self.ctx.lib.fgpu_sys_solution_grad(self.h, g), "solution_grad")
        return g.reshape(-1, 3)


class Batch:
    """Continuous batching: systems join a running batch (their solve starts at once) and
    leave when finished; every step runs `iterations` CG iterations of all running members."""

    def __init__(self, ctx: Context, o: SolveOpts):
        self.ctx = ctx
        h = _P()
        ctx.check(ctx.lib.fgpu_batch_create(ctx.h, C.byref(o), C.byref(h)), "fgpu_batch_create")
        self.h = h
        self.members = {}

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.fgpu_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def __len__(self):
        return self.ctx.lib.fgpu_batch_size(self.h)

    def add(self, system: "System"):
        self.ctx.check(self.ctx.lib.fgpu_batch_add(self.h, system.h), "fgpu_batch_add")
        self.members[system.h.value] = system

    def step(self, iterations: int = 32) -> int:
        n = C.c_int()
        self.ctx.check(self.ctx.lib.fgpu_batch_step(self.h, int(iterations), C.byref(n)), "fgpu_batch_step")
        return n.value

    def step_async(self, iterations: int = 32):
        """Enqueues one step of every member and returns at once."""
        self.ctx.check(self.ctx.lib.fgpu_batch_step_async(self.h, int(iterations)), "fgpu_batch_step_async")

    def wait(self) -> int:
        """Waits for the step in flight; returns how many of its members still run."""
        n = C.c_int()
        self.ctx.check(self.ctx.lib.fgpu_batch_wait(self.h, C.byref(n)), "fgpu_batch_wait")
        return n.value

    def time_kernel(self, which=0, reps=50) -> float:
        ms = C.c_double()
        self.ctx.check(self.ctx.lib.fgpu_batch_time_kernel(self.h, int(which), int(reps), C.byref(ms)),
                       "fgpu_batch_time_kernel")
        return ms.value

    def retire(self):
        """Removes the finished members; returns [(System, SolveStats)]."""
        cap = max(1, len(self.members))
        out = (_P * cap)()
        st = (SolveStats * cap)()
        n = C.c_int()
        self.ctx.check(self.ctx.lib.fgpu_batch_retire(self.h, cap, out, st, C.byref(n)), "fgpu_batch_retire")
        res = []
        for i in range(n.value):
            res.append((self.members.pop(out[i]), st[i]))
        return res


class Q2System:
    """The Laplace system with Q2 elements on a mesh's hexes (fgpu_q2_*): dofs are the mesh's
    vertices (same order as Mesh.dofs), then edge midpoints, face centres and cell centres."""

    def __init__(self, mesh: Mesh):
        self.ctx = mesh.ctx
        self.mesh = mesh
        h = _P()
        self.ctx.check(self.ctx.lib.fgpu_q2_create(mesh.h, C.byref(h)), "fgpu_q2_create")
        self.h = h
        n, c = C.c_int(), C.c_int()
        self.ctx.lib.fgpu_q2_sizes(h, C.byref(n), C.byref(c))
        self.n_dofs, self.n_cells = n.value, c.value

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.fgpu_q2_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _c(self, rc, what):
        self.ctx.check(rc, what)

    def dofs(self):
        """(support points (n, 3), 27 dofs per cell in import cell order)."""
        xyz = np.zeros((self.n_dofs, 3))
        cells = np.zeros((self.n_cells, 27), np.int32)
        self._c(self.ctx.lib.fgpu_q2_export_dofs(self.h, xyz.ctypes.data, cells.ctypes.data), "fgpu_q2_export_dofs")
        return xyz, cells

    def reset(self):
        self._c(self.ctx.lib.fgpu_q2_reset(self.h), "fgpu_q2_reset")

    def add_dirichlet(self, bid, value):
        self._c(self.ctx.lib.fgpu_q2_add_dirichlet(self.h, int(bid), float(value)), "fgpu_q2_add_dirichlet")

    def add_neumann(self, bid, value):
        self._c(self.ctx.lib.fgpu_q2_add_neumann(self.h, int(bid), float(value)), "fgpu_q2_add_neumann")

    def apply_dirichlet(self):
        self._c(self.ctx.lib.fgpu_q2_apply_dirichlet(self.h), "fgpu_q2_apply_dirichlet")

    def fill_solution(self, v):
        self._c(self.ctx.lib.fgpu_q2_fill_solution(self.h, float(v)), "fgpu_q2_fill_solution")

    def solve(self, opts: SolveOpts) -> SolveStats:
        st = SolveStats()
        self._c(self.ctx.lib.fgpu_q2_solve(self.h, C.byref(opts), C.byref(st)), "fgpu_q2_solve")
        return st

    def solution(self):
        x = np.zeros(self.n_dofs)
        self._c(self.ctx.lib.fgpu_q2_get_solution(self.h, x), "fgpu_q2_get_solution")
        return x

    def rhs(self):
        b = np.zeros(self.n_dofs)
        self._c(self.ctx.lib.fgpu_q2_get_rhs(self.h, b), "fgpu_q2_get_rhs")
        return b

    def apply(self, x):
        y = np.zeros(self.n_dofs)
        self._c(self.ctx.lib.fgpu_q2_apply(self.h, np.ascontiguousarray(x, np.float64), y), "fgpu_q2_apply")
        return y

    def time_apply(self, reps=20) -> float:
        ms = C.c_double()
        self._c(self.ctx.lib.fgpu_q2_time_apply(self.h, int(reps), C.byref(ms)), "fgpu_q2_time_apply")
        return ms.value

    def uses_tet_split(self) -> bool:
        """Whether the apply runs the tet-split form (per-cell metric)."""
        on = C.c_int()
        self._c(self.ctx.lib.fgpu_q2_uses_tet_split(self.h, C.byref(on)), "fgpu_q2_uses_tet_split")
        return bool(on.value)


def tethex_split(ctx: "Context", nodes, tets, tris=None):
    """The reference's tet -> 4-hex split on the device (Tethex Mesh::convert_3D): (all nodes
    (n, 3), node kind (1 vertex, 2 edge, 3 face, 4 tet), hexes (4T, 8) in femocs/UCD order,
    quadrangles (3F, 4) of the boundary triangles or None, number of edges, number of faces)."""
    x = np.ascontiguousarray(nodes, np.float64).reshape(-1)
    t = np.ascontiguousarray(tets, np.int32).reshape(-1)
    tr = None if tris is None else np.ascontiguousarray(tris, np.int32).reshape(-1)
    h = _P()
    ctx.check(ctx.lib.fgpu_tethex_split(ctx.h, x.size // 3, x, t.size // 4, t, 0 if tr is None else tr.size // 3,
                                        None if tr is None else tr.ctypes.data, C.byref(h)), "fgpu_tethex_split")
    try:
        nn, ne, nf = C.c_int(), C.c_int(), C.c_int()
        ctx.check(ctx.lib.fgpu_tethex_sizes(h, C.byref(nn), C.byref(ne), C.byref(nf)), "fgpu_tethex_sizes")
        # every entry is written by the export
        xyz = np.empty((nn.value, 3))
        hexes = np.empty((t.size, 8), np.int32)
        quads = None if tr is None else np.empty((tr.size, 4), np.int32)
        ctx.check(ctx.lib.fgpu_tethex_export(h, xyz.ctypes.data, hexes.ctypes.data,
                                             None if quads is None else quads.ctypes.data), "fgpu_tethex_export")
    finally:
        ctx.lib.fgpu_tethex_destroy(h)
    nv = x.size // 3
    kind = np.concatenate([np.full(nv, 1), np.full(ne.value, 2), np.full(nf.value, 3),
                           np.full(t.size // 4, 4)]).astype(np.int8)
    return xyz, kind, hexes, quads, ne.value, nf.value


def opts(max_it=10000, tol=1e-9, relative=True, precond=FGPU_PC_JACOBI, precision=FGPU_FP64, inner_tol=1e-4,
         check_every=32, use_graph=True, cheb_degree=0) -> SolveOpts:
    return SolveOpts(max_it, tol, int(relative), precond, precision, inner_tol, check_every, int(use_graph),
                     int(cheb_degree))

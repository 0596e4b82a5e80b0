"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and by the
CPU-baseline leg of bench.py as the checker; the product library never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

BID_COPPER_SURFACE, BID_VACUUM_TOP, BID_COPPER_BOTTOM, BID_SIDES = 2, 8, 7, 4


def build(force: bool = False):
    subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


class _Mesh(C.Structure):
    _fields_ = [("n_dofs", C.c_int), ("n_cells", C.c_int), ("x", C.POINTER(C.c_double)),
                ("vertex_of_dof", C.POINTER(C.c_int)), ("cells", C.POINTER(C.c_int)),
                ("n_bfaces", C.c_int), ("bface_cell", C.POINTER(C.c_int)),
                ("bface_face", C.POINTER(C.c_int)), ("bface_id", C.POINTER(C.c_int))]


class _Csr(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int64), ("rowptr", C.POINTER(C.c_int64)),
                ("col", C.POINTER(C.c_int)), ("val", C.POINTER(C.c_double)),
                ("diag", C.POINTER(C.c_int64))]


class _Lintet(C.Structure):
    _fields_ = [("n_tets", C.c_int), ("det0", C.POINTER(C.c_double)), ("det", C.POINTER(C.c_double)),
                ("centroid", C.POINTER(C.c_double)), ("nbr_ptr", C.POINTER(C.c_int)),
                ("nbr", C.POINTER(C.c_int)), ("marker", C.POINTER(C.c_int))]


_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_U8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class _Lintri(C.Structure):
    _fields_ = [("n", C.c_int)] + [(k, C.POINTER(C.c_double)) for k in ("v0", "e1", "e2", "pv", "nrm", "maxd", "cen")] + \
               [(k, C.POINTER(C.c_int)) for k in ("nbr_ptr", "nbr", "tri2tet", "tris", "quads", "quad2hex", "tri6")] + \
               [("qcen", C.POINTER(C.c_double))]


def _declare(L):
    L.orc_mesh_import.argtypes = [C.POINTER(_Mesh), C.c_int, _D, C.c_int, _I, C.c_int, C.c_int, C.c_int, C.c_int]
    L.orc_mesh_free.argtypes = [C.POINTER(_Mesh)]
    L.orc_csr_pattern.argtypes = [C.POINTER(_Csr), C.POINTER(_Mesh)]
    L.orc_csr_free.argtypes = [C.POINTER(_Csr)]
    L.orc_csr_spmv.argtypes = [C.POINTER(_Csr), _D, _D]
    L.orc_assemble.argtypes = [C.POINTER(_Csr), C.POINTER(_Mesh), C.c_void_p, C.c_void_p]
    L.orc_assemble_cell_rhs.argtypes = [_D, C.POINTER(_Mesh), _D]
    L.orc_dof_volumes.argtypes = [C.POINTER(_Mesh), _D]
    L.orc_charge_density.argtypes = [C.POINTER(_Mesh), _D, _D]
    L.orc_assemble_face_rhs.argtypes = [_D, C.POINTER(_Mesh), C.c_int, C.c_void_p, C.c_double]
    L.orc_dirichlet_mask.argtypes = [C.POINTER(_Mesh), C.c_int, _U8]
    L.orc_apply_dirichlet.argtypes = [C.POINTER(_Csr), _D, _D, _U8, C.c_void_p]
    L.orc_pcg.argtypes = [C.POINTER(_Csr), _D, _D, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                          C.POINTER(C.c_double)]
    L.orc_pcg.restype = C.c_int
    L.orc_pcg_omp.argtypes = [C.POINTER(_Csr), _D, _D, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                              C.POINTER(C.c_double)]
    L.orc_pcg_omp.restype = C.c_int
    L.orc_omp_max_threads.restype = C.c_int
    L.orc_export_solution_grad.argtypes = [C.POINTER(_Mesh), _D, _D]
    L.orc_element_matrix.argtypes = [_D, C.c_void_p, C.c_void_p, _D]
    L.orc_nodal_field.argtypes = [C.c_int, _D, C.c_int, _I, _I, C.c_int, _D, _D]
    L.orc_lintet_build.argtypes = [C.POINTER(_Lintet), C.c_int, _D, C.c_int, _I, _I, _I, C.c_int]
    L.orc_lintet_free.argtypes = [C.POINTER(_Lintet)]
    L.orc_lintet_locate.argtypes = [C.POINTER(_Lintet), _D, C.c_int]
    L.orc_lintet_locate.restype = C.c_int
    L.orc_lintet_bcc.argtypes = [C.POINTER(_Lintet), _D, C.c_int, _D]
    L.orc_locate_chain.argtypes = [C.POINTER(_Lintet), C.c_int, _D, C.c_int, _I]
    L.orc_linhex_shape_functions.argtypes = [_D, _D, _D]
    L.orc_linhex_gradient.argtypes = [_D, _I, _D, C.c_int, _D, _I, _D]
    L.orc_update_cells.argtypes = [C.POINTER(_Lintet), _I, _I, C.c_int, _D, _I]
    L.orc_space_charge.argtypes = [C.POINTER(_Mesh), _D, _I, _I, C.c_int, _D, _I, C.c_double, _D]
    L.orc_qp_fields.argtypes = [C.POINTER(_Mesh), _D, C.c_void_p, C.c_void_p]
    L.orc_quadtet_cells.argtypes = [C.c_int, _I, _I, _I, _I]
    L.orc_extract_nodal.argtypes = [C.c_int, _D, C.c_int, _I, _I, C.c_int, _U8, _D, C.c_void_p, C.c_double, _D]
    L.orc_decay_factor.argtypes = [_D, C.c_int, _I]
    L.orc_decay_factor.restype = C.c_double
    L.orc_smooth_nodal.argtypes = [C.c_int, C.c_int, _D, _I, C.c_int, _I, _I, C.c_int, C.c_double, _D]
    L.orc_interp_points.argtypes = [C.POINTER(_Lintet), _D, _I, _I, _I, _D, C.c_int, _D, C.c_int, _I, _D]
    L.orc_interp_cells.argtypes = [C.POINTER(_Lintet), _D, _I, _I, _I, _D, C.c_int, _D, C.c_int, _I, _D]
    L.orc_lintri_build.argtypes = [C.POINTER(_Lintri), C.c_int, _D, C.c_int, _I, _I]
    L.orc_lintri_free.argtypes = [C.POINTER(_Lintri)]
    L.orc_lintri_point_in_cell.argtypes = [C.POINTER(_Lintri), _D, C.c_int]
    L.orc_lintri_locate.argtypes = [C.POINTER(_Lintri), _D, C.c_int]
    L.orc_lintri_shape_functions.argtypes = [C.POINTER(_Lintri), _D, C.c_int, _D]
    L.orc_lintri_fast_distance.argtypes = [C.POINTER(_Lintri), _D, C.c_int]
    L.orc_lintri_fast_distance.restype = C.c_double
    L.orc_lintri_set_quads.argtypes = [C.POINTER(_Lintri), C.c_int, _D, _I, _I, _I, _I]
    L.orc_interp_surface.argtypes = [C.POINTER(_Lintri), C.POINTER(_Lintet), _D, _I, _I, _I, _D, C.c_int, _D,
                                     C.c_int, C.c_int, _I, _D]
    L.orc_interp_surface_cells.argtypes = [C.POINTER(_Lintri), C.POINTER(_Lintet), _D, _I, _I, _I, _D, C.c_int, _D,
                                           C.c_int, C.c_int, _I, _D]
    for f in ("orc_resistivity", "orc_sigma"):
        getattr(L, f).argtypes = [C.c_double]
        getattr(L, f).restype = C.c_double
    L.orc_kappa.argtypes = [C.c_double, C.c_double]
    L.orc_kappa.restype = C.c_double
    L.orc_sigma_kappa_array.argtypes = [C.c_int, _D, C.c_double, _D, _D]


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Mesh:
    """DealSolver::import_mesh restated: compaction, UCD->lexicographic, orientation fix,
    boundary faces marked by the bounding box of their centres."""

    def __init__(self, vertices, ucd_cells, top, bottom, sides, other):
        self._m = _Mesh()
        v = np.ascontiguousarray(vertices, np.float64).reshape(-1)
        c = np.ascontiguousarray(ucd_cells, np.int32).reshape(-1)
        rc = lib().orc_mesh_import(C.byref(self._m), len(v) // 3, v, len(c) // 8, c, top, bottom, sides, other)
        if rc != 0:
            raise ValueError("mesh import failed")
        m = self._m
        self.n_dofs, self.n_cells = m.n_dofs, m.n_cells
        self.x = _arr(m.x, 3 * m.n_dofs, np.float64).reshape(-1, 3)
        self.vertex_of_dof = _arr(m.vertex_of_dof, m.n_dofs, np.int32)
        self.cells = _arr(m.cells, 8 * m.n_cells, np.int32).reshape(-1, 8)
        self.bface_cell = _arr(m.bface_cell, m.n_bfaces, np.int32)
        self.bface_face = _arr(m.bface_face, m.n_bfaces, np.int32)
        self.bface_id = _arr(m.bface_id, m.n_bfaces, np.int32)
        self._csr = None

    # vertices of deal.II face f (lexicographic cell numbering): the 4 cell vertices with bit
    # f // 2 equal to f % 2
    FACE_VERTICES = [[k for k in range(8) if (k >> (f // 2)) & 1 == f % 2] for f in range(6)]

    def face_centroids(self, bid):
        """DealSolver::export_surface_centroids (src/DealSolver.cpp:229-245): the centre (mean of
        the 4 vertices) of every boundary face with id bid, in active-cell x face order."""
        sel = np.nonzero(self.bface_id == bid)[0]
        fv = np.array(self.FACE_VERTICES)[self.bface_face[sel]]
        dofs = self.cells[self.bface_cell[sel][:, None], fv]
        return self.x[dofs].mean(axis=1)

    def __del__(self):
        try:
            if self._csr is not None:
                lib().orc_csr_free(C.byref(self._csr))
            lib().orc_mesh_free(C.byref(self._m))
        except Exception:
            pass

    # ------------------------------------------------------------------------------
    def assemble(self, kappa_q=None, gamma_q=None):
        if self._csr is None:
            self._csr = _Csr()
            lib().orc_csr_pattern(C.byref(self._csr), C.byref(self._m))
        kq = None if kappa_q is None else np.ascontiguousarray(kappa_q, np.float64)
        gq = None if gamma_q is None else np.ascontiguousarray(gamma_q, np.float64)
        lib().orc_assemble(C.byref(self._csr), C.byref(self._m),
                           None if kq is None else kq.ctypes.data, None if gq is None else gq.ctypes.data)
        return self

    def spmv(self, x):
        y = np.zeros(self.n_dofs)
        lib().orc_csr_spmv(C.byref(self._csr), np.ascontiguousarray(x, np.float64), y)
        return y

    def csr(self):
        A = self._csr
        n, nnz = A.n, A.nnz
        return (_arr(A.rowptr, n + 1, np.int64), _arr(A.col, nnz, np.int32), _arr(A.val, nnz, np.float64))

    def cell_rhs(self, f_q):
        b = np.zeros(self.n_dofs)
        lib().orc_assemble_cell_rhs(b, C.byref(self._m), np.ascontiguousarray(f_q, np.float64))
        return b

    def dof_volumes(self):
        v = np.zeros(self.n_dofs)
        lib().orc_dof_volumes(C.byref(self._m), v)
        return v

    def charge_density(self, b):
        rho = np.zeros(self.n_dofs)
        lib().orc_charge_density(C.byref(self._m), np.ascontiguousarray(b, np.float64), rho)
        return rho

    def face_rhs(self, bid, values=None, constant=0.0, b=None):
        b = np.zeros(self.n_dofs) if b is None else b
        fv = None if values is None else np.ascontiguousarray(values, np.float64)
        lib().orc_assemble_face_rhs(b, C.byref(self._m), bid, None if fv is None else fv.ctypes.data, constant)
        return b

    def dirichlet_mask(self, bid, mask=None):
        mask = np.zeros(self.n_dofs, np.uint8) if mask is None else mask
        lib().orc_dirichlet_mask(C.byref(self._m), bid, mask)
        return mask

    def apply_dirichlet(self, b, x, mask, values=None):
        vals = None if values is None else np.ascontiguousarray(values, np.float64)
        lib().orc_apply_dirichlet(C.byref(self._csr), b, x, mask, None if vals is None else vals.ctypes.data)

    def pcg(self, b, x, max_it=10000, tol=1e-9, relative=False, precond="ssor", omega=1.2):
        kind = {"none": 0, "jacobi": 1, "ssor": 2}[precond]
        res = C.c_double()
        it = lib().orc_pcg(C.byref(self._csr), b, x, max_it, tol, int(relative), kind, omega, C.byref(res))
        return it, res.value

    def pcg_threads(self, b, x, threads, max_it=10000, tol=1e-9, relative=False, precond="ssor", omega=1.2):
        """pcg() with the matrix-vector product and vector operations on `threads` OpenMP
        threads (SSOR sweeps sequential); for the CPU reference timing."""
        kind = {"none": 0, "jacobi": 1, "ssor": 2}[precond]
        res = C.c_double()
        it = lib().orc_pcg_omp(C.byref(self._csr), b, x, max_it, tol, int(relative), kind, omega, int(threads),
                               C.byref(res))
        return it, res.value

    def qp_fields(self, u):
        """(values 8 per cell, gradients 8x3 per cell) of a dof vector at the Gauss points."""
        val = np.zeros(8 * self.n_cells)
        grad = np.zeros(24 * self.n_cells)
        lib().orc_qp_fields(C.byref(self._m), np.ascontiguousarray(u, np.float64), val.ctypes.data, grad.ctypes.data)
        return val.reshape(-1, 8), grad.reshape(-1, 8, 3)

    def export_solution_grad(self, u):
        g = np.zeros(3 * self.n_dofs)
        lib().orc_export_solution_grad(C.byref(self._m), np.ascontiguousarray(u, np.float64), g)
        return g.reshape(-1, 3)


def element_matrix(xe_lex, kappa_q=None, gamma_q=None):
    K = np.zeros(64)
    kq = None if kappa_q is None else np.ascontiguousarray(kappa_q, np.float64)
    gq = None if gamma_q is None else np.ascontiguousarray(gamma_q, np.float64)
    lib().orc_element_matrix(np.ascontiguousarray(xe_lex, np.float64).reshape(-1),
                             None if kq is None else kq.ctypes.data, None if gq is None else gq.ctypes.data, K)
    return K.reshape(8, 8)


def _pcg(m, b, x, threads, **kw):
    """m.pcg, or its multi-threaded form (CPU reference timing) when threads is given."""
    if threads:
        return m.pcg_threads(b, x, threads, **kw)
    return m.pcg(b, x, **kw)


def poisson_laplace(vertices, ucd_cells, applied_field, tol=1e-9, relative=False, precond="ssor",
                    omega=1.2, max_it=10000, anode="neumann", potential=0.0, threads=None):
    """PoissonSolver setup/assemble/solve for field_mode = laplace (src/PoissonSolver.cpp:162-210,
    src/DealSolver.cpp:443-458): phi = 0 on the copper surface (id 2), Neumann -E0 on the top
    (id 8).  Returns (mesh, solution per dof, iterations, residual)."""
    m = Mesh(vertices, ucd_cells, BID_VACUUM_TOP, BID_COPPER_SURFACE, BID_SIDES, BID_COPPER_SURFACE)
    m.assemble()
    b = np.zeros(m.n_dofs)
    x = np.zeros(m.n_dofs)
    mask = m.dirichlet_mask(BID_COPPER_SURFACE)
    vals = np.zeros(m.n_dofs)
    if anode == "neumann":
        m.face_rhs(BID_VACUUM_TOP, constant=applied_field, b=b)
    else:
        m.dirichlet_mask(BID_VACUUM_TOP, mask)
        top = np.zeros(m.n_dofs, np.uint8)
        m.dirichlet_mask(BID_VACUUM_TOP, top)
        vals[top > 0] = potential
        vals[m.dirichlet_mask(BID_COPPER_SURFACE) > 0] = 0.0
    m.apply_dirichlet(b, x, mask, vals)
    it, res = _pcg(m, b, x, threads, max_it=max_it, tol=tol, relative=relative, precond=precond, omega=omega)
    return m, x, it, res, b


CU_RHO_CP = 3.4496e-24      # volumetric heat capacity of copper [J/(K A^3)], include/CurrentHeatSolver.h:110
LORENTZ = 2.44e-8           # Lorenz number [W Ohm/K^2], src/Config.cpp:75
BID_COPPER_BOTTOM = 7


def current_solve(vertices, bulk_cells, face_j, tol=1e-12, precond="ssor", max_it=10000, sigma_q=None, threads=None):
    """CurrentSolver::assemble + solve (src/CurrentHeatSolver.cpp:421-484): sigma = 1 (or the
    given conductivity per cell and Gauss point), Neumann flux face_j on the copper surface (one
    value per id-2 face), phi = 0 on the bottom (id 7)."""
    m = Mesh(vertices, bulk_cells, BID_COPPER_SURFACE, BID_COPPER_BOTTOM, BID_SIDES, BID_COPPER_SURFACE)
    m.assemble(kappa_q=None if sigma_q is None else np.ascontiguousarray(sigma_q, np.float64).reshape(-1))
    b = m.face_rhs(BID_COPPER_SURFACE, values=face_j)
    x = np.zeros(m.n_dofs)
    m.apply_dirichlet(b, x, m.dirichlet_mask(BID_COPPER_BOTTOM), None)
    it, res = _pcg(m, b, x, threads, max_it=max_it, tol=tol, relative=True, precond=precond)
    return m, x, it


def heat_step(m: "Mesh", T_old, phi_c, dt, face_nottingham, T_amb, tol=1e-12, lorentz=LORENTZ, precond="ssor",
              theta=1.0, threads=None):
    """HeatSolver::assemble (implicit Euler, src/CurrentHeatSolver.cpp:116-152, 347-393) and solve
    on the bulk mesh m: A = gamma M + K_kappa(T_old), b = int phi_i (gamma T_old + sigma(T_old)
    |grad phi_c|^2) + Nottingham flux on id 2, T = T_amb on id 7; T_old is the initial guess.
    theta < 1 (extension, Crank-Nicolson at 1/2): A = gamma M + theta K, and b gains
    -(1 - theta) K T_old."""
    gamma = CU_RHO_CP / dt
    Tq, _ = m.qp_fields(T_old)
    _, gq = m.qp_fields(phi_c)
    flat = np.ascontiguousarray(Tq, np.float64).reshape(-1)
    kq, sq = np.empty(flat.size), np.empty(flat.size)
    lib().orc_sigma_kappa_array(flat.size, flat, lorentz, sq, kq)
    kq, sq = kq.reshape(Tq.shape), sq.reshape(Tq.shape)
    explicit = None
    if theta < 1.0:
        m.assemble(kappa_q=kq.reshape(-1), gamma_q=np.zeros(Tq.size))
        explicit = m.spmv(np.ascontiguousarray(T_old, np.float64))
    m.assemble(kappa_q=(theta * kq).reshape(-1), gamma_q=np.full(Tq.size, gamma))
    f = gamma * Tq + sq * (gq ** 2).sum(-1)
    b = m.cell_rhs(f.reshape(-1))
    if explicit is not None:
        b -= (1.0 - theta) * explicit
    m.face_rhs(BID_COPPER_SURFACE, values=face_nottingham, b=b)
    x = np.ascontiguousarray(T_old, np.float64).copy()
    mask = m.dirichlet_mask(BID_COPPER_BOTTOM)
    vals = np.where(mask > 0, T_amb, 0.0)
    m.apply_dirichlet(b, x, mask, vals)
    it, res = _pcg(m, b, x, threads, tol=tol, relative=True, precond=precond)
    return x, it


def sigma_T(Tq, sigma0, alpha):
    """Temperature-dependent conductivity of the current solve (extension):
    sigma0 / (1 + alpha (T - 300 K))."""
    return sigma0 / (1.0 + alpha * (np.asarray(Tq) - 300.0))


def nodal_field(nodes, hexes, hex_marker, region_sign, phi_nodes):
    n = len(nodes)
    out = np.zeros(3 * n)
    lib().orc_nodal_field(n, np.ascontiguousarray(nodes, np.float64).reshape(-1), len(hexes),
                          np.ascontiguousarray(hexes, np.int32).reshape(-1),
                          np.ascontiguousarray(hex_marker, np.int32), region_sign,
                          np.ascontiguousarray(phi_nodes, np.float64), out)
    return out.reshape(-1, 3)


class LinTet:
    def __init__(self, nodes, tets, tet_nbr, tet_marker, search_region=3):
        self._L = _Lintet()
        lib().orc_lintet_build(C.byref(self._L), len(nodes), np.ascontiguousarray(nodes, np.float64).reshape(-1),
                               len(tets), np.ascontiguousarray(tets, np.int32).reshape(-1),
                               np.ascontiguousarray(tet_nbr, np.int32).reshape(-1),
                               np.ascontiguousarray(tet_marker, np.int32), search_region)
        L = self._L
        nt = L.n_tets
        self.det0 = _arr(L.det0, nt, np.float64)
        self.det = _arr(L.det, 16 * nt, np.float64).reshape(-1, 4, 4)
        self.centroid = _arr(L.centroid, 3 * nt, np.float64).reshape(-1, 3)
        self.nbr_ptr = _arr(L.nbr_ptr, nt + 1, np.int32)
        self.nbr = _arr(L.nbr, int(self.nbr_ptr[-1]), np.int32)
        self.marker = _arr(L.marker, nt, np.int32)

    def __del__(self):
        try:
            lib().orc_lintet_free(C.byref(self._L))
        except Exception:
            pass

    def locate(self, p, guess):
        return lib().orc_lintet_locate(C.byref(self._L), np.ascontiguousarray(p, np.float64), guess)

    def bcc(self, p, tet):
        b = np.zeros(4)
        lib().orc_lintet_bcc(C.byref(self._L), np.ascontiguousarray(p, np.float64), tet, b)
        return b

    def locate_chain(self, points, rank=1):
        pts = np.ascontiguousarray(points, np.float64).reshape(-1)
        out = np.zeros(len(pts) // 3, np.int32)
        lib().orc_locate_chain(C.byref(self._L), len(out), pts, rank, out)
        return out


def _i32(a):
    return np.ascontiguousarray(a, np.int32).reshape(-1)


def _f64(a):
    return np.ascontiguousarray(a, np.float64).reshape(-1)


def quadtet_cells(tets, hexes, node_kind):
    out = np.zeros((len(tets), 10), np.int32)
    lib().orc_quadtet_cells(len(tets), _i32(tets), _i32(hexes), _i32(node_kind), out.reshape(-1))
    return out


def extract_nodal(nodes, hexes, hex_marker, region_sign, in_region, phi, rho=None, empty=0.0):
    """Nodal Solution rows (Ex, Ey, Ez, scalar1, scalar2) of Interpolator::extract_solution."""
    n = len(nodes)
    out = np.zeros(5 * n)
    r = None if rho is None else _f64(rho)
    lib().orc_extract_nodal(n, _f64(nodes), len(hexes), _i32(hexes), _i32(hex_marker), region_sign,
                            np.ascontiguousarray(in_region, np.uint8), _f64(phi),
                            None if r is None else r.ctypes.data, float(empty), out)
    return out.reshape(-1, 5)


def decay_factor(nodes, tets):
    return lib().orc_decay_factor(_f64(nodes), len(tets), _i32(tets))


def smooth_nodal(sol5, nodes, n_tetnodes, node_kind, hexes, hex_marker, vacuum=True, decay=None, tets=None):
    sol = np.ascontiguousarray(sol5, np.float64).copy()
    if decay is None:
        decay = decay_factor(nodes, tets)
    lib().orc_smooth_nodal(len(nodes), n_tetnodes, _f64(nodes), _i32(node_kind), len(hexes), _i32(hexes),
                           _i32(hex_marker), int(bool(vacuum)), float(decay), sol.reshape(-1))
    return sol


def interp_points(L: "LinTet", nodes, tets, cells10, hexes, sol5, points, rank):
    """SolutionReader::calc_full_interpolation (dim 3): (cells, Solution rows)."""
    pts = _f64(points)
    n = len(pts) // 3
    cells = np.zeros(n, np.int32)
    out = np.zeros(5 * n)
    lib().orc_interp_points(C.byref(L._L), _f64(nodes), _i32(tets), _i32(cells10), _i32(hexes), _f64(sol5), n, pts,
                            rank, cells, out)
    return cells, out.reshape(-1, 5)


def interp_cells(L: "LinTet", nodes, tets, cells10, hexes, sol5, points, rank, cells):
    pts = _f64(points)
    n = len(pts) // 3
    out = np.zeros(5 * n)
    lib().orc_interp_cells(C.byref(L._L), _f64(nodes), _i32(tets), _i32(cells10), _i32(hexes), _f64(sol5), n, pts,
                           rank, _i32(cells), out)
    return out.reshape(-1, 5)


class LinTri:
    """LinearTriangles on the surface triangles (3 nodes each) with their two tets."""

    def __init__(self, nodes, tris, tri2tet, quads=None, node_kind=None, tets=None):
        self._T = _Lintri()
        lib().orc_lintri_build(C.byref(self._T), len(nodes), _f64(nodes), len(tris), _i32(tris), _i32(tri2tet))
        T = self._T
        self.n = T.n
        self.centroid = _arr(T.cen, 3 * T.n, np.float64).reshape(-1, 3)
        self.nbr_ptr = _arr(T.nbr_ptr, T.n + 1, np.int32)
        self.nbr = _arr(T.nbr, int(self.nbr_ptr[-1]), np.int32)
        if quads is not None:
            rc = lib().orc_lintri_set_quads(C.byref(self._T), len(nodes), _f64(nodes), _i32(node_kind), _i32(tets),
                                            _i32(tris), _i32(quads))
            if rc != 0:
                raise ValueError("a quad's vertex is not a vertex of the face's tets")
            self.quad_centroid = _arr(T.qcen, 9 * T.n, np.float64).reshape(-1, 3)
            self.quad2hex = _arr(T.quad2hex, 6 * T.n, np.int32).reshape(-1, 2)
            self.tri6 = _arr(T.tri6, 6 * T.n, np.int32).reshape(-1, 6)

    def __del__(self):
        try:
            lib().orc_lintri_free(C.byref(self._T))
        except Exception:
            pass

    def point_in_cell(self, p, tri):
        return bool(lib().orc_lintri_point_in_cell(C.byref(self._T), _f64(p), tri))

    def locate(self, p, guess):
        return lib().orc_lintri_locate(C.byref(self._T), _f64(p), guess)

    def shape_functions(self, p, tri):
        w = np.zeros(3)
        lib().orc_lintri_shape_functions(C.byref(self._T), _f64(p), tri, w)
        return w

    def fast_distance(self, p, tri):
        return lib().orc_lintri_fast_distance(C.byref(self._T), _f64(p), tri)

    def interp(self, L: "LinTet", tets, sol5, points, cells=None, rank=1, centroids=False, xyz=None, cells10=None,
               hexes=None):
        """dim 2: (cells, Solution rows); with cells given, the mapped-atom path.  Ranks 2 and 3
        need the quadratic tets (cells10) and the hexes with their nodes (xyz)."""
        pts = _f64(points)
        n = len(pts) // 3
        out = np.zeros(5 * n)
        z = np.zeros(1)
        zi = np.zeros(1, np.int32)
        args = (C.byref(self._T), C.byref(L._L), z if xyz is None else _f64(xyz), _i32(tets),
                zi if cells10 is None else _i32(cells10), zi if hexes is None else _i32(hexes), _f64(sol5), n, pts,
                int(rank), int(bool(centroids)))
        if cells is None:
            cells = np.zeros(n, np.int32)
            lib().orc_interp_surface(*args, cells, out)
        else:
            cells = _i32(cells)
            lib().orc_interp_surface_cells(*args, cells, out)
        return cells, out.reshape(-1, 5)


def linhex_shape_functions(xe_ucd, p):
    sf = np.zeros(8)
    lib().orc_linhex_shape_functions(np.ascontiguousarray(xe_ucd, np.float64).reshape(-1),
                                     np.ascontiguousarray(p, np.float64), sf)
    return sf


def linhex_gradient(nodes, hexes, sol5, points, hex):
    """-grad of the nodal scalar2 at the points in the given hexes (interp_gradient)."""
    pts = _f64(points)
    n = len(pts) // 3
    out = np.zeros(3 * n)
    lib().orc_linhex_gradient(_f64(nodes), _i32(hexes), _f64(sol5), n, pts, _i32(hex), out)
    return out.reshape(-1, 3)


def update_cells(L: "LinTet", deal2femocs, femocs2deal, points, cells):
    """Pic::update_point_cell for every particle (solver cells in, solver cells or -1 out)."""
    pts = _f64(points)
    c = np.array(cells, np.int32).reshape(-1)
    lib().orc_update_cells(C.byref(L._L), _i32(deal2femocs), _i32(femocs2deal), len(c), pts, c)
    return c


def space_charge(m: "Mesh", nodes, hexes, deal2femocs, points, cells, factor, b=None):
    """b += the particles' charge deposit (PoissonSolver::assemble_space_charge_fast)."""
    pts = _f64(points)
    b = np.zeros(m.n_dofs) if b is None else b
    lib().orc_space_charge(C.byref(m._m), _f64(nodes), _i32(hexes), _i32(deal2femocs), len(pts) // 3, pts,
                           _i32(cells), float(factor), b)
    return b


def resistivity(T):
    return lib().orc_resistivity(float(T))


def sigma(T):
    return lib().orc_sigma(float(T))


def kappa(T, lorentz=2.44e-8):
    return lib().orc_kappa(float(T), float(lorentz))

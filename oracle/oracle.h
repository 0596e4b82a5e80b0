/*
 * CPU oracle for the FEMOCS field-solve path.  TEST INFRASTRUCTURE ONLY: the product
 * (libfemocs_gpu.so) never links, loads or calls anything here; tests, the smoke check and
 * the CPU-baseline leg of the benchmark use it as the checker.
 *
 * It restates, in plain C, the arithmetic the reference performs through deal.II 9.2 and its
 * own interpolators.  deal.II is not available, so the deal.II parts (Q1 assembly, boundary
 * handling, SolverCG + PreconditionSSOR) are restated from their published algorithms, and
 * are pinned by closed forms, invariants and brute force in tests/ (see each function).
 *
 * Compiled with -ffp-contract=off so that every expression rounds like the reference built
 * for x86-64 without FMA.
 */
#ifndef FEMOCS_ORACLE_H
#define FEMOCS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- mesh import (DealSolver::import_mesh, src/DealSolver.cpp:192-209; mark_boundary
 *      src/DealSolver.cpp:461-518) ------------------------------------------------------- */
typedef struct {
    int n_dofs;            /* vertices used by cells, compacted in increasing input order */
    int n_cells;
    double *x;             /* 3*n_dofs SoA-interleaved xyz of the used vertices           */
    int *vertex_of_dof;    /* n_dofs: input vertex index of each dof                      */
    int *cells;            /* 8*n_cells, deal.II lexicographic order, dof indices         */
    int n_bfaces;          /* boundary faces, in active-cell x face 0..5 order            */
    int *bface_cell;       /* n_bfaces                                                    */
    int *bface_face;       /* n_bfaces: deal.II face number 0..5                          */
    int *bface_id;         /* n_bfaces: boundary id                                       */
} orc_mesh;

int  orc_mesh_import(orc_mesh *m, int n_vertices, const double *xyz, int n_cells,
                     const int *ucd_cells, int top, int bottom, int sides, int other);
void orc_mesh_free(orc_mesh *m);

/* ---- CSR system (SparseMatrix<double>, DynamicSparsityPattern) ------------------------ */
typedef struct {
    int n;
    int64_t nnz;
    int64_t *rowptr;
    int *col;
    double *val;
    int64_t *diag;         /* position of the diagonal in each row */
} orc_csr;

int  orc_csr_pattern(orc_csr *A, const orc_mesh *m);
void orc_csr_free(orc_csr *A);
void orc_csr_spmv(const orc_csr *A, const double *x, double *y);

/* Element matrices: K_e(i,j) = sum_q JxW (kappa_q grad phi_i . grad phi_j + gamma_q phi_i phi_j),
 * kappa_q/gamma_q per cell and quadrature point (8 per cell, lexicographic Gauss points), or
 * NULL for kappa = 1 / gamma = 0.  PoissonSolver::assemble_local_cell
 * (src/PoissonSolver.cpp:239-263), HeatSolver::assemble_local_cell
 * (src/CurrentHeatSolver.cpp:347-393). */
void orc_element_matrix(const double xe[8][3], const double kappa_q[8], const double gamma_q[8],
                        double Ke[8][8]);
void orc_assemble(orc_csr *A, const orc_mesh *m, const double *kappa_q, const double *gamma_q);
/* b_i += sum_q JxW phi_i f_q (per cell and quadrature point) */
void orc_assemble_cell_rhs(double *b, const orc_mesh *m, const double *f_q);

/* DealSolver::assemble_rhs (src/DealSolver.cpp:390-430): Neumann load on faces with id bid.
 * face_values: one value per matching face in cell x face order, or NULL with a constant. */
void orc_assemble_face_rhs(double *b, const orc_mesh *m, int bid, const double *face_values,
                           double constant);

/* dofs on faces with the given id (VectorTools::interpolate_boundary_values).  mask[i]=1 */
int  orc_dirichlet_mask(const orc_mesh *m, int bid, unsigned char *mask);
/* MatrixTools::apply_boundary_values with column elimination (deal.II 9.2 algorithm) */
void orc_apply_dirichlet(orc_csr *A, double *b, double *x, const unsigned char *mask,
                         const double *values);

/* SolverCG + PreconditionSSOR / Jacobi / identity (src/DealSolver.cpp:443-458).
 * precond: 0 identity, 1 Jacobi, 2 SSOR(omega).  relative != 0: stop on |r|/|b| <= tol.
 * Returns iterations, negative on non-convergence. final residual in *res. */
int  orc_pcg(const orc_csr *A, const double *b, double *x, int max_it, double tol, int relative,
             int precond, double omega, double *res);

/* -grad u at every vertex, taken at the Gauss point "vertex2node" of the last cell touching it
 * (DealSolver::export_solution_grad, src/DealSolver.cpp:281-301). */
void orc_export_solution_grad(const orc_mesh *m, const double *u, double *grad);

/* ---- interpolation (src/Interpolator.cpp, src/InterpolatorCells.cpp) ------------------ */
/* Nodal field: mean over the region's hexes of the exact -grad phi at the hex corner
 * (Interpolator::store_elfield, src/Interpolator.cpp:125-140 with
 * LinearHexahedra::shape_fun_grads(hex,node), src/InterpolatorCells.cpp:1462-1505).
 * hexes: femocs/UCD order over all nodes; region_sign: +1 vacuum (marker>0), -1 bulk.
 * phi: per node (nodes outside the region ignored).  out: 3*n_nodes (zero if no hex). */
void orc_nodal_field(int n_nodes, const double *xyz, int n_hexes, const int *hexes,
                     const int *hex_marker, int region_sign, const double *phi, double *out);

typedef struct {
    int n_tets;
    double *det0;          /* 1/d0            */
    double *det;           /* 16 per tet: det1..det4 as Vec4 */
    double *centroid;      /* 3 per tet       */
    int *nbr_ptr;          /* neighbours CSR  */
    int *nbr;
    int *marker;           /* search marker: 0 = searchable */
} orc_lintet;

int  orc_lintet_build(orc_lintet *L, int n_nodes, const double *xyz, int n_tets, const int *tets,
                      const int *tet_nbr, const int *tet_marker, int search_region);
void orc_lintet_free(orc_lintet *L);
int  orc_lintet_point_in_cell(const orc_lintet *L, const double p[3], int tet);
void orc_lintet_bcc(const orc_lintet *L, const double p[3], int tet, double bcc[4]);
int  orc_lintet_locate(const orc_lintet *L, const double p[3], int guess);
/* locate every point with the chained guess of SolutionReader::calc_full_interpolation
 * (src/SolutionReader.cpp:136-165).  rank 1: tet index; rank 3: hex index */
void orc_locate_chain(const orc_lintet *L, int n_points, const double *xyz, int rank, int *cells);
/* LinearHexahedra::locate_cell given a tet locate (src/InterpolatorCells.cpp:1536-1562) */
int  orc_linhex_locate(const orc_lintet *L, const double p[3], int guess);

/* trilinear hex shape functions in femocs order via Newton inversion
 * (src/InterpolatorCells.cpp:1237-1244, 1272-1317, 1334-1353) */
void orc_linhex_shape_functions(const double xe[8][3], const double p[3], double sf[8]);

/* physical quantities: resistivity table, sigma, kappa (src/PhysicalQuantities.cpp:31-64,
 * 173-185; table src/PhysicalQuantityData.cpp:47-76) */
double orc_resistivity(double T);
double orc_sigma(double T);
double orc_kappa(double T, double lorentz);

#ifdef __cplusplus
}
#endif
#endif

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

/* DealSolver::calc_dof_volumes (src/DealSolver.cpp:344-366): vol_i = sum over the cells of
 * dof i of the cell volume (sum_q JxW, not shape-weighted).  orc_charge_density: the
 * write-step charge density of PoissonSolver::assemble (src/PoissonSolver.cpp:196-207),
 * rho_i = b_i / vol_i with b the right-hand side before the Dirichlet conditions. */
void orc_dof_volumes(const orc_mesh *m, double *vol);
void orc_charge_density(const orc_mesh *m, const double *b, double *rho);

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
/* orc_pcg with the matrix-vector product and vector operations on nthreads OpenMP threads
 * (SSOR sweeps sequential), for the CPU reference timing (refcpu.c) */
int  orc_pcg_omp(const orc_csr *A, const double *b, double *x, int max_it, double tol, int relative,
                 int precond, double omega, int nthreads, double *res_out);
int  orc_omp_max_threads(void);

/* -grad u at every vertex, taken at the Gauss point "vertex2node" of the last cell touching it
 * (DealSolver::export_solution_grad, src/DealSolver.cpp:281-301). */
void orc_export_solution_grad(const orc_mesh *m, const double *u, double *grad);
/* FEValues::get_function_values / get_function_gradients at the QGauss(2) points of every
 * cell (as the heat and current assemblies read them, src/CurrentHeatSolver.cpp:355-356,
 * 462): val 8 per cell, grad 24 per cell (point-major, x y z); either may be NULL */
void orc_qp_fields(const orc_mesh *m, const double *u, double *val, double *grad);

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
/* LinearHexahedra::interp_gradient(point, hex) (src/InterpolatorCells.cpp:407-421, 1363-1416):
 * out = -sum_i grad N_i(p) * scalar2 of node i, for point q in hex hex[q] (3 per point) */
void orc_linhex_gradient(const double *nodes, const int *hexes, const double *sol5, int n, const double *pts,
                         const int *hex, double *out);

/* Pic::update_point_cell (src/Pic.cpp:185-195) for each particle: cells[i] (solver cell) ->
 * LinearHexahedra::locate_cell from hex deal2femocs[cells[i]] -> femocs2deal, -1 outside */
void orc_update_cells(const orc_lintet *L, const int *deal2femocs, const int *femocs2deal, int n, const double *pts,
                      int *cells);

/* PoissonSolver::assemble_space_charge_fast (src/PoissonSolver.cpp:300-319): particles in order,
 * b[dofs of solver cell c_p] += shape functions of femocs hex deal2femocs[c_p] at x_p, in the
 * cell's lexicographic order (shape_funs_dealii, src/InterpolatorCells.cpp:1355-1358), times
 * factor (q/eps0 * W_sp).  nodes: the femocs node coordinates the hexes index (UCD order) */
void orc_space_charge(const orc_mesh *m, const double *nodes, const int *hexes, const int *deal2femocs, int n,
                      const double *pts, const int *cells, double factor, double *b);

/* QuadraticTetrahedra::calc_cell (src/InterpolatorCells.cpp:1151-1174): per tet its four
 * vertices and the six edge-centre nodes n5..n10 found as the first common EDGECENTROID node
 * (node_kind 2) of the hexes 4t+i, 4t+j (pairs 01, 12, 20, 03, 13, 23).  cells10: 10 per tet */
void orc_quadtet_cells(int n_tets, const int *tets, const int *hexes, const int *node_kind, int *cells10);

/* Interpolator::extract_solution(PoissonSolver&) without smoothing (src/Interpolator.cpp:79-140,
 * 172-190): Solution{E, charge density, phi} per node, 5 doubles (vector x y z, scalar1,
 * scalar2).  in_region[n] != 0 for nodes of the solver mesh (femocs2deal >= 0): E = nodal
 * field (zero if no hex of the region touches the node), scalar1 = rho[n], scalar2 = phi[n];
 * other nodes get empty in all five. */
void orc_extract_nodal(int n_nodes, const double *xyz, int n_hexes, const int *hexes, const int *hex_marker,
                       int region_sign, const unsigned char *in_region, const double *phi, const double *rho,
                       double empty, double *sol5);

/* Interpolator::average_nodal_fields + TetgenMesh::calc_pseudo_3D_vorocells
 * (src/Interpolator.cpp:142-170, src/TetgenMesh.cpp:819-837): for each tet node i < n_tetnodes,
 * E_i <- sum_w E_nbr / sum w, w = exp(decay * |x_i - x_nbr|), over the nodes of the region's
 * hexes tied to i (|hex marker| - 1 == i) with node_kind >= 2, repeats kept.  decay =
 * -1 / (longest tet edge) (LinearTetrahedra::precompute, src/InterpolatorCells.cpp:534) */
double orc_decay_factor(const double *xyz, int n_tets, const int *tets);
void orc_smooth_nodal(int n_nodes, int n_tetnodes, const double *xyz, const int *node_kind, int n_hexes,
                      const int *hexes, const int *hex_marker, int vacuum, double decay, double *sol5);

/* SolutionReader::calc_full_interpolation for dim 3, unsorted atoms (src/SolutionReader.cpp:43-165,
 * InterpolatorCells::locate_interpolate / interp_solution src/InterpolatorCells.cpp:331-355,
 * 441-445): chained locate from guess -1, then sum_i w_i Solution_i with rank 1 the linear tet
 * weights, rank 2 the quadratic tet weights (src/InterpolatorCells.cpp:795-812) over cells10,
 * rank 3 the trilinear hex weights over hexes.  cells: tet (ranks 1, 2) or hex (rank 3) per
 * point, negative outside; out: 5 per point. */
void orc_interp_points(const orc_lintet *L, const double *xyz, const int *tets, const int *cells10,
                       const int *hexes, const double *sol5, int n_points, const double *pts, int rank,
                       int *cells, double *out);
/* SolutionReader::calc_interpolation with mapped atoms: interpolation in |cells[i]| */
void orc_interp_cells(const orc_lintet *L, const double *xyz, const int *tets, const int *cells10,
                      const int *hexes, const double *sol5, int n_points, const double *pts, int rank,
                      const int *cells, double *out);

/* LinearTriangles (src/InterpolatorCells.cpp:1585-1747): surface triangles of the tet mesh,
 * tri2tet the two tets of each triangle (TetgenFaces::to_tets) */
typedef struct {
    int n;
    double *v0, *e1, *e2, *pv, *nrm, *maxd, *cen;   /* 3 per triangle (maxd 1) */
    int *nbr_ptr, *nbr, *tri2tet, *tris;
    /* set by orc_lintri_set_quads: 3 quadrangles per triangle (quad 3f+k around vertex k),
     * their centroids, the two hexes of each quad (TetgenMesh::calc_quad2hex2quad_mapping,
     * hex 4t+j around vertex j of tet t) and the 6-node triangles of QuadraticTriangles */
    int *quads, *quad2hex, *tri6;
    double *qcen;
} orc_lintri;
int  orc_lintri_build(orc_lintri *T, int n_nodes, const double *xyz, int n_tris, const int *tris, const int *tri2tet);
void orc_lintri_free(orc_lintri *T);
int  orc_lintri_point_in_cell(const orc_lintri *T, const double p[3], int tri);
void orc_lintri_shape_functions(const orc_lintri *T, const double p[3], int tri, double w[3]);
double orc_lintri_fast_distance(const orc_lintri *T, const double p[3], int tri);
int  orc_lintri_locate(const orc_lintri *T, const double p[3], int guess);
/* Quadrangles of the surface triangles (Tethex convert_triangles, src/Tethex.cpp:1104-1170) with
 * the tets of the mesh and the node kinds (edge centroids = 2) */
int  orc_lintri_set_quads(orc_lintri *T, int n_nodes, const double *xyz, const int *node_kind, const int *tets,
                          const int *tris, const int *quads);
/* SolutionReader::calc_full_interpolation, dim 2 (src/SolutionReader.cpp:43-90, 136-163).
 * rank 1 LinearTriangles, rank 2 QuadraticTriangles, rank 3 LinearQuadrangles
 * (src/InterpolatorCells.cpp:1662-1686, 1838-1860, 1925-1992): chained face (or quad) locate,
 * then interpolation in the tet (hex) the face picks: tets[0] when on the face plane or inside
 * tets[0], else the tet (hex) search from tets[1].  centroids != 0: locate_interp_centroid
 * (:309-356, 453-461): exact centroid match, else Solution(-1) and the previous cell kept. */
void orc_interp_surface(const orc_lintri *T, const orc_lintet *L, const double *xyz, const int *tets,
                        const int *cells10, const int *hexes, const double *sol5, int n_points, const double *pts,
                        int rank, int centroids, int *cells, double *out);
void orc_interp_surface_cells(const orc_lintri *T, const orc_lintet *L, const double *xyz, const int *tets,
                              const int *cells10, const int *hexes, const double *sol5, int n_points,
                              const double *pts, int rank, int centroids, const int *cells, double *out);

/* physical quantities: resistivity table, sigma, kappa (src/PhysicalQuantities.cpp:31-64,
 * 173-185; table src/PhysicalQuantityData.cpp:47-76) */
double orc_resistivity(double T);
double orc_sigma(double T);
double orc_kappa(double T, double lorentz);
void orc_sigma_kappa_array(int n, const double *T, double lorentz, double *sigma, double *kappa);

#ifdef __cplusplus
}
#endif
#endif

/*
 * femocs_gpu: MI355X (gfx950) field-solve path of FEMOCS behind a C ABI.
 *
 * Plain C types only; no torch, no C++ in the signatures.  Every entry point returns a status
 * (FGPU_OK or a negative code) and never throws; fgpu_ctx_error() gives the message of the
 * last failure on a context.  Host arrays are caller-owned and copied; device memory is owned
 * by the handle that allocated it.  Handles are independent: one context per thread or
 * per client is re-entrant, calls on one context are serialised on its stream.
 *
 * The reference path these entry points replace is the deal.II-based DealSolver family and
 * the interpolators; each declaration cites the reference function it stands in for.
 */
#ifndef FEMOCS_GPU_H
#define FEMOCS_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FGPU_OK 0
#define FGPU_EINVAL (-1)
#define FGPU_EHIP (-2)
#define FGPU_ENOMEM (-3)
#define FGPU_ESTATE (-4)
#define FGPU_EMESH (-5) /* moved vertices would invert a cell: rebuild the mesh instead */

/* boundary ids, include/Globals.h:60-68 */
#define FGPU_BID_COPPER_SURFACE 2
#define FGPU_BID_VACUUM_TOP 8
#define FGPU_BID_COPPER_BOTTOM 7
#define FGPU_BID_SIDES 4

/* precision of a system's iterations */
#define FGPU_FP64 0   /* FP64 CG (the reference's arithmetic)                            */
#define FGPU_MIXED 1  /* FP32 operator and CG inside FP64 defect correction               */

/* preconditioners.  FGPU_PC_NONE: identity, as PreconditionIdentity when ssor_param <= 0
 * (src/DealSolver.cpp:447-452).  FGPU_PC_JACOBI: the diagonal, plus the two-level coarse
 * correction of a system that has one (fgpu_sys_set_coarse), in either precision.
 * FGPU_PC_CHEBYSHEV (FGPU_FP64 only): Chebyshev-Jacobi polynomial preconditioner
 * z = p(D^-1 A) D^-1 r on [lambda_max / 100, lambda_max], lambda_max from a Lanczos estimate
 * (the first Jacobi-PCG iterations of the solve, repeated when the system changed).  All
 * three stand in for the SSOR preconditioner of the reference, whose sweeps are sequential. */
#define FGPU_PC_NONE 0
#define FGPU_PC_JACOBI 1
#define FGPU_PC_CHEBYSHEV 2

typedef struct fgpu_ctx fgpu_ctx;
typedef struct fgpu_mesh fgpu_mesh;
typedef struct fgpu_sys fgpu_sys;
typedef struct fgpu_interp fgpu_interp;
typedef struct fgpu_tethex fgpu_tethex;

typedef struct {
    int max_it;        /* max CG iterations (inner + outer for FGPU_MIXED)                    */
    double tol;        /* stop when |r| <= tol (absolute) or |r|/|b| <= tol (relative)         */
    int relative;      /* 0: absolute, as SolverControl in src/DealSolver.cpp:444; 1: relative */
    int precond;       /* FGPU_PC_*                                                           */
    int precision;     /* FGPU_FP64 or FGPU_MIXED                                             */
    double inner_tol;  /* FGPU_MIXED: relative reduction asked of each FP32 inner solve        */
    int check_every;   /* iterations between host convergence checks (0: default)             */
    int use_graph;     /* capture the iterations between checks in a hipGraph                 */
    int cheb_degree;   /* FGPU_PC_CHEBYSHEV: polynomial degree (0: chosen by the system size)  */
} fgpu_solve_opts;

typedef struct {
    int iterations;       /* CG iterations (negative: not converged), as solve_cg's return   */
    int outer;            /* FP64 residual evaluations (defect-correction steps)             */
    double residual;      /* final |b - Ax| (FP64 true residual for FGPU_MIXED)              */
    double rhs_norm;      /* |b| of the lifted system                                        */
} fgpu_solve_stats;

/* ---- context ------------------------------------------------------------------------ */
int fgpu_ctx_create(int device, fgpu_ctx **out);
void fgpu_ctx_destroy(fgpu_ctx *ctx);
const char *fgpu_ctx_error(const fgpu_ctx *ctx);
int fgpu_ctx_sync(fgpu_ctx *ctx);
/* device memory owned by the caller on the context's device (e.g. for
 * fgpu_sys_solution_to_device checkpoints); no counterpart in the reference */
int fgpu_ctx_alloc(fgpu_ctx *ctx, size_t bytes, void **dev);
int fgpu_ctx_free(fgpu_ctx *ctx, void *dev);

/* ---- mesh ---------------------------------------------------------------------------
 * DealSolver::import_mesh(vector<Point<3>>, vector<CellData<3>>) (src/DealSolver.cpp:192-209)
 * followed by mark_boundary(top, bottom, sides, other) (src/DealSolver.cpp:461-518):
 * unused vertices are dropped (order kept), UCD cells are converted to lexicographic order,
 * negatively oriented cells are inverted, boundary faces are marked by the bounding box of
 * their centres with eps = 1e-6.  xyz: 3*n_vertices, ucd_cells: 8*n_cells. */
int fgpu_mesh_import(fgpu_ctx *ctx, int n_vertices, const double *xyz, int n_cells,
                     const int *ucd_cells, int top, int bottom, int sides, int other,
                     fgpu_mesh **out);
/* the same import with flags.  The tables (compaction, lexicographic cells, orientation,
 * boundary faces and ids, vertex -> cell, the internal order, patches and slots) are built on
 * the device by default; FGPU_IMPORT_HOST builds them on the host.  Both give the same bits
 * (fgpu_mesh_digest).  No counterpart in the reference, which imports through deal.II
 * (DealSolver::import_mesh, src/DealSolver.cpp:192-209). */
#define FGPU_IMPORT_HOST 1
int fgpu_mesh_import_ex(fgpu_ctx *ctx, int n_vertices, const double *xyz, int n_cells,
                        const int *ucd_cells, int top, int bottom, int sides, int other, int flags,
                        fgpu_mesh **out);
/* 64-bit FNV-1a hashes of every import table, host mirrors and device arrays (n >=
 * FGPU_MESH_DIGEST_TABLES): two imports of one mesh agree table by table.  No counterpart in
 * the reference. */
#define FGPU_MESH_DIGEST_TABLES 32
int fgpu_mesh_digest(const fgpu_mesh *mesh, int n, unsigned long long *out);
void fgpu_mesh_destroy(fgpu_mesh *mesh);
/* DealSolver::size / get_n_cells; n_bfaces: boundary faces */
int fgpu_mesh_sizes(const fgpu_mesh *mesh, int *n_dofs, int *n_cells, int *n_bfaces);
/* DealSolver::export_vertices/export_dofs (src/DealSolver.cpp:248-267): dof coordinates
 * (3*n_dofs) and the input vertex index of every dof; either pointer may be NULL */
int fgpu_mesh_export_dofs(const fgpu_mesh *mesh, double *xyz, int *vertex_of_dof);
/* boundary faces in active-cell x face order: cell index, deal.II face 0..5, boundary id */
int fgpu_mesh_export_bfaces(const fgpu_mesh *mesh, int *cell, int *face, int *bid);
/* DealSolver::export_surface_centroids (src/DealSolver.cpp:229-245) for a boundary id;
 * returns the count in *n (xyz may be NULL to query it) */
int fgpu_mesh_export_face_centroids(const fgpu_mesh *mesh, int bid, double *xyz, int *n);
/* moved vertices, same topology (3*n_vertices of the import numbering).  The reference
 * rebuilds the mesh every step (src/ProjectRunaway.cpp:214-228); this keeps it resident.
 * FGPU_EMESH, coordinates unchanged, when a cell's centre Jacobian changes sign: the
 * orientation chosen at import would no longer be the one a fresh import picks. */
int fgpu_mesh_update_vertices(fgpu_mesh *mesh, const double *xyz);
/* the same for a subset: n vertices (import numbering) get new positions xyz (3*n); for
 * callers that move few vertices per step (surface atoms), no counterpart in the reference */
int fgpu_mesh_move_vertices(fgpu_mesh *mesh, int n, const int *vertex_ids, const double *xyz);
/* linear mesh motion computed on the device, no counterpart in the reference: `rows` vertices
 * (import numbering) follow n_ctrl control points, vertex_ids[r] sitting at its position at
 * this call plus sum_k w[k] * disp[ctrl[k]] for k in ptr[r]..ptr[r+1]-1 (e.g. the Tethex
 * edge, face and tet centres following the surface atoms they average).  fgpu_mesh_displace
 * sets the displacements (3*n_ctrl, relative to the positions at fgpu_mesh_set_motion);
 * FGPU_EMESH as for fgpu_mesh_update_vertices. */
int fgpu_mesh_set_motion(fgpu_mesh *mesh, int rows, const int *vertex_ids, const int *ptr, const int *ctrl,
                         const double *w, int n_ctrl);
int fgpu_mesh_displace(fgpu_mesh *mesh, const double *disp);
/* fgpu_mesh_displace without any allocation, no counterpart in the reference: for callers that
 * cannot allocate device memory (code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same inputs, results and errors as fgpu_mesh_displace; the partial
 * minima of the validity check live in a fixed device array of the library (512 KiB per GPU,
 * held until the process ends), so meshes of more than 16777216 cells are refused with
 * FGPU_EINVAL.  Calls on one GPU share that array: calls that overlap on one GPU (several
 * host threads on different contexts of the same device) race. */
int fgpu_mesh_displace_noalloc(fgpu_mesh *mesh, const double *disp);
/* device memory held by a mesh / a system (bytes), for admission by memory budget; no
 * counterpart in the reference */
int fgpu_mesh_device_bytes(const fgpu_mesh *mesh, size_t *bytes);
int fgpu_sys_device_bytes(const fgpu_sys *sys, size_t *bytes);

/* Tuning, no counterpart in the reference: the element kernel works on patches of
 * consecutive cells whose nodes are staged in LDS; cells per patch (8..256) and patches per
 * workgroup (1..64) trade LDS reuse against parallelism.  Changing the patch size rebuilds
 * the mesh's patch tables; systems on the mesh follow at their next call. */
int fgpu_mesh_set_patch_cells(fgpu_mesh *mesh, int cells);
int fgpu_mesh_patch_info(const fgpu_mesh *mesh, int *cells, int *n_patches, int *node_cap, int *n_slots);
int fgpu_sys_set_patches_per_block(fgpu_sys *sys, int patches_per_block);

/* ---- linear system on a mesh ----------------------------------------------------------
 * One fgpu_sys = the system_matrix/system_rhs/solution triple of a DealSolver, held on the
 * GPU matrix-free. */
int fgpu_sys_create(fgpu_mesh *mesh, fgpu_sys **out);
void fgpu_sys_destroy(fgpu_sys *sys);
/* Cell coefficients per quadrature point (8 per cell, lexicographic QGauss(2) points, in the
 * mesh's cell order): stiffness kappa and mass gamma.  NULL: kappa = 1, gamma = 0
 * (PoissonSolver::assemble_local_cell src/PoissonSolver.cpp:239-263;
 *  HeatSolver::assemble_local_cell src/CurrentHeatSolver.cpp:347-393). */
int fgpu_sys_set_coefficients(fgpu_sys *sys, const double *kappa_q, const double *gamma_q);
/* system_rhs = 0 and boundary_values.clear() (src/PoissonSolver.cpp:176, DealSolver.h:151) */
int fgpu_sys_reset(fgpu_sys *sys);
/* DealSolver::assemble_rhs(bid) (src/DealSolver.cpp:390-430): b_i += int phi_i g dS over the
 * faces with id bid; g = face_values[k] for the k-th such face in cell x face order, or the
 * constant when face_values is NULL */
int fgpu_sys_add_neumann(fgpu_sys *sys, int bid, const double *face_values, double constant);
/* b_i += int phi_i f over cells, f per quadrature point (8 per cell) */
int fgpu_sys_add_cell_rhs(fgpu_sys *sys, const double *f_q);
/* DealSolver::append_dirichlet(bid, value) (src/DealSolver.cpp:433-435) */
int fgpu_sys_add_dirichlet(fgpu_sys *sys, int bid, double value);
/* PoissonSolver::assemble on a write step (src/PoissonSolver.cpp:196-207): charge density =
 * the right-hand side assembled so far divided per dof by its volume (DealSolver::calc_dof_volumes,
 * src/DealSolver.cpp:344-366: the volumes of the cells holding the dof).  Call before
 * fgpu_sys_apply_dirichlet (FGPU_EINVAL after it); fgpu_sys_reassemble recomputes it until the
 * next fgpu_sys_reset, which clears it */
int fgpu_sys_save_charge_density(fgpu_sys *sys);
/* PoissonSolver::assemble_space_charge_fast (src/PoissonSolver.cpp:300-319), the field mode with
 * particles: b_i += factor * (shape function of dof i of the particle's cell at its position),
 * added in particle order (bit-identical to the sequential loop).  cells: the solver cell of
 * each particle (deal.II active cell index = cell index of the imported mesh); the shape
 * functions are those of the interpolator's hex for that cell (Interpolator::deal2femocs,
 * LinearHexahedra::shape_funs_dealii), so the system's mesh must have been imported from the
 * interpolator's region hexes in hex order.  factor = q/eps0 * W_sp.  Before
 * fgpu_sys_apply_dirichlet; fgpu_sys_reassemble refuses a system holding a deposit (reset and
 * deposit again after a move).  FGPU_EINVAL for a cell out of range. */
int fgpu_sys_add_space_charge(fgpu_sys *sys, fgpu_interp *interp, int n, const double *xyz, const int *cells,
                              double charge_factor);
/* PoissonSolver::export_charge_dens (src/PoissonSolver.cpp:141-149), dof order: the saved
 * charge density, zeros when none was saved since the last reset (charge_density.reinit) */
int fgpu_sys_export_charge_density(fgpu_sys *sys, double *rho);
/* DealSolver::apply_dirichlet (src/DealSolver.cpp:438-440): lifting with column elimination;
 * sets solution = value on the constrained dofs */
int fgpu_sys_apply_dirichlet(fgpu_sys *sys);
/* after the vertices moved: rebuilds the right-hand side terms added since the last reset
 * (Neumann loads, cell sources), the diagonal and the Dirichlet lifting on the device, as the
 * same calls would after a reset; no counterpart in the reference (it reassembles from scratch,
 * src/PoissonSolver.cpp:170-210) */
int fgpu_sys_reassemble(fgpu_sys *sys);
/* solution = value everywhere (DealSolver::setup_system, src/DealSolver.cpp:383-384) */
int fgpu_sys_fill_solution(fgpu_sys *sys, double value);
/* DealSolver::import_solution / export_solution (src/DealSolver.cpp:270-315), dof order */
int fgpu_sys_set_solution(fgpu_sys *sys, const double *x);
int fgpu_sys_get_solution(fgpu_sys *sys, double *x);
/* device-to-device copies of the solution (n_dofs doubles in the library's internal dof
 * order, meaningful only for the same mesh) to and from caller-owned device memory, ordered on
 * the context's stream: checkpoints of served results without host copies.  No counterpart in
 * the reference. */
int fgpu_sys_solution_to_device(fgpu_sys *sys, double *dev);
int fgpu_sys_solution_from_device(fgpu_sys *sys, const double *dev);
/* lifted right-hand side (dof order), for checks */
int fgpu_sys_get_rhs(fgpu_sys *sys, double *b);
/* y = A x with the un-lifted operator (the matrix before apply_dirichlet), FP64 */
int fgpu_sys_apply(fgpu_sys *sys, const double *x, double *y);
/* DealSolver::solve_cg (src/DealSolver.cpp:443-458) */
int fgpu_sys_solve(fgpu_sys *sys, const fgpu_solve_opts *opts, fgpu_solve_stats *stats);
/* many systems in one set of launches; stats: one per system */
int fgpu_solve_batch(fgpu_ctx *ctx, int n, fgpu_sys *const *systems, const fgpu_solve_opts *opts,
                     fgpu_solve_stats *stats);
/* Continuous batching, no counterpart in the reference (it solves one system at a time):
 * callers that serve many independent solves keep one batch running and let systems join
 * (fgpu_batch_add: the solve of that system starts at once) and leave (fgpu_batch_retire: the
 * finished ones, with their stats) between steps (fgpu_batch_step: `iterations` CG iterations of
 * every running member; *n_running: members not finished).  A system's result is the same
 * as when solved alone.  Each member keeps the options given at creation. */
typedef struct fgpu_batch fgpu_batch;
int fgpu_batch_create(fgpu_ctx *ctx, const fgpu_solve_opts *opts, fgpu_batch **out);
void fgpu_batch_destroy(fgpu_batch *batch);
int fgpu_batch_size(const fgpu_batch *batch);
int fgpu_batch_add(fgpu_batch *batch, fgpu_sys *sys);
int fgpu_batch_step(fgpu_batch *batch, int iterations, int *n_running);
int fgpu_batch_retire(fgpu_batch *batch, int max, fgpu_sys **out, fgpu_solve_stats *stats, int *n_out);
/* Pipelined steps (no reference counterpart): fgpu_batch_step_async enqueues one step of every
 * member and returns at once; fgpu_batch_wait waits for it and gives the number still running.
 * Between the two, the caller may retire the members a previous waited step finished and add
 * new ones (they join the next step); fgpu_batch_retire then uses the finished flags of the
 * last waited step.  Members that finished keep their results while later steps run (their
 * kernels skip finished systems).  One step in flight at a time (FGPU_EINVAL otherwise). */
int fgpu_batch_step_async(fgpu_batch *batch, int iterations);
int fgpu_batch_wait(fgpu_batch *batch, int *n_running);
/* average duration (ms) of one launch of the element apply (which 0), the dof gather
 * (which 1) or the coarse product of the two-level preconditioner (which 2, mixed batches
 * with a coarse level) over all members, in the batch's precision; for measurement, no
 * counterpart in the reference.  Overwrites the members' slot buffers and coarse corrections. */
int fgpu_batch_time_kernel(fgpu_batch *batch, int which, int reps, double *avg_ms);
/* ---- transfer to atoms (Interpolator, InterpolatorCells, SolutionReader) ---------------
 * An interpolator holds the tet mesh, its Tethex hexes (4 per tet, hex 4t+k around vertex k,
 * femocs/UCD node order) and the nodal solution, 5 doubles per node (Solution: vector x y z,
 * scalar1, scalar2). */
#define FGPU_TYPE_BULK 1
#define FGPU_TYPE_VACUUM 3
/* Interpolator::initialize (src/Interpolator.cpp:28-77) with LinearTetrahedra /
 * QuadraticTetrahedra / LinearHexahedra::precompute (src/InterpolatorCells.cpp:523-629,
 * 761-781, 1151-1174, 1205-1267): node_kind 1 tet vertex, 2 edge, 3 face, 4 tet centre;
 * tet_nbr: 4 face neighbours per tet (-1 none); search_region FGPU_TYPE_VACUUM or
 * FGPU_TYPE_BULK; every node starts at empty_value.  Cells are found with a bounding-volume
 * tree on the device instead of the reference's linear scan, with the same results. */
int fgpu_interp_create(fgpu_ctx *ctx, int n_nodes, const double *xyz, const int *node_kind, int n_tets,
                       const int *tets, const int *tet_nbr, const int *tet_marker, int n_hexes, const int *hexes,
                       const int *hex_marker, int search_region, double empty_value, fgpu_interp **out);
/* the same with flags.  The tables (tet determinants and centroids, neighbour lists, search
 * tree, 10-node tets, node -> hex and pseudo-Voronoi lists) are built on the device by default;
 * FGPU_INTERP_HOST builds them on the host.  Both give the same bits (fgpu_interp_digest).  No
 * counterpart in the reference, whose InterpolatorCells::precompute runs on the host
 * (src/InterpolatorCells.cpp:523-742). */
#define FGPU_INTERP_HOST 1
/* located transfers resolve the chained guess on the host, point after point as the reference
 * does (the candidates still come from the device tree search), instead of the device prefix
 * scan; the same cells either way */
#define FGPU_INTERP_HOST_CHAIN 2
int fgpu_interp_create_ex(fgpu_ctx *ctx, int n_nodes, const double *xyz, const int *node_kind, int n_tets,
                          const int *tets, const int *tet_nbr, const int *tet_marker, int n_hexes, const int *hexes,
                          const int *hex_marker, int search_region, double empty_value, int flags,
                          fgpu_interp **out);
/* 64-bit FNV-1a hashes of every interpolator table (host copy of the tet tables, device
 * tables, the smoothing decay), n >= FGPU_INTERP_DIGEST_TABLES.  No counterpart in the
 * reference. */
#define FGPU_INTERP_DIGEST_TABLES 23
int fgpu_interp_digest(fgpu_interp *interp, int n, unsigned long long *out);
void fgpu_interp_destroy(fgpu_interp *interp);
/* Interpolator::extract_solution(PoissonSolver&, smoothen) (src/Interpolator.cpp:142-190): from
 * a solved system whose mesh was imported with these nodes as vertices: Solution{nodal field
 * (store_elfield), charge density (fgpu_sys_export_charge_density), potential}; nodes outside the
 * region get empty_value; smoothen: average_nodal_fields over the pseudo-Voronoi cells */
int fgpu_interp_extract(fgpu_interp *interp, fgpu_sys *sys, int smoothen);
/* Interpolator::extract_solution(CurrentHeatSolver&) (src/Interpolator.cpp:207-216): with
 * the current and heat systems on one mesh imported from these nodes: Solution{current
 * density as DealSolver::export_solution_grad gives it (src/DealSolver.cpp:281-301), current
 * potential, temperature}; nodes outside the region get {0, 0, empty_value} */
int fgpu_interp_extract_heat(fgpu_interp *interp, fgpu_sys *current, fgpu_sys *heat);
/* the nodal solution (5 * n_nodes), as Interpolator::store_solution / nodes hold it */
int fgpu_interp_get_nodal(fgpu_interp *interp, double *sol5);
/* LinearHexahedra::interp_gradient(point, hex) (src/InterpolatorCells.cpp:407-421, shape_fun_grads
 * :1363-1416): grad = -sum_i grad N_i(x) * scalar2 of node i, for each point in the given hex
 * (femocs hex index: 4t + k), 3 per point; the particle field of the PIC push
 * (src/Pic.cpp:39,204).  FGPU_EINVAL for a hex out of range. */
int fgpu_interp_hex_gradient(fgpu_interp *interp, int n, const double *xyz, const int *hexes, double *grad);
/* Pic::update_point_cell (src/Pic.cpp:185-195): cells[i] holds particle i's solver cell
 * (deal.II index, the region's hexes in hex order) and is replaced by the cell holding xyz[i],
 * found by LinearHexahedra::locate_cell from that guess (src/InterpolatorCells.cpp:1536-1562,
 * 269-307), or -1 when the point is outside every tet of the region.  Each particle uses its
 * own guess (no chain).  FGPU_EINVAL for a cell out of range. */
int fgpu_interp_update_cells(fgpu_interp *interp, int n, const double *xyz, int *cells);
int fgpu_interp_set_nodal(fgpu_interp *interp, const double *sol5);
/* SolutionReader::calc_full_interpolation, dim 3 (src/SolutionReader.cpp:43-165): every point
 * located with the guess chained from the previous one (starting at -1), then interpolated;
 * rank 1 linear tets, 2 quadratic tets, 3 linear hexes.  cells: tet (ranks 1, 2) or hex
 * (rank 3) per point, negative when outside the region (cell of the nearest centroid);
 * out: 5 per point */
int fgpu_interp_points(fgpu_interp *interp, int n, const double *xyz, int rank, int *cells, double *out);
/* SolutionReader::calc_interpolation with the atoms already mapped (cells from a previous
 * fgpu_interp_points call) */
int fgpu_interp_cells(fgpu_interp *interp, int n, const double *xyz, int rank, const int *cells, double *out);
/* Surface faces for dim 2 interpolation: LinearTriangles::precompute on the faces between the
 * vacuum and the bulk (src/InterpolatorCells.cpp:1585-1637; faces and their two tets as
 * TetgenFaces holds them, src/TetgenCells.cpp:387-409).  tris: 3 nodes per face; tri2tet: the
 * face's 2 tets in TetgenFaces::to_tets order (the first is tried first), both tets of the
 * mesh.  quads (may be NULL): 3 quadrangles per face, 4 nodes each, quad 3f+k starting at
 * vertex k of face f (Tethex convert_triangles, src/Tethex.cpp:1104-1170); needed by rank 3
 * (LinearQuadrangles) and the rank 2 centroids (QuadraticTriangles cells,
 * src/InterpolatorCells.cpp:1866-1886).  n_tris = 0 drops the faces. */
int fgpu_interp_set_triangles(fgpu_interp *interp, int n_tris, const int *tris, const int *tri2tet,
                              const int *quads);
/* SolutionReader::calc_full_interpolation, dim 2 (src/SolutionReader.cpp:43-90, 136-163).
 * rank 1 LinearTriangles, rank 2 QuadraticTriangles, rank 3 LinearQuadrangles
 * (src/InterpolatorCells.cpp:1662-1686, 1838-1860, 1925-1992): each point located among the
 * faces (quads for rank 3) with the chained guess (negative: nearest centroid), then
 * interpolated in the tet (hex) the face picks.  centroids != 0 (interp_centroids):
 * locate_interp_centroid (:309-356, 453-461), an exact centroid match (squared distance
 * < 1e-15) giving the mean of the cell's nodes, else -1 in all 5 values with the previous
 * cell kept. */
int fgpu_interp_surface_points(fgpu_interp *interp, int n, const double *xyz, int rank, int centroids, int *cells,
                               double *out);
/* SolutionReader::calc_interpolation, dim 2, atoms already mapped to faces (quads); with
 * centroids every cell must be >= 0 (InterpolatorCells::interp_centroid) */
int fgpu_interp_surface_cells(fgpu_interp *interp, int n, const double *xyz, int rank, int centroids,
                              const int *cells, double *out);

/* HeatSolver::assemble cell terms, implicit Euler (src/CurrentHeatSolver.cpp:116-152, 347-393;
 * PhysicalQuantities::sigma/kappa src/PhysicalQuantities.cpp:31-64, 173-185): with T_old the
 * heat system's current solution and phi_c the solution of `current` on the same mesh, sets the
 * coefficients kappa(T_old) (stiffness) and gamma (mass) at every Gauss point and adds
 * int phi_i (gamma T_old + sigma(T_old) |grad phi_c|^2) to the right-hand side.  Call after
 * fgpu_sys_reset; gamma = 3.4496e-24 / dt [J/(K A^3 s)], lorentz [W Ohm/K^2] (2.44e-8).
 * fgpu_sys_reassemble replays the step: T_old is then the heat system's solution at that time
 * (the last step's temperature) and phi_c the current system's solution at that time; a
 * current system destroyed while a heat system holds the term is released with that term. */
int fgpu_sys_heat_terms(fgpu_sys *heat, const fgpu_sys *current, double gamma, double lorentz);
/* The theta scheme of the same step (extension: the reference's HeatSolver::assemble_crank_nicolson,
 * src/CurrentHeatSolver.cpp:155-272, is disabled): stiffness coefficient theta kappa(T_old), and
 * b -= (1 - theta) K_kappa(T_old) T_old on top of the terms above; theta = 1/2 is Crank-Nicolson,
 * theta = 1 is fgpu_sys_heat_terms.  0 < theta <= 1.  A system holding the explicit term is not
 * replayed by fgpu_sys_reassemble. */
int fgpu_sys_heat_terms_theta(fgpu_sys *heat, const fgpu_sys *current, double gamma, double lorentz, double theta);
/* Temperature-dependent conductivity of the current solve (extension: the reference's current
 * matrix has sigma = 1, src/CurrentHeatSolver.cpp:469): stiffness coefficient sigma0 /
 * (1 + alpha (T - 300 K)) at every Gauss point, T the heat system's solution on the same mesh.
 * Before fgpu_sys_apply_dirichlet. */
int fgpu_sys_sigma_terms(fgpu_sys *current, const fgpu_sys *heat, double sigma0, double alpha);
/* Two-level preconditioner for the FP32 CG of the mixed solver, no counterpart in the
 * reference (which uses SSOR, DealSolver::solve_cg, src/DealSolver.cpp:443-458):
 * M^-1 r = D^-1 r + Z E^-1 Z^T r, Z the indicator of aggregates of consecutive internal dofs
 * (free dofs only), E = Z^T A Z.  enable: 0 removes it; 256, 512 or 1024 sets the aggregate
 * size; 1 picks it by mesh size (256 up to 700k dofs, 512 up to 1.5M, else 1024).  Builds E
 * from the system's current operator and Dirichlet set (call after fgpu_sys_apply_dirichlet)
 * and stores E^-1 in FP32 (nc^2 * 4 bytes, nc = aggregates); the FP64 solver and
 * fgpu_sys_apply ignore it.  E may go stale as the mesh moves or the conditions change: the
 * preconditioner stays symmetric positive definite, so the solve converges to the same
 * tolerance, in more iterations; call again to rebuild.  colors (may be NULL): the number of
 * operator applications the build took. */
int fgpu_sys_set_coarse(fgpu_sys *sys, int enable, int *colors);
/* for checks: nc (0 without a coarse level), the block of every dof (API order, n_dofs ints,
 * may be NULL) and E^-1 (nc*nc floats, row-major, may be NULL) */
int fgpu_sys_get_coarse(fgpu_sys *sys, int *nc, int *block, float *einv);
/* y = A x with the FP32 operator of the mixed solver (inputs rounded to FP32, result widened);
 * for checks, no counterpart in the reference */
int fgpu_sys_apply32(fgpu_sys *sys, const double *x, double *y);
/* FP32 operator of tet-split meshes (every hex the Tethex sub-hex of a tet, as mesh import
 * orients them): a 6-number metric per cell instead of the node coordinates.  Detected per
 * coordinate generation; allow = 0 forces the general trilinear kernel.  *active: whether the
 * mesh uses it now.  No counterpart in the reference. */
int fgpu_mesh_use_affine(fgpu_mesh *mesh, int allow, int *active);
/* DealSolver::check_limits / stat (src/DealSolver.cpp:158-167) */
int fgpu_sys_minmax(fgpu_sys *sys, double *min, double *max);
/* DealSolver::export_solution_grad (src/DealSolver.cpp:281-301): -grad u at the Gauss point
 * vertex2node of the last cell touching each dof; 3*n_dofs */
int fgpu_sys_solution_grad(fgpu_sys *sys, double *grad);

/* Benchmark helper, no counterpart in the reference: average duration in ms of `reps`
 * back-to-back launches of one kernel of the CG step on this system's data (which = 0: the
 * element apply, 1: the dof gather, in the given precision), timed with HIP events on the
 * context stream. */
int fgpu_sys_time_kernel(fgpu_sys *sys, int which, int precision, int reps, double *avg_ms);

/* ---- the reference's tet -> 4-hex split (Tethex Mesh::convert_3D, src/Tethex.cpp:965-1169) ----
 * On the device, with the reference's numbering: nodes = the n_vertices vertices, then edge
 * midpoints (edges ordered by (larger, smaller) vertex, IncidenceMatrix :374-425), face
 * centroids (faces in order of first encounter over tet, local edge, opposite vertex,
 * Mesh::face_numeration :1272-1340), tet centroids (Mesh::set_new_vertices :897-919); hex
 * 4t + k around tet vertex k, femocs (UCD) vertex order, front and back faces swapped where
 * the deal.II-ordered cell would be negative (Mesh::convert_tetrahedra :1038-1100); three
 * quadrangles per boundary triangle (Mesh::convert_triangles :1104-1169).  tets: 4 per tet,
 * tris: 3 per boundary triangle (faces of the tets; may be 0).  At most 2097151 vertices.
 * fgpu_tethex_export writes 3 * n_nodes coordinates, 32 * n_tets hex vertices and
 * 12 * n_tris quadrangle vertices (any pointer may be NULL); the results stay in device memory
 * (about 230 bytes per tet) until fgpu_tethex_destroy. */
int fgpu_tethex_split(fgpu_ctx *ctx, int n_vertices, const double *xyz, int n_tets, const int *tets, int n_tris,
                      const int *tris, fgpu_tethex **out);
int fgpu_tethex_sizes(const fgpu_tethex *split, int *n_nodes, int *n_edges, int *n_faces);
int fgpu_tethex_export(const fgpu_tethex *split, double *xyz, int *hexes, int *quads);
void fgpu_tethex_destroy(fgpu_tethex *split);

/* ---- Q2 Laplace solve --------------------------------------------------------------
 * The reference's Poisson path with the element degree raised to 2: a recompile of
 * include/DealSolver.h:130-131 with shape_degree = 2 (FE_Q(2), QGauss(3) cells, QGauss<2>(3)
 * faces, trilinear geometry) on the hexes of an imported mesh; no such build exists in the
 * reference.  Dofs in the API order: the mesh's dofs (vertices) first, in fgpu_mesh_export_dofs
 * order, then edge midpoints, face centres and cell centres.  The mesh stays alive while a Q2
 * system built on it does. */
typedef struct fgpu_q2 fgpu_q2;
int fgpu_q2_create(fgpu_mesh *mesh, fgpu_q2 **out);
void fgpu_q2_destroy(fgpu_q2 *q2);
int fgpu_q2_sizes(const fgpu_q2 *q2, int *n_dofs, int *n_cells);
/* support points (3*n_dofs: vertices, then the mean of the entity's vertices) and the 27 dofs
 * of every cell (import cell order, lexicographic 3x3x3 nodes i + 3j + 9k); either may be NULL */
int fgpu_q2_export_dofs(const fgpu_q2 *q2, double *xyz, int *cells27);
/* DealSolver::assemble pieces (src/DealSolver.cpp:390-440, src/PoissonSolver.cpp:170-210):
 * clear, Dirichlet value on every dof of the faces with an id, constant Neumann load on the
 * faces with an id (QGauss<2>(3)), the lifting; then the solution from its current value */
int fgpu_q2_reset(fgpu_q2 *q2);
int fgpu_q2_add_dirichlet(fgpu_q2 *q2, int bid, double value);
int fgpu_q2_add_neumann(fgpu_q2 *q2, int bid, double value);
int fgpu_q2_apply_dirichlet(fgpu_q2 *q2);
int fgpu_q2_fill_solution(fgpu_q2 *q2, double value);
/* DealSolver::solve_cg (src/DealSolver.cpp:443-458): FGPU_FP64 only; FGPU_PC_NONE, _JACOBI
 * or _CHEBYSHEV */
int fgpu_q2_solve(fgpu_q2 *q2, const fgpu_solve_opts *opts, fgpu_solve_stats *stats);
int fgpu_q2_get_solution(fgpu_q2 *q2, double *x);
int fgpu_q2_get_rhs(fgpu_q2 *q2, double *b);
/* y = A x with the un-lifted Q2 operator (API dof order) */
int fgpu_q2_apply(fgpu_q2 *q2, const double *x, double *y);
/* average duration of reps element applies (HIP events on the context's stream), benchmark */
int fgpu_q2_time_apply(fgpu_q2 *q2, int reps, double *avg_ms);
/* whether the apply runs the tet-split form (every hex the affine image of the reference
 * sub-hex of a tet split: a per-cell metric instead of the Jacobian at each Gauss point; the
 * mesh's fgpu_mesh_use_affine switch turns it off); no counterpart in the reference */
int fgpu_q2_uses_tet_split(fgpu_q2 *q2, int *on);

#ifdef __cplusplus
}
#endif
#endif

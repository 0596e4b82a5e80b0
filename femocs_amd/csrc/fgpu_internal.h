// Internal structures of libfemocs_gpu: host-side mesh topology, device buffers, and the
// per-system descriptor the batched kernels read.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/femocs_gpu.h"

namespace fgpu {

constexpr int kBlock = 256;   // threads per workgroup (4 wave64)
constexpr int kDofIter = 4;   // dofs per thread in the per-dof kernels
constexpr int kDofBlock = kBlock * kDofIter;
// coarse aggregates of the two-level preconditioner: 2^agg_shift consecutive internal dofs, a
// multiple of kAggMin (one round of a dof workgroup: its sums of r are kept per round)
constexpr int kAggMin = kBlock;
constexpr int kAggPerBlock = kDofIter;
constexpr int kPatchCells = 256;  // largest patch: cells per element workgroup (one per thread)
constexpr int kPatchNodes = 640;  // largest number of distinct nodes a patch may touch
constexpr int kPatchPerBlock = 1; // default patches processed in turn by one element workgroup
// default patch size: kPatchCells, halved (down to 64) while the element launch would have
// fewer than about a thousand workgroups, so that small meshes still spread over the chip
// the eta flip of a lexicographic hex (an involution): vertex j <-> kEtaFlip[j]
constexpr int kEtaFlip[8] = {2, 3, 0, 1, 6, 7, 4, 5};
inline int default_patch_cells(int n_cells) {
    int c = kPatchCells;
    while (c > 64 && (n_cells + c - 1) / c < 1024) c /= 2;
    return c;
}

// Cells are processed in patches of consecutive (space-filling-curve ordered) cells.  A patch
// stages the coordinates and values of its distinct nodes in LDS, applies its cells, sums the
// contributions per node in LDS in a fixed order and writes one output per (node, patch).
struct Patches {
    int n = 0;
    int cells = kPatchCells;      // cells per patch (the last of a run may have fewer)
    int node_cap = 0;             // largest node count of a patch, rounded up to 8
    std::vector<int> cell0;       // n+1: patch p covers internal cells [cell0[p], cell0[p+1])
    std::vector<int> node0;       // n+1: its nodes are nodes[node0[p] .. node0[p+1])
    // the lists below are built on the host, uploaded, then released (the device keeps them)
    std::vector<int> nodes;       // internal dofs, first-appearance order within each patch
    std::vector<int> slot;        // slot-buffer position of every patch node
    std::vector<uint16_t> lidx;   // 8 per cell: patch-local index of each cell vertex
    std::vector<uint16_t> lptr;   // n_p + 1 per patch (at node0[p] + p): CSR over local nodes
    std::vector<uint16_t> lent;   // 8 per cell (at 8 * cell0[p]): entries cl*8+k, ascending
};
constexpr int kScal = 24;     // doubles of per-system scalars

// indices into the per-system scalar block
enum Scal : int {
    S_RZ = 0,      // r.z of the current iterate
    S_RZOLD = 1,   // r.z of the previous iterate (beta denominator)
    S_PQ = 2,      // p.q
    S_ALPHA = 3,
    S_BETA = 4,
    S_RR = 5,      // |r|^2
    S_THR2 = 6,    // stop threshold on |r|^2
    S_IT = 7,      // iterations done
    S_ACTIVE = 8,  // CG: 1 while iterating; FP64 check of the mixed solver: residual requested
    S_MAXIT = 9,
    S_BB = 10,     // |b|^2
    S_CONV = 11,   // 1 once converged
    S_TOL = 12,    // tolerance (absolute, or relative to |b| when S_REL != 0)
    S_REL = 13,
    S_RRREF = 14,  // |r|^2 at the last reliable update (mixed precision)
    S_PEND = 15,   // FP32 CG waits for a reliable update
    S_NUPD = 16,   // reliable updates done
    S_DELTA = 17,  // reliable update when |r|^2 < S_DELTA * S_RRREF
    S_DONE = 18,   // solve finished (converged or out of iterations)
};

// Device descriptor of one system for one launch family.  Pointers are typed by the
// launch's precision (double or float) where marked T.
struct Desc {
    int n_dofs, n_cells, nb_cell, nb_dof;
    const int *cells;             // 8 per cell, lexicographic, internal dof ids
    const void *X;                // T4 coordinates per dof
    const int *gptr;              // dof i owns slot-buffer positions [gptr[i], gptr[i+1])
    int wide_slots;               // dofs own many slots (Q2): each wave gathers its dofs' slot
                                  // range cooperatively, staged through LDS
    int n_patch;                  // patches (see Patches)
    int pcells, node_cap, ppb;    // patch size, LDS node capacity, patches per workgroup
    const int *pcell0, *pnode0, *pnodes, *pslot;
    const uint16_t *plidx, *plptr, *plent;
    const unsigned char *dmask;   // 1 on constrained dofs
    const void *kq, *gq;          // T, 8 per cell (stiffness / mass coefficient) or null
    const void *ua, *ub;          // T: element input u = ua + beta*ub (ub may be null)
    void *Ye;                     // T: element outputs, dof-major (see gptr/spos)
    void *q;                      // T: gathered A u
    void *p;                      // T: search direction (written by the CG gather)
    const void *z;                // T: preconditioned residual
    void *x, *r, *zw;             // T: iterate, residual, preconditioned residual (written)
    const void *invdiag;          // T: 1/diag, 0 on constrained dofs
    const double *b;              // FP64 rhs (residual mode)
    double *scal;                 // kScal doubles
    double *part;                 // kPartRegions regions of nb_dof per-block partial sums
    const float *metric;          // FP32 apply of tet-split meshes: 6 per cell, then |det A| per cell, or null
    // two-level preconditioner of the FP32 CG (null: Jacobi only): coarse correction c per
    // dof block (T), nc x nc inverse of the coarse operator (row-major), nc = dof blocks
    void *cc;
    const _Float16 *cinv;         // FP16, scaled: the inverse is cinv x cscale; rows of cld
    double cscale;                // a power of two
    int cld;                      // row stride of cinv: nc rounded up to 8 (zero padding)
    int nc, agg_shift;            // aggregates and log2 of their size
    int cc_in;                    // the element input is z: add its coarse part (CG sets only;
                                  // a residual set's input is the iterate)
    // preconditioner of this descriptor's CG: 0 identity, 1 Jacobi, 2 Chebyshev-Jacobi
    int pc, cheb_deg;
    double *hist;                 // alpha_j, beta_j of the first kHist iterations, or null
    double *cd, *cres;            // FP64 Chebyshev search vector and residual
    const double *ccoef;          // 1/theta, then (c1_j, c2_j) for j = 1..cheb_deg-1
};
// Element-kernel work item: one workgroup, starting at patch p of system sys, with that
// patch's first cell and node and its cell and node counts (nc = 0: nothing to do), so the
// first patch needs no table lookups before its cell and node lists are read.
struct ETab {
    int sys, p, c0, n0, nc, np, pad0, pad1;
};
constexpr int kHist = 32;         // CG coefficients recorded per system for the Lanczos estimate
constexpr int kChebMax = 12;      // largest Chebyshev degree
constexpr int kPartRegions = 5 + kAggPerBlock;   // per descriptor set: p.q, r.z, r.r, b.b, s.c,
                                                 // aggregate sums of r (kAggPerBlock regions)
constexpr int kCoarseRows = 16;   // coarse rows per workgroup of the coarse product
constexpr int kCoarseMax = 8000;  // aggregates per system (their sums are staged in LDS)

struct Mesh {
    fgpu_ctx *ctx = nullptr;
    int refs = 0;              // live systems on this mesh
    bool closing = false;      // destroy requested while systems were alive
    int n_vertices = 0, n_dofs = 0, n_cells = 0;
    // host topology in the API numbering
    std::vector<int> vertex_of_dof;        // n_dofs
    std::vector<int> dof_of_vertex;        // n_vertices, -1 unused
    std::vector<int> cells_lex;            // 8*n_cells, API dofs, lexicographic
    std::vector<double> x;                 // 3*n_dofs (API order)
    std::vector<int> bf_cell, bf_face, bf_id;
    // internal numbering: cells sorted along a space-filling curve, dofs by first touch
    std::vector<int> icell_of_cell, cell_of_icell;
    std::vector<int> idof_of_dof, dof_of_idof;
    std::vector<int> v2c_cell, v2c_node;   // last cell (API) touching each dof and its node
    Patches patches;
    int n_slots = 0;                       // slot-buffer entries (sum of patch node counts)
    int patch_gen = 0;                     // bumped whenever the patches are rebuilt
    int coord_gen = 0;                     // bumped whenever the vertices move
    unsigned long long uid = 0;            // unique per imported mesh (caches keyed on it)
    // every cell an affine image of the tet-split reference sub-hex: the FP32 operator reads
    // a per-cell metric instead of the node coordinates (valid for coordinate generation
    // metric_gen)
    float *d_metric = nullptr;             // 6 per cell, then |det A| per cell (internal order)
    int metric_gen = -1;
    bool affine = false;
    bool affine_allowed = true;
    bool x_stale = false;                  // device coordinates moved, host copy `x` not yet
    // linear mesh motion (fgpu_mesh_set_motion): row r (dof idof[r]) sits at base[r] +
    // sum_k w[k] * disp[ctrl[k]] for k in ptr[r]..ptr[r+1]
    struct Motion {
        int rows = 0, n_ctrl = 0;
        int *idof = nullptr, *ptr = nullptr, *ctrl = nullptr;
        double *w = nullptr, *base = nullptr, *disp = nullptr, *prev = nullptr;
        bool have_prev = false;
    } motion;
    std::vector<int> icells;               // 8*n_cells, internal order and dofs (host copy)
    // per internal cell: 1 when its lexicographic vertices hold the tet-split sub-hex with the
    // eta flip (Tethex's orientation fix swaps its front and back faces), 0 otherwise
    std::vector<unsigned char> layout;
    unsigned char *d_layout = nullptr;
    // device arrays (internal numbering)
    int *d_cells = nullptr;                // 8*n_cells
    double *d_X64 = nullptr;               // 4*n_dofs
    float *d_X32 = nullptr;                // 4*n_dofs
    int *d_gptr = nullptr;
    int *d_pcell0 = nullptr, *d_pnode0 = nullptr, *d_pnodes = nullptr, *d_pslot = nullptr;
    uint16_t *d_plidx = nullptr, *d_plptr = nullptr, *d_plent = nullptr;
    int *d_bf_icell = nullptr, *d_bf_face = nullptr;       // boundary faces (API order)
    int *d_bf_id = nullptr;                                  // boundary id per face
    int *d_fgptr = nullptr, *d_fgidx = nullptr;            // dof -> boundary-face slots
    int *d_v2c = nullptr, *d_v2n = nullptr;                 // gradient export (internal)
    ~Mesh();
};

struct Sys {
    Mesh *mesh = nullptr;
    fgpu_ctx *ctx = nullptr;
    // host-side boundary data (API dof order)
    std::vector<unsigned char> dmask;
    std::vector<double> gval;
    bool have_coef = false;
    bool diag_valid = false;
    // FP64 arrays
    double *b = nullptr, *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
    double *Ye = nullptr, *invd = nullptr, *diag = nullptr, *g = nullptr, *tmp = nullptr;
    double *kq = nullptr, *gq = nullptr, *fslot = nullptr, *fval = nullptr;
    unsigned char *d_dmask = nullptr;
    // FP32 arrays (mixed precision inner solve)
    float *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *q32 = nullptr, *d32 = nullptr;
    float *Ye32 = nullptr, *invd32 = nullptr, *kq32 = nullptr, *gq32 = nullptr;
    // scalars and reduction scratch: [0] FP64 CG / outer, [1] FP32 inner
    double *scal = nullptr;       // 2*kScal
    double *part = nullptr;       // 2 * kPartRegions*nb_dof
    int nb_cell = 0, nb_dof = 0;
    int ppb = kPatchPerBlock;     // patches per element workgroup (tunable)
    int patch_gen = -1;           // mesh patch generation the slot buffers are sized for
    int diag_gen = -1;            // mesh coordinate generation the diagonal was computed for
    // right-hand-side terms since the last reset, replayed by fgpu_sys_reassemble after the
    // vertices moved: Neumann face values (one per boundary face) or cell values (8 per cell)
    struct RhsTerm {
        int kind;                 // 0: Neumann on boundary id `bid`, 1: cell source, 2: heat terms
        int bid;
        double *dval;             // face values (kind 0) or cell source, 8 per cell (kinds 1, 2)
        // kind 2 (implicit-Euler heat step): coefficients and Joule source recomputed from the
        // system's own solution (the previous temperature) and the current potential of `src`
        Sys *src = nullptr;
        double gamma = 0, lorentz = 0;
    };
    std::vector<RhsTerm> rhs_terms;
    bool lifted = false;          // fgpu_sys_apply_dirichlet ran since the last reset
    bool g_zero = false;          // every Dirichlet value is +0.0 (the lifting's A g vanishes)
    // charge density of a write step (fgpu_sys_save_charge_density), internal order
    double *rho = nullptr;
    bool rho_saved = false;
    // a right-hand-side term since the last reset that fgpu_sys_reassemble cannot replay
    // (particle deposits, the explicit stiffness term of a theta step): the reason, or null
    const char *no_replay = nullptr;
    // two-level preconditioner of the mixed solver (fgpu_sys_set_coarse)
    bool coarse = false;
    int batches = 0;              // running batches holding the system (their tables point here)
    int nc = 0;                   // coarse size: aggregates of 2^agg_shift internal dofs
    int agg_shift = 10;
    // nc x nc, FP16 scaled by cscale (a power of two: the largest entry, on the diagonal of
    // the SPD inverse, stays below 2^14): the coarse product reads half the bytes of FP32
    _Float16 *cinv = nullptr;     // nc rows of cld entries (zero past nc)
    double cscale = 1.0;
    int cld = 0;
    float *cc = nullptr;          // nc
    bool desc32_affine = false;   // the last FP32 descriptor built uses the tet-split form
    // replayable heat terms of other systems that read this one's solution: while any is
    // recorded, fgpu_sys_destroy only marks the system and the last such term frees it
    int term_refs = 0;
    bool closing = false;
    // two-level preconditioner in the FP64 CG (coarse correction in FP64, same E^-1)
    double *cc64 = nullptr;
    // Chebyshev-Jacobi preconditioner of the FP64 CG: search vectors d and the Chebyshev
    // residual, coefficients [1/theta, (c1_j, c2_j) for j = 1..degree-1] on the device, the
    // estimate of lambda_max(D^-1 A) and the coordinate generation it belongs to
    double *cd = nullptr, *cres = nullptr, *ccoef = nullptr;
    double *hist = nullptr;       // CG alpha_j, beta_j of the first kHist iterations (Lanczos)
    double cheb_lmax = 0;
    int cheb_gen = -1;
    void clear_terms();
    ~Sys();
};

}  // namespace fgpu

struct fgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    double *scratch = nullptr;    // device address of the library's scratch array on `device`
    std::string err;
    // batch tables reused between solves (grown on demand)
    void *d_tab = nullptr;
    size_t tab_bytes = 0;
    double *h_flags = nullptr;    // pinned, per system
    size_t flags_cap = 0;
    int refs = 0;                 // live meshes and systems
    bool closing = false;         // destroy requested while children were alive
    bool affine_table = false;    // reference sub-hex constants uploaded to this device
    // pinned staging buffers of asynchronous uploads from short-lived host data, each reused
    // once the copy that last read it has completed (its event)
    static constexpr int kStage = 4;
    void *stage_host[kStage] = {};
    size_t stage_cap[kStage] = {};
    hipEvent_t stage_ev[kStage] = {};
    bool stage_busy[kStage] = {};
    int stage_next = 0;
};

namespace fgpu {
// Temporaries of one device build (import, tet split, interpolator tables): a bump allocator
// over a few large device blocks freed together (one allocation instead of dozens; each
// hipFree costs a fraction of a millisecond), plus one reusable region for the temporary
// storage of the stream-ordered hipcub sorts and scans (each call may reuse what the previous
// one used).  Fail::raise(hipError_t, const char *) reports a failed allocation.
template <class Fail> struct DeviceArena {
    std::vector<void *> blocks;
    char *cur = nullptr;
    size_t left = 0, chunk;
    void *scr = nullptr;
    size_t scr_bytes = 0;
    explicit DeviceArena(size_t chunk_bytes) : chunk(chunk_bytes) {}
    void *alloc(size_t b) {
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, b);
        if (e != hipSuccess) Fail::raise(e, "hipMalloc (device arena block)");
        blocks.push_back(q);
        return q;
    }
    template <class T> T *get(size_t n) {
        const size_t bytes = (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~(size_t)255;
        if (bytes > left) {
            const size_t b = std::max(bytes, blocks.empty() ? chunk : chunk / 4);
            cur = static_cast<char *>(alloc(b));
            left = b;
        }
        T *r = reinterpret_cast<T *>(cur);
        cur += bytes;
        left -= bytes;
        return r;
    }
    // hipcub temporary storage of at least `bytes` (valid until the next call)
    void *scratch(size_t bytes) {
        if (bytes > scr_bytes) {
            scr_bytes = std::max(bytes, 2 * scr_bytes);
            scr = alloc(scr_bytes);
        }
        return scr;
    }
    ~DeviceArena() {
        for (void *q : blocks) (void)hipFree(q);
    }
};
}  // namespace fgpu

struct fgpu_mesh : fgpu::Mesh {};
struct fgpu_sys : fgpu::Sys {};

// drops a reference on the context (frees it when closed and unreferenced)
void fgpu_ctx_release_ref(fgpu_ctx *c);

namespace fgpu {
// mesh topology built on the host (fgpu_mesh.cpp)
int build_topology(Mesh &m, int nv, const double *xyz, int nc, const int *ucd, int top, int bottom,
                   int sides, int other, std::string &err);
// internal order and gather lists
void build_internal_order(Mesh &m, std::vector<int> &icells, std::vector<double> &X,
                          std::vector<int> &gptr, std::vector<int> &spos, std::vector<int> &fgptr,
                          std::vector<int> &fgidx, std::vector<int> &bf_icell);
void build_patches(int nc, int nd, const std::vector<int> &icells, int cells_per_patch, Patches &P);
// dof-major slot layout of the patch outputs: gptr (n_dofs+1) and Patches::slot
void build_slots(int nd, Patches &P, std::vector<int> &gptr);
void cell_layouts(int nc, const std::vector<int> &icells, const std::vector<double> &X, std::vector<unsigned char> &layout);
// the same import on the device (fgpu_import.hip): fills every host table above and the mesh's
// device arrays (cells, coordinates, patches, slots, boundary faces, gradient export, layouts);
// nv > 0 and nc > 0
int device_import(Mesh &m, hipStream_t st, int nv, const double *xyz, int nc, const int *ucd, int top, int bottom,
                  int sides, int other, std::string &err);
}  // namespace fgpu

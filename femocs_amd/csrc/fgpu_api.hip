// C ABI of libfemocs_gpu and the batched CG drivers.
#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "fgpu_internal.h"
#include "fgpu_kernels.h"
#include "fgpu_cheb.h"

using namespace fgpu;

namespace {

// error plumbing: every entry point catches and converts to a status code
struct HipError {
    hipError_t e;
    const char *what;
};

#define HIPCHECK(call)                                        \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) throw HipError{e_, #call};      \
    } while (0)

struct MeshError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <class F> int guarded(fgpu_ctx *ctx, F &&f) {
    try {
        return f();
    } catch (const HipError &h) {
        if (ctx) {
            ctx->err = std::string(h.what) + ": " + hipGetErrorString(h.e);
        }
        return h.e == hipErrorOutOfMemory ? FGPU_ENOMEM : FGPU_EHIP;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->err = "host allocation failed";
        return FGPU_ENOMEM;
    } catch (const MeshError &ex) {
        if (ctx) ctx->err = ex.what();
        return FGPU_EMESH;
    } catch (const std::exception &ex) {
        if (ctx) ctx->err = ex.what();
        return FGPU_EINVAL;
    }
}

template <class T> T *dalloc(size_t n) {
    void *p = nullptr;
    HIPCHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return static_cast<T *>(p);
}
template <class T> void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}
template <class T> void h2d(T *dst, const T *src, size_t n, hipStream_t s) {
    if (n) HIPCHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <class T> void d2h(T *dst, const T *src, size_t n, hipStream_t s) {
    if (n) HIPCHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

constexpr int kXcd = 8;  // MI355X: workgroups are dealt round-robin over 8 XCDs

int padded_blocks(int n, int per_block) {
    const int nb = (n + per_block - 1) / per_block;
    return ((std::max(nb, 1) + kXcd - 1) / kXcd) * kXcd;
}

// Block table of one system: position i (relative to the system's base, a multiple of 8)
// runs local block (i % 8) * (nb / 8) + i / 8, so each XCD gets one contiguous slice of the
// system's cells or dofs and its L2 sees neighbouring data.
void append_table(std::vector<int2> &tab, int sys, int nb) {
    const int per = nb / kXcd;
    for (int i = 0; i < nb; ++i) tab.push_back(make_int2(sys, (i % kXcd) * per + i / kXcd));
}

}  // namespace

fgpu::Mesh::~Mesh() {
    dfree(d_cells);
    dfree(d_X64);
    dfree(d_X32);
    dfree(d_gptr);
    dfree(d_pcell0);
    dfree(d_pnode0);
    dfree(d_pnodes);
    dfree(d_pslot);
    dfree(d_plidx);
    dfree(d_plptr);
    dfree(d_plent);
    dfree(d_layout);
    dfree(d_bf_icell);
    dfree(d_bf_face);
    dfree(d_bf_id);
    dfree(d_metric);
    dfree(motion.idof); dfree(motion.ptr); dfree(motion.ctrl); dfree(motion.w);
    dfree(motion.base); dfree(motion.disp); dfree(motion.prev);
    dfree(d_fgptr);
    dfree(d_fgidx);
    dfree(d_v2c);
    dfree(d_v2n);
}

void fgpu::Sys::clear_terms() {
    for (auto &t : rhs_terms) dfree(t.dval);
    rhs_terms.clear();
    lifted = false;
    rho_saved = false;
    no_replay = nullptr;
}

fgpu::Sys::~Sys() {
    clear_terms();
    dfree(b); dfree(x); dfree(r); dfree(z); dfree(p); dfree(q); dfree(Ye); dfree(invd); dfree(diag);
    dfree(g); dfree(tmp); dfree(kq); dfree(gq); dfree(fslot); dfree(fval); dfree(d_dmask);
    dfree(r32); dfree(z32); dfree(p32); dfree(q32); dfree(d32); dfree(Ye32); dfree(invd32);
    dfree(kq32); dfree(gq32); dfree(scal); dfree(part); dfree(cinv); dfree(cc); dfree(rho);
    dfree(cc64); dfree(cd); dfree(cres); dfree(ccoef); dfree(hist);
}

// ------------------------------------------------------------------------------------------
// scratch of the no-allocation entry points: one array per GPU, in the library's code object.
// Its address on a device comes from hipGetSymbolAddress with that device current (the name
// itself is only a host-side stand-in); every call on a device shares that device's array.
// ------------------------------------------------------------------------------------------
constexpr int kScratchMinBlocks = 65536;   // per-block minima of fgpu_mesh_displace_noalloc
constexpr int kScratchDoubles = kScratchMinBlocks + 64;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// the device address of `scratch` on the current device `device` (once per device)
static double *scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress gave no address for the scratch array");
        g_scratch[device] = static_cast<double *>(p);
    }
    return g_scratch[device];
}

// ------------------------------------------------------------------------------------------
// context
// ------------------------------------------------------------------------------------------
extern "C" int fgpu_ctx_create(int device, fgpu_ctx **out) {
    if (!out) return FGPU_EINVAL;
    *out = nullptr;
    fgpu_ctx *c = new (std::nothrow) fgpu_ctx();
    if (!c) return FGPU_ENOMEM;
    const int rc = guarded(c, [&] {
        int n = 0;
        HIPCHECK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw std::invalid_argument("no such device");
        c->device = device;
        HIPCHECK(hipSetDevice(device));
        HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->scratch = scratch_ptr(device);
        return FGPU_OK;
    });
    if (rc != FGPU_OK) {
        fprintf(stderr, "fgpu_ctx_create: %s\n", c->err.c_str());
        delete c;
        return rc;
    }
    *out = c;
    return FGPU_OK;
}

static void ctx_free(fgpu_ctx *c) {
    (void)hipSetDevice(c->device);
    if (c->stream) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
    }
    if (c->d_tab) (void)hipFree(c->d_tab);
    if (c->h_flags) (void)hipHostFree(c->h_flags);
    for (int k = 0; k < fgpu_ctx::kStage; ++k) {
        if (c->stage_host[k]) (void)hipHostFree(c->stage_host[k]);
        if (c->stage_ev[k]) (void)hipEventDestroy(c->stage_ev[k]);
    }
    delete c;
}

// A context outlives its meshes and systems: destroying it while they are alive only marks
// it, and the last child to go frees it.  Same for a mesh and its systems.
extern "C" void fgpu_ctx_destroy(fgpu_ctx *c) {
    if (!c) return;
    if (c->refs > 0) {
        c->closing = true;
        return;
    }
    ctx_free(c);
}

static void ctx_release(fgpu_ctx *c) {
    if (--c->refs == 0 && c->closing) ctx_free(c);
}

// for the other translation units holding a context reference
void fgpu_ctx_release_ref(fgpu_ctx *c) { ctx_release(c); }

static void mesh_free(fgpu_mesh *m);
// for the other translation units holding a mesh reference (a live Q2 system)
void fgpu_mesh_release_ref(fgpu_mesh *m) {
    if (--m->refs == 0 && m->closing) mesh_free(m);
}

extern "C" const char *fgpu_ctx_error(const fgpu_ctx *c) { return c ? c->err.c_str() : "null context"; }

// device memory for the caller (e.g. checkpoints of solutions), freed with fgpu_ctx_free
extern "C" int fgpu_ctx_alloc(fgpu_ctx *c, size_t bytes, void **dev) {
    if (!c || !dev) return FGPU_EINVAL;
    *dev = nullptr;
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        HIPCHECK(hipMalloc(dev, std::max<size_t>(bytes, 1)));
        return FGPU_OK;
    });
}

extern "C" int fgpu_ctx_free(fgpu_ctx *c, void *dev) {
    if (!c) return FGPU_EINVAL;
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (dev) HIPCHECK(hipFree(dev));
        return FGPU_OK;
    });
}

extern "C" int fgpu_ctx_sync(fgpu_ctx *c) {
    if (!c) return FGPU_EINVAL;
    return guarded(c, [&] {
        HIPCHECK(hipStreamSynchronize(c->stream));
        return FGPU_OK;
    });
}

// ------------------------------------------------------------------------------------------
// mesh
// ------------------------------------------------------------------------------------------
static void upload_coords(fgpu_mesh *m, const std::vector<double> &X) {
    hipStream_t s = m->ctx->stream;
    h2d(m->d_X64, X.data(), X.size(), s);
    std::vector<float> X32(X.size());
    for (size_t i = 0; i < X.size(); ++i) X32[i] = (float)X[i];
    h2d(m->d_X32, X32.data(), X32.size(), s);
    HIPCHECK(hipStreamSynchronize(s));
}

// (re)uploads the patch structures and the dof-major slot layout
static void upload_patches(fgpu_mesh *m, const std::vector<int> &gptr) {
    hipStream_t s = m->ctx->stream;
    HIPCHECK(hipStreamSynchronize(s));
    dfree(m->d_gptr); dfree(m->d_pcell0); dfree(m->d_pnode0); dfree(m->d_pnodes); dfree(m->d_pslot);
    dfree(m->d_plidx); dfree(m->d_plptr); dfree(m->d_plent);
    const Patches &P = m->patches;
    m->n_slots = P.node0.back();
    m->d_gptr = dalloc<int>(gptr.size());
    h2d(m->d_gptr, gptr.data(), gptr.size(), s);
    m->d_pcell0 = dalloc<int>(P.cell0.size());
    h2d(m->d_pcell0, P.cell0.data(), P.cell0.size(), s);
    m->d_pnode0 = dalloc<int>(P.node0.size());
    h2d(m->d_pnode0, P.node0.data(), P.node0.size(), s);
    m->d_pnodes = dalloc<int>(P.nodes.size());
    h2d(m->d_pnodes, P.nodes.data(), P.nodes.size(), s);
    m->d_pslot = dalloc<int>(P.slot.size());
    h2d(m->d_pslot, P.slot.data(), P.slot.size(), s);
    // bit 15 of a cell's first local index carries its sub-hex layout (local indices < 2^15)
    std::vector<uint16_t> lidx(P.lidx);
    for (size_t c = 0; c < m->layout.size() && 8 * c < lidx.size(); ++c)
        if (m->layout[c]) lidx[8 * c] |= 0x8000;
    m->d_plidx = dalloc<uint16_t>(lidx.size());
    h2d(m->d_plidx, lidx.data(), lidx.size(), s);
    m->d_plptr = dalloc<uint16_t>(P.lptr.size());
    h2d(m->d_plptr, P.lptr.data(), P.lptr.size(), s);
    m->d_plent = dalloc<uint16_t>(P.lent.size());
    h2d(m->d_plent, P.lent.data(), P.lent.size(), s);
    HIPCHECK(hipStreamSynchronize(s));
    m->patch_gen++;
    // the device holds the lists now (as after a device import): keep only the patch bounds
    Patches &Q = m->patches;
    std::vector<int>().swap(Q.nodes);
    std::vector<int>().swap(Q.slot);
    std::vector<uint16_t>().swap(Q.lidx);
    std::vector<uint16_t>().swap(Q.lptr);
    std::vector<uint16_t>().swap(Q.lent);
}

extern "C" int fgpu_mesh_import(fgpu_ctx *ctx, int nv, const double *xyz, int nc, const int *ucd, int top, int bottom,
                                int sides, int other, fgpu_mesh **out) {
    return fgpu_mesh_import_ex(ctx, nv, xyz, nc, ucd, top, bottom, sides, other, 0, out);
}

extern "C" int fgpu_mesh_import_ex(fgpu_ctx *ctx, int nv, const double *xyz, int nc, const int *ucd, int top, int bottom,
                                   int sides, int other, int flags, fgpu_mesh **out) {
    if (!ctx || !out || (flags & ~FGPU_IMPORT_HOST)) return FGPU_EINVAL;
    *out = nullptr;
    fgpu_mesh *m = new (std::nothrow) fgpu_mesh();
    if (!m) return FGPU_ENOMEM;
    m->ctx = ctx;
    static std::atomic<unsigned long long> next_uid{1};
    // the device builds the tables of any non-empty mesh; the host path gives the same bits
    const bool on_device = !(flags & FGPU_IMPORT_HOST) && nv > 0 && nc > 0 && xyz && ucd;
    const int rc = guarded(ctx, [&] {
        HIPCHECK(hipSetDevice(ctx->device));
        if (on_device) {
            const int r = device_import(*m, ctx->stream, nv, xyz, nc, ucd, top, bottom, sides, other, ctx->err);
            if (r != FGPU_OK) return r;
            m->uid = next_uid++;
            return FGPU_OK;
        }
        const int r = build_topology(*m, nv, xyz, nc, ucd, top, bottom, sides, other, ctx->err);
        if (r != FGPU_OK) return r;
        if (m->n_cells == 0 || m->n_dofs == 0) {
            ctx->err = "mesh import: empty mesh";
            return FGPU_EINVAL;
        }
        std::vector<int> icells, gptr, spos, fgptr, fgidx, bf_icell;
        std::vector<double> X;
        build_internal_order(*m, icells, X, gptr, spos, fgptr, fgidx, bf_icell);
        cell_layouts(m->n_cells, icells, X, m->layout);
        hipStream_t s = ctx->stream;
        m->d_layout = dalloc<unsigned char>(m->layout.size());
        h2d(m->d_layout, m->layout.data(), m->layout.size(), s);
        m->d_cells = dalloc<int>(icells.size());
        h2d(m->d_cells, icells.data(), icells.size(), s);
        HIPCHECK(hipStreamSynchronize(s));
        m->uid = next_uid++;
        m->d_X64 = dalloc<double>(X.size());
        m->d_X32 = dalloc<float>(X.size());
        upload_patches(m, gptr);
        m->icells = std::move(icells);
        const size_t nb = m->bf_cell.size();
        m->d_bf_icell = dalloc<int>(nb);
        h2d(m->d_bf_icell, bf_icell.data(), nb, s);
        m->d_bf_face = dalloc<int>(nb);
        h2d(m->d_bf_face, m->bf_face.data(), nb, s);
        m->d_bf_id = dalloc<int>(nb);
        h2d(m->d_bf_id, m->bf_id.data(), nb, s);
        m->d_fgptr = dalloc<int>(fgptr.size());
        h2d(m->d_fgptr, fgptr.data(), fgptr.size(), s);
        m->d_fgidx = dalloc<int>(fgidx.size());
        h2d(m->d_fgidx, fgidx.data(), fgidx.size(), s);
        // gradient export: last cell (API order) touching each dof, in internal numbering
        std::vector<int> v2c(m->n_dofs), v2n(m->n_dofs);
        for (int i = 0; i < m->n_dofs; ++i) {
            const int d = m->dof_of_idof[i];
            v2c[i] = m->icell_of_cell[m->v2c_cell[d]];
            v2n[i] = m->v2c_node[d];
        }
        std::vector<int>().swap(m->v2c_cell);   // read only here
        std::vector<int>().swap(m->v2c_node);
        m->d_v2c = dalloc<int>(m->n_dofs);
        h2d(m->d_v2c, v2c.data(), v2c.size(), s);
        m->d_v2n = dalloc<int>(m->n_dofs);
        h2d(m->d_v2n, v2n.data(), v2n.size(), s);
        upload_coords(m, X);
        return FGPU_OK;
    });
    if (rc != FGPU_OK) {
        delete m;
        return rc;
    }
    ctx->refs++;
    *out = m;
    return FGPU_OK;
}

static void mesh_free(fgpu_mesh *m) {
    fgpu_ctx *c = m->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete m;
    ctx_release(c);
}

extern "C" void fgpu_mesh_destroy(fgpu_mesh *m) {
    if (!m) return;
    if (m->refs > 0) {
        m->closing = true;
        return;
    }
    mesh_free(m);
}

extern "C" int fgpu_mesh_sizes(const fgpu_mesh *m, int *nd, int *nc, int *nb) {
    if (!m) return FGPU_EINVAL;
    if (nd) *nd = m->n_dofs;
    if (nc) *nc = m->n_cells;
    if (nb) *nb = (int)m->bf_cell.size();
    return FGPU_OK;
}

// the host coordinate copy after device-side motion (fgpu_mesh_displace)
static void refresh_host_x(const fgpu_mesh *cm) {
    if (!cm->x_stale) return;
    fgpu_mesh *m = const_cast<fgpu_mesh *>(cm);
    std::vector<double> X(4 * (size_t)m->n_dofs);
    d2h(X.data(), m->d_X64, X.size(), m->ctx->stream);
    HIPCHECK(hipStreamSynchronize(m->ctx->stream));
    for (int i = 0; i < m->n_dofs; ++i)
        for (int a = 0; a < 3; ++a) m->x[3 * (size_t)m->dof_of_idof[i] + a] = X[4 * (size_t)i + a];
    m->x_stale = false;
}

extern "C" int fgpu_mesh_export_dofs(const fgpu_mesh *m, double *xyz, int *vod) {
    if (!m) return FGPU_EINVAL;
    if (xyz) {
        const int rc = guarded(m->ctx, [&] {
            HIPCHECK(hipSetDevice(m->ctx->device));
            refresh_host_x(m);
            return FGPU_OK;
        });
        if (rc != FGPU_OK) return rc;
        std::memcpy(xyz, m->x.data(), m->x.size() * sizeof(double));
    }
    if (vod) std::memcpy(vod, m->vertex_of_dof.data(), m->vertex_of_dof.size() * sizeof(int));
    return FGPU_OK;
}

extern "C" int fgpu_mesh_export_bfaces(const fgpu_mesh *m, int *cell, int *face, int *bid) {
    if (!m) return FGPU_EINVAL;
    const size_t nb = m->bf_cell.size();
    if (cell) std::memcpy(cell, m->bf_cell.data(), nb * sizeof(int));
    if (face) std::memcpy(face, m->bf_face.data(), nb * sizeof(int));
    if (bid) std::memcpy(bid, m->bf_id.data(), nb * sizeof(int));
    return FGPU_OK;
}

extern "C" int fgpu_mesh_export_face_centroids(const fgpu_mesh *m, int bid, double *xyz, int *n) {
    static const int kFV[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
    if (!m || !n) return FGPU_EINVAL;
    if (xyz) {
        const int rc = guarded(m->ctx, [&] {
            HIPCHECK(hipSetDevice(m->ctx->device));
            refresh_host_x(m);
            return FGPU_OK;
        });
        if (rc != FGPU_OK) return rc;
    }
    int k = 0;
    for (size_t f = 0; f < m->bf_cell.size(); ++f) {
        if (m->bf_id[f] != bid) continue;
        if (xyz)
            for (int a = 0; a < 3; ++a) {
                double s = 0;
                for (int j = 0; j < 4; ++j) s += m->x[3 * (size_t)m->cells_lex[8 * (size_t)m->bf_cell[f] + kFV[m->bf_face[f]][j]] + a];
                xyz[3 * (size_t)k + a] = s / 4.0;
            }
        ++k;
    }
    *n = k;
    return FGPU_OK;
}

static double min_center_det(fgpu_mesh *m);

extern "C" int fgpu_mesh_update_vertices(fgpu_mesh *m, const double *xyz) {
    if (!m || !xyz) return FGPU_EINVAL;
    return guarded(m->ctx, [&] {
        HIPCHECK(hipSetDevice(m->ctx->device));
        refresh_host_x(m);
        std::vector<double> X(4 * (size_t)m->n_dofs, 0.0);
        for (int i = 0; i < m->n_dofs; ++i) {
            const size_t v = m->vertex_of_dof[m->dof_of_idof[i]];
            for (int a = 0; a < 3; ++a) X[4 * (size_t)i + a] = xyz[3 * v + a];
        }
        h2d(m->d_X64, X.data(), X.size(), m->ctx->stream);
        // the import fixed each cell's orientation by the sign of its centre Jacobian; a move
        // that flips one would change the operator the import semantics define
        const double mn = min_center_det(m);
        if (!(mn >= 0.0)) {
            std::vector<double> Xold(4 * (size_t)m->n_dofs, 0.0);
            for (int i = 0; i < m->n_dofs; ++i)
                for (int a = 0; a < 3; ++a) Xold[4 * (size_t)i + a] = m->x[3 * (size_t)m->dof_of_idof[i] + a];
            upload_coords(m, Xold);
            throw MeshError("moved vertices turn a cell inside out (centre Jacobian " + std::to_string(mn) +
                            "); coordinates left unchanged, the mesh must be rebuilt");
        }
        for (int d = 0; d < m->n_dofs; ++d)
            for (int a = 0; a < 3; ++a) m->x[3 * (size_t)d + a] = xyz[3 * (size_t)m->vertex_of_dof[d] + a];
        upload_coords(m, X);
        m->coord_gen++;
        return FGPU_OK;
    });
}

// smallest centre Jacobian over the cells at the current device coordinates
static double min_center_det(fgpu_mesh *m) {
    const int nb = center_det_blocks(m->n_cells);
    double *dmin = dalloc<double>(nb);
    launch_center_det_min(m->ctx->stream, m->n_cells, m->d_cells, m->d_X64, dmin);
    std::vector<double> bmin(nb);
    d2h(bmin.data(), dmin, nb, m->ctx->stream);
    HIPCHECK(hipStreamSynchronize(m->ctx->stream));
    dfree(dmin);
    double mn = 1e300;
    for (double v : bmin) mn = std::min(mn, v);
    return mn;
}

extern "C" int fgpu_mesh_move_vertices(fgpu_mesh *m, int n, const int *ids, const double *xyz) {
    if (!m || n < 0 || (n > 0 && (!ids || !xyz))) return FGPU_EINVAL;
    return guarded(m->ctx, [&] {
        HIPCHECK(hipSetDevice(m->ctx->device));
        hipStream_t st = m->ctx->stream;
        refresh_host_x(m);
        std::vector<int> idx;
        std::vector<double> nx, ox;
        idx.reserve(n);
        for (int k = 0; k < n; ++k) {
            if (ids[k] < 0 || ids[k] >= m->n_vertices) throw std::invalid_argument("vertex id out of range");
            const int d = m->dof_of_vertex[ids[k]];
            if (d < 0) continue;                          // vertex not used by any cell
            idx.push_back(m->idof_of_dof[d]);
            for (int a = 0; a < 3; ++a) {
                nx.push_back(xyz[3 * (size_t)k + a]);
                ox.push_back(m->x[3 * (size_t)d + a]);
            }
        }
        const int cnt = (int)idx.size();
        if (cnt == 0) return FGPU_OK;
        int *d_idx = dalloc<int>(cnt);
        double *d_xyz = dalloc<double>(3 * (size_t)cnt);
        h2d(d_idx, idx.data(), cnt, st);
        h2d(d_xyz, nx.data(), nx.size(), st);
        launch_scatter_coords(st, cnt, d_idx, d_xyz, m->d_X64, m->d_X32);
        const double mn = min_center_det(m);
        if (!(mn >= 0.0)) {
            h2d(d_xyz, ox.data(), ox.size(), st);
            launch_scatter_coords(st, cnt, d_idx, d_xyz, m->d_X64, m->d_X32);
            HIPCHECK(hipStreamSynchronize(st));
            dfree(d_idx);
            dfree(d_xyz);
            throw MeshError("moved vertices turn a cell inside out (centre Jacobian " + std::to_string(mn) +
                            "); coordinates left unchanged, the mesh must be rebuilt");
        }
        dfree(d_idx);
        dfree(d_xyz);
        for (int k = 0, j = 0; k < n; ++k) {
            const int d = m->dof_of_vertex[ids[k]];
            if (d < 0) continue;
            for (int a = 0; a < 3; ++a) m->x[3 * (size_t)d + a] = nx[3 * (size_t)j + a];
            ++j;
        }
        m->coord_gen++;
        return FGPU_OK;
    });
}

extern "C" int fgpu_mesh_set_motion(fgpu_mesh *m, int rows, const int *vertex_ids, const int *ptr, const int *ctrl,
                                    const double *w, int n_ctrl) {
    if (!m || rows < 0 || n_ctrl < 0 || (rows > 0 && (!vertex_ids || !ptr || !ctrl || !w))) return FGPU_EINVAL;
    return guarded(m->ctx, [&] {
        HIPCHECK(hipSetDevice(m->ctx->device));
        hipStream_t st = m->ctx->stream;
        refresh_host_x(m);
        // rows of unused vertices are dropped
        std::vector<int> idof, p{0}, cc;
        std::vector<double> ww, base;
        for (int r = 0; r < rows; ++r) {
            if (vertex_ids[r] < 0 || vertex_ids[r] >= m->n_vertices) throw std::invalid_argument("vertex id out of range");
            if (ptr[r + 1] < ptr[r]) throw std::invalid_argument("motion rows must be ordered");
            const int d = m->dof_of_vertex[vertex_ids[r]];
            if (d < 0) continue;
            idof.push_back(m->idof_of_dof[d]);
            for (int a = 0; a < 3; ++a) base.push_back(m->x[3 * (size_t)d + a]);
            for (int k = ptr[r]; k < ptr[r + 1]; ++k) {
                if (ctrl[k] < 0 || ctrl[k] >= n_ctrl) throw std::invalid_argument("control index out of range");
                cc.push_back(ctrl[k]);
                ww.push_back(w[k]);
            }
            p.push_back((int)cc.size());
        }
        Mesh::Motion &M = m->motion;
        dfree(M.idof); dfree(M.ptr); dfree(M.ctrl); dfree(M.w); dfree(M.base); dfree(M.disp); dfree(M.prev);
        M = Mesh::Motion{};
        M.rows = (int)idof.size();
        M.n_ctrl = n_ctrl;
        M.idof = dalloc<int>(idof.size());
        M.ptr = dalloc<int>(p.size());
        M.ctrl = dalloc<int>(cc.size());
        M.w = dalloc<double>(ww.size());
        M.base = dalloc<double>(base.size());
        M.disp = dalloc<double>(3 * (size_t)n_ctrl);
        M.prev = dalloc<double>(3 * (size_t)n_ctrl);
        h2d(M.idof, idof.data(), idof.size(), st);
        h2d(M.ptr, p.data(), p.size(), st);
        h2d(M.ctrl, cc.data(), cc.size(), st);
        h2d(M.w, ww.data(), ww.size(), st);
        h2d(M.base, base.data(), base.size(), st);
        HIPCHECK(hipStreamSynchronize(st));
        return FGPU_OK;
    });
}

extern "C" int fgpu_mesh_displace(fgpu_mesh *m, const double *disp) {
    if (!m || !disp) return FGPU_EINVAL;
    return guarded(m->ctx, [&] {
        HIPCHECK(hipSetDevice(m->ctx->device));
        hipStream_t st = m->ctx->stream;
        Mesh::Motion &M = m->motion;
        if (!M.idof) throw std::invalid_argument("no motion map set (fgpu_mesh_set_motion)");
        h2d(M.disp, disp, 3 * (size_t)M.n_ctrl, st);
        launch_motion(st, M.rows, M.idof, M.ptr, M.ctrl, M.w, M.base, M.disp, m->d_X64, m->d_X32);
        const double mn = min_center_det(m);
        if (!(mn >= 0.0)) {
            // back to the last accepted displacement (the base positions before the first)
            if (!M.have_prev) launch_set_const(st, 3 * M.n_ctrl, 0.0, M.prev);
            launch_motion(st, M.rows, M.idof, M.ptr, M.ctrl, M.w, M.base, M.prev, m->d_X64, m->d_X32);
            HIPCHECK(hipStreamSynchronize(st));
            throw MeshError("displacement turns a cell inside out (centre Jacobian " + std::to_string(mn) +
                            "); coordinates left unchanged, the mesh must be rebuilt");
        }
        std::swap(M.disp, M.prev);
        M.have_prev = true;
        m->x_stale = true;
        m->coord_gen++;
        return FGPU_OK;
    });
}

// fgpu_mesh_displace without allocation: the per-block minima of the centre Jacobian and their
// minimum live in the device's scratch array (kScratchMinBlocks blocks of kBlock cells at
// most).  Calls that overlap on one GPU race for that array.
extern "C" int fgpu_mesh_displace_noalloc(fgpu_mesh *m, const double *disp) {
    if (!m || !disp) return FGPU_EINVAL;
    return guarded(m->ctx, [&] {
        HIPCHECK(hipSetDevice(m->ctx->device));
        hipStream_t st = m->ctx->stream;
        Mesh::Motion &M = m->motion;
        if (!M.idof) throw std::invalid_argument("no motion map set (fgpu_mesh_set_motion)");
        const int nb = center_det_blocks(m->n_cells);
        if (nb > kScratchMinBlocks)
            throw std::invalid_argument("mesh too large for fgpu_mesh_displace_noalloc (more than 16777216 cells)");
        double *sc = m->ctx->scratch;
        h2d(M.disp, disp, 3 * (size_t)M.n_ctrl, st);
        launch_motion(st, M.rows, M.idof, M.ptr, M.ctrl, M.w, M.base, M.disp, m->d_X64, m->d_X32);
        launch_center_det_min(st, m->n_cells, m->d_cells, m->d_X64, sc);
        launch_min_reduce(st, nb, sc, sc + kScratchMinBlocks);
        double mn = 0.0;
        d2h(&mn, sc + kScratchMinBlocks, 1, st);
        HIPCHECK(hipStreamSynchronize(st));
        if (!(mn >= 0.0)) {
            if (!M.have_prev) launch_set_const(st, 3 * M.n_ctrl, 0.0, M.prev);
            launch_motion(st, M.rows, M.idof, M.ptr, M.ctrl, M.w, M.base, M.prev, m->d_X64, m->d_X32);
            HIPCHECK(hipStreamSynchronize(st));
            throw MeshError("displacement turns a cell inside out (centre Jacobian " + std::to_string(mn) +
                            "); coordinates left unchanged, the mesh must be rebuilt");
        }
        std::swap(M.disp, M.prev);
        M.have_prev = true;
        m->x_stale = true;
        m->coord_gen++;
        return FGPU_OK;
    });
}

static void ensure_metric(fgpu_mesh *m);

extern "C" int fgpu_mesh_use_affine(fgpu_mesh *m, int allow, int *active) {
    if (!m) return FGPU_EINVAL;
    return guarded(m->ctx, [&] {
        HIPCHECK(hipSetDevice(m->ctx->device));
        m->affine_allowed = allow != 0;
        ensure_metric(m);
        if (active) *active = m->affine && m->affine_allowed;
        return FGPU_OK;
    });
}

extern "C" int fgpu_mesh_set_patch_cells(fgpu_mesh *m, int cells) {
    if (!m || cells < 8 || cells > kPatchCells) return FGPU_EINVAL;
    return guarded(m->ctx, [&] {
        HIPCHECK(hipSetDevice(m->ctx->device));
        if (cells == m->patches.cells) return FGPU_OK;
        std::vector<int> gptr;
        if (m->icells.empty()) {   // a device import keeps the internal cells on the device only
            m->icells.resize(8 * (size_t)m->n_cells);
            d2h(m->icells.data(), m->d_cells, m->icells.size(), m->ctx->stream);
            HIPCHECK(hipStreamSynchronize(m->ctx->stream));
        }
        build_patches(m->n_cells, m->n_dofs, m->icells, cells, m->patches);
        build_slots(m->n_dofs, m->patches, gptr);
        upload_patches(m, gptr);
        return FGPU_OK;
    });
}

extern "C" int fgpu_mesh_device_bytes(const fgpu_mesh *m, size_t *bytes) {
    if (!m || !bytes) return FGPU_EINVAL;
    const size_t nd = m->n_dofs, nc = m->n_cells, ns = m->n_slots, nb = m->bf_cell.size();
    const Patches &P = m->patches;
    // patch lists: nodes and slots (4 + 4 per slot), local indices and entries (2 + 2 per cell
    // vertex), entry pointers (2 per slot and patch)
    *bytes = nd * (32 + 16 + 4 + 4 + 4) + nc * 32 + (P.cell0.size() + P.node0.size()) * 4 + ns * 8 +
             (16 * nc + ns + P.n) * 2 + nb * 8 + (nd + 1 + 4 * nb) * 4;
    return FGPU_OK;
}

extern "C" int fgpu_sys_device_bytes(const fgpu_sys *s, size_t *bytes) {
    if (!s || !bytes) return FGPU_EINVAL;
    const fgpu_mesh *m = static_cast<const fgpu_mesh *>(s->mesh);
    const size_t nd = m->n_dofs, ns = m->n_slots, nb = m->bf_cell.size(), nc = m->n_cells;
    *bytes = nd * (11 * 8 + 1 + 6 * 4) + ns * (8 + 4) + nb * 5 * 8 + 2 * kPartRegions * (size_t)s->nb_dof * 8 +
             2 * kScal * 8 + (s->kq ? nc * 8 * (8 + 8 + 4 + 4) : 0) + (size_t)s->nc * s->cld * 2 + (size_t)s->nc * 12;
    return FGPU_OK;
}

extern "C" int fgpu_mesh_patch_info(const fgpu_mesh *m, int *cells, int *n_patches, int *node_cap, int *n_slots) {
    if (!m) return FGPU_EINVAL;
    if (cells) *cells = m->patches.cells;
    if (n_patches) *n_patches = m->patches.n;
    if (node_cap) *node_cap = m->patches.node_cap;
    if (n_slots) *n_slots = m->n_slots;
    return FGPU_OK;
}

// ------------------------------------------------------------------------------------------
// systems
// ------------------------------------------------------------------------------------------
// slot buffers and element block counts follow the mesh's current patches and the system's
// patches-per-workgroup setting
static void ensure_layout(fgpu_sys *s) {
    fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
    s->nb_cell = padded_blocks(m->patches.n, s->ppb);
    if (s->patch_gen == m->patch_gen) return;
    HIPCHECK(hipStreamSynchronize(m->ctx->stream));
    dfree(s->Ye);
    dfree(s->Ye32);
    s->Ye = dalloc<double>(m->n_slots);
    s->Ye32 = dalloc<float>(m->n_slots);
    s->patch_gen = m->patch_gen;
    s->diag_valid = false;
}
static void ensure_layout(fgpu_sys *s);

static Desc make_desc64(fgpu_sys *s, bool residual) {
    ensure_layout(s);
    Desc d{};
    fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
    d.n_dofs = m->n_dofs;
    d.n_cells = m->n_cells;
    d.nb_cell = s->nb_cell;
    d.nb_dof = s->nb_dof;
    d.cells = m->d_cells;
    d.X = m->d_X64;
    d.gptr = m->d_gptr;
    d.n_patch = m->patches.n;
    d.pcells = m->patches.cells;
    d.node_cap = m->patches.node_cap;
    d.ppb = s->ppb;
    d.pcell0 = m->d_pcell0;
    d.pnode0 = m->d_pnode0;
    d.pnodes = m->d_pnodes;
    d.pslot = m->d_pslot;
    d.plidx = m->d_plidx;
    d.plptr = m->d_plptr;
    d.plent = m->d_plent;
    d.dmask = s->d_dmask;
    d.kq = s->have_coef ? s->kq : nullptr;
    d.gq = s->have_coef ? s->gq : nullptr;
    d.ua = residual ? s->x : s->z;
    d.ub = residual ? nullptr : s->p;
    d.Ye = s->Ye;
    d.q = s->q;
    d.p = s->p;
    d.z = s->z;
    d.x = s->x;
    d.r = s->r;
    d.zw = s->z;
    d.invdiag = s->invd;
    d.b = s->b;
    d.scal = s->scal;
    d.part = s->part;
    d.pc = FGPU_PC_JACOBI;
    return d;
}

// the per-cell metric of the tet-split form for the current coordinates, and whether every
// cell has that form (nodes within 1e-9 of it, relative to the cell size)
static void ensure_metric(fgpu_mesh *m) {
    if (m->metric_gen == m->coord_gen) return;
    hipStream_t st = m->ctx->stream;
    if (!m->ctx->affine_table) {
        HIPCHECK(upload_affine_table(st));
        m->ctx->affine_table = true;
    }
    if (!m->d_metric) m->d_metric = dalloc<float>(7 * (size_t)std::max(1, m->n_cells));
    const int nb = metric_blocks(m->n_cells);
    double *dev = dalloc<double>(nb);
    launch_metric(st, m->n_cells, m->d_cells, m->d_layout, m->d_X64, m->d_metric, dev);
    std::vector<double> h(nb);
    d2h(h.data(), dev, nb, st);
    HIPCHECK(hipStreamSynchronize(st));
    dfree(dev);
    double mx = 0;
    for (double v : h) mx = std::max(mx, v);
    m->affine = m->n_cells > 0 && mx <= 1e-9;
    m->metric_gen = m->coord_gen;
}

// coarse: the two-level preconditioner of the system's FP32 CG, when set up (only the CG
// iteration descriptors carry it; plain applies must not add the coarse correction)
static Desc make_desc32(fgpu_sys *s, bool coarse = false) {
    ensure_layout(s);
    Desc d{};
    fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
    ensure_metric(m);
    d.metric = (m->affine && m->affine_allowed) ? m->d_metric : nullptr;
    s->desc32_affine = d.metric != nullptr;   // sizes the LDS of the launches using it
    d.n_dofs = m->n_dofs;
    d.n_cells = m->n_cells;
    d.nb_cell = s->nb_cell;
    d.nb_dof = s->nb_dof;
    d.cells = m->d_cells;
    d.X = m->d_X32;
    d.gptr = m->d_gptr;
    d.n_patch = m->patches.n;
    d.pcells = m->patches.cells;
    d.node_cap = m->patches.node_cap;
    d.ppb = s->ppb;
    d.pcell0 = m->d_pcell0;
    d.pnode0 = m->d_pnode0;
    d.pnodes = m->d_pnodes;
    d.pslot = m->d_pslot;
    d.plidx = m->d_plidx;
    d.plptr = m->d_plptr;
    d.plent = m->d_plent;
    d.dmask = s->d_dmask;
    d.kq = s->have_coef ? s->kq32 : nullptr;
    d.gq = s->have_coef ? s->gq32 : nullptr;
    d.ua = s->z32;
    d.ub = s->p32;
    d.Ye = s->Ye32;
    d.q = s->q32;
    d.p = s->p32;
    d.z = s->z32;
    d.x = s->d32;
    d.r = s->r32;
    d.zw = s->z32;
    d.invdiag = s->invd32;
    d.b = nullptr;
    d.scal = s->scal + kScal;
    d.part = s->part + kPartRegions * (size_t)s->nb_dof;
    if (coarse && s->coarse) {
        d.cc = s->cc;
        d.cinv = s->cinv;
        d.cscale = s->cscale;
        d.cld = s->cld;
        d.nc = s->nc;
        d.agg_shift = s->agg_shift;
        d.cc_in = 1;
    }
    d.pc = FGPU_PC_JACOBI;
    return d;
}

extern "C" int fgpu_sys_create(fgpu_mesh *m, fgpu_sys **out) {
    if (!m || !out) return FGPU_EINVAL;
    *out = nullptr;
    fgpu_sys *s = new (std::nothrow) fgpu_sys();
    if (!s) return FGPU_ENOMEM;
    s->mesh = m;
    s->ctx = m->ctx;
    const int rc = guarded(m->ctx, [&] {
        HIPCHECK(hipSetDevice(m->ctx->device));
        const size_t n = m->n_dofs;
        s->nb_cell = padded_blocks(m->patches.n, s->ppb);
        s->nb_dof = padded_blocks(m->n_dofs, kDofBlock);
        s->b = dalloc<double>(n); s->x = dalloc<double>(n); s->r = dalloc<double>(n); s->z = dalloc<double>(n);
        s->p = dalloc<double>(n); s->q = dalloc<double>(n);
        s->invd = dalloc<double>(n); s->diag = dalloc<double>(n); s->g = dalloc<double>(n); s->tmp = dalloc<double>(n);
        s->fslot = dalloc<double>(4 * m->bf_cell.size()); s->fval = dalloc<double>(m->bf_cell.size());
        s->d_dmask = dalloc<unsigned char>(n);
        s->r32 = dalloc<float>(n); s->z32 = dalloc<float>(n); s->p32 = dalloc<float>(n); s->q32 = dalloc<float>(n);
        s->d32 = dalloc<float>(n); s->invd32 = dalloc<float>(n);
        ensure_layout(s);
        s->scal = dalloc<double>(2 * kScal);
        s->part = dalloc<double>(2 * kPartRegions * (size_t)s->nb_dof);
        hipStream_t st = m->ctx->stream;
        HIPCHECK(hipMemsetAsync(s->scal, 0, 2 * kScal * sizeof(double), st));
        HIPCHECK(hipMemsetAsync(s->part, 0, 2 * kPartRegions * (size_t)s->nb_dof * sizeof(double), st));
        HIPCHECK(hipMemsetAsync(s->b, 0, n * sizeof(double), st));
        HIPCHECK(hipMemsetAsync(s->x, 0, n * sizeof(double), st));
        HIPCHECK(hipMemsetAsync(s->g, 0, n * sizeof(double), st));
        HIPCHECK(hipMemsetAsync(s->d_dmask, 0, n, st));
        s->dmask.assign(n, 0);
        s->gval.assign(n, 0.0);
        HIPCHECK(hipStreamSynchronize(st));
        return FGPU_OK;
    });
    if (rc != FGPU_OK) {
        delete s;
        return rc;
    }
    m->refs++;
    m->ctx->refs++;
    *out = s;
    return FGPU_OK;
}

extern "C" int fgpu_sys_set_patches_per_block(fgpu_sys *s, int ppb) {
    if (!s || ppb < 1 || ppb > 64) return FGPU_EINVAL;
    s->ppb = ppb;
    return FGPU_OK;
}

static void sys_free(fgpu_sys *s) {
    fgpu_ctx *c = s->ctx;
    fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    std::vector<Sys *> srcs;
    for (auto &t : s->rhs_terms)
        if (t.src) srcs.push_back(t.src);
    delete s;
    for (Sys *src : srcs)
        if (--src->term_refs == 0 && src->closing) sys_free(static_cast<fgpu_sys *>(src));
    if (--m->refs == 0 && m->closing) mesh_free(m);
    ctx_release(c);
}

// a system whose solution a recorded heat term of another system reads stays alive (marked)
// until that term is dropped
extern "C" void fgpu_sys_destroy(fgpu_sys *s) {
    if (!s) return;
    if (s->term_refs > 0) {
        s->closing = true;
        return;
    }
    sys_free(s);
}

// single-system tables for setup launches: [Desc][elem tab][dof tab] in a scratch buffer
struct OneSys {
    Desc *D;
    int2 *etab, *dtab;
    int ne, nd;
};

// Asynchronous host-to-device copy from host data that may die when the caller returns: the
// bytes go through one of the context's pinned staging buffers, on the context's stream
static void stage_h2d(fgpu_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!bytes) return;
    const int k = c->stage_next;
    c->stage_next = (k + 1) % fgpu_ctx::kStage;
    if (c->stage_busy[k]) HIPCHECK(hipEventSynchronize(c->stage_ev[k]));
    c->stage_busy[k] = false;
    if (bytes > c->stage_cap[k]) {
        if (c->stage_host[k]) HIPCHECK(hipHostFree(c->stage_host[k]));
        c->stage_host[k] = nullptr;
        c->stage_cap[k] = 0;
        const size_t cap = std::max<size_t>(2 * bytes, 1 << 16);
        HIPCHECK(hipHostMalloc(&c->stage_host[k], cap, hipHostMallocDefault));
        c->stage_cap[k] = cap;
    }
    if (!c->stage_ev[k]) HIPCHECK(hipEventCreateWithFlags(&c->stage_ev[k], hipEventDisableTiming));
    std::memcpy(c->stage_host[k], src, bytes);
    HIPCHECK(hipMemcpyAsync(dst, c->stage_host[k], bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipEventRecord(c->stage_ev[k], c->stream));
    c->stage_busy[k] = true;
}

static void *ensure_tab(fgpu_ctx *c, size_t bytes) {
    if (bytes > c->tab_bytes) {
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (c->d_tab) HIPCHECK(hipFree(c->d_tab));
        c->d_tab = nullptr;
        c->tab_bytes = 0;
        const size_t nb = std::max<size_t>(bytes * 2, 1 << 16);
        HIPCHECK(hipMalloc(&c->d_tab, nb));
        c->tab_bytes = nb;
    }
    return c->d_tab;
}

// Uploads descriptors and tables of a batch into the context's table buffer.  Layout:
// [ndesc_sets x n Desc][elem table][dof table]
struct BatchTabs {
    std::vector<const Desc *> sets;  // device pointers of each descriptor set
    const ETab *etab = nullptr;
    const int2 *dtab = nullptr, *ctab = nullptr;
    int ne = 0, nd = 0, nct = 0;   // nct: coarse-product workgroups (systems with a coarse level)
    int max_nc = 0;                // largest coarse size of the batch (LDS of the coarse product)
    bool affine32 = true;          // every system's FP32 descriptor uses the tet-split form
    size_t lds64 = 0, lds32 = 0;  // dynamic LDS of the element kernel, largest patch of the batch
};

// coordinates and inputs per node (inputs only for the tet-split FP32 form); FP32: one 64-bit
// fixed-point sum per node and four floats; FP64: outputs per cell, entry lists
static size_t patch_lds(int node_cap, int cells, size_t tsize, bool affine = false) {
    if (tsize == 4) return (affine ? 1 : 4) * (size_t)node_cap * 4 + 8 * (size_t)node_cap + 16;
    return ((affine ? 1 : 4) * (size_t)node_cap + 8 * (size_t)cells) * tsize + 16 * (size_t)cells;
}

static BatchTabs upload_batch(fgpu_ctx *c, const std::vector<std::vector<Desc>> &sets, const std::vector<fgpu_sys *> &sys,
                              void **own_buf = nullptr, size_t *own_cap = nullptr) {
    std::vector<int2> et0, dt, ct;
    int max_nc_b = 0;
    std::vector<ETab> et;
    for (size_t i = 0; i < sys.size(); ++i) {
        et0.clear();
        append_table(et0, (int)i, sys[i]->nb_cell);
        const Patches &P = static_cast<const fgpu_mesh *>(sys[i]->mesh)->patches;
        for (const int2 &e : et0) {
            const int p = e.y * sys[i]->ppb;
            ETab w{(int)i, p, 0, 0, 0, 0, 0, 0};
            if (p < P.n) {
                w.c0 = P.cell0[p];
                w.nc = P.cell0[p + 1] - w.c0;
                w.n0 = P.node0[p];
                w.np = P.node0[p + 1] - w.n0;
            }
            et.push_back(w);
        }
        append_table(dt, (int)i, sys[i]->nb_dof);
        if (sys[i]->coarse) {
            for (int b = 0; b < (sys[i]->nc + kCoarseRows - 1) / kCoarseRows; ++b) ct.push_back(make_int2((int)i, b));
            max_nc_b = std::max(max_nc_b, sys[i]->nc);
        }
    }
    const size_t n = sys.size();
    const size_t dbytes = sets.size() * n * sizeof(Desc);
    const size_t align = 256;
    const size_t off_e = (dbytes + align - 1) / align * align;
    const size_t off_d = off_e + (et.size() * sizeof(ETab) + align - 1) / align * align;
    const size_t off_c = off_d + (dt.size() * sizeof(int2) + align - 1) / align * align;
    const size_t total = off_c + ct.size() * sizeof(int2);
    std::vector<char> host(total);
    for (size_t k = 0; k < sets.size(); ++k) std::memcpy(host.data() + k * n * sizeof(Desc), sets[k].data(), n * sizeof(Desc));
    std::memcpy(host.data() + off_e, et.data(), et.size() * sizeof(ETab));
    std::memcpy(host.data() + off_d, dt.data(), dt.size() * sizeof(int2));
    if (!ct.empty()) std::memcpy(host.data() + off_c, ct.data(), ct.size() * sizeof(int2));
    char *dev;
    if (own_buf) {
        if (total > *own_cap) {
            HIPCHECK(hipStreamSynchronize(c->stream));
            if (*own_buf) HIPCHECK(hipFree(*own_buf));
            *own_buf = nullptr;
            *own_cap = 0;
            HIPCHECK(hipMalloc(own_buf, 2 * total));
            *own_cap = 2 * total;
        }
        dev = static_cast<char *>(*own_buf);
    } else {
        dev = static_cast<char *>(ensure_tab(c, total));
    }
    // stream-ordered: kernels queued before still read the previous tables of this buffer
    stage_h2d(c, dev, host.data(), total);
    BatchTabs t;
    for (size_t k = 0; k < sets.size(); ++k) t.sets.push_back(reinterpret_cast<const Desc *>(dev + k * n * sizeof(Desc)));
    t.etab = reinterpret_cast<const ETab *>(dev + off_e);
    t.dtab = reinterpret_cast<const int2 *>(dev + off_d);
    t.ctab = reinterpret_cast<const int2 *>(dev + off_c);
    t.ne = (int)et.size();
    t.nd = (int)dt.size();
    t.nct = (int)ct.size();
    t.max_nc = max_nc_b;
    for (auto *x : sys) {
        const fgpu_mesh *m = static_cast<const fgpu_mesh *>(x->mesh);
        t.lds64 = std::max(t.lds64, patch_lds(m->patches.node_cap, m->patches.cells, sizeof(double)));
        // the tet-split FP32 form needs no coordinates: as decided by the system's FP32
        // descriptor, which every caller builds before uploading it
        t.lds32 = std::max(t.lds32, patch_lds(m->patches.node_cap, m->patches.cells, sizeof(float), x->desc32_affine));
        t.affine32 = t.affine32 && x->desc32_affine;
    }
    return t;
}

static void ensure_diag(fgpu_sys *s) {
    if (s->diag_valid && s->diag_gen == static_cast<fgpu_mesh *>(s->mesh)->coord_gen) return;
    fgpu_ctx *c = s->ctx;
    fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
    Desc d = make_desc64(s, true);
    d.q = s->diag;
    BatchTabs t = upload_batch(c, {{d}}, {s});
    launch_elem_diag<double>(c->stream, t.sets[0], t.etab, t.ne, s->have_coef, t.lds64);
    launch_gather_plain<double>(c->stream, t.sets[0], t.dtab, t.nd, false);
    launch_invdiag(c->stream, m->n_dofs, s->d_dmask, s->diag, s->invd, s->invd32);
    s->diag_valid = true;
    s->diag_gen = m->coord_gen;
}

extern "C" int fgpu_sys_sigma_terms(fgpu_sys *cur, const fgpu_sys *heat, double sigma0, double alpha) {
    if (!cur || !heat || !(sigma0 > 0) || !std::isfinite(alpha)) return FGPU_EINVAL;
    return guarded(cur->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(cur->mesh);
        if (heat->mesh != cur->mesh) throw std::invalid_argument("current and heat systems must share the mesh");
        if (cur->lifted) throw std::invalid_argument("coefficients are set before fgpu_sys_apply_dirichlet");
        hipStream_t st = cur->ctx->stream;
        HIPCHECK(hipSetDevice(cur->ctx->device));
        const size_t nc = m->n_cells;
        if (!cur->kq) {
            cur->kq = dalloc<double>(8 * nc); cur->gq = dalloc<double>(8 * nc);
            cur->kq32 = dalloc<float>(8 * nc); cur->gq32 = dalloc<float>(8 * nc);
        }
        launch_sigma_terms(st, (int)nc, m->d_cells, heat->x, sigma0, alpha, cur->kq, cur->gq, cur->kq32, cur->gq32);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(st));
        cur->have_coef = true;
        cur->diag_valid = false;
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_set_coefficients(fgpu_sys *s, const double *kq, const double *gq) {
    if (!s) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        hipStream_t st = s->ctx->stream;
        HIPCHECK(hipSetDevice(s->ctx->device));
        if (!kq && !gq) {
            s->have_coef = false;
        } else {
            const size_t nc = m->n_cells;
            std::vector<double> k(8 * nc, 1.0), g(8 * nc, 0.0);
            for (size_t i = 0; i < nc; ++i) {
                const size_t c = m->cell_of_icell[i];
                for (int q = 0; q < 8; ++q) {
                    if (kq) k[8 * i + q] = kq[8 * c + q];
                    if (gq) g[8 * i + q] = gq[8 * c + q];
                }
            }
            if (!s->kq) {
                s->kq = dalloc<double>(8 * nc); s->gq = dalloc<double>(8 * nc);
                s->kq32 = dalloc<float>(8 * nc); s->gq32 = dalloc<float>(8 * nc);
            }
            std::vector<float> k32(k.begin(), k.end()), g32(g.begin(), g.end());
            h2d(s->kq, k.data(), k.size(), st);
            h2d(s->gq, g.data(), g.size(), st);
            h2d(s->kq32, k32.data(), k32.size(), st);
            h2d(s->gq32, g32.data(), g32.size(), st);
            HIPCHECK(hipStreamSynchronize(st));
            s->have_coef = true;
        }
        s->diag_valid = false;
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_heat_terms(fgpu_sys *h, const fgpu_sys *cur, double gamma, double lorentz) {
    return fgpu_sys_heat_terms_theta(h, cur, gamma, lorentz, 1.0);
}

extern "C" int fgpu_sys_heat_terms_theta(fgpu_sys *h, const fgpu_sys *cur, double gamma, double lorentz, double theta) {
    if (!h || !cur || !(gamma >= 0) || !(theta > 0 && theta <= 1)) return FGPU_EINVAL;
    return guarded(h->ctx, [&] {
        if (h->lifted) throw std::invalid_argument("heat terms are added before fgpu_sys_apply_dirichlet");
        fgpu_mesh *m = static_cast<fgpu_mesh *>(h->mesh);
        if (cur->mesh != h->mesh) throw std::invalid_argument("heat and current systems must share the mesh");
        hipStream_t st = h->ctx->stream;
        HIPCHECK(hipSetDevice(h->ctx->device));
        const size_t nc = m->n_cells;
        if (!h->kq) {
            h->kq = dalloc<double>(8 * nc); h->gq = dalloc<double>(8 * nc);
            h->kq32 = dalloc<float>(8 * nc); h->gq32 = dalloc<float>(8 * nc);
        }
        double *f = dalloc<double>(8 * nc);
        if (theta < 1) {
            h->rhs_terms.push_back({1, 0, f});
        } else {
            // replayed by fgpu_sys_reassemble from the solution and potential at that time
            Sys::RhsTerm term{2, 0, f};
            term.src = const_cast<fgpu_sys *>(cur);
            term.gamma = gamma;
            term.lorentz = lorentz;
            h->rhs_terms.push_back(term);
            term.src->term_refs++;
        }
        double *kx = theta < 1 ? dalloc<double>(8 * nc) : nullptr;
        launch_heat_terms(st, (int)nc, m->d_cells, m->d_X64, h->x, cur->x, gamma, lorentz, theta, h->kq, h->gq, h->kq32,
                          h->gq32, f, kx);
        h->have_coef = true;
        h->diag_valid = false;
        Desc d = make_desc64(h, true);
        d.q = h->b;
        BatchTabs t = upload_batch(h->ctx, {{d}}, {h});
        launch_elem_rhs(st, t.sets[0], t.etab, t.ne, f, t.lds64);
        launch_gather_plain<double>(st, t.sets[0], t.dtab, t.nd, true);
        if (kx) {
            // explicit part of the theta step: b -= (1 - theta) K_kappa(T_old) T_old, every dof
            double *gz = dalloc<double>(8 * nc);
            launch_set_const(st, (int)(8 * nc), 0.0, gz);
            Desc dx = make_desc64(h, true);
            dx.kq = kx;
            dx.gq = gz;
            dx.ua = h->x;
            dx.ub = nullptr;
            dx.q = h->b;
            BatchTabs tx = upload_batch(h->ctx, {{dx}}, {h});
            launch_elem<double>(st, tx.sets[0], tx.etab, tx.ne, true, false, false, S_BETA, tx.lds64);
            launch_gather_plain<double>(st, tx.sets[0], tx.dtab, tx.nd, true);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipStreamSynchronize(st));
            dfree(gz);
            dfree(kx);
            h->no_replay = "the explicit stiffness term of a theta step depends on the previous temperature: reset "
                           "and assemble again";
        }
        HIPCHECK(hipStreamSynchronize(st));
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_reset(fgpu_sys *s) {
    if (!s) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        HIPCHECK(hipSetDevice(s->ctx->device));
        launch_set_const(s->ctx->stream, m->n_dofs, 0.0, s->b);
        std::fill(s->dmask.begin(), s->dmask.end(), 0);
        std::fill(s->gval.begin(), s->gval.end(), 0.0);
        HIPCHECK(hipStreamSynchronize(s->ctx->stream));
        std::vector<Sys *> srcs;
        for (auto &t : s->rhs_terms)
            if (t.src) srcs.push_back(t.src);
        s->clear_terms();
        for (Sys *src : srcs)
            if (--src->term_refs == 0 && src->closing) sys_free(static_cast<fgpu_sys *>(src));
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_add_neumann(fgpu_sys *s, int bid, const double *fv, double cst) {
    if (!s) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        hipStream_t st = s->ctx->stream;
        HIPCHECK(hipSetDevice(s->ctx->device));
        const size_t nb = m->bf_cell.size();
        std::vector<double> val(nb, 0.0);
        int k = 0;
        for (size_t f = 0; f < nb; ++f)
            if (m->bf_id[f] == bid) val[f] = fv ? fv[k++] : cst;
        double *dval = dalloc<double>(nb);
        s->rhs_terms.push_back({0, bid, dval});
        h2d(dval, val.data(), nb, st);
        launch_face_load(st, (int)nb, m->d_bf_icell, m->d_bf_face, m->d_bf_id, bid, dval, m->d_cells, m->d_X64, s->fslot);
        launch_face_gather(st, m->n_dofs, m->d_fgptr, m->d_fgidx, s->fslot, s->b);
        HIPCHECK(hipStreamSynchronize(st));
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_add_cell_rhs(fgpu_sys *s, const double *fq) {
    if (!s || !fq) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        hipStream_t st = s->ctx->stream;
        HIPCHECK(hipSetDevice(s->ctx->device));
        const size_t nc = m->n_cells;
        std::vector<double> f(8 * nc);
        for (size_t i = 0; i < nc; ++i)
            for (int q = 0; q < 8; ++q) f[8 * i + q] = fq[8 * (size_t)m->cell_of_icell[i] + q];
        double *d_f = dalloc<double>(8 * nc);
        s->rhs_terms.push_back({1, 0, d_f});
        h2d(d_f, f.data(), f.size(), st);
        Desc d = make_desc64(s, true);
        d.q = s->b;
        BatchTabs t = upload_batch(s->ctx, {{d}}, {s});
        launch_elem_rhs(st, t.sets[0], t.etab, t.ne, d_f, t.lds64);
        launch_gather_plain<double>(st, t.sets[0], t.dtab, t.nd, true);
        HIPCHECK(hipStreamSynchronize(st));
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_add_dirichlet(fgpu_sys *s, int bid, double value) {
    static const int kFV[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
    if (!s) return FGPU_EINVAL;
    fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
    for (size_t f = 0; f < m->bf_cell.size(); ++f)
        if (m->bf_id[f] == bid)
            for (int j = 0; j < 4; ++j) {
                const int d = m->cells_lex[8 * (size_t)m->bf_cell[f] + kFV[m->bf_face[f]][j]];
                s->dmask[d] = 1;
                s->gval[d] = value;
            }
    return FGPU_OK;
}

// Jacobi diagonal and the Dirichlet lifting from the device mask and values: A g with the
// un-lifted operator, then b_i -= (A g)_i on free rows, b_i = A_ii g_i and x_i = g_i on
// constrained ones
static void lift(fgpu_sys *s) {
    fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
    hipStream_t st = s->ctx->stream;
    const int n = m->n_dofs;
    ensure_diag(s);
    launch_invdiag(st, n, s->d_dmask, s->diag, s->invd, s->invd32);
    if (s->g_zero) {
        launch_lift(st, n, s->d_dmask, s->g, nullptr, s->diag, s->b, s->x);
        return;
    }
    Desc d = make_desc64(s, true);
    d.ua = s->g;
    d.q = s->tmp;
    BatchTabs t = upload_batch(s->ctx, {{d}}, {s});
    launch_elem<double>(st, t.sets[0], t.etab, t.ne, s->have_coef, false, false, S_BETA, t.lds64);
    launch_gather_plain<double>(st, t.sets[0], t.dtab, t.nd, false);
    launch_lift(st, n, s->d_dmask, s->g, s->tmp, s->diag, s->b, s->x);
}

// rho = b / dof volume from the right-hand side as assembled (before the lifting); the dof
// volumes come from the element kernel's volume mode gathered like any slot buffer
static void save_charge(fgpu_sys *s) {
    fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
    hipStream_t st = s->ctx->stream;
    if (!s->rho) s->rho = dalloc<double>(std::max(1, m->n_dofs));
    Desc d = make_desc64(s, true);
    d.q = s->tmp;
    BatchTabs t = upload_batch(s->ctx, {{d}}, {s});
    launch_elem_vol(st, t.sets[0], t.etab, t.ne, t.lds64);
    launch_gather_plain<double>(st, t.sets[0], t.dtab, t.nd, false);
    launch_charge(st, m->n_dofs, s->b, s->tmp, s->rho);
    s->rho_saved = true;
}

extern "C" int fgpu_sys_save_charge_density(fgpu_sys *s) {
    if (!s) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        if (s->lifted) throw std::invalid_argument("the charge density is taken before fgpu_sys_apply_dirichlet");
        HIPCHECK(hipSetDevice(s->ctx->device));
        save_charge(s);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(s->ctx->stream));
        return FGPU_OK;
    });
}

// Rebuilds the right-hand side, the diagonal and the lifting for the current vertex positions
// from the terms recorded since the last reset (same kernels, same order, so the result is
// the one a reset and the same calls would give)
extern "C" int fgpu_sys_reassemble(fgpu_sys *s) {
    if (!s) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        if (s->no_replay) throw std::invalid_argument(s->no_replay);
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        hipStream_t st = s->ctx->stream;
        HIPCHECK(hipSetDevice(s->ctx->device));
        launch_set_const(st, m->n_dofs, 0.0, s->b);
        const size_t nb = m->bf_cell.size();
        for (const auto &term : s->rhs_terms) {
            if (term.kind == 0) {
                launch_face_load(st, (int)nb, m->d_bf_icell, m->d_bf_face, m->d_bf_id, term.bid, term.dval, m->d_cells,
                                 m->d_X64, s->fslot);
                launch_face_gather(st, m->n_dofs, m->d_fgptr, m->d_fgidx, s->fslot, s->b);
            } else {
                if (term.kind == 2) {
                    // heat step: coefficients and source from the present solution (the
                    // previous temperature) and the present current potential
                    if (term.src->mesh != s->mesh) throw std::invalid_argument("heat term of another mesh");
                    launch_heat_terms(st, m->n_cells, m->d_cells, m->d_X64, s->x, term.src->x, term.gamma, term.lorentz,
                                      1.0, s->kq, s->gq, s->kq32, s->gq32, term.dval, nullptr);
                    s->diag_valid = false;
                }
                Desc d = make_desc64(s, true);
                d.q = s->b;
                BatchTabs t = upload_batch(s->ctx, {{d}}, {s});
                launch_elem_rhs(st, t.sets[0], t.etab, t.ne, term.dval, t.lds64);
                launch_gather_plain<double>(st, t.sets[0], t.dtab, t.nd, true);
            }
        }
        if (s->rho_saved) save_charge(s);
        if (s->lifted) lift(s);
        HIPCHECK(hipGetLastError());
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_apply_dirichlet(fgpu_sys *s) {
    if (!s) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        hipStream_t st = s->ctx->stream;
        HIPCHECK(hipSetDevice(s->ctx->device));
        const int n = m->n_dofs;
        std::vector<unsigned char> mk(n);
        std::vector<double> gv(n);
        for (int i = 0; i < n; ++i) {
            const int d = m->dof_of_idof[i];
            mk[i] = s->dmask[d];
            gv[i] = s->dmask[d] ? s->gval[d] : 0.0;
        }
        s->g_zero = true;
        for (int i = 0; i < n; ++i) s->g_zero = s->g_zero && gv[i] == 0.0 && !std::signbit(gv[i]);
        h2d(s->d_dmask, mk.data(), n, st);
        h2d(s->g, gv.data(), n, st);
        lift(s);
        HIPCHECK(hipStreamSynchronize(st));   // mk and gv are read until the copies end
        s->lifted = true;
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_fill_solution(fgpu_sys *s, double v) {
    if (!s) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        HIPCHECK(hipSetDevice(s->ctx->device));
        launch_set_const(s->ctx->stream, m->n_dofs, v, s->x);
        HIPCHECK(hipStreamSynchronize(s->ctx->stream));
        return FGPU_OK;
    });
}

static void permute_to_internal(const fgpu_mesh *m, const double *api, std::vector<double> &out) {
    out.resize(m->n_dofs);
    for (int i = 0; i < m->n_dofs; ++i) out[i] = api[m->dof_of_idof[i]];
}

extern "C" int fgpu_sys_set_solution(fgpu_sys *s, const double *x) {
    if (!s || !x) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        HIPCHECK(hipSetDevice(s->ctx->device));
        std::vector<double> v;
        permute_to_internal(m, x, v);
        h2d(s->x, v.data(), v.size(), s->ctx->stream);
        HIPCHECK(hipStreamSynchronize(s->ctx->stream));
        return FGPU_OK;
    });
}

static int get_vec(fgpu_sys *s, const double *dev, double *out) {
    return guarded(s->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        HIPCHECK(hipSetDevice(s->ctx->device));
        std::vector<double> v(m->n_dofs);
        d2h(v.data(), dev, v.size(), s->ctx->stream);
        HIPCHECK(hipStreamSynchronize(s->ctx->stream));
        for (int i = 0; i < m->n_dofs; ++i) out[m->dof_of_idof[i]] = v[i];
        return FGPU_OK;
    });
}

// device-to-device copies of the solution in the library's internal dof order
extern "C" int fgpu_sys_solution_to_device(fgpu_sys *s, double *dev) {
    if (!s || !dev) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        HIPCHECK(hipSetDevice(s->ctx->device));
        const int n = static_cast<fgpu_mesh *>(s->mesh)->n_dofs;
        HIPCHECK(hipMemcpyAsync(dev, s->x, n * sizeof(double), hipMemcpyDeviceToDevice, s->ctx->stream));
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_solution_from_device(fgpu_sys *s, const double *dev) {
    if (!s || !dev) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        if (s->batches > 0) throw std::invalid_argument("the system is in a running batch");
        HIPCHECK(hipSetDevice(s->ctx->device));
        const int n = static_cast<fgpu_mesh *>(s->mesh)->n_dofs;
        HIPCHECK(hipMemcpyAsync(s->x, dev, n * sizeof(double), hipMemcpyDeviceToDevice, s->ctx->stream));
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_get_solution(fgpu_sys *s, double *x) {
    if (!s || !x) return FGPU_EINVAL;
    return get_vec(s, s->x, x);
}

extern "C" int fgpu_sys_get_rhs(fgpu_sys *s, double *b) {
    if (!s || !b) return FGPU_EINVAL;
    return get_vec(s, s->b, b);
}

extern "C" int fgpu_sys_export_charge_density(fgpu_sys *s, double *rho) {
    if (!s || !rho) return FGPU_EINVAL;
    if (s->rho_saved) return get_vec(s, s->rho, rho);
    return guarded(s->ctx, [&] {
        std::fill(rho, rho + static_cast<fgpu_mesh *>(s->mesh)->n_dofs, 0.0);
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_apply32(fgpu_sys *s, const double *x, double *y) {
    if (!s || !x || !y) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        hipStream_t st = s->ctx->stream;
        HIPCHECK(hipSetDevice(s->ctx->device));
        std::vector<double> v;
        permute_to_internal(m, x, v);
        std::vector<float> v32(v.begin(), v.end());
        h2d(s->p32, v32.data(), v32.size(), st);
        Desc d = make_desc32(s);
        d.ua = s->p32;
        d.ub = nullptr;
        d.q = s->q32;
        BatchTabs t = upload_batch(s->ctx, {{d}}, {s});
        launch_elem<float>(st, t.sets[0], t.etab, t.ne, s->have_coef, false, false, S_BETA, t.lds32, t.affine32);
        launch_gather_plain<float>(st, t.sets[0], t.dtab, t.nd, false);
        std::vector<float> o(m->n_dofs);
        d2h(o.data(), s->q32, o.size(), st);
        HIPCHECK(hipStreamSynchronize(st));
        for (int i = 0; i < m->n_dofs; ++i) y[m->dof_of_idof[i]] = o[i];
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_apply(fgpu_sys *s, const double *x, double *y) {
    if (!s || !x || !y) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        hipStream_t st = s->ctx->stream;
        HIPCHECK(hipSetDevice(s->ctx->device));
        std::vector<double> v;
        permute_to_internal(m, x, v);
        h2d(s->tmp, v.data(), v.size(), st);
        Desc d = make_desc64(s, true);
        d.ua = s->tmp;
        d.q = s->q;
        BatchTabs t = upload_batch(s->ctx, {{d}}, {s});
        launch_elem<double>(st, t.sets[0], t.etab, t.ne, s->have_coef, false, false, S_BETA, t.lds64);
        launch_gather_plain<double>(st, t.sets[0], t.dtab, t.nd, false);
        HIPCHECK(hipStreamSynchronize(st));
        return get_vec(s, s->q, y);
    });
}

// ------------------------------------------------------------------------------------------
// two-level preconditioner of the mixed solver
// ------------------------------------------------------------------------------------------
// E = Z^T A Z for the system's current operator and constraints, Z the free-dof indicator of
// the aggregates of 2^shift internal dofs.  The block graph (blocks sharing a cell) is coloured
// so that no two blocks of a colour are within distance 2; A applied to the indicator of all
// blocks of one colour then gives, summed over the free dofs of a block b1, E[b1][b2] for the
// only block b2 of that colour in b1's closed neighbourhood.  One FP64 apply per colour, all
// on the device; E^-1 by blocked Gauss-Jordan exchanges on the device (32 pivots per step),
// stored in FP16 with a power-of-two scale.
static int build_coarse(fgpu_sys *s, int shift) {
    fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
    hipStream_t st = s->ctx->stream;
    const int n = m->n_dofs;
    const int agg = 1 << shift, g = agg / kAggMin;
    const int nc = (n + agg - 1) / agg;
    const int nq = (n + kAggMin - 1) / kAggMin;   // 256-dof rounds
    if (nc > kCoarseMax) throw std::invalid_argument("too many aggregates: choose larger ones");
    // block graph
    std::vector<int64_t> pairs;
    for (int c = 0; c < m->n_cells; ++c) {
        int b[8], nb = 0;
        for (int k = 0; k < 8; ++k) {
            const int x = m->idof_of_dof[m->cells_lex[8 * (size_t)c + k]] >> shift;
            bool seen = false;
            for (int j = 0; j < nb; ++j) seen = seen || b[j] == x;
            if (!seen) b[nb++] = x;
        }
        for (int k = 0; k < nb; ++k)
            for (int l = 0; l < nb; ++l)
                if (k != l) pairs.push_back((int64_t)b[k] * nc + b[l]);
    }
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    std::vector<int> nptr(nc + 1, 0), nbr;
    for (int64_t pr : pairs) nptr[pr / nc + 1]++;
    for (int b = 0; b < nc; ++b) nptr[b + 1] += nptr[b];
    nbr.resize(pairs.size());
    for (size_t i = 0; i < pairs.size(); ++i) nbr[i] = (int)(pairs[i] % nc);   // sorted: rows in order
    // greedy colouring of the square of the graph
    std::vector<int> color(nc, -1), stamp;
    int ncol = 0;
    for (int b = 0; b < nc; ++b) {
        auto mark = [&](int x) {
            if (color[x] >= 0) {
                if ((int)stamp.size() <= color[x]) stamp.resize(color[x] + 1, -1);
                stamp[color[x]] = b;
            }
        };
        mark(b);
        for (int e = nptr[b]; e < nptr[b + 1]; ++e) {
            const int a = nbr[e];
            mark(a);
            for (int f = nptr[a]; f < nptr[a + 1]; ++f) mark(nbr[f]);
        }
        int c = 0;
        while (c < (int)stamp.size() && stamp[c] == b) ++c;
        color[b] = c;
        ncol = std::max(ncol, c + 1);
    }
    // block sums of A (indicator of each colour)
    int *d_color = dalloc<int>(nc);
    double *d_sums = dalloc<double>((size_t)ncol * nq);
    h2d(d_color, color.data(), nc, st);
    Desc d = make_desc64(s, true);
    d.ua = s->tmp;
    d.q = s->q;
    BatchTabs t = upload_batch(s->ctx, {{d}}, {s});
    for (int c = 0; c < ncol; ++c) {
        launch_coarse_indicator(st, n, s->d_dmask, d_color, shift, c, s->tmp);
        launch_elem<double>(st, t.sets[0], t.etab, t.ne, s->have_coef, false, false, S_BETA, t.lds64);
        launch_gather_plain<double>(st, t.sets[0], t.dtab, t.nd, false);
        launch_coarse_sums(st, n, s->d_dmask, s->q, d_sums + (size_t)c * nq);
    }
    HIPCHECK(hipGetLastError());
    std::vector<double> qsums((size_t)ncol * nq);
    d2h(qsums.data(), d_sums, qsums.size(), st);
    std::vector<unsigned char> mk(n);
    d2h(mk.data(), s->d_dmask, n, st);
    HIPCHECK(hipStreamSynchronize(st));
    dfree(d_color);
    dfree(d_sums);
    std::vector<int> nfree(nc, 0);
    for (int i = 0; i < n; ++i) nfree[i >> shift] += mk[i] ? 0 : 1;
    // per colour, the sums over each aggregate's rounds in round order
    std::vector<double> sums((size_t)ncol * nc, 0.0);
    for (int c = 0; c < ncol; ++c)
        for (int a = 0; a < nc; ++a) {
            double v = 0.0;
            for (int u = 0; u < g && a * g + u < nq; ++u) v += qsums[(size_t)c * nq + a * g + u];
            sums[(size_t)c * nc + a] = v;
        }
    std::vector<double> E((size_t)nc * nc, 0.0);
    for (int b1 = 0; b1 < nc; ++b1) {
        E[(size_t)b1 * nc + b1] = sums[(size_t)color[b1] * nc + b1];
        for (int e = nptr[b1]; e < nptr[b1 + 1]; ++e) {
            const int b2 = nbr[e];
            E[(size_t)b1 * nc + b2] = sums[(size_t)color[b2] * nc + b1];
        }
    }
    for (int b1 = 0; b1 < nc; ++b1)
        for (int b2 = b1 + 1; b2 < nc; ++b2) {
            const double v = 0.5 * (E[(size_t)b1 * nc + b2] + E[(size_t)b2 * nc + b1]);
            E[(size_t)b1 * nc + b2] = E[(size_t)b2 * nc + b1] = v;
        }
    // blocks without free dofs: a decoupled unit row (their block sums of r are 0)
    for (int b = 0; b < nc; ++b)
        if (nfree[b] == 0) {
            for (int j = 0; j < nc; ++j) E[(size_t)b * nc + j] = E[(size_t)j * nc + b] = 0.0;
            E[(size_t)b * nc + b] = 1.0;
        } else if (!(E[(size_t)b * nc + b] > 0.0)) {
            throw std::runtime_error("coarse operator is not positive definite");
        }
    double *dA = dalloc<double>((size_t)nc * nc), *dB = dalloc<double>((size_t)nc * nc);
    const int gb = gj_block();
    double *dPi = dalloc<double>((size_t)gb * gb), *dV = dalloc<double>((size_t)nc * gb),
           *dW = dalloc<double>((size_t)nc * gb);
    h2d(dA, E.data(), E.size(), st);
    for (int k = 0; k < nc; k += gb) {
        launch_gj_block(st, nc, k, std::min(gb, nc - k), dA, dB, dPi, dV, dW);
        std::swap(dA, dB);
    }
    HIPCHECK(hipGetLastError());
    // the inverse of an SPD matrix has a positive diagonal
    std::vector<double> dg(nc);
    HIPCHECK(hipMemcpy2DAsync(dg.data(), sizeof(double), dA, (nc + 1) * sizeof(double), sizeof(double), nc,
                              hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    dfree(dPi);
    dfree(dV);
    dfree(dW);
    double dmax = 0.0;
    for (double v : dg) {
        if (!(v > 0.0) || !std::isfinite(v)) {
            dfree(dA);
            dfree(dB);
            throw std::runtime_error("coarse operator inversion failed");
        }
        dmax = std::max(dmax, v);
    }
    // FP16 storage: every entry of an SPD matrix is at most its largest diagonal entry, which
    // the power-of-two scale brings into [2^13, 2^14) (FP16 holds up to 65504)
    int ex = 0;
    (void)std::frexp(dmax, &ex);
    const double cscale = std::ldexp(1.0, ex - 14);
    dfree(s->cinv);
    dfree(s->cc);
    const int ld = (nc + 7) / 8 * 8;
    s->cinv = dalloc<_Float16>((size_t)nc * ld);
    s->cscale = cscale;
    s->cld = ld;
    s->cc = dalloc<float>(nc);
    dfree(s->cc64);
    s->cc64 = dalloc<double>(nc);
    launch_to_half(st, nc, ld, dA, cscale, s->cinv);
    HIPCHECK(hipMemsetAsync(s->cc, 0, nc * sizeof(float), st));
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    dfree(dA);
    dfree(dB);
    s->nc = nc;
    s->agg_shift = shift;
    s->coarse = true;
    return ncol;
}

extern "C" int fgpu_sys_set_coarse(fgpu_sys *s, int enable, int *colors) {
    if (!s) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        HIPCHECK(hipSetDevice(s->ctx->device));
        if (s->batches > 0) throw std::invalid_argument("the system is in a running batch");
        HIPCHECK(hipStreamSynchronize(s->ctx->stream));
        s->coarse = false;
        s->nc = 0;
        dfree(s->cinv);
        dfree(s->cc);
        if (colors) *colors = 0;
        if (!enable) return FGPU_OK;
        if (!s->lifted) throw std::invalid_argument("apply the Dirichlet conditions before building the coarse level");
        int shift;
        if (enable == 1) {
            // the coarse product reads nc^2 floats per iteration: aggregates grow with the mesh
            const int n = static_cast<fgpu_mesh *>(s->mesh)->n_dofs;
            shift = n <= 700000 ? 8 : (n <= 1500000 ? 9 : 10);
        } else if (enable == 256 || enable == 512 || enable == 1024) {
            shift = enable == 256 ? 8 : (enable == 512 ? 9 : 10);
        } else {
            throw std::invalid_argument("aggregate size must be 256, 512 or 1024 (or 1: by mesh size)");
        }
        const int ncol = build_coarse(s, shift);
        if (colors) *colors = ncol;
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_get_coarse(fgpu_sys *s, int *nc, int *block, float *einv) {
    if (!s || !nc) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        const fgpu_mesh *m = static_cast<const fgpu_mesh *>(s->mesh);
        *nc = s->coarse ? s->nc : 0;
        if (block)
            for (int d = 0; d < m->n_dofs; ++d) block[d] = m->idof_of_dof[d] >> s->agg_shift;
        if (einv && s->coarse) {
            HIPCHECK(hipSetDevice(s->ctx->device));
            std::vector<_Float16> h((size_t)s->nc * s->cld);
            d2h(h.data(), s->cinv, h.size(), s->ctx->stream);
            HIPCHECK(hipStreamSynchronize(s->ctx->stream));
            for (int r = 0; r < s->nc; ++r)
                for (int c = 0; c < s->nc; ++c)
                    einv[(size_t)r * s->nc + c] = (float)((double)(float)h[(size_t)r * s->cld + c] * s->cscale);
        }
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_minmax(fgpu_sys *s, double *mn, double *mx) {
    if (!s) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        HIPCHECK(hipSetDevice(s->ctx->device));
        launch_minmax(s->ctx->stream, m->n_dofs, s->x, s->tmp);
        double h[2];
        d2h(h, s->tmp, 2, s->ctx->stream);
        HIPCHECK(hipStreamSynchronize(s->ctx->stream));
        if (mn) *mn = h[0];
        if (mx) *mx = h[1];
        return FGPU_OK;
    });
}

extern "C" int fgpu_sys_solution_grad(fgpu_sys *s, double *grad) {
    if (!s || !grad) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        hipStream_t st = s->ctx->stream;
        HIPCHECK(hipSetDevice(s->ctx->device));
        const int n = m->n_dofs;
        double *dg = dalloc<double>(3 * (size_t)n);
        launch_grad_export(st, n, m->d_v2c, m->d_v2n, m->d_cells, m->d_X64, s->x, dg);
        std::vector<double> h(3 * (size_t)n);
        d2h(h.data(), dg, h.size(), st);
        HIPCHECK(hipStreamSynchronize(st));
        dfree(dg);
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) grad[3 * (size_t)m->dof_of_idof[i] + a] = h[3 * (size_t)i + a];
        return FGPU_OK;
    });
}

// ------------------------------------------------------------------------------------------
// batched solve
// ------------------------------------------------------------------------------------------
static double *flags_buffer(fgpu_ctx *c, size_t n) {
    if (n > c->flags_cap) {
        if (c->h_flags) HIPCHECK(hipHostFree(c->h_flags));
        c->h_flags = nullptr;
        HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&c->h_flags), n * sizeof(double), hipHostMallocDefault));
        c->flags_cap = n;
    }
    return c->h_flags;
}

// true while some system of the set has scalar `slot` equal to `want`
static bool any_flag(fgpu_ctx *c, const Desc *D, int n, int slot, double want, double *dflags) {
    double *h = flags_buffer(c, n);
    launch_collect(c->stream, D, n, slot, dflags);
    HIPCHECK(hipMemcpyAsync(h, dflags, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i)
        if (h[i] == want) return true;
    return false;
}

// preconditioner pieces of one solve's CG sets
struct SolveKind {
    bool mixed = false, coef = false;
    int cheb_deg = 0;        // FP64 CG with the Chebyshev-Jacobi preconditioner of this degree
    bool coarse64 = false;   // FP64 CG with the two-level correction
};

// the deg - 1 element applies and updates of the Chebyshev preconditioner (FP64 CG; D: the
// systems' CG or residual descriptors, DC: the same with the Chebyshev search vector as input)
static void cheb_steps(hipStream_t st, const Desc *D, const Desc *DC, const BatchTabs &t, bool coef, int deg) {
    for (int j = 1; j < deg; ++j) {
        launch_elem<double>(st, DC, t.etab, t.ne, coef, false, true, S_BETA, t.lds64);
        launch_gather_cheb(st, D, t.dtab, t.nd, j);
    }
}

// one CG iteration of every system of the set
template <class T>
static void cg_iter(hipStream_t st, const Desc *D, const Desc *DC, const BatchTabs &t, int n, const SolveKind &k,
                    int reliable) {
    launch_elem<T>(st, D, t.etab, t.ne, k.coef, false, true, S_BETA, sizeof(T) == 8 ? t.lds64 : t.lds32, t.affine32);
    launch_gather_cg<T>(st, D, t.dtab, t.nd);
    launch_scal_alpha(st, D, n);
    launch_update_cg<T>(st, D, t.dtab, t.nd);
    if (sizeof(T) == 4) {
        launch_coarse(st, D, nullptr, t.ctab, t.nct, 0, t.max_nc);
    } else {
        if (k.coarse64) launch_coarse64(st, D, t.ctab, t.nct, t.max_nc);
        if (k.cheb_deg > 1) cheb_steps(st, D, DC, t, k.coef, k.cheb_deg);
    }
    launch_scal_beta(st, D, n, reliable);
}

// FP64 residual of the lifted system (R64: residual descriptors, u = x with the constrained
// dofs zeroed) and the resulting state; in an FP64 solve also the first preconditioned
// residual (coarse correction or Chebyshev steps)
static void fp64_residual(hipStream_t st, const Desc *R64, const Desc *D32, const Desc *DC, const BatchTabs &t, int n,
                          const SolveKind &k, int mixed) {
    launch_elem<double>(st, R64, t.etab, t.ne, k.coef, true, true, S_BETA, t.lds64);
    launch_gather_res<double>(st, R64, t.dtab, t.nd);
    if (!mixed) {
        if (k.coarse64) launch_coarse64(st, R64, t.ctab, t.nct, t.max_nc);
        if (k.cheb_deg > 1) cheb_steps(st, R64, DC, t, k.coef, k.cheb_deg);
    }
    launch_scal_res(st, R64, D32, n, mixed);
}

// reliable update of the systems whose FP32 CG waits: fold the FP32 correction into x64,
// recompute the FP64 residual, resume the FP32 CG from it
static void reliable_update(hipStream_t st, const Desc *D32, const Desc *R64, const BatchTabs &t, int n,
                            const SolveKind &k) {
    launch_mixed_acc(st, D32, R64, t.dtab, t.nd);
    fp64_residual(st, R64, D32, nullptr, t, n, k, 1);
    launch_mixed_restart(st, D32, R64, t.dtab, t.nd, 0);
    launch_coarse(st, D32, R64, t.ctab, t.nct, 2, t.max_nc);
    launch_scal_restart(st, D32, R64, n, 0);
}

struct ChunkArgs {
    SolveKind kind;
    int k, u, n;
    const Desc *D64, *R64, *D32, *DC;
    const BatchTabs *t;
};

static void chunk(hipStream_t st, const ChunkArgs &a) {
    for (int j = 0; j < a.k; ++j) {
        if (!a.kind.mixed) {
            cg_iter<double>(st, a.D64, a.DC, *a.t, a.n, a.kind, 0);
        } else {
            cg_iter<float>(st, a.D32, nullptr, *a.t, a.n, a.kind, 1);
            if ((j + 1) % a.u == 0) reliable_update(st, a.D32, a.R64, *a.t, a.n, a.kind);
        }
    }
}

static bool batch_coef(const std::vector<fgpu_sys *> &sys, fgpu_ctx *c) {
    bool coef = false;
    for (auto *s : sys) {
        if (!s || s->ctx != c) throw std::invalid_argument("system belongs to another context");
        coef = coef || s->have_coef;
    }
    for (auto *s : sys)
        if (s->have_coef != coef) throw std::invalid_argument("a batch mixes coefficient and plain systems");
    return coef;
}

static void check_precond(const fgpu_solve_opts *o) {
    if (o->precond != FGPU_PC_NONE && o->precond != FGPU_PC_JACOBI && o->precond != FGPU_PC_CHEBYSHEV)
        throw std::invalid_argument("unknown preconditioner");
    if (o->precond == FGPU_PC_CHEBYSHEV && o->precision != FGPU_FP64)
        throw std::invalid_argument("the Chebyshev-Jacobi preconditioner is for FGPU_FP64 solves");
}

// Descriptor sets of a solve: [0] FP64 CG (input z + beta p), [1] FP64 residual (input x),
// [2] FP32 CG of the mixed solver, [3] FP64 Chebyshev apply (input: the Chebyshev search
// vector).  The FP64 sets carry the solve's preconditioner: identity, Jacobi (plus the
// two-level correction of systems that have one) or Chebyshev-Jacobi of degree cheb_deg; the
// FP32 CG uses Jacobi plus the two-level correction unless no preconditioner is asked for.
// hist: record the first CG coefficients (Lanczos estimate).
static SolveKind make_sets(const std::vector<fgpu_sys *> &sys, std::vector<std::vector<Desc>> &sets,
                           const fgpu_solve_opts *o, int cheb_deg = 0, bool hist = false) {
    SolveKind k;
    k.mixed = o->precision == FGPU_MIXED;
    sets.assign(4, {});
    for (auto *s : sys) {
        Desc d0 = make_desc64(s, false), d1 = make_desc64(s, true), d2 = make_desc32(s, true), dc = make_desc64(s, false);
        if (!k.mixed) {
            d0.pc = d1.pc = o->precond;
            if (o->precond == FGPU_PC_CHEBYSHEV) {
                for (Desc *d : {&d0, &d1}) {
                    d->cd = s->cd;
                    d->cres = s->cres;
                    d->ccoef = s->ccoef;
                    d->cheb_deg = cheb_deg;
                }
                k.cheb_deg = cheb_deg;
            }
            if (o->precond == FGPU_PC_JACOBI && s->coarse) {
                for (Desc *d : {&d0, &d1}) {
                    d->cc = s->cc64;
                    d->cinv = s->cinv;
                    d->cscale = s->cscale;
                    d->cld = s->cld;
                    d->nc = s->nc;
                    d->agg_shift = s->agg_shift;
                }
                d0.cc_in = 1;          // the residual set's element input is x, not z
                k.coarse64 = true;
            }
            if (hist) d0.hist = s->hist;
        } else if (o->precond == FGPU_PC_NONE) {
            d2.pc = FGPU_PC_NONE;
            d2.cc = nullptr;
        }
        dc.ua = s->cd;
        dc.ub = nullptr;
        dc.cc = nullptr;
        sets[0].push_back(d0);
        sets[1].push_back(d1);
        sets[2].push_back(d2);
        sets[3].push_back(dc);
    }
    return k;
}

// Starts the solves of a set of systems: scalars, x = g on constrained dofs, FP64 residual of
// the lifted system and the first preconditioned residual (FP64), or the FP32 CG state
// (mixed).  t: the set's uploaded tables.
static void start_solves(fgpu_ctx *c, const std::vector<fgpu_sys *> &sys, const fgpu_solve_opts *o, const BatchTabs &t,
                         const SolveKind &k) {
    const int n = (int)sys.size();
    const bool mixed = k.mixed;
    const int max_it = std::max(0, o->max_it);
    const double delta = o->inner_tol > 0 && o->inner_tol < 1 ? o->inner_tol : 0.1;
    std::vector<double> h((size_t)2 * kScal * n, 0.0);
    for (int i = 0; i < n; ++i) {
        double *s64 = &h[(size_t)2 * kScal * i], *s32 = s64 + kScal;
        s64[S_ACTIVE] = 1.0;
        s64[S_MAXIT] = max_it;
        s64[S_TOL] = o->tol;
        s64[S_REL] = o->relative ? 1.0 : 0.0;
        s32[S_MAXIT] = max_it;
        s32[S_DELTA] = delta * delta;
    }
    for (int i = 0; i < n; ++i) {
        fgpu_sys *s = sys[i];
        stage_h2d(c, s->scal, &h[(size_t)2 * kScal * i], 2 * kScal * sizeof(double));
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        launch_enforce_dirichlet(c->stream, m->n_dofs, s->d_dmask, s->g, s->x);
    }
    const Desc *R64 = t.sets[1], *D32 = t.sets[2], *DC = t.sets[3];
    fp64_residual(c->stream, R64, D32, DC, t, n, k, mixed ? 1 : 0);
    if (mixed) {
        launch_mixed_restart(c->stream, D32, R64, t.dtab, t.nd, 1);
        launch_coarse(c->stream, D32, R64, t.ctab, t.nct, 1, t.max_nc);
        launch_scal_restart(c->stream, D32, R64, n, 1);
    }
    HIPCHECK(hipGetLastError());
}

static void read_stats(fgpu_sys *s, bool mixed, fgpu_solve_stats *st) {
    double hs[2 * kScal];
    HIPCHECK(hipMemcpy(hs, s->scal, sizeof(hs), hipMemcpyDeviceToHost));
    const double *s64 = hs, *s32 = hs + kScal;
    const int it = (int)(mixed ? s32[S_IT] : s64[S_IT]);
    st->iterations = s64[S_CONV] != 0.0 ? it : -std::max(it, 1);
    st->outer = mixed ? (int)s32[S_NUPD] + 1 : 1;
    st->residual = std::sqrt(std::max(0.0, s64[S_RR]));
    st->rhs_norm = std::sqrt(std::max(0.0, s64[S_BB]));
}

// CG of every system to its own convergence, chunks of iterations between host checks
static void run_solves(fgpu_ctx *c, const std::vector<fgpu_sys *> &sys, const fgpu_solve_opts *o, int cheb_deg, bool hist,
                       fgpu_solve_stats *stats) {
    const int n = (int)sys.size();
    if (n == 0) return;
    const bool coef = batch_coef(sys, c);
    const int k = std::max(2, o->check_every > 0 ? o->check_every : 32);
    const int max_it = std::max(0, o->max_it);
    for (auto *s : sys) ensure_diag(s);
    std::vector<std::vector<Desc>> sets;
    SolveKind kind = make_sets(sys, sets, o, cheb_deg, hist);
    kind.coef = coef;
    BatchTabs t = upload_batch(c, sets, sys);
    start_solves(c, sys, o, t, kind);
    double *dflags = nullptr;
    HIPCHECK(hipMalloc(&dflags, std::max(1, n) * sizeof(double)));
    ChunkArgs a{kind, k, 8, n, t.sets[0], t.sets[1], t.sets[2], t.sets[3], &t};
    // a system waiting for its reliable update idles at most u iterations per update
    const int max_chunks = 2 * (max_it / k) + 16;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    bool running = any_flag(c, t.sets[1], n, S_DONE, 0.0, dflags);
    for (int ch = 0; running && ch < max_chunks; ++ch) {
        if (o->use_graph) {
            if (!exec) {
                HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                chunk(c->stream, a);
                HIPCHECK(hipStreamEndCapture(c->stream, &graph));
                HIPCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            }
            HIPCHECK(hipGraphLaunch(exec, c->stream));
        } else {
            chunk(c->stream, a);
        }
        running = any_flag(c, t.sets[1], n, S_DONE, 0.0, dflags);
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipFree(dflags));
    if (stats)
        for (int i = 0; i < n; ++i) read_stats(sys[i], kind.mixed, &stats[i]);
}

static int solve_batch(fgpu_ctx *c, int n, fgpu_sys *const *systems, const fgpu_solve_opts *o, fgpu_solve_stats *stats) {
    HIPCHECK(hipSetDevice(c->device));
    check_precond(o);
    std::vector<fgpu_sys *> sys(systems, systems + n);
    if (o->precond != FGPU_PC_CHEBYSHEV) {
        run_solves(c, sys, o, 0, false, stats);
        return FGPU_OK;
    }
    // Chebyshev-Jacobi: lambda_max(D^-1 A) of every system whose coordinates changed since its
    // last estimate, from the Lanczos matrix of kLanczosSteps Jacobi-PCG iterations (which
    // also start its solve), then the CG with the Chebyshev preconditioner from that iterate
    std::vector<fgpu_solve_stats> st(n), pre(n);
    std::vector<char> done(n, 0);
    std::vector<int> est;
    for (int i = 0; i < n; ++i) {
        fgpu_sys *s = sys[i];
        const fgpu_mesh *m = static_cast<const fgpu_mesh *>(s->mesh);
        if (!(s->cheb_lmax > 0) || s->cheb_gen != m->coord_gen || !s->diag_valid) est.push_back(i);
        if (!s->cd) {
            s->cd = dalloc<double>(m->n_dofs);
            s->cres = dalloc<double>(m->n_dofs);
            s->ccoef = dalloc<double>(2 * kChebMax + 1);
        }
    }
    if (!est.empty()) {
        std::vector<fgpu_sys *> es;
        for (int i : est) {
            if (!sys[i]->hist) sys[i]->hist = dalloc<double>(2 * kHist);
            es.push_back(sys[i]);
        }
        fgpu_solve_opts oj = *o;
        oj.precond = FGPU_PC_JACOBI;
        oj.max_it = std::min(o->max_it, kLanczosSteps);
        oj.check_every = kLanczosSteps;
        std::vector<fgpu_solve_stats> es_st(es.size());
        run_solves(c, es, &oj, 0, true, es_st.data());
        for (size_t e = 0; e < es.size(); ++e) {
            const int i = est[e];
            fgpu_sys *s = sys[i];
            pre[i] = es_st[e];
            const int m_its = std::abs(es_st[e].iterations);
            if (es_st[e].iterations > 0 || m_its >= o->max_it) {
                done[i] = 1;       // converged (or out of iterations) in the Jacobi steps
                continue;
            }
            double h[2 * kHist];
            HIPCHECK(hipMemcpy(h, s->hist, sizeof(h), hipMemcpyDeviceToHost));
            const double lm = lanczos_lmax(h, m_its);
            if (!(lm > 0) || !std::isfinite(lm)) throw std::runtime_error("Lanczos estimate of lambda_max failed");
            s->cheb_lmax = lm;
            s->cheb_gen = static_cast<const fgpu_mesh *>(s->mesh)->coord_gen;
        }
    }
    std::vector<fgpu_sys *> rs;
    std::vector<int> ri;
    int nmax = 0;
    for (int i = 0; i < n; ++i)
        if (!done[i]) {
            rs.push_back(sys[i]);
            ri.push_back(i);
            nmax = std::max(nmax, static_cast<const fgpu_mesh *>(sys[i]->mesh)->n_dofs);
        }
    const int deg = o->cheb_degree > 0 ? std::min(o->cheb_degree, kChebMax) : cheb_auto_degree(nmax);
    for (fgpu_sys *s : rs) {
        std::vector<double> cf;
        cheb_coefficients(s->cheb_lmax, deg, cf);
        stage_h2d(c, s->ccoef, cf.data(), cf.size() * sizeof(double));
    }
    if (!rs.empty()) {
        fgpu_solve_opts oc = *o;
        int used = 0;
        for (int i : ri) used = std::max(used, std::abs(pre[i].iterations));
        oc.max_it = std::max(1, o->max_it - used);
        std::vector<fgpu_solve_stats> rs_st(rs.size());
        run_solves(c, rs, &oc, deg, false, rs_st.data());
        for (size_t e = 0; e < rs.size(); ++e) st[ri[e]] = rs_st[e];
    }
    if (stats)
        for (int i = 0; i < n; ++i) {
            if (done[i]) {
                stats[i] = pre[i];
                continue;
            }
            fgpu_solve_stats r = st[i];
            const int p = std::abs(pre[i].iterations);
            r.iterations = r.iterations > 0 ? r.iterations + p : r.iterations - p;
            r.outer += pre[i].outer;
            stats[i] = r;
        }
    return FGPU_OK;
}

// ------------------------------------------------------------------------------------------
// continuous batching: systems join and leave a running batch between steps
// ------------------------------------------------------------------------------------------
static double time_launches(fgpu_ctx *c, const Desc *D, const BatchTabs &t, bool coef, int which, int precision, int reps);

struct fgpu_batch {
    fgpu_ctx *ctx = nullptr;
    fgpu_solve_opts o{};
    std::vector<fgpu_sys *> sys;
    int coef = -1;                // -1: no member yet
    bool dirty = true;            // membership changed since the tables were uploaded
    void *dtab = nullptr;
    size_t cap = 0;
    BatchTabs t;
    double *dflags = nullptr;
    int flags_cap = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_k = 0;
    std::vector<double> done;     // S_DONE of each member after the last step (valid while
    bool done_valid = false;      // the membership has not changed since)
    // an asynchronous step in flight: its members and their S_DONE, copied to pinned memory
    // when the step ends (event); members that finished are known once it is waited for
    bool pending = false;
    std::vector<fgpu_sys *> pend_sys;
    double *h_pend = nullptr;
    int pend_cap = 0;
    hipEvent_t pend_ev = nullptr;
    std::vector<std::pair<fgpu_sys *, double>> last;   // (member, S_DONE) of the last waited step
    bool last_valid = false;
    std::vector<fgpu_sys *> joined;   // added while a step was in flight (not its members, even
                                      // at the address of one that left)
    SolveKind kind;                   // preconditioner pieces of the uploaded sets
    // the batch's own stream: its steps run beside other batches' (their kernels fill each
    // other's idle CUs: scalar kernels, launch tails); set-up work stays on the context's
    // stream, and every step first waits for what was enqueued there (ev_ctx)
    hipStream_t st = nullptr;
    hipEvent_t ev_ctx = nullptr;
};

static void batch_drop_graph(fgpu_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

extern "C" int fgpu_batch_create(fgpu_ctx *c, const fgpu_solve_opts *o, fgpu_batch **out) {
    if (!c || !o || !out) return FGPU_EINVAL;
    if (o->precond != FGPU_PC_NONE && o->precond != FGPU_PC_JACOBI) {
        c->err = "continuous batches take FGPU_PC_NONE or FGPU_PC_JACOBI (the two-level correction of members "
                 "that have one included)";
        return FGPU_EINVAL;
    }
    fgpu_batch *b = new (std::nothrow) fgpu_batch();
    if (!b) return FGPU_ENOMEM;
    const int rc = guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        HIPCHECK(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));      // default priority
        HIPCHECK(hipEventCreateWithFlags(&b->ev_ctx, hipEventDisableTiming));
        return FGPU_OK;
    });
    if (rc != FGPU_OK) {
        if (b->st) (void)hipStreamDestroy(b->st);
        delete b;
        return rc;
    }
    b->ctx = c;
    b->o = *o;
    c->refs++;
    *out = b;
    return FGPU_OK;
}

extern "C" void fgpu_batch_destroy(fgpu_batch *b) {
    if (!b) return;
    fgpu_ctx *c = b->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (b->st) (void)hipStreamSynchronize(b->st);
    batch_drop_graph(b);
    for (auto *s : b->sys) s->batches--;
    if (b->st) (void)hipStreamDestroy(b->st);
    if (b->ev_ctx) (void)hipEventDestroy(b->ev_ctx);
    if (b->dtab) (void)hipFree(b->dtab);
    if (b->dflags) (void)hipFree(b->dflags);
    if (b->h_pend) (void)hipHostFree(b->h_pend);
    if (b->pend_ev) (void)hipEventDestroy(b->pend_ev);
    delete b;
    ctx_release(c);
}

extern "C" int fgpu_batch_size(const fgpu_batch *b) { return b ? (int)b->sys.size() : 0; }

extern "C" int fgpu_batch_add(fgpu_batch *b, fgpu_sys *s) {
    if (!b || !s) return FGPU_EINVAL;
    fgpu_ctx *c = b->ctx;
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        for (auto *x : b->sys)
            if (x == s) throw std::invalid_argument("system already in the batch");
        const bool coef = batch_coef({s}, c);
        if (b->coef >= 0 && (int)coef != b->coef) throw std::invalid_argument("a batch mixes coefficient and plain systems");
        ensure_diag(s);
        std::vector<std::vector<Desc>> sets;
        SolveKind kind = make_sets({s}, sets, &b->o);
        kind.coef = coef;
        BatchTabs t = upload_batch(c, sets, {s});
        start_solves(c, {s}, &b->o, t, kind);
        for (auto &e : b->last)
            if (e.first == s) e.first = nullptr;
        if (b->pending) b->joined.push_back(s);
        b->coef = coef;
        b->sys.push_back(s);
        s->batches++;
        b->dirty = true;
        b->done_valid = false;
        return FGPU_OK;
    });
}

// enqueues one step of every member and the collection of their S_DONE into b->dflags
static void batch_launch(fgpu_batch *b, int iterations) {
    fgpu_ctx *c = b->ctx;
    const int n = (int)b->sys.size();
    {
        const bool mixed = b->o.precision == FGPU_MIXED;
        const int k = mixed ? (iterations + 7) / 8 * 8 : iterations;
        if (b->dirty) {
            batch_drop_graph(b);
            std::vector<std::vector<Desc>> sets;
            b->kind = make_sets(b->sys, sets, &b->o);
            b->kind.coef = b->coef > 0;
            b->t = upload_batch(c, sets, b->sys, &b->dtab, &b->cap);
            if (n > b->flags_cap) {
                if (b->dflags) HIPCHECK(hipFree(b->dflags));
                b->dflags = nullptr;
                HIPCHECK(hipMalloc(&b->dflags, 2 * n * sizeof(double)));
                b->flags_cap = 2 * n;
            }
            b->dirty = false;
        }
        ChunkArgs a{b->kind, k, 8, n, b->t.sets[0], b->t.sets[1], b->t.sets[2], b->t.sets[3], &b->t};
        // the members' set-up and the tables were enqueued on the context's stream
        HIPCHECK(hipEventRecord(b->ev_ctx, c->stream));
        HIPCHECK(hipStreamWaitEvent(b->st, b->ev_ctx, 0));
        if (b->o.use_graph) {
            if (!b->exec || b->graph_k != k) {
                batch_drop_graph(b);
                HIPCHECK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
                chunk(b->st, a);
                HIPCHECK(hipStreamEndCapture(b->st, &b->graph));
                HIPCHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
                b->graph_k = k;
            }
            HIPCHECK(hipGraphLaunch(b->exec, b->st));
        } else {
            chunk(b->st, a);
        }
        launch_collect(b->st, b->t.sets[1], n, S_DONE, b->dflags);
    }
}

static void batch_wait(fgpu_batch *b);

extern "C" int fgpu_batch_step(fgpu_batch *b, int iterations, int *n_running) {
    if (!b || iterations <= 0) return FGPU_EINVAL;
    fgpu_ctx *c = b->ctx;
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        batch_wait(b);
        const int n = (int)b->sys.size();
        if (n_running) *n_running = 0;
        if (n == 0) return FGPU_OK;
        batch_launch(b, iterations);
        double *h = flags_buffer(c, n);
        HIPCHECK(hipMemcpyAsync(h, b->dflags, n * sizeof(double), hipMemcpyDeviceToHost, b->st));
        HIPCHECK(hipStreamSynchronize(b->st));
        int running = 0;
        for (int i = 0; i < n; ++i) running += h[i] == 0.0;
        if (n_running) *n_running = running;
        b->done.assign(h, h + n);
        b->done_valid = true;
        b->last_valid = false;
        return FGPU_OK;
    });
}

// waits for the asynchronous step in flight, if any, and keeps its members' S_DONE
static void batch_wait(fgpu_batch *b) {
    if (!b->pending) return;
    HIPCHECK(hipEventSynchronize(b->pend_ev));
    b->pending = false;
    b->last.clear();
    for (size_t i = 0; i < b->pend_sys.size(); ++i)
        if (std::find(b->joined.begin(), b->joined.end(), b->pend_sys[i]) == b->joined.end())
            b->last.emplace_back(b->pend_sys[i], b->h_pend[i]);
    b->joined.clear();
    b->last_valid = true;
}

extern "C" int fgpu_batch_step_async(fgpu_batch *b, int iterations) {
    if (!b || iterations <= 0) return FGPU_EINVAL;
    fgpu_ctx *c = b->ctx;
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        if (b->pending) throw std::invalid_argument("a step is in flight: fgpu_batch_wait first");
        const int n = (int)b->sys.size();
        if (n == 0) return FGPU_OK;
        if (n > b->pend_cap) {
            if (b->h_pend) HIPCHECK(hipHostFree(b->h_pend));
            b->h_pend = nullptr;
            b->pend_cap = 0;
            HIPCHECK(hipHostMalloc(reinterpret_cast<void **>(&b->h_pend), 2 * n * sizeof(double), hipHostMallocDefault));
            b->pend_cap = 2 * n;
        }
        if (!b->pend_ev) HIPCHECK(hipEventCreateWithFlags(&b->pend_ev, hipEventDisableTiming));
        batch_launch(b, iterations);
        HIPCHECK(hipMemcpyAsync(b->h_pend, b->dflags, n * sizeof(double), hipMemcpyDeviceToHost, b->st));
        HIPCHECK(hipEventRecord(b->pend_ev, b->st));
        b->pend_sys = b->sys;
        b->joined.clear();
        b->pending = true;
        b->done_valid = false;
        return FGPU_OK;
    });
}

extern "C" int fgpu_batch_wait(fgpu_batch *b, int *n_running) {
    if (!b) return FGPU_EINVAL;
    fgpu_ctx *c = b->ctx;
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        const bool had = b->pending;
        batch_wait(b);
        if (n_running) {
            *n_running = 0;
            if (had)
                for (auto &e : b->last) *n_running += e.second == 0.0;
        }
        return FGPU_OK;
    });
}

extern "C" int fgpu_batch_time_kernel(fgpu_batch *b, int which, int reps, double *avg_ms) {
    if (!b || !avg_ms || reps <= 0) return FGPU_EINVAL;
    fgpu_ctx *c = b->ctx;
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        if (b->sys.empty()) throw std::invalid_argument("empty batch");
        std::vector<std::vector<Desc>> sets;
        make_sets(b->sys, sets, &b->o);
        BatchTabs t = upload_batch(c, sets, b->sys);
        const bool mixed = b->o.precision == FGPU_MIXED;
        *avg_ms = time_launches(c, mixed ? t.sets[2] : t.sets[1], t, b->coef > 0, which, b->o.precision, reps);
        return FGPU_OK;
    });
}

extern "C" int fgpu_batch_retire(fgpu_batch *b, int max, fgpu_sys **out, fgpu_solve_stats *stats, int *n_out) {
    if (!b || !n_out || (max > 0 && !out)) return FGPU_EINVAL;
    fgpu_ctx *c = b->ctx;
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        *n_out = 0;
        std::vector<fgpu_sys *> keep;
        const bool mixed = b->o.precision == FGPU_MIXED;
        const bool known = b->done_valid && b->done.size() == b->sys.size();
        for (size_t i = 0; i < b->sys.size(); ++i) {
            fgpu_sys *s = b->sys[i];
            double done = 0;
            if (known) {
                done = b->done[i];
            } else if (b->last_valid) {
                // the last waited asynchronous step; a member that joined after it is running
                for (auto &e : b->last)
                    if (e.first == s) done = e.second;
            } else {
                HIPCHECK(hipMemcpy(&done, s->scal + S_DONE, sizeof(double), hipMemcpyDeviceToHost));
            }
            if (done != 0.0 && *n_out < max) {
                out[*n_out] = s;
                s->batches--;
                for (auto &e : b->last)
                    if (e.first == s) e.first = nullptr;   // a later add of s starts afresh
                if (stats) read_stats(s, mixed, &stats[*n_out]);
                ++*n_out;
            } else {
                keep.push_back(s);
            }
        }
        if (keep.size() != b->sys.size()) {
            b->sys.swap(keep);
            b->dirty = true;
            b->done_valid = false;
            if (b->sys.empty()) b->coef = -1;
        }
        return FGPU_OK;
    });
}

extern "C" int fgpu_solve_batch(fgpu_ctx *c, int n, fgpu_sys *const *systems, const fgpu_solve_opts *o, fgpu_solve_stats *stats) {
    if (!c || n <= 0 || !systems || !o) return FGPU_EINVAL;
    return guarded(c, [&] { return solve_batch(c, n, systems, o, stats); });
}

extern "C" int fgpu_sys_solve(fgpu_sys *s, const fgpu_solve_opts *o, fgpu_solve_stats *stats) {
    if (!s || !o) return FGPU_EINVAL;
    return guarded(s->ctx, [&] { return solve_batch(s->ctx, 1, &s, o, stats); });
}

// ------------------------------------------------------------------------------------------
// kernel timing (benchmark helper): average duration of `reps` back-to-back launches of one
// kernel of the solve, measured with HIP events on the context stream
// ------------------------------------------------------------------------------------------
// average duration of `reps` back-to-back launches of the element apply (which 0), the dof
// gather (which 1) or the coarse product of the two-level FP32 CG (which 2) over descriptor
// set D, HIP events on the context's stream
static double time_launches(fgpu_ctx *c, const Desc *D, const BatchTabs &t, bool coef, int which, int precision, int reps) {
    if (which == 2 && (precision != FGPU_MIXED || t.nct == 0)) throw std::invalid_argument("no coarse level in the batch");
    auto one = [&] {
        if (which == 2) {
            launch_coarse(c->stream, D, nullptr, t.ctab, t.nct, 0, t.max_nc);
        } else if (which == 0) {
            if (precision == FGPU_MIXED)
                launch_elem<float>(c->stream, D, t.etab, t.ne, coef, false, false, S_BETA, t.lds32, t.affine32);
            else launch_elem<double>(c->stream, D, t.etab, t.ne, coef, false, false, S_BETA, t.lds64);
        } else {
            if (precision == FGPU_MIXED) launch_gather_plain<float>(c->stream, D, t.dtab, t.nd, false);
            else launch_gather_plain<double>(c->stream, D, t.dtab, t.nd, false);
        }
    };
    one();
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, c->stream));
    for (int i = 0; i < reps; ++i) one();
    HIPCHECK(hipEventRecord(e1, c->stream));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}

extern "C" int fgpu_sys_time_kernel(fgpu_sys *s, int which, int precision, int reps, double *avg_ms) {
    if (!s || !avg_ms || reps <= 0) return FGPU_EINVAL;
    return guarded(s->ctx, [&] {
        fgpu_ctx *c = s->ctx;
        HIPCHECK(hipSetDevice(c->device));
        ensure_diag(s);
        Desc d = precision == FGPU_MIXED ? make_desc32(s) : make_desc64(s, true);
        // inputs: the current z / x vectors; outputs: the slot buffer and q
        BatchTabs t = upload_batch(c, {{d}}, {s});
        *avg_ms = time_launches(c, t.sets[0], t, s->have_coef, which, precision, reps);
        return FGPU_OK;
    });
}

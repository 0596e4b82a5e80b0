// Transfer of the field solution to atoms on the GPU: nodal solution (Interpolator), point
// location and interpolation (InterpolatorCells, SolutionReader).  Compiled without
// floating-point contraction: cell indices and interpolated values are bit-identical to the
// reference's sequential code on the same inputs.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "femocs_gpu.h"
#include "fgpu_internal.h"
#include "fgpu_interp.h"
#include "fgpu_kernels.h"

using namespace fgpu;

struct fgpu_interp {
    fgpu_ctx *ctx = nullptr;
    int n_nodes = 0, n_tetnodes = 0, n_tets = 0, n_hexes = 0, region = 3;
    double empty = 0, decay = 0;
    LinTetHost L;
    Bvh B;
    // device tables
    double *xyz = nullptr, *det0 = nullptr, *det = nullptr, *cen = nullptr;
    int *tets = nullptr, *c10 = nullptr, *hexes = nullptr, *marker = nullptr;
    int *child = nullptr, *prim = nullptr;
    int *esc = nullptr;                     // escape links of the tet tree
    float *box = nullptr, *cbox = nullptr, *leafbox = nullptr;
    int *n2c_ptr = nullptr, *n2c = nullptr, *vor_ptr = nullptr, *vor = nullptr;
    double *sol = nullptr;                  // 5 per node: vector xyz, scalar1, scalar2
    int *node_idof = nullptr;               // node -> internal dof of `map_mesh`, -1 outside
    unsigned long long map_mesh = 0;        // uid of that mesh
    // point buffers (grown on demand)
    double *pts = nullptr, *out = nullptr;
    int4 *cand = nullptr, *cand2 = nullptr;
    int *tet = nullptr, *cells = nullptr, *cells2 = nullptr;
    size_t cap = 0;
    // device chained guess (grown on demand): containing cells and tables per point, per
    // workgroup of 256 points its composition and starting position, the fallback flag
    int *chain_lst = nullptr, *chain_entry = nullptr, *chain_flag = nullptr;
    unsigned char *chain_tab = nullptr, *chain_bt = nullptr;
    size_t chain_cap = 0;
    bool host_chain = false;                // FGPU_INTERP_HOST_CHAIN
    // Morton order of the query points (grown on demand): keys, point indices, sort storage
    unsigned long long *skey = nullptr, *skey2 = nullptr;
    int *sidx = nullptr, *sperm = nullptr;
    void *stmp = nullptr;
    size_t sort_cap = 0, stmp_bytes = 0;
    // surface faces (LinearTriangles), set by fgpu_interp_set_triangles
    std::vector<double> hxyz;               // host copy of the coordinates, fetched for the faces
    LinTriHost T;
    Bvh TB;
    double *tri_tab = nullptr, *tri_cen = nullptr;
    int *tri_child = nullptr, *tri_prim = nullptr, *tri_nodes = nullptr, *tri_esc = nullptr;
    int *tri_nbr_ptr = nullptr, *tri_nbr = nullptr;   // face neighbour lists (LinearTriangles)
    int *tri2tet = nullptr, *quad2hex = nullptr;       // the two tets of a face, hexes of a quad
    float *tri_box = nullptr, *tri_cbox = nullptr, *tri_leafbox = nullptr;
    // quadrangles: nodes, 6-node faces, centroids and their tree
    Bvh QB;
    double *quad_cen = nullptr;
    int *quad_nodes = nullptr, *tri6 = nullptr, *quad_child = nullptr, *quad_prim = nullptr;
    float *quad_cbox = nullptr;
    std::vector<int> hkind, htets;
    // the region's hexes in hex order: solver cell c is hex region_hex[c]
    // (Interpolator::deal2femocs); host copy and device table
    std::vector<int> region_hex;
    int *d_region_hex = nullptr;
    int *d_hex_cell = nullptr;              // inverse: solver cell of each hex, -1 outside
    int *d_nbr_ptr = nullptr, *d_nbr = nullptr;   // tet neighbour lists (LinearTetrahedra)
    int n_nbr = 0, n_n2c = 0, n_vor = 0;          // lengths of the neighbour, node -> hex, Voronoi lists
    bool L_on_host = false;                       // host copy L of the tet tables is filled
    // space-charge deposit: API cell -> internal cell of mesh `sc_mesh`, checked against
    // the region's hexes once per mesh
    int *sc_icell = nullptr;
    unsigned long long sc_mesh = 0;
};

namespace {

struct HipErr : std::runtime_error {
    using std::runtime_error::runtime_error;
};
#define IPCHECK(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) throw HipErr(std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

template <class T> T *upload(const T *h, size_t n, hipStream_t s) {
    void *p = nullptr;
    IPCHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    if (n) IPCHECK(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    return static_cast<T *>(p);
}
template <class T> void release(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <class F> int guard(fgpu_interp *ip, F &&f) {
    try {
        return f();
    } catch (const HipErr &e) {
        if (ip && ip->ctx) ip->ctx->err = e.what();
        return FGPU_EHIP;
    } catch (const std::bad_alloc &) {
        if (ip && ip->ctx) ip->ctx->err = "host allocation failed";
        return FGPU_ENOMEM;
    } catch (const std::exception &e) {
        if (ip && ip->ctx) ip->ctx->err = e.what();
        return FGPU_EINVAL;
    }
}

// ------------------------------------------------------------------------------------------
// device helpers: same operation order as the reference (no contraction in this file)
// ------------------------------------------------------------------------------------------
struct D3 {
    double x, y, z;
};
__device__ inline D3 mk(double x, double y, double z) { return {x, y, z}; }
__device__ inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3 scl(D3 a, double r) { return {a.x * r, a.y * r, a.z * r}; }
__device__ inline D3 dvd(D3 a, double r) { return {a.x / r, a.y / r, a.z / r}; }
__device__ inline D3 ldx(const double *xyz, int i) { return {xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2]}; }
__device__ inline double det3(D3 a, D3 b, D3 c) {
    return a.x * (b.y * c.z - c.y * b.z) - b.x * (a.y * c.z - c.y * a.z) + c.x * (a.y * b.z - b.y * a.z);
}
__device__ inline double dot4(const double *p, const double *d) { return p[0] * d[0] + p[1] * d[1] + p[2] * d[2] + 1.0 * d[3]; }

// LinearTetrahedra::point_in_cell (src/InterpolatorCells.cpp:631-647)
__device__ inline bool pic(const double *det0, const double *det, const double *p, int t) {
    const double *D = det + 16 * (size_t)t;
    const double d0 = det0[t];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (d0 * dot4(p, D + 4 * k) < -kZero) return false;
    return true;
}
// LinearTetrahedra::shape_functions (src/InterpolatorCells.cpp:649-661): bcc + 1e-15
__device__ inline void bcc(const double *det0, const double *det, const double *p, int t, double b[4]) {
    const double *D = det + 16 * (size_t)t;
#pragma unroll
    for (int k = 0; k < 4; ++k) b[k] = kZero + det0[t] * dot4(p, D + 4 * k);
}

__device__ inline bool in_box(const float *b, const double *p) {
    return p[0] >= (double)b[0] && p[1] >= (double)b[1] && p[2] >= (double)b[2] && p[0] <= (double)b[3] &&
           p[1] <= (double)b[4] && p[2] <= (double)b[5];
}
// lower bound of the squared distance from p to any point of the box, rounded like the
// centroid distances it is compared with
__device__ inline double box_d2(const float *b, const double *p) {
    double g[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double lo = (double)b[a], hi = (double)b[3 + a];
        g[a] = p[a] < lo ? lo - p[a] : (p[a] > hi ? p[a] - hi : 0.0);
    }
    return g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
}

// ------------------------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------------------------
// Cell families the point search runs over: linear tets (searchable by the marker of the
// search region) and the surface faces (all searchable).
struct TetCells {
    const double *det0, *det;
    const int *marker;
    __device__ bool contains(const double *p, int t) const { return pic(det0, det, p, t); }
    __device__ bool searchable(int t) const { return marker[t] == 0; }
};

// LinearTriangles::point_in_cell (src/InterpolatorCells.cpp:1639-1650) on the 16-double rows
struct TriCells {
    const double *tab;
    __device__ bool contains(const double *p, int f) const {
        const double *t = tab + 16 * (size_t)f;
        const D3 tv = sub(mk(p[0], p[1], p[2]), mk(t[0], t[1], t[2]));
        const double u = tv.x * t[9] + tv.y * t[10] + tv.z * t[11];
        if (u < -kZero || u > 1 + kZero) return false;
        const D3 e1 = mk(t[3], t[4], t[5]);
        const D3 q = mk(tv.y * e1.z - tv.z * e1.y, tv.z * e1.x - tv.x * e1.z, tv.x * e1.y - tv.y * e1.x);
        const double v = q.x * t[12] + q.y * t[13] + q.z * t[14];
        if (v < -kZero || u + v > 1 + kZero) return false;
        return fabs(q.x * t[6] + q.y * t[7] + q.z * t[8]) < t[15];
    }
    __device__ bool searchable(int) const { return true; }
};

// Escape links of a tree (child pairs of ni internal nodes, leaves as ~leaf): for each node,
// the node a depth-first walk visits next when it skips that node's subtree (the right sibling
// of the nearest left child among the node and its ancestors), kBvhEnd after the last one.
// esc[v] for internal node v, esc[ni + l] for leaf l.  The point queries walk the tree along
// them without a stack.
constexpr int kBvhEnd = INT_MAX;
__global__ __launch_bounds__(256) void kb_parents(int ni, const int *__restrict__ child, int *__restrict__ par) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= ni) return;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int c = child[2 * (size_t)v + k];
        par[c >= 0 ? c : ni + ~c] = 2 * v + k;      // position in the child array
    }
}
__global__ __launch_bounds__(256) void kb_escape(int ni, int nl, const int *__restrict__ child, const int *__restrict__ par,
                                                 int *__restrict__ esc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni + nl) return;
    int pe = t == 0 ? -1 : par[t];
    int e = kBvhEnd;
    while (pe >= 0) {
        if ((pe & 1) == 0) {           // a left child: its right sibling is next
            e = child[pe + 1];
            break;
        }
        const int v = pe >> 1;
        pe = v == 0 ? -1 : par[v];
    }
    esc[t] = e;
}

// Per point: the lowest searchable cell containing it (x, -1 if none), the number of cells of
// any marker containing it (y), the cell of the nearest centroid, lowest index among equal
// distances (z, only when x < 0) and a traversal-overflow flag (w, always 0: the walk along the
// escape links needs no stack). With `lst`, also the cells containing the point, ascending, in
// its row of kChainW entries (the first kChainW of them).
template <class Cells>
__global__ __launch_bounds__(256) void k_candidates(int n, const double *__restrict__ pts, int n_leaves,
                                                    const int *__restrict__ child, const int *__restrict__ esc,
                                                    const float *__restrict__ box,
                                                    const float *__restrict__ cbox, const float *__restrict__ leafbox,
                                                    const int *__restrict__ prim, Cells cells,
                                                    const double *__restrict__ cen, int4 *__restrict__ cand,
                                                    int *__restrict__ lst, const int *__restrict__ perm) {
    // with `perm`, thread s takes point perm[s] (points in Morton order: a wave's points lie
    // close together and walk the tree alike); results go to the point's own index
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int i = perm ? perm[s] : s;
    const double p[3] = {pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]};
    const int ni = n_leaves - 1;
    int t0 = INT_MAX, cnt = 0;
    int *row = lst ? lst + (size_t)i * kChainW : nullptr;
    // nearest centroid: strict < over increasing indices in the reference's scan, i.e. the
    // lowest index among the minima
    double best = 1e100;
    int amin = 0;
    auto consider = [&](int leaf) {
        const int t = prim[leaf];
        const double xx = p[0] - cen[3 * (size_t)t];
        const double yy = p[1] - cen[3 * (size_t)t + 1];
        const double zz = p[2] - cen[3 * (size_t)t + 2];
        const double d2 = xx * xx + yy * yy + zz * zz;
        if (d2 < best || (d2 == best && t < amin)) {
            best = d2;
            amin = t;
        }
    };
    auto leaf_hit = [&](int leaf) {
        if (!in_box(leafbox + 6 * (size_t)leaf, p)) return;
        consider(leaf);                 // the cells near the point seed the nearest-centroid search
        const int t = prim[leaf];
        if (cells.contains(p, t)) {
            if (row && cnt < kChainW) {
                // insertion into the ascending row
                int k = cnt;
                while (k > 0 && row[k - 1] > t) {
                    row[k] = row[k - 1];
                    --k;
                }
                row[k] = t;
            }
            ++cnt;
            if (cells.searchable(t) && t < t0) t0 = t;
        }
    };
    if (n_leaves == 1) {
        leaf_hit(0);
    } else {
        int ref = 0;                    // the root is always entered
        while (ref != kBvhEnd) {
            if (ref >= 0) {
                ref = ref == 0 || in_box(box + 6 * (size_t)ref, p) ? child[2 * (size_t)ref] : esc[ref];
            } else {
                leaf_hit(~ref);
                ref = esc[ni + ~ref];
            }
        }
    }
    if (t0 == INT_MAX) {
        if (n_leaves == 1) {
            consider(0);
        } else {
            // subtrees whose centroid box lies farther than the best distance so far are
            // skipped; a lower bound, so no minimum (and no tie of lower index) is missed
            int ref = 0;
            while (ref != kBvhEnd) {
                if (ref >= 0) {
                    ref = box_d2(cbox + 6 * (size_t)ref, p) > best ? esc[ref] : child[2 * (size_t)ref];
                } else {
                    consider(~ref);
                    ref = esc[ni + ~ref];
                }
            }
        }
    }
    cand[i] = make_int4(t0 == INT_MAX ? -1 : t0, cnt, t0 == INT_MAX ? amin : 0, 0);
}

// ------------------------------------------------------------------------------------------
// chained guess on the device
// ------------------------------------------------------------------------------------------
// The reference locates point after point, each from the previous point's result
// (LinearTetrahedra::locate_cell, src/InterpolatorCells.cpp:268-307, driven by
// src/SolutionReader.cpp:136-165): r_i = f_i(|r_{i-1}|), f_i(g) = g if g contains p_i, else
// the first neighbour of g (in list order) that contains it, else the lowest searchable cell
// containing it (cand.x) or, if none, minus the nearest centroid (-cand.z). f_i is constant
// when no cell contains p_i or exactly one searchable cell does. Otherwise r_i is one of
// C_i = (the cells containing p_i, ascending) followed by -cand.z when cand.x < 0, and the
// chain becomes a composition of tables T_i: position in C_{i-1} -> position in C_i (at most
// kChainW entries each), evaluated as a prefix scan: per workgroup of 256 points, then over
// the workgroups, then back to every point.
__device__ __forceinline__ bool chain_fixed(int4 c) { return c.y == 0 || (c.y == 1 && c.x >= 0); }
__device__ __forceinline__ int chain_size(int4 c) { return chain_fixed(c) ? 1 : c.y + (c.x < 0 ? 1 : 0); }
// value at position k of C_i (k < min(chain_size, kChainW))
__device__ __forceinline__ int chain_value(int4 c, const int *row, int k) {
    if (chain_fixed(c)) return c.y == 0 ? -c.z : c.x;
    return k < c.y ? row[k] : -c.z;
}
// position of t in the ascending row of m entries, -1 if absent
__device__ __forceinline__ int chain_find(const int *row, int m, int t) {
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (row[mid] < t) lo = mid + 1;
        else hi = mid;
    }
    return lo < m && row[lo] == t ? lo : -1;
}

// T_i for every point whose f_i is not constant (the rows of the others are never read);
// flag = 1 when a point's candidates are incomplete (traversal overflow, or more than kChainW
// possible results): the caller then resolves the chain on the host
__global__ __launch_bounds__(256) void k_chain_tables(int n, const int4 *__restrict__ cand, const int *__restrict__ lst,
                                                      const int *__restrict__ nbr_ptr, const int *__restrict__ nbr,
                                                      int n_cells, int g0, unsigned char *__restrict__ tab,
                                                      int *__restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 c = cand[i];
    if (c.w || chain_size(c) > kChainW) atomicOr(flag, 1);
    if (chain_fixed(c)) return;
    unsigned char *T = tab + (size_t)i * kChainW;
    const int *row = lst + (size_t)i * kChainW;
    const int m = min(c.y, kChainW);
    const int fb = c.x >= 0 ? chain_find(row, m, c.x) : m;
    int nd = 1;
    int4 cp = make_int4(0, 0, 0, 0);
    const int *prow = nullptr;
    if (i > 0) {
        cp = cand[i - 1];
        prow = lst + (size_t)(i - 1) * kChainW;
        nd = min(chain_size(cp), kChainW);
    }
    for (int k = 0; k < kChainW; ++k) {
        int pos = 0;
        if (k < nd) {
            const int g = i == 0 ? g0 : abs(chain_value(cp, prow, k));
            pos = -1;
            if (g >= 0 && g < n_cells) {
                pos = chain_find(row, m, g);
                for (int e = nbr_ptr[g]; pos < 0 && e < nbr_ptr[g + 1]; ++e) pos = chain_find(row, m, nbr[e]);
            }
            if (pos < 0) pos = fb;
        }
        T[k] = (unsigned char)min(max(pos, 0), kChainW - 1);
    }
}

// LDS rows of a table: kChainW bytes, padded to spread consecutive rows over the banks
constexpr int kChainRow = kChainW + 4;

// inclusive scan of the tables of 256 consecutive functions in LDS (F[cur] in, result in the
// returned buffer): row j becomes T_j o ... o T_0
__device__ __forceinline__ int chain_scan256(unsigned char (*F)[256 * kChainRow], int cur) {
    const int j = threadIdx.x;
    for (int d = 1; d < 256; d <<= 1) {
        __syncthreads();
        const unsigned char *src = F[cur] + j * kChainRow;
        unsigned *dst = reinterpret_cast<unsigned *>(F[cur ^ 1] + j * kChainRow);
        if (j >= d) {
            const unsigned *prev = reinterpret_cast<const unsigned *>(F[cur] + (j - d) * kChainRow);
#pragma unroll 4
            for (int w = 0; w < kChainW / 4; ++w) {
                const unsigned q = prev[w];
                dst[w] = (unsigned)src[q & 255] | (unsigned)src[(q >> 8) & 255] << 8 |
                         (unsigned)src[(q >> 16) & 255] << 16 | (unsigned)src[q >> 24] << 24;
            }
        } else {
            const unsigned *s4 = reinterpret_cast<const unsigned *>(src);
#pragma unroll 4
            for (int w = 0; w < kChainW / 4; ++w) dst[w] = s4[w];
        }
        cur ^= 1;
    }
    __syncthreads();
    return cur;
}

// per workgroup of 256 points: tab row i becomes the prefix T_i o ... o T_(first point of the
// workgroup); the workgroup's whole composition goes to bt
__global__ __launch_bounds__(256) void k_chain_scan(int n, const int4 *__restrict__ cand, unsigned char *__restrict__ tab,
                                                    unsigned char *__restrict__ bt) {
    __shared__ __align__(16) unsigned char F[2][256 * kChainRow];
    const int j = threadIdx.x;
    const int i = blockIdx.x * 256 + j;
    unsigned *r = reinterpret_cast<unsigned *>(F[0] + j * kChainRow);
    if (i >= n) {
        for (int w = 0; w < kChainW / 4; ++w)      // identity
            r[w] = (unsigned)(4 * w) | (unsigned)(4 * w + 1) << 8 | (unsigned)(4 * w + 2) << 16 |
                   (unsigned)(4 * w + 3) << 24;
    } else if (chain_fixed(cand[i])) {
        for (int w = 0; w < kChainW / 4; ++w) r[w] = 0;
    } else {
        const uint4 *t = reinterpret_cast<const uint4 *>(tab + (size_t)i * kChainW);
        for (int w = 0; w < kChainW / 16; ++w) {
            const uint4 v = t[w];
            r[4 * w] = v.x;
            r[4 * w + 1] = v.y;
            r[4 * w + 2] = v.z;
            r[4 * w + 3] = v.w;
        }
    }
    const int cur = chain_scan256(F, 0);
    const unsigned *o = reinterpret_cast<const unsigned *>(F[cur] + j * kChainRow);
    if (i < n) {
        uint4 *t = reinterpret_cast<uint4 *>(tab + (size_t)i * kChainW);
        for (int w = 0; w < kChainW / 16; ++w) t[w] = make_uint4(o[4 * w], o[4 * w + 1], o[4 * w + 2], o[4 * w + 3]);
    }
    if (j == 255) {
        uint4 *t = reinterpret_cast<uint4 *>(bt + (size_t)blockIdx.x * kChainW);
        for (int w = 0; w < kChainW / 16; ++w) t[w] = make_uint4(o[4 * w], o[4 * w + 1], o[4 * w + 2], o[4 * w + 3]);
    }
}

// one workgroup: the position each workgroup of points starts from (in C of the point before
// it), by scanning the workgroups' compositions 256 at a time; the chain starts at position 0
// of the one-entry domain {initial guess}
__global__ __launch_bounds__(256) void k_chain_top(int nb, const unsigned char *__restrict__ bt, int *__restrict__ entry) {
    __shared__ __align__(16) unsigned char F[2][256 * kChainRow];
    const int j = threadIdx.x;
    int s = 0;
    for (int base = 0; base < nb; base += 256) {
        unsigned *r = reinterpret_cast<unsigned *>(F[0] + j * kChainRow);
        if (base + j < nb) {
            const uint4 *t = reinterpret_cast<const uint4 *>(bt + (size_t)(base + j) * kChainW);
            for (int w = 0; w < kChainW / 16; ++w) {
                const uint4 v = t[w];
                r[4 * w] = v.x;
                r[4 * w + 1] = v.y;
                r[4 * w + 2] = v.z;
                r[4 * w + 3] = v.w;
            }
        } else {
            for (int w = 0; w < kChainW / 4; ++w)
                r[w] = (unsigned)(4 * w) | (unsigned)(4 * w + 1) << 8 | (unsigned)(4 * w + 2) << 16 |
                       (unsigned)(4 * w + 3) << 24;
        }
        const int cur = chain_scan256(F, 0);
        if (base + j < nb) entry[base + j] = j == 0 ? s : F[cur][(j - 1) * kChainRow + s];
        s = F[cur][255 * kChainRow + s];
        __syncthreads();
    }
}

// each point's result from its workgroup's starting position and its prefix table
__global__ __launch_bounds__(256) void k_chain_finish(int n, const int4 *__restrict__ cand, const int *__restrict__ lst,
                                                      const unsigned char *__restrict__ tab,
                                                      const int *__restrict__ entry, int *__restrict__ res) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 c = cand[i];
    const int pos = tab[(size_t)i * kChainW + entry[i >> 8]];
    res[i] = chain_value(c, lst + (size_t)i * kChainW, min(pos, kChainW - 1));
}

// LinearHexahedra::project_to_nat_coords (src/InterpolatorCells.cpp:1237-1244, 1272-1317):
// natural coordinates of a point by Newton iterations, vertices in femocs order
__device__ void linhex_uvw(const D3 x[8], const double *pp, double &u_out, double &v_out, double &w_out) {
    const D3 f0s = dvd(add(add(add(add(add(add(add(x[0], x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    const D3 f1 = dvd(sub(add(add(sub(sub(add(add(scl(x[0], -1), x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    const D3 f2 = dvd(add(add(sub(sub(add(add(sub(scl(x[0], -1), x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    const D3 f3 = dvd(add(add(add(add(sub(sub(sub(scl(x[0], -1), x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    const D3 f4 = dvd(sub(add(sub(add(sub(add(sub(x[0], x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    const D3 f5 = dvd(sub(add(add(sub(add(sub(sub(x[0], x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    const D3 f6 = dvd(add(add(sub(sub(sub(sub(add(x[0], x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    const D3 f7 = dvd(sub(add(sub(add(add(sub(add(scl(x[0], -1), x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    const D3 f0 = sub(mk(pp[0], pp[1], pp[2]), f0s);
    double u = 0, v = 0, w = 0;
    for (int i = 0; i < 20; ++i) {
        const D3 f = sub(sub(sub(sub(sub(sub(sub(f0, scl(f1, u)), scl(f2, v)), scl(f3, w)), scl(f4, u * v)), scl(f5, u * w)),
                             scl(f6, v * w)),
                         scl(f7, u * v * w));
        const D3 fu = add(add(add(f1, scl(f4, v)), scl(f5, w)), scl(f7, v * w));
        const D3 fv = add(add(add(f2, scl(f4, u)), scl(f6, w)), scl(f7, u * w));
        const D3 fw = add(add(add(f3, scl(f5, u)), scl(f6, v)), scl(f7, u * v));
        double D = det3(fu, fv, fw);
        D = 1.0 / D;
        const double du = det3(f, fv, fw) * D;
        const double dv = det3(fu, f, fw) * D;
        const double dw = det3(fu, fv, f) * D;
        u += du;
        v += dv;
        w += dw;
        if (du * du + dv * dv + dw * dw < kZero) break;
    }
    u_out = u;
    v_out = v;
    w_out = w;
}

// LinearHexahedra::shape_functions (src/InterpolatorCells.cpp:1334-1353)
__device__ void linhex_sf(const D3 x[8], const double *pp, double sf[8]) {
    double u, v, w;
    linhex_uvw(x, pp, u, v, w);
    sf[0] = (1 - u) * (1 - v) * (1 - w) / 8.0;
    sf[1] = (1 + u) * (1 - v) * (1 - w) / 8.0;
    sf[2] = (1 + u) * (1 + v) * (1 - w) / 8.0;
    sf[3] = (1 - u) * (1 + v) * (1 - w) / 8.0;
    sf[4] = (1 - u) * (1 - v) * (1 + w) / 8.0;
    sf[5] = (1 + u) * (1 - v) * (1 + w) / 8.0;
    sf[6] = (1 + u) * (1 + v) * (1 + w) / 8.0;
    sf[7] = (1 - u) * (1 + v) * (1 + w) / 8.0;
}

// LinearHexahedra::shape_fun_grads (src/InterpolatorCells.cpp:1363-1416): gradients of the
// femocs-order shape functions at a point, through the Jacobian at its natural coordinates
__device__ void linhex_sfg(const D3 x[8], const double *pp, D3 sfg[8]) {
    double u, v, w;
    linhex_uvw(x, pp, u, v, w);
    D3 dN[8] = {mk(-(1 - v) * (1 - w), -(1 - u) * (1 - w), -(1 - u) * (1 - v)),
                mk((1 - v) * (1 - w), -(1 + u) * (1 - w), -(1 + u) * (1 - v)),
                mk((1 + v) * (1 - w), (1 + u) * (1 - w), -(1 + u) * (1 + v)),
                mk(-(1 + v) * (1 - w), (1 - u) * (1 - w), -(1 - u) * (1 + v)),
                mk(-(1 - v) * (1 + w), -(1 - u) * (1 + w), (1 - u) * (1 - v)),
                mk((1 - v) * (1 + w), -(1 + u) * (1 + w), (1 + u) * (1 - v)),
                mk((1 + v) * (1 + w), (1 + u) * (1 + w), (1 + u) * (1 + v)),
                mk(-(1 + v) * (1 + w), (1 - u) * (1 + w), (1 - u) * (1 + v))};
    for (int i = 0; i < 8; ++i) dN[i] = scl(dN[i], 0.125);
    // J[i] = sum_k x_k dN_k,i
    D3 J[3] = {mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)};
    for (int k = 0; k < 8; ++k) {
        J[0] = add(J[0], scl(x[k], dN[k].x));
        J[1] = add(J[1], scl(x[k], dN[k].y));
        J[2] = add(J[2], scl(x[k], dN[k].z));
    }
    double Jdet = det3(J[0], J[1], J[2]);
    Jdet = 1.0 / Jdet;
    const D3 Ji[3] = {scl(mk(J[1].y * J[2].z - J[2].y * J[1].z, J[2].x * J[1].z - J[1].x * J[2].z,
                             J[1].x * J[2].y - J[2].x * J[1].y), Jdet),
                      scl(mk(J[2].y * J[0].z - J[0].y * J[2].z, J[0].x * J[2].z - J[2].x * J[0].z,
                             J[2].x * J[0].y - J[0].x * J[2].y), Jdet),
                      scl(mk(J[0].y * J[1].z - J[1].y * J[0].z, J[1].x * J[0].z - J[0].x * J[1].z,
                             J[0].x * J[1].y - J[1].x * J[0].y), Jdet)};
    for (int k = 0; k < 8; ++k) {
        D3 g = mk(0, 0, 0);
        g = add(g, scl(Ji[0], dN[k].x));
        g = add(g, scl(Ji[1], dN[k].y));
        g = add(g, scl(Ji[2], dN[k].z));
        sfg[k] = g;
    }
}

// InterpolatorCells::interp_gradient(point, cell) of LinearHexahedra (src/InterpolatorCells.cpp:
// 407-421): minus the gradient of the nodal scalar2 at the point, in the given hex
__global__ __launch_bounds__(256) void k_hex_gradient(int n, const double *__restrict__ pts, const int *__restrict__ hex,
                                                      const int *__restrict__ hexes, const double *__restrict__ xyz,
                                                      const double *__restrict__ sol, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int *h = hexes + 8 * (size_t)hex[i];
    D3 x[8], g[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = ldx(xyz, h[k]);
    linhex_sfg(x, pts + 3 * (size_t)i, g);
    D3 e = mk(0, 0, 0);
    for (int k = 0; k < 8; ++k) e = sub(e, scl(g[k], sol[5 * (size_t)h[k] + 4]));
    out[3 * (size_t)i] = e.x;
    out[3 * (size_t)i + 1] = e.y;
    out[3 * (size_t)i + 2] = e.z;
}

// InterpolatorCells::interp_solution (src/InterpolatorCells.cpp:331-355) for ranks 1-3.
// from_tet != 0 (rank 3): cell holds the signed tet of the locate step and the hex is chosen
// by the largest barycentric coordinate (LinearHexahedra::locate_cell, :1536-1562).
__global__ __launch_bounds__(256) void k_interp(int n, const double *__restrict__ pts, int rank, int from_tet,
                                                const int *__restrict__ cell_in, const double *__restrict__ det0,
                                                const double *__restrict__ det, const int *__restrict__ tets,
                                                const int *__restrict__ c10, const int *__restrict__ hexes,
                                                const double *__restrict__ xyz, const double *__restrict__ sol,
                                                int *__restrict__ cell_out, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double p[3] = {pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]};
    int cell = cell_in[i];
    double w[10];
    int nodes[10];
    int nn;
    if (rank == 3) {
        if (from_tet) {
            const int sign = cell < 0 ? -1 : 1;
            const int t = cell < 0 ? -cell : cell;
            double b[4];
            bcc(det0, det, p, t, b);
            int k = -1;
            if (b[0] >= b[1] && b[0] >= b[2] && b[0] >= b[3]) k = 0;
            else if (b[1] >= b[0] && b[1] >= b[2] && b[1] >= b[3]) k = 1;
            else if (b[2] >= b[0] && b[2] >= b[1] && b[2] >= b[3]) k = 2;
            else if (b[3] >= b[0] && b[3] >= b[1] && b[3] >= b[2]) k = 3;
            cell = k < 0 ? -1 : sign * (4 * t + k);
        }
        const int h = cell < 0 ? -cell : cell;
        D3 x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            nodes[k] = hexes[8 * (size_t)h + k];
            x[k] = ldx(xyz, nodes[k]);
        }
        linhex_sf(x, p, w);
        nn = 8;
    } else {
        const int t = cell < 0 ? -cell : cell;
        double b[4];
        bcc(det0, det, p, t, b);
        if (rank == 1) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                w[k] = b[k];
                nodes[k] = tets[4 * (size_t)t + k];
            }
            nn = 4;
        } else {
            const double b1 = b[0], b2 = b[1], b3 = b[2], b4 = b[3];
            w[0] = b1 * (2 * b1 - 1);
            w[1] = b2 * (2 * b2 - 1);
            w[2] = b3 * (2 * b3 - 1);
            w[3] = b4 * (2 * b4 - 1);
            w[4] = 4 * b1 * b2;
            w[5] = 4 * b2 * b3;
            w[6] = 4 * b3 * b1;
            w[7] = 4 * b1 * b4;
            w[8] = 4 * b2 * b4;
            w[9] = 4 * b3 * b4;
#pragma unroll
            for (int k = 0; k < 10; ++k) nodes[k] = c10[10 * (size_t)t + k];
            nn = 10;
        }
    }
    double vx = 0, vy = 0, vz = 0, s1 = 0, s2 = 0;
    for (int k = 0; k < nn; ++k) {
        const double *s = sol + 5 * (size_t)nodes[k];
        vx += s[0] * w[k];
        vy += s[1] * w[k];
        vz += s[2] * w[k];
        s1 += s[3] * w[k];
        s2 += s[4] * w[k];
    }
    double *o = out + 5 * (size_t)i;
    o[0] = vx;
    o[1] = vy;
    o[2] = vz;
    o[3] = s1;
    o[4] = s2;
    cell_out[i] = cell;
}

// Per point: the lowest cell whose centroid lies within squared distance 1e-15 (x, -1 if none)
// and a traversal-overflow flag (w); InterpolatorCells::is_centroid (src/InterpolatorCells.cpp:
// 330-333) on every centroid the tree does not rule out.
__global__ __launch_bounds__(256) void k_centroid_match(int n, const double *__restrict__ pts, int n_leaves,
                                                        const int *__restrict__ child, const float *__restrict__ cbox,
                                                        const int *__restrict__ prim, const double *__restrict__ cen,
                                                        int4 *__restrict__ cand) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double p[3] = {pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]};
    int best = INT_MAX, overflow = 0;
    auto leaf = [&](int l) {
        const int c = prim[l];
        const double xx = p[0] - cen[3 * (size_t)c];
        const double yy = p[1] - cen[3 * (size_t)c + 1];
        const double zz = p[2] - cen[3 * (size_t)c + 2];
        if (xx * xx + yy * yy + zz * zz < kZero && c < best) best = c;
    };
    if (n_leaves == 1) {
        leaf(0);
    } else {
        int stack[kBvhStack];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const int v = stack[--sp];
            // box_d2 never exceeds the squared distance to a centroid inside the box
            if (box_d2(cbox + 6 * (size_t)v, p) >= kZero) continue;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int c = child[2 * (size_t)v + k];
                if (c < 0) leaf(~c);
                else if (sp < kBvhStack) stack[sp++] = c;
                else overflow = 1;
            }
        }
    }
    cand[i] = make_int4(best == INT_MAX ? -1 : best, 0, 0, overflow);
}

// InterpolatorCells::interp_centroid (src/InterpolatorCells.cpp:336-356): the cell's nodes
// weighted alike; cell < 0 (no centroid matched) gives Solution(-1)
__global__ __launch_bounds__(256) void k_interp_nodes(int n, const int *__restrict__ cell, const int *__restrict__ nodes,
                                                      int nn, double w, const double *__restrict__ sol,
                                                      double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *o = out + 5 * (size_t)i;
    const int c = cell[i];
    if (c < 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) o[k] = -1.0;
        return;
    }
    double vx = 0, vy = 0, vz = 0, s1 = 0, s2 = 0;
    for (int k = 0; k < nn; ++k) {
        const double *s = sol + 5 * (size_t)nodes[nn * (size_t)c + k];
        vx += s[0] * w;
        vy += s[1] * w;
        vz += s[2] * w;
        s1 += s[3] * w;
        s2 += s[4] * w;
    }
    o[0] = vx;
    o[1] = vy;
    o[2] = vz;
    o[3] = s1;
    o[4] = s2;
}

// Interpolator::extract_solution(PoissonSolver&) with store_elfield (src/Interpolator.cpp:79-140,
// 172-190; LinearHexahedra::shape_fun_grads(hex, node), src/InterpolatorCells.cpp:1462-1505)
__global__ __launch_bounds__(256) void k_extract(int n_nodes, const int *__restrict__ node_idof,
                                                 const double *__restrict__ u, const int *__restrict__ n2c_ptr,
                                                 const int *__restrict__ n2c, const int *__restrict__ hexes,
                                                 const double *__restrict__ xyz, const double *__restrict__ rho,
                                                 double empty, double *__restrict__ sol) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    double *s = sol + 5 * (size_t)n;
    const int id = node_idof[n];
    if (id < 0) {
        for (int k = 0; k < 5; ++k) s[k] = empty;
        return;
    }
    D3 mean = mk(0, 0, 0);
    const int e0 = n2c_ptr[n], e1 = n2c_ptr[n + 1];
    for (int e = e0; e < e1; ++e) {
        const int h = n2c[e] >> 3, node = n2c[e] & 7;
        const int *sh = hexes + 8 * (size_t)h;
        // corner `node` of the femocs hex at natural coordinates uvw (corners 0-3 counter-clockwise
        // at w = -1, 4-7 above them); its neighbours along u, v and w are node ^ 1, ^ 3 and ^ 4
        const int q = node & 3;
        const double uvw[3] = {(q == 1 || q == 2) ? 1.0 : -1.0, q >= 2 ? 1.0 : -1.0, node >= 4 ? 1.0 : -1.0};
        const int nn[3] = {node ^ 1, node ^ 3, node ^ 4};
        const D3 v0 = ldx(xyz, sh[node]);
        double J[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const D3 r = scl(sub(v0, ldx(xyz, sh[nn[a]])), uvw[a] * 0.5);
            J[a][0] = r.x;
            J[a][1] = r.y;
            J[a][2] = r.z;
        }
        double Jdet = det3(mk(J[0][0], J[0][1], J[0][2]), mk(J[1][0], J[1][1], J[1][2]), mk(J[2][0], J[2][1], J[2][2]));
        Jdet = 1.0 / Jdet;
        double Ji[3][3] = {{J[1][1] * J[2][2] - J[2][1] * J[1][2], J[2][1] * J[0][2] - J[0][1] * J[2][2],
                            J[0][1] * J[1][2] - J[1][1] * J[0][2]},
                           {J[2][0] * J[1][2] - J[1][0] * J[2][2], J[0][0] * J[2][2] - J[2][0] * J[0][2],
                            J[1][0] * J[0][2] - J[0][0] * J[1][2]},
                           {J[1][0] * J[2][1] - J[2][0] * J[1][1], J[2][0] * J[0][1] - J[0][0] * J[2][1],
                            J[0][0] * J[1][1] - J[1][0] * J[0][1]}};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) Ji[a][b] *= Jdet;
        // shape-function gradients at the corner: nonzero for the corner and its three
        // neighbours only; every corner's term is subtracted in corner order, zero rows included
        double gn[3], gb[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            gn[a] = 0.5 * (Ji[a][0] * uvw[0] + Ji[a][1] * uvw[1] + Ji[a][2] * uvw[2]);
#pragma unroll
            for (int b = 0; b < 3; ++b) gb[b][a] = -0.5 * Ji[a][b] * uvw[b];
        }
        D3 vi = mk(0, 0, 0);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            D3 gk = mk(0.0, 0.0, 0.0);
            if (k == node) gk = mk(gn[0], gn[1], gn[2]);
#pragma unroll
            for (int b = 0; b < 3; ++b)
                if (k == nn[b]) gk = mk(gb[b][0], gb[b][1], gb[b][2]);
            vi = sub(vi, scl(gk, u[node_idof[sh[k]]]));
        }
        mean = add(mean, vi);
    }
    if (e1 > e0) mean = scl(mean, 1.0 / (e1 - e0));
    s[0] = mean.x;
    s[1] = mean.y;
    s[2] = mean.z;
    s[3] = rho ? rho[id] : 0.0;   // charge density of a write step, else zeros
    s[4] = u[id];
}

// Interpolator::average_nodal_fields (src/Interpolator.cpp:142-170)
__global__ __launch_bounds__(256) void k_smooth(int n_tetnodes, const int *__restrict__ ptr, const int *__restrict__ lst,
                                                const double *__restrict__ xyz, double decay, double *__restrict__ sol) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_tetnodes) return;
    const int e0 = ptr[i], e1 = ptr[i + 1];
    if (e0 == e1) return;
    const D3 xi = ldx(xyz, i);
    D3 vec = mk(0, 0, 0);
    double w_sum = 0;
    for (int e = e0; e < e1; ++e) {
        const int nb = lst[e];
        const D3 xn = ldx(xyz, nb);
        const double xx = xi.x - xn.x, yy = xi.y - xn.y, zz = xi.z - xn.z;
        const double w = exp(decay * sqrt(xx * xx + yy * yy + zz * zz));
        w_sum += w;
        vec = add(vec, scl(mk(sol[5 * (size_t)nb], sol[5 * (size_t)nb + 1], sol[5 * (size_t)nb + 2]), w));
    }
    if (w_sum > 0) {
        vec = scl(vec, 1.0 / w_sum);
        sol[5 * (size_t)i] = vec.x;
        sol[5 * (size_t)i + 1] = vec.y;
        sol[5 * (size_t)i + 2] = vec.z;
    }
}

// Interpolator::extract_solution(CurrentHeatSolver&) (src/Interpolator.cpp:207-216):
// Solution{-grad phi_c (export_solution_grad), phi_c, T}; nodes outside: {0, 0, empty}
__global__ __launch_bounds__(256) void k_extract_heat(int n_nodes, const int *__restrict__ node_idof,
                                                      const double *__restrict__ grad, const double *__restrict__ phic,
                                                      const double *__restrict__ T, double empty, double *__restrict__ sol) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    double *s = sol + 5 * (size_t)n;
    const int id = node_idof[n];
    if (id < 0) {
        s[0] = s[1] = s[2] = s[3] = 0.0;
        s[4] = empty;
        return;
    }
    s[0] = grad[3 * (size_t)id];
    s[1] = grad[3 * (size_t)id + 1];
    s[2] = grad[3 * (size_t)id + 2];
    s[3] = phic[id];
    s[4] = T[id];
}

inline int grid(int n) { return (n + 255) / 256; }

void ensure_points(fgpu_interp *ip, size_t n) {
    if (n <= ip->cap) return;
    IPCHECK(hipStreamSynchronize(ip->ctx->stream));
    release(ip->pts);
    release(ip->out);
    release(ip->cand);
    release(ip->cand2);
    release(ip->tet);
    release(ip->cells);
    release(ip->cells2);
    const size_t c = std::max<size_t>(n, 1024);
    IPCHECK(hipMalloc(&ip->cand2, c * sizeof(int4)));
    IPCHECK(hipMalloc(&ip->pts, 3 * c * sizeof(double)));
    IPCHECK(hipMalloc(&ip->out, 5 * c * sizeof(double)));
    IPCHECK(hipMalloc(&ip->cand, c * sizeof(int4)));
    IPCHECK(hipMalloc(&ip->tet, c * sizeof(int)));
    IPCHECK(hipMalloc(&ip->cells, c * sizeof(int)));
    IPCHECK(hipMalloc(&ip->cells2, c * sizeof(int)));
    ip->cap = c;
}

void ensure_chain(fgpu_interp *ip, size_t n) {
    if (n <= ip->chain_cap) return;
    IPCHECK(hipStreamSynchronize(ip->ctx->stream));
    for (int **p : {&ip->chain_lst, &ip->chain_entry, &ip->chain_flag}) release(*p);
    release(ip->chain_tab);
    release(ip->chain_bt);
    const size_t c = std::max<size_t>(n, 1024), nb = (c + 255) / 256;
    IPCHECK(hipMalloc(&ip->chain_lst, c * kChainW * sizeof(int)));
    IPCHECK(hipMalloc(&ip->chain_tab, c * kChainW));
    IPCHECK(hipMalloc(&ip->chain_bt, nb * kChainW));
    IPCHECK(hipMalloc(&ip->chain_entry, nb * sizeof(int)));
    IPCHECK(hipMalloc(&ip->chain_flag, sizeof(int)));
    ip->chain_cap = c;
}

void free_triangles(fgpu_interp *ip) {
    for (double **p : {&ip->tri_tab, &ip->tri_cen, &ip->quad_cen}) release(*p);
    for (int **p : {&ip->tri_child, &ip->tri_prim, &ip->tri_nodes, &ip->quad_nodes, &ip->tri6, &ip->quad_child,
                    &ip->quad_prim, &ip->tri_esc, &ip->tri_nbr_ptr, &ip->tri_nbr, &ip->tri2tet, &ip->quad2hex})
        release(*p);
    for (float **p : {&ip->tri_box, &ip->tri_cbox, &ip->tri_leafbox, &ip->quad_cbox}) release(*p);
    ip->T = LinTriHost{};
    ip->TB = Bvh{};
    ip->QB = Bvh{};
}

void free_interp(fgpu_interp *ip) {
    for (double **p : {&ip->xyz, &ip->det0, &ip->det, &ip->cen, &ip->sol, &ip->pts, &ip->out}) release(*p);
    for (int **p : {&ip->tets, &ip->c10, &ip->hexes, &ip->marker, &ip->child, &ip->prim, &ip->n2c_ptr, &ip->n2c,
                    &ip->vor_ptr, &ip->vor, &ip->node_idof, &ip->tet, &ip->cells, &ip->cells2, &ip->d_region_hex,
                    &ip->sc_icell,
                    &ip->d_hex_cell, &ip->d_nbr_ptr, &ip->d_nbr, &ip->esc})
        release(*p);
    for (float **p : {&ip->box, &ip->cbox, &ip->leafbox}) release(*p);
    release(ip->cand);
    release(ip->cand2);
    for (int **p : {&ip->chain_lst, &ip->chain_entry, &ip->chain_flag}) release(*p);
    release(ip->chain_tab);
    release(ip->chain_bt);
    ip->chain_cap = 0;
    for (unsigned long long **p : {&ip->skey, &ip->skey2}) release(*p);
    for (int **p : {&ip->sidx, &ip->sperm}) release(*p);
    release(ip->stmp);
    ip->sort_cap = ip->stmp_bytes = 0;
    free_triangles(ip);
}

#include "fgpu_interp_build.inc"

// Morton code of each query point in the box `root` (a tree's root box, 6 floats on the
// device; points outside it are clamped to its faces)
__global__ __launch_bounds__(256) void k_point_codes(int n, const double *__restrict__ pts, const float *__restrict__ root,
                                                     unsigned long long *__restrict__ key, int *__restrict__ idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long m = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double lo = root[a], hi = root[3 + a];
        const double t = hi > lo ? (pts[3 * (size_t)i + a] - lo) / (hi - lo) : 0.0;
        const unsigned long long q = (unsigned long long)fmin(2097151.0, fmax(0.0, t * 2097151.0));
        m |= spread21_d(q) << (2 - a);
    }
    key[i] = m;
    idx[i] = i;
}

// the order in which the point searches take the n points of ip->pts (their Morton order in the
// tree's root box), or nullptr where sorting does not pay (few points, a one-leaf tree)
const int *point_order(fgpu_interp *ip, int n, int n_leaves, const float *root_box) {
    if (n < 2048 || n_leaves < 2) return nullptr;
    hipStream_t st = ip->ctx->stream;
    if ((size_t)n > ip->sort_cap) {
        IPCHECK(hipStreamSynchronize(st));
        for (unsigned long long **p : {&ip->skey, &ip->skey2}) release(*p);
        for (int **p : {&ip->sidx, &ip->sperm}) release(*p);
        const size_t c = std::max<size_t>(n, 1024);
        IPCHECK(hipMalloc(&ip->skey, c * sizeof(unsigned long long)));
        IPCHECK(hipMalloc(&ip->skey2, c * sizeof(unsigned long long)));
        IPCHECK(hipMalloc(&ip->sidx, c * sizeof(int)));
        IPCHECK(hipMalloc(&ip->sperm, c * sizeof(int)));
        ip->sort_cap = c;
    }
    size_t bytes = 0;
    IPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, ip->skey, ip->skey2, ip->sidx, ip->sperm, n, 0, 63, st));
    if (bytes > ip->stmp_bytes) {
        IPCHECK(hipStreamSynchronize(st));
        release(ip->stmp);
        IPCHECK(hipMalloc(&ip->stmp, bytes));
        ip->stmp_bytes = bytes;
    }
    hipLaunchKernelGGL(k_point_codes, dim3(grid(n)), dim3(256), 0, st, n, ip->pts, root_box, ip->skey, ip->sidx);
    IPCHECK(hipcub::DeviceRadixSort::SortPairs(ip->stmp, bytes, ip->skey, ip->skey2, ip->sidx, ip->sperm, n, 0, 63, st));
    return ip->sperm;
}

void tet_candidates(fgpu_interp *ip, int n, int4 *cand, int *lst = nullptr) {
    const int *perm = point_order(ip, n, ip->B.n_leaves, ip->box);
    hipLaunchKernelGGL(k_candidates<TetCells>, dim3(grid(n)), dim3(256), 0, ip->ctx->stream, n, ip->pts,
                       ip->B.n_leaves, ip->child, ip->esc, ip->box, ip->cbox, ip->leafbox, ip->prim,
                       TetCells{ip->det0, ip->det, ip->marker}, ip->cen, cand, lst, perm);
}

// the escape links of a tree of nl leaves (device child array) into a new array
int *tree_escapes(hipStream_t s, int nl, const int *child) {
    const int ni = std::max(nl - 1, 0);
    int *esc = nullptr, *par = nullptr;
    IPCHECK(hipMalloc(&esc, std::max<size_t>((size_t)ni + nl, 1) * sizeof(int)));
    if (ni > 0) {
        IPCHECK(hipMalloc(&par, ((size_t)ni + nl) * sizeof(int)));
        hipLaunchKernelGGL(kb_parents, dim3(grid(ni)), dim3(256), 0, s, ni, child, par);
        hipLaunchKernelGGL(kb_escape, dim3(grid(ni + nl)), dim3(256), 0, s, ni, nl, child, par, esc);
        IPCHECK(hipGetLastError());
        IPCHECK(hipStreamSynchronize(s));
        IPCHECK(hipFree(par));
    }
    return esc;
}

// the chained guess over the located points (ip->pts, n of them) into ip->tet, in the
// reference's numbering of the tets; g0 is the first point's guess. Leaves *ip->chain_flag = 1
// when some point's candidates are incomplete (the host chain is then needed).
// the chain over candidates `cand` with their containing-cell rows in ip->chain_lst (cells
// with neighbour lists nbr_ptr/nbr, n_cells of them) into res; the flag was cleared before
void run_chain(fgpu_interp *ip, int n, const int4 *cand, const int *nbr_ptr, const int *nbr, int n_cells, int g0,
               int *res) {
    hipStream_t st = ip->ctx->stream;
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_chain_tables, dim3(grid(n)), dim3(256), 0, st, n, cand, ip->chain_lst, nbr_ptr, nbr, n_cells,
                       g0, ip->chain_tab, ip->chain_flag);
    hipLaunchKernelGGL(k_chain_scan, dim3(nb), dim3(256), 0, st, n, cand, ip->chain_tab, ip->chain_bt);
    hipLaunchKernelGGL(k_chain_top, dim3(1), dim3(256), 0, st, nb, ip->chain_bt, ip->chain_entry);
    hipLaunchKernelGGL(k_chain_finish, dim3(grid(n)), dim3(256), 0, st, n, cand, ip->chain_lst, ip->chain_tab,
                       ip->chain_entry, res);
    IPCHECK(hipGetLastError());
}

void device_chain(fgpu_interp *ip, int n, int g0) {
    ensure_chain(ip, n);
    IPCHECK(hipMemsetAsync(ip->chain_flag, 0, sizeof(int), ip->ctx->stream));
    tet_candidates(ip, n, ip->cand, ip->chain_lst);
    run_chain(ip, n, ip->cand, ip->d_nbr_ptr, ip->d_nbr, ip->n_tets, g0, ip->tet);
}

}  // namespace

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------
extern "C" int fgpu_interp_create(fgpu_ctx *c, int n_nodes, const double *xyz, const int *node_kind, int n_tets,
                                  const int *tets, const int *tet_nbr, const int *tet_marker, int n_hexes,
                                  const int *hexes, const int *hex_marker, int search_region, double empty_value,
                                  fgpu_interp **out) {
    return fgpu_interp_create_ex(c, n_nodes, xyz, node_kind, n_tets, tets, tet_nbr, tet_marker, n_hexes, hexes,
                                 hex_marker, search_region, empty_value, 0, out);
}

extern "C" int fgpu_interp_create_ex(fgpu_ctx *c, int n_nodes, const double *xyz, const int *node_kind, int n_tets,
                                     const int *tets, const int *tet_nbr, const int *tet_marker, int n_hexes,
                                     const int *hexes, const int *hex_marker, int search_region, double empty_value,
                                     int flags, fgpu_interp **out) {
    if (!c || !out || n_nodes <= 0 || n_tets <= 0 || !xyz || !node_kind || !tets || !tet_nbr || !tet_marker ||
        (n_hexes > 0 && (!hexes || !hex_marker)) || (n_hexes != 0 && n_hexes != 4 * n_tets) ||
        (flags & ~(FGPU_INTERP_HOST | FGPU_INTERP_HOST_CHAIN)))
        return FGPU_EINVAL;
    *out = nullptr;
    fgpu_interp *ip = new (std::nothrow) fgpu_interp();
    if (!ip) return FGPU_ENOMEM;
    ip->ctx = c;
    const int rc = guard(ip, [&] {
        IPCHECK(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        for (size_t i = 0; i < 4 * (size_t)n_tets; ++i)
            if (tets[i] < 0 || tets[i] >= n_nodes) throw std::invalid_argument("tet node out of range");
        for (size_t i = 0; i < 8 * (size_t)n_hexes; ++i)
            if (hexes[i] < 0 || hexes[i] >= n_nodes) throw std::invalid_argument("hex node out of range");
        for (size_t i = 0; i < 4 * (size_t)n_tets; ++i)
            if (tet_nbr[i] < -1 || tet_nbr[i] >= n_tets) throw std::invalid_argument("tet neighbour out of range");
        ip->n_nodes = n_nodes;
        ip->n_tets = n_tets;
        ip->n_hexes = n_hexes;
        ip->region = search_region;
        ip->empty = empty_value;
        int tn = 0;
        for (size_t i = 0; i < 4 * (size_t)n_tets; ++i) tn = std::max(tn, tets[i] + 1);
        ip->n_tetnodes = tn;
        ip->hkind.assign(node_kind, node_kind + n_nodes);
        ip->htets.assign(tets, tets + 4 * (size_t)n_tets);
        ip->xyz = upload(xyz, 3 * (size_t)n_nodes, s);
        ip->tets = upload(tets, 4 * (size_t)n_tets, s);
        ip->hexes = upload(hexes, 8 * (size_t)n_hexes, s);
        // the region's hexes in hex order and their solver cells (Interpolator::deal2femocs)
        auto in_region = [&](int h) { return search_region == 3 ? hex_marker[h] > 0 : hex_marker[h] < 0; };
        if (n_hexes > 0) {
            for (int h = 0; h < n_hexes; ++h)
                if (in_region(h)) ip->region_hex.push_back(h);
            ip->d_region_hex = upload(ip->region_hex.data(), ip->region_hex.size(), s);
            std::vector<int> hex_cell(n_hexes, -1);
            for (size_t k = 0; k < ip->region_hex.size(); ++k) hex_cell[ip->region_hex[k]] = (int)k;
            ip->d_hex_cell = upload(hex_cell.data(), hex_cell.size(), s);
            IPCHECK(hipStreamSynchronize(s));       // hex_cell is read until the copy ends
        }
        ip->host_chain = (flags & FGPU_INTERP_HOST_CHAIN) != 0;
        if (!(flags & FGPU_INTERP_HOST)) {
            device_build(ip, tet_nbr, tet_marker, node_kind, hex_marker);
        } else {
            build_lintet(ip->L, n_nodes, xyz, n_tets, tets, tet_nbr, tet_marker, search_region);
            ip->L_on_host = true;
            ip->n_nbr = (int)ip->L.nbr.size();
            ip->d_nbr_ptr = upload(ip->L.nbr_ptr.data(), ip->L.nbr_ptr.size(), s);
            ip->d_nbr = upload(ip->L.nbr.data(), ip->L.nbr.size(), s);
            // LinearTetrahedra::precompute: decay factor of the smoothing, -1 / longest tet edge
            double emax = -1e308;
            for (int t = 0; t < n_tets; ++t) {
                const double *v[4];
                for (int k = 0; k < 4; ++k) v[k] = xyz + 3 * (size_t)tets[4 * (size_t)t + k];
                const int P[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
                for (auto &pq : P) {
                    const double xx = v[pq[0]][0] - v[pq[1]][0], yy = v[pq[0]][1] - v[pq[1]][1], zz = v[pq[0]][2] - v[pq[1]][2];
                    emax = std::max(emax, xx * xx + yy * yy + zz * zz);
                }
            }
            ip->decay = -1.0 / std::sqrt(emax);
            // bounding boxes of the tets for the search tree
            std::vector<double> lo(3 * (size_t)n_tets), hi(3 * (size_t)n_tets);
            for (int t = 0; t < n_tets; ++t)
                for (int a = 0; a < 3; ++a) {
                    double l = 1e300, h = -1e300;
                    for (int k = 0; k < 4; ++k) {
                        const double x = xyz[3 * (size_t)tets[4 * (size_t)t + k] + a];
                        l = std::min(l, x);
                        h = std::max(h, x);
                    }
                    lo[3 * (size_t)t + a] = l;
                    hi[3 * (size_t)t + a] = h;
                }
            build_bvh(n_tets, lo, hi, ip->L.centroid, ip->B);
            std::vector<int> c10;
            if (n_hexes > 0) build_quadtet(n_tets, tets, hexes, node_kind, c10);
            else c10.assign(10 * (size_t)n_tets, 0);
            // node -> (hex, corner) of the region's hexes in hex order (Interpolator::initialize)
            std::vector<int> nptr(n_nodes + 1, 0), n2c;
            std::vector<int> vptr(tn + 1, 0), vor;
            if (n_hexes > 0) {
                for (int h = 0; h < n_hexes; ++h)
                    if (in_region(h))
                        for (int k = 0; k < 8; ++k) nptr[hexes[8 * (size_t)h + k] + 1]++;
                for (int i = 0; i < n_nodes; ++i) nptr[i + 1] += nptr[i];
                n2c.resize(nptr[n_nodes]);
                std::vector<int> pos(nptr.begin(), nptr.end());
                for (int h = 0; h < n_hexes; ++h)
                    if (in_region(h))
                        for (int k = 0; k < 8; ++k) n2c[pos[hexes[8 * (size_t)h + k]]++] = 8 * h + k;
                // TetgenMesh::calc_pseudo_3D_vorocells (src/TetgenMesh.cpp:819-837)
                const int mult = search_region == 3 ? 1 : -1;
                for (int pass = 0; pass < 2; ++pass) {
                    std::vector<int> vpos(vptr.begin(), vptr.end());
                    for (int h = 0; h < n_hexes; ++h) {
                        int t = mult * hex_marker[h];
                        if (t <= 0) continue;
                        t -= 1;
                        if (t >= tn) continue;
                        for (int k = 0; k < 8; ++k) {
                            const int node = hexes[8 * (size_t)h + k];
                            if (node != t && node_kind[node] >= 2) {
                                if (pass == 0) vptr[t + 1]++;
                                else vor[vpos[t]++] = node;
                            }
                        }
                    }
                    if (pass == 0) {
                        for (int i = 0; i < tn; ++i) vptr[i + 1] += vptr[i];
                        vor.resize(vptr[tn]);
                    }
                }
            }
            ip->n_n2c = (int)n2c.size();
            ip->n_vor = (int)vor.size();
            ip->det0 = upload(ip->L.det0.data(), ip->L.det0.size(), s);
            ip->det = upload(ip->L.det.data(), ip->L.det.size(), s);
            ip->cen = upload(ip->L.centroid.data(), ip->L.centroid.size(), s);
            ip->marker = upload(ip->L.marker.data(), ip->L.marker.size(), s);
            ip->c10 = upload(c10.data(), c10.size(), s);
            ip->child = upload(ip->B.child.data(), ip->B.child.size(), s);
            ip->box = upload(ip->B.box.data(), ip->B.box.size(), s);
            ip->cbox = upload(ip->B.cbox.data(), ip->B.cbox.size(), s);
            ip->leafbox = upload(ip->B.leafbox.data(), ip->B.leafbox.size(), s);
            ip->prim = upload(ip->B.prim.data(), ip->B.prim.size(), s);
            ip->n2c_ptr = upload(nptr.data(), nptr.size(), s);
            ip->n2c = upload(n2c.data(), n2c.size(), s);
            ip->vor_ptr = upload(vptr.data(), vptr.size(), s);
            ip->vor = upload(vor.data(), vor.size(), s);
            IPCHECK(hipStreamSynchronize(s));   // the host vectors are read until the copies end
        }
        ip->esc = tree_escapes(s, ip->B.n_leaves, ip->child);
        // the nodal solution starts at the empty value everywhere (filled on the device)
        IPCHECK(hipMalloc(&ip->sol, 5 * (size_t)n_nodes * sizeof(double)));
        launch_set_const(s, 5 * n_nodes, empty_value, ip->sol);
        IPCHECK(hipGetLastError());
        IPCHECK(hipStreamSynchronize(s));
        return FGPU_OK;
    });
    if (rc != FGPU_OK) {
        free_interp(ip);
        delete ip;
        return rc;
    }
    c->refs++;
    *out = ip;
    return FGPU_OK;
}

namespace {
unsigned long long fnv64(const void *p, size_t bytes) {
    unsigned long long h = 1469598103934665603ULL;
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < bytes; ++i) {
        h ^= b[i];
        h *= 1099511628211ULL;
    }
    return h;
}
template <class T> unsigned long long fnv_host(const std::vector<T> &v) { return fnv64(v.data(), v.size() * sizeof(T)); }
template <class T> unsigned long long fnv_device(const T *d, size_t n, hipStream_t st) {
    std::vector<T> h(n);
    if (n) IPCHECK(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, st));
    IPCHECK(hipStreamSynchronize(st));
    return fnv_host(h);
}
}  // namespace

extern "C" int fgpu_interp_digest(fgpu_interp *ip, int n, unsigned long long *out) {
    if (!ip || !out || n < FGPU_INTERP_DIGEST_TABLES) return FGPU_EINVAL;
    return guard(ip, [&] {
        IPCHECK(hipSetDevice(ip->ctx->device));
        hipStream_t st = ip->ctx->stream;
        ensure_host_tets(ip);
        const size_t nt = ip->n_tets, ni = nt > 0 ? nt - 1 : 0;
        const LinTetHost &L = ip->L;
        int k = 0;
        out[k++] = fnv_host(L.det0);
        out[k++] = fnv_host(L.det);
        out[k++] = fnv_host(L.centroid);
        out[k++] = fnv_host(L.marker);
        out[k++] = fnv_host(L.nbr_ptr);
        out[k++] = fnv_host(L.nbr);
        out[k++] = fnv_device(ip->det0, nt, st);
        out[k++] = fnv_device(ip->det, 16 * nt, st);
        out[k++] = fnv_device(ip->cen, 3 * nt, st);
        out[k++] = fnv_device(ip->marker, nt, st);
        out[k++] = fnv_device(ip->d_nbr_ptr, nt + 1, st);
        out[k++] = fnv_device(ip->d_nbr, (size_t)ip->n_nbr, st);
        out[k++] = fnv_device(ip->child, 2 * ni, st);
        out[k++] = fnv_device(ip->box, 6 * ni, st);
        out[k++] = fnv_device(ip->cbox, 6 * ni, st);
        out[k++] = fnv_device(ip->leafbox, 6 * nt, st);
        out[k++] = fnv_device(ip->prim, nt, st);
        out[k++] = fnv_device(ip->c10, 10 * nt, st);
        out[k++] = fnv_device(ip->n2c_ptr, (size_t)ip->n_nodes + 1, st);
        out[k++] = fnv_device(ip->n2c, (size_t)ip->n_n2c, st);
        out[k++] = fnv_device(ip->vor_ptr, (size_t)ip->n_tetnodes + 1, st);
        out[k++] = fnv_device(ip->vor, (size_t)ip->n_vor, st);
        out[k++] = fnv64(&ip->decay, sizeof(double));
        return FGPU_OK;
    });
}

extern "C" void fgpu_interp_destroy(fgpu_interp *ip) {
    if (!ip) return;
    fgpu_ctx *c = ip->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_interp(ip);
    delete ip;
    fgpu_ctx_release_ref(c);
}

static void node_map(fgpu_interp *ip, const fgpu_mesh *m);

extern "C" int fgpu_interp_extract(fgpu_interp *ip, fgpu_sys *s, int smoothen) {
    if (!ip || !s) return FGPU_EINVAL;
    return guard(ip, [&] {
        IPCHECK(hipSetDevice(ip->ctx->device));
        const fgpu_mesh *m = static_cast<const fgpu_mesh *>(s->mesh);
        if (s->ctx != ip->ctx) throw std::invalid_argument("system of another context");
        if (m->n_vertices != ip->n_nodes)
            throw std::invalid_argument("the system's mesh was not imported from these nodes");
        hipStream_t st = ip->ctx->stream;
        node_map(ip, m);
        hipLaunchKernelGGL(k_extract, dim3(grid(ip->n_nodes)), dim3(256), 0, st, ip->n_nodes, ip->node_idof, s->x,
                           ip->n2c_ptr, ip->n2c, ip->hexes, ip->xyz, s->rho_saved ? s->rho : nullptr, ip->empty, ip->sol);
        if (smoothen && ip->n_tetnodes > 0)
            hipLaunchKernelGGL(k_smooth, dim3(grid(ip->n_tetnodes)), dim3(256), 0, st, ip->n_tetnodes, ip->vor_ptr,
                               ip->vor, ip->xyz, ip->decay, ip->sol);
        IPCHECK(hipGetLastError());
        IPCHECK(hipStreamSynchronize(st));
        return FGPU_OK;
    });
}

// node -> internal dof of mesh m (cached per mesh); every corner of the region's hexes must
// be a dof of m
static void node_map(fgpu_interp *ip, const fgpu_mesh *m) {
    {
        hipStream_t st = ip->ctx->stream;
        if (ip->map_mesh != m->uid) {
            std::vector<int> map(ip->n_nodes, -1);
            for (int v = 0; v < ip->n_nodes; ++v) {
                const int d = m->dof_of_vertex[v];
                if (d >= 0) map[v] = m->idof_of_dof[d];
            }
            // every corner of the region's hexes must be a dof of the system
            std::vector<int> nptr(ip->n_nodes + 1);
            IPCHECK(hipMemcpy(nptr.data(), ip->n2c_ptr, nptr.size() * sizeof(int), hipMemcpyDeviceToHost));
            std::vector<int> n2c(nptr.back());
            if (!n2c.empty()) IPCHECK(hipMemcpy(n2c.data(), ip->n2c, n2c.size() * sizeof(int), hipMemcpyDeviceToHost));
            std::vector<int> hx(8 * (size_t)ip->n_hexes);
            if (!hx.empty()) IPCHECK(hipMemcpy(hx.data(), ip->hexes, hx.size() * sizeof(int), hipMemcpyDeviceToHost));
            for (int e : n2c)
                for (int k = 0; k < 8; ++k)
                    if (map[hx[8 * (size_t)(e >> 3) + k]] < 0)
                        throw std::invalid_argument("a hex of the search region has a node outside the system's mesh");
            release(ip->node_idof);
            ip->node_idof = upload(map.data(), map.size(), st);
            IPCHECK(hipStreamSynchronize(st));
            ip->map_mesh = m->uid;
        }
    }
}

extern "C" int fgpu_interp_extract_heat(fgpu_interp *ip, fgpu_sys *cur, fgpu_sys *heat) {
    if (!ip || !cur || !heat) return FGPU_EINVAL;
    return guard(ip, [&] {
        IPCHECK(hipSetDevice(ip->ctx->device));
        const fgpu_mesh *m = static_cast<const fgpu_mesh *>(cur->mesh);
        if (cur->ctx != ip->ctx || heat->ctx != ip->ctx) throw std::invalid_argument("system of another context");
        if (heat->mesh != cur->mesh) throw std::invalid_argument("current and heat systems must share the mesh");
        if (m->n_vertices != ip->n_nodes) throw std::invalid_argument("the systems' mesh was not imported from these nodes");
        hipStream_t st = ip->ctx->stream;
        node_map(ip, m);
        double *grad = nullptr;
        IPCHECK(hipMalloc(&grad, 3 * (size_t)std::max(1, m->n_dofs) * sizeof(double)));
        launch_grad_export(st, m->n_dofs, m->d_v2c, m->d_v2n, m->d_cells, m->d_X64, cur->x, grad);
        hipLaunchKernelGGL(k_extract_heat, dim3(grid(ip->n_nodes)), dim3(256), 0, st, ip->n_nodes, ip->node_idof, grad,
                           cur->x, heat->x, ip->empty, ip->sol);
        const hipError_t e = hipGetLastError();
        const hipError_t e2 = hipStreamSynchronize(st);
        (void)hipFree(grad);
        IPCHECK(e);
        IPCHECK(e2);
        return FGPU_OK;
    });
}

// ------------------------------------------------------------------------------------------
// space-charge deposit (PoissonSolver::assemble_space_charge_fast, src/PoissonSolver.cpp:300-319)
// ------------------------------------------------------------------------------------------
// per particle p: the shape functions of its cell's hex at x_p (LinearHexahedra::shape_functions,
// Newton inversion) in the solver cell's lexicographic order (shape_funs_dealii,
// src/InterpolatorCells.cpp:1355-1358), times factor, keyed by internal dof; val = 8p + k
// keeps the reference's accumulation order through a stable sort
__global__ __launch_bounds__(256) void k_sc_shape(int n, const double *__restrict__ pts, const int *__restrict__ cells,
                                                  const int *__restrict__ region_hex, const int *__restrict__ hexes,
                                                  const double *__restrict__ xyz, const int *__restrict__ icell,
                                                  const int *__restrict__ mcells, double factor,
                                                  double *__restrict__ contrib, int *__restrict__ key,
                                                  int *__restrict__ val) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    constexpr int kUcdOfLex[8] = {0, 1, 4, 5, 3, 2, 7, 6};
    const int c = cells[p];
    const int *h = hexes + 8 * (size_t)region_hex[c];
    D3 x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = ldx(xyz, h[k]);
    double sf[8];
    linhex_sf(x, pts + 3 * (size_t)p, sf);
    const int *cl = mcells + 8 * (size_t)icell[c];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const size_t j = 8 * (size_t)p + k;
        contrib[j] = sf[kUcdOfLex[k]] * factor;
        key[j] = cl[k];
        val[j] = (int)j;
    }
}

// b[dof] += its contributions in particle order, one thread per run of equal sorted keys
__global__ __launch_bounds__(256) void k_sc_fold(int m, const int *__restrict__ key, const int *__restrict__ val,
                                                 const double *__restrict__ contrib, double *__restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || (i > 0 && key[i] == key[i - 1])) return;
    const int d = key[i];
    double acc = b[d];
    for (int j = i; j < m && key[j] == d; ++j) acc += contrib[val[j]];
    b[d] = acc;
}

extern "C" int fgpu_sys_add_space_charge(fgpu_sys *s, fgpu_interp *ip, int n, const double *xyz, const int *cells,
                                         double charge_factor) {
    if (!s || !ip || n < 0 || (n > 0 && (!xyz || !cells))) return FGPU_EINVAL;
    if (n == 0) return FGPU_OK;
    if (n > INT_MAX / 8) return FGPU_EINVAL;
    return guard(ip, [&] {
        IPCHECK(hipSetDevice(ip->ctx->device));
        fgpu_mesh *m = static_cast<fgpu_mesh *>(s->mesh);
        if (s->ctx != ip->ctx) throw std::invalid_argument("system of another context");
        if (s->lifted) throw std::invalid_argument("space charge is added before fgpu_sys_apply_dirichlet");
        if (m->n_vertices != ip->n_nodes || (int)ip->region_hex.size() != m->n_cells)
            throw std::invalid_argument("the system's mesh was not imported from the interpolator's region hexes");
        for (int p = 0; p < n; ++p)
            if (cells[p] < 0 || cells[p] >= m->n_cells) throw std::invalid_argument("particle cell out of range");
        hipStream_t st = ip->ctx->stream;
        if (ip->sc_mesh != m->uid) {
            // solver cell c must hold the dofs of hex region_hex[c]'s nodes (as sets: the
            // import may have flipped the cell)
            std::vector<int> hx(8 * (size_t)ip->n_hexes);
            IPCHECK(hipMemcpy(hx.data(), ip->hexes, hx.size() * sizeof(int), hipMemcpyDeviceToHost));
            for (int c = 0; c < m->n_cells; ++c) {
                int a[8], b[8];
                for (int k = 0; k < 8; ++k) {
                    a[k] = m->dof_of_vertex[hx[8 * (size_t)ip->region_hex[c] + k]];
                    b[k] = m->cells_lex[8 * (size_t)c + k];
                }
                std::sort(a, a + 8);
                std::sort(b, b + 8);
                if (!std::equal(a, a + 8, b)) throw std::invalid_argument("solver cell and region hex differ");
            }
            release(ip->sc_icell);
            ip->sc_icell = upload(m->icell_of_cell.data(), m->icell_of_cell.size(), st);
            IPCHECK(hipStreamSynchronize(st));
            ip->sc_mesh = m->uid;
        }
        const int nk = 8 * n;
        int bits = 1;
        while (bits < 31 && (1 << bits) < m->n_dofs) ++bits;
        double *d_pts = nullptr, *contrib = nullptr;
        int *d_cells = nullptr, *key = nullptr, *val = nullptr, *skey = nullptr, *sval = nullptr;
        void *tmp = nullptr;
        size_t tmp_bytes = 0;
        auto cleanup = [&] {
            (void)hipStreamSynchronize(st);
            for (void *q : {(void *)d_pts, (void *)contrib, (void *)d_cells, (void *)key, (void *)val, (void *)skey,
                            (void *)sval, tmp})
                if (q) (void)hipFree(q);
        };
        try {
            d_pts = upload(xyz, 3 * (size_t)n, st);
            d_cells = upload(cells, (size_t)n, st);
            for (int **q : {&key, &val, &skey, &sval}) IPCHECK(hipMalloc(q, nk * sizeof(int)));
            IPCHECK(hipMalloc(&contrib, nk * sizeof(double)));
            hipLaunchKernelGGL(k_sc_shape, dim3(grid(n)), dim3(256), 0, st, n, d_pts, d_cells, ip->d_region_hex,
                               ip->hexes, ip->xyz, ip->sc_icell, m->d_cells, charge_factor, contrib, key, val);
            IPCHECK(hipGetLastError());
            // LSD radix sort: stable, so equal dofs keep the particle order
            unsigned *ukey = reinterpret_cast<unsigned *>(key), *uskey = reinterpret_cast<unsigned *>(skey);
            IPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, ukey, uskey, val, sval, nk, 0, bits, st));
            IPCHECK(hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
            IPCHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, ukey, uskey, val, sval, nk, 0, bits, st));
            hipLaunchKernelGGL(k_sc_fold, dim3(grid(nk)), dim3(256), 0, st, nk, skey, sval, contrib, s->b);
            IPCHECK(hipGetLastError());
            IPCHECK(hipStreamSynchronize(st));   // the host inputs are read until the copies end
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
        s->no_replay = "particles are deposited again after a move: reset and assemble";
        return FGPU_OK;
    });
}

// LinearTriangles::fast_distance (src/InterpolatorCells.cpp:1743-1747) on the 16-double rows
__device__ inline double tri_fast_distance(const double *tab, const double *p, int f) {
    const double *t = tab + 16 * (size_t)f;
    const D3 tv = sub(mk(p[0], p[1], p[2]), mk(t[0], t[1], t[2]));
    const D3 q = mk(tv.y * t[5] - tv.z * t[4], tv.z * t[3] - tv.x * t[5], tv.x * t[4] - tv.y * t[3]);
    return t[6] * q.x + t[7] * q.y + t[8] * q.z;
}
// LinearTetrahedra::locate_cell from guess g (src/InterpolatorCells.cpp:268-307) with the
// point's tree candidates standing in for the linear scan
__device__ inline int tet_locate_from(const double *det0, const double *det, const int *nbr_ptr, const int *nbr,
                                      int n_tets, const double *p, int g, int4 c) {
    if (g >= 0 && g < n_tets) {
        if (pic(det0, det, p, g)) return g;
        for (int e = nbr_ptr[g]; e < nbr_ptr[g + 1]; ++e)
            if (pic(det0, det, p, nbr[e])) return nbr[e];
    }
    return c.x >= 0 ? c.x : -c.z;
}

// The cells the surface transfer interpolates in (SolutionReader::calc_interpolation for dim 2,
// src/SolutionReader.cpp:136-190): per point its face (rank 1, 2) or quadrangle (rank 3) in
// cells[] (located: the chained face in, the quadrangle around the face's largest barycentric
// coordinate out, LinearQuadrangles::locate_cell, src/InterpolatorCells.cpp:1969-1990, with
// LinearTriangles::shape_functions, :1652-1660), then the tet the face picks
// (LinearTriangles::interp_solution, :1662-1686: the face's first tet when the point lies on its
// plane or inside that tet, else the search from its second tet) or the hex the quadrangle picks
// (LinearQuadrangles::interp_solution, :1922-1940: the first hex when on the plane or when it
// holds the point, else LinearHexahedra::locate_cell from the second).  flag |= 1 when a
// quadrangle pick has no largest coordinate (NaN weights): the host then resolves the call.
__global__ __launch_bounds__(256) void k_surface_pick(int n, const double *__restrict__ pts, int rank, int locate,
                                                      const double *__restrict__ tab, const int *__restrict__ tri2tet,
                                                      const int *__restrict__ quad2hex, const int4 *__restrict__ tcand,
                                                      const double *__restrict__ det0, const double *__restrict__ det,
                                                      const int *__restrict__ nbr_ptr, const int *__restrict__ nbr,
                                                      int n_tets, int *__restrict__ cells, int *__restrict__ use,
                                                      int *__restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double p[3] = {pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]};
    int r = cells[i];
    if (locate && rank == 3) {
        // LinearTriangles::shape_functions of the face, the quadrangle around its largest weight
        const int sign = r < 0 ? -1 : 1, f = abs(r);
        const double *t = tab + 16 * (size_t)f;
        const D3 tv = sub(mk(p[0], p[1], p[2]), mk(t[0], t[1], t[2]));
        const D3 q = mk(tv.y * t[5] - tv.z * t[4], tv.z * t[3] - tv.x * t[5], tv.x * t[4] - tv.y * t[3]);
        const double v = tv.x * t[9] + tv.y * t[10] + tv.z * t[11];
        const double ww = q.x * t[12] + q.y * t[13] + q.z * t[14];
        const double u = 1.0 - v - ww;
        const double b0 = kZero + u, b1 = kZero + v, b2 = kZero + ww;
        if (b0 >= b1 && b0 >= b2) r = sign * (3 * f);
        else if (b1 >= b0 && b1 >= b2) r = sign * (3 * f + 1);
        else if (b2 >= b0 && b2 >= b1) r = sign * (3 * f + 2);
        else {
            r = -1;
            atomicOr(flag, 1);
        }
        cells[i] = r;
    }
    const int4 tc = tcand[i];
    int out;
    if (rank == 3) {
        const int qd = abs(r);
        const int h0 = quad2hex[2 * (size_t)qd], h1 = quad2hex[2 * (size_t)qd + 1];
        out = INT_MIN;
        if (fabs(tri_fast_distance(tab, p, qd / 3)) <= 100.0 * kZero) out = h0;
        double b[4];
        if (out == INT_MIN) {
            bcc(det0, det, p, h0 / 4, b);
            if (b[0] >= 0 && b[1] >= 0 && b[2] >= 0 && b[3] >= 0) {
                const int k = h0 % 4;
                bool top = true;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (j != k && !(b[k] >= b[j])) top = false;
                if (top) out = h0;
            }
        }
        if (out == INT_MIN) {
            int tt = tet_locate_from(det0, det, nbr_ptr, nbr, n_tets, p, h1 / 4, tc);
            const int sign = tt < 0 ? -1 : 1;
            tt = abs(tt);
            bcc(det0, det, p, tt, b);
            if (b[0] >= b[1] && b[0] >= b[2] && b[0] >= b[3]) out = sign * (4 * tt);
            else if (b[1] >= b[0] && b[1] >= b[2] && b[1] >= b[3]) out = sign * (4 * tt + 1);
            else if (b[2] >= b[0] && b[2] >= b[1] && b[2] >= b[3]) out = sign * (4 * tt + 2);
            else if (b[3] >= b[0] && b[3] >= b[1] && b[3] >= b[2]) out = sign * (4 * tt + 3);
            else out = -1;
        }
    } else {
        const int f = abs(r);
        const int t0 = tri2tet[2 * (size_t)f], t1 = tri2tet[2 * (size_t)f + 1];
        if (fabs(tri_fast_distance(tab, p, f)) <= 100.0 * kZero) out = t0;
        else if (pic(det0, det, p, t0)) out = t0;
        else out = tet_locate_from(det0, det, nbr_ptr, nbr, n_tets, p, t1, tc);
    }
    use[i] = out;
}

// Pic::update_point_cell (src/Pic.cpp:185-195): each particle located from its own previous
// cell, LinearHexahedra::locate_cell (src/InterpolatorCells.cpp:1536-1562) over
// LinearTetrahedra::locate_cell (:269-307): the guess tet, its neighbours in order, else the
// first containing region tet in index order (cand.x) or, if none, minus the nearest centroid
// (cand.z); the hex by the largest barycentric coordinate; back to the solver cell, -1 when
// the point lies outside every region tet
__global__ __launch_bounds__(256) void k_update_cells(int n, const double *__restrict__ pts, const int4 *__restrict__ cand,
                                                      const int *__restrict__ region_hex, const int *__restrict__ hex_cell,
                                                      const int *__restrict__ nbr_ptr, const int *__restrict__ nbr,
                                                      const double *__restrict__ det0, const double *__restrict__ det,
                                                      int n_tets, int *__restrict__ cells) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double p[3] = {pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]};
    const int4 c = cand[i];
    if (c.w) {                 // the tree walk overflowed: the candidates are incomplete
        cells[i] = INT_MIN;
        return;
    }
    const int g = region_hex[cells[i]] / 4;
    int r;
    if (c.y == 0) {
        r = -c.z;
    } else if (c.y == 1 && c.x >= 0) {
        r = c.x;
    } else {
        r = INT_MIN;
        if (g >= 0 && g < n_tets) {
            if (pic(det0, det, p, g)) {
                r = g;
            } else {
                for (int e = nbr_ptr[g]; e < nbr_ptr[g + 1]; ++e)
                    if (pic(det0, det, p, nbr[e])) {
                        r = nbr[e];
                        break;
                    }
            }
        }
        if (r == INT_MIN) r = c.x >= 0 ? c.x : -c.z;
    }
    if (r < 0) {
        cells[i] = -1;
        return;
    }
    double b[4];
    bcc(det0, det, p, r, b);
    int k = -1;
    if (b[0] >= b[1] && b[0] >= b[2] && b[0] >= b[3]) k = 0;
    else if (b[1] >= b[0] && b[1] >= b[2] && b[1] >= b[3]) k = 1;
    else if (b[2] >= b[0] && b[2] >= b[1] && b[2] >= b[3]) k = 2;
    else if (b[3] >= b[0] && b[3] >= b[1] && b[3] >= b[2]) k = 3;
    cells[i] = k < 0 ? -1 : hex_cell[4 * r + k];
}

extern "C" int fgpu_interp_update_cells(fgpu_interp *ip, int n, const double *xyz, int *cells) {
    if (!ip || n < 0 || (n > 0 && (!xyz || !cells))) return FGPU_EINVAL;
    if (n == 0) return FGPU_OK;
    if (ip->region_hex.empty()) return FGPU_EINVAL;
    return guard(ip, [&] {
        for (int i = 0; i < n; ++i)
            if (cells[i] < 0 || cells[i] >= (int)ip->region_hex.size())
                throw std::invalid_argument("particle cell out of range");
        IPCHECK(hipSetDevice(ip->ctx->device));
        hipStream_t st = ip->ctx->stream;
        ensure_points(ip, n);
        IPCHECK(hipMemcpyAsync(ip->pts, xyz, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
        IPCHECK(hipMemcpyAsync(ip->cells, cells, (size_t)n * sizeof(int), hipMemcpyHostToDevice, st));
        tet_candidates(ip, n, ip->cand);
        hipLaunchKernelGGL(k_update_cells, dim3(grid(n)), dim3(256), 0, st, n, ip->pts, ip->cand, ip->d_region_hex,
                           ip->d_hex_cell, ip->d_nbr_ptr, ip->d_nbr, ip->det0, ip->det, ip->n_tets, ip->cells);
        IPCHECK(hipGetLastError());
        std::vector<int> res(n);
        IPCHECK(hipMemcpyAsync(res.data(), ip->cells, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, st));
        IPCHECK(hipStreamSynchronize(st));
        for (int v : res)
            if (v == INT_MIN) throw std::runtime_error("point search stack overflow");
        std::copy(res.begin(), res.end(), cells);
        return FGPU_OK;
    });
}

extern "C" int fgpu_interp_hex_gradient(fgpu_interp *ip, int n, const double *xyz, const int *hexes, double *grad) {
    if (!ip || n < 0 || (n > 0 && (!xyz || !hexes || !grad))) return FGPU_EINVAL;
    if (n == 0) return FGPU_OK;
    if (ip->n_hexes == 0) return FGPU_EINVAL;
    return guard(ip, [&] {
        for (int i = 0; i < n; ++i)
            if (hexes[i] < 0 || hexes[i] >= ip->n_hexes) throw std::invalid_argument("hex index out of range");
        IPCHECK(hipSetDevice(ip->ctx->device));
        hipStream_t st = ip->ctx->stream;
        ensure_points(ip, n);
        IPCHECK(hipMemcpyAsync(ip->pts, xyz, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
        IPCHECK(hipMemcpyAsync(ip->cells, hexes, (size_t)n * sizeof(int), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_hex_gradient, dim3(grid(n)), dim3(256), 0, st, n, ip->pts, ip->cells, ip->hexes, ip->xyz,
                           ip->sol, ip->out);
        IPCHECK(hipGetLastError());
        IPCHECK(hipMemcpyAsync(grad, ip->out, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
        IPCHECK(hipStreamSynchronize(st));
        return FGPU_OK;
    });
}

extern "C" int fgpu_interp_get_nodal(fgpu_interp *ip, double *sol5) {
    if (!ip || !sol5) return FGPU_EINVAL;
    return guard(ip, [&] {
        IPCHECK(hipSetDevice(ip->ctx->device));
        IPCHECK(hipMemcpyAsync(sol5, ip->sol, 5 * (size_t)ip->n_nodes * sizeof(double), hipMemcpyDeviceToHost,
                               ip->ctx->stream));
        IPCHECK(hipStreamSynchronize(ip->ctx->stream));
        return FGPU_OK;
    });
}

extern "C" int fgpu_interp_set_nodal(fgpu_interp *ip, const double *sol5) {
    if (!ip || !sol5) return FGPU_EINVAL;
    return guard(ip, [&] {
        IPCHECK(hipSetDevice(ip->ctx->device));
        IPCHECK(hipMemcpyAsync(ip->sol, sol5, 5 * (size_t)ip->n_nodes * sizeof(double), hipMemcpyHostToDevice,
                               ip->ctx->stream));
        IPCHECK(hipStreamSynchronize(ip->ctx->stream));
        return FGPU_OK;
    });
}

static int interp_run(fgpu_interp *ip, int n, const double *xyz, int rank, int *cells, double *out, bool locate) {
    if (!ip || n < 0 || (n > 0 && (!xyz || !cells || !out)) || rank < 1 || rank > 3) return FGPU_EINVAL;
    if (rank >= 2 && ip->n_hexes == 0) return FGPU_EINVAL;
    if (n == 0) return FGPU_OK;
    return guard(ip, [&] {
        IPCHECK(hipSetDevice(ip->ctx->device));
        hipStream_t st = ip->ctx->stream;
        ensure_points(ip, n);
        IPCHECK(hipMemcpyAsync(ip->pts, xyz, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
        if (locate) {
            int flag = 1;
            if (!ip->host_chain) {
                // the first point's guess: |-1| as a tet, or |-1| / 4 from the hex numbering
                device_chain(ip, n, rank == 3 ? 0 : 1);
                hipLaunchKernelGGL(k_interp, dim3(grid(n)), dim3(256), 0, st, n, ip->pts, rank, rank == 3 ? 1 : 0,
                                   ip->tet, ip->det0, ip->det, ip->tets, ip->c10, ip->hexes, ip->xyz, ip->sol,
                                   ip->cells, ip->out);
                IPCHECK(hipGetLastError());
                IPCHECK(hipMemcpyAsync(&flag, ip->chain_flag, sizeof(int), hipMemcpyDeviceToHost, st));
                IPCHECK(hipMemcpyAsync(cells, ip->cells, n * sizeof(int), hipMemcpyDeviceToHost, st));
                IPCHECK(hipMemcpyAsync(out, ip->out, 5 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
                IPCHECK(hipStreamSynchronize(st));
            } else {
                tet_candidates(ip, n, ip->cand);
                IPCHECK(hipGetLastError());
            }
            if (flag) {
                // the chain on the host, point after point (FGPU_INTERP_HOST_CHAIN, or a
                // traversal overflow, or a point in more than kChainW - 1 cells)
                std::vector<int4> cand(n);
                IPCHECK(hipMemcpyAsync(cand.data(), ip->cand, n * sizeof(int4), hipMemcpyDeviceToHost, st));
                IPCHECK(hipStreamSynchronize(st));
                for (int i = 0; i < n; ++i)
                    if (cand[i].w) throw std::runtime_error("point search stack overflow");
                std::vector<int> tet(n);
                ensure_host_tets(ip);
                resolve_chain(ip->L, n, xyz, cand.data(), rank, tet.data());
                IPCHECK(hipMemcpyAsync(ip->tet, tet.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
                hipLaunchKernelGGL(k_interp, dim3(grid(n)), dim3(256), 0, st, n, ip->pts, rank, rank == 3 ? 1 : 0,
                                   ip->tet, ip->det0, ip->det, ip->tets, ip->c10, ip->hexes, ip->xyz, ip->sol,
                                   ip->cells, ip->out);
                IPCHECK(hipGetLastError());
                IPCHECK(hipMemcpyAsync(cells, ip->cells, n * sizeof(int), hipMemcpyDeviceToHost, st));
                IPCHECK(hipMemcpyAsync(out, ip->out, 5 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
                IPCHECK(hipStreamSynchronize(st));
            }
        } else {
            const int lim = rank == 3 ? ip->n_hexes : ip->n_tets;
            for (int i = 0; i < n; ++i)
                if (cells[i] <= -lim || cells[i] >= lim) throw std::invalid_argument("cell index out of range");
            IPCHECK(hipMemcpyAsync(ip->tet, cells, n * sizeof(int), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_interp, dim3(grid(n)), dim3(256), 0, st, n, ip->pts, rank, 0, ip->tet, ip->det0, ip->det,
                               ip->tets, ip->c10, ip->hexes, ip->xyz, ip->sol, ip->cells, ip->out);
            IPCHECK(hipGetLastError());
            IPCHECK(hipMemcpyAsync(out, ip->out, 5 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
            IPCHECK(hipStreamSynchronize(st));
        }
        return FGPU_OK;
    });
}

extern "C" int fgpu_interp_points(fgpu_interp *ip, int n, const double *xyz, int rank, int *cells, double *out) {
    return interp_run(ip, n, xyz, rank, cells, out, true);
}

extern "C" int fgpu_interp_cells(fgpu_interp *ip, int n, const double *xyz, int rank, const int *cells, double *out) {
    return interp_run(ip, n, xyz, rank, const_cast<int *>(cells), out, false);
}

// ------------------------------------------------------------------------------------------
// surface faces (dim 2)
// ------------------------------------------------------------------------------------------
extern "C" int fgpu_interp_set_triangles(fgpu_interp *ip, int n_tris, const int *tris, const int *tri2tet,
                                         const int *quads) {
    if (!ip || n_tris < 0 || (n_tris > 0 && (!tris || !tri2tet))) return FGPU_EINVAL;
    return guard(ip, [&] {
        for (size_t i = 0; i < 3 * (size_t)n_tris; ++i)
            if (tris[i] < 0 || tris[i] >= ip->n_nodes) throw std::invalid_argument("triangle node out of range");
        for (size_t i = 0; i < 2 * (size_t)n_tris; ++i)
            if (tri2tet[i] < 0 || tri2tet[i] >= ip->n_tets)
                throw std::invalid_argument("every triangle needs two tets of the mesh");
        if (quads)
            for (size_t i = 0; i < 12 * (size_t)n_tris; ++i)
                if (quads[i] < 0 || quads[i] >= ip->n_nodes) throw std::invalid_argument("quad node out of range");
        IPCHECK(hipSetDevice(ip->ctx->device));
        hipStream_t s = ip->ctx->stream;
        IPCHECK(hipStreamSynchronize(s));
        free_triangles(ip);
        if (n_tris == 0) return FGPU_OK;
        if (ip->hxyz.empty()) {
            ip->hxyz.resize(3 * (size_t)ip->n_nodes);
            IPCHECK(hipMemcpy(ip->hxyz.data(), ip->xyz, ip->hxyz.size() * sizeof(double), hipMemcpyDeviceToHost));
        }
        LinTriHost T;
        build_lintri(T, ip->n_nodes, ip->hxyz.data(), n_tris, tris, tri2tet);
        if (quads) set_quads(T, ip->hxyz.data(), ip->hkind.data(), ip->htets.data(), quads);
        // search boxes: the face widened by |edge2| on every side contains the prism of
        // points within that distance of the face plane that point_in_cell accepts
        std::vector<double> lo(3 * (size_t)n_tris), hi(3 * (size_t)n_tris);
        for (int f = 0; f < n_tris; ++f) {
            const double m = T.tab[16 * (size_t)f + 15];
            for (int a = 0; a < 3; ++a) {
                double l = 1e300, h = -1e300;
                for (int k = 0; k < 3; ++k) {
                    const double x = ip->hxyz[3 * (size_t)tris[3 * (size_t)f + k] + a];
                    l = std::min(l, x);
                    h = std::max(h, x);
                }
                lo[3 * (size_t)f + a] = l - m;
                hi[3 * (size_t)f + a] = h + m;
            }
        }
        ip->T = std::move(T);
        build_bvh(n_tris, lo, hi, ip->T.centroid, ip->TB);
        ip->tri_tab = upload(ip->T.tab.data(), ip->T.tab.size(), s);
        ip->tri_cen = upload(ip->T.centroid.data(), ip->T.centroid.size(), s);
        ip->tri_nodes = upload(ip->T.tris.data(), ip->T.tris.size(), s);
        ip->tri_child = upload(ip->TB.child.data(), ip->TB.child.size(), s);
        ip->tri_prim = upload(ip->TB.prim.data(), ip->TB.prim.size(), s);
        ip->tri_box = upload(ip->TB.box.data(), ip->TB.box.size(), s);
        ip->tri_cbox = upload(ip->TB.cbox.data(), ip->TB.cbox.size(), s);
        ip->tri_leafbox = upload(ip->TB.leafbox.data(), ip->TB.leafbox.size(), s);
        ip->tri_esc = tree_escapes(s, ip->TB.n_leaves, ip->tri_child);
        ip->tri_nbr_ptr = upload(ip->T.nbr_ptr.data(), ip->T.nbr_ptr.size(), s);
        ip->tri_nbr = upload(ip->T.nbr.data(), ip->T.nbr.size(), s);
        ip->tri2tet = upload(ip->T.tri2tet.data(), ip->T.tri2tet.size(), s);
        if (quads) ip->quad2hex = upload(ip->T.quad2hex.data(), ip->T.quad2hex.size(), s);
        if (quads) {
            build_bvh(3 * n_tris, ip->T.qcen, ip->T.qcen, ip->T.qcen, ip->QB);
            ip->quad_cen = upload(ip->T.qcen.data(), ip->T.qcen.size(), s);
            ip->quad_nodes = upload(ip->T.quads.data(), ip->T.quads.size(), s);
            ip->tri6 = upload(ip->T.tri6.data(), ip->T.tri6.size(), s);
            ip->quad_child = upload(ip->QB.child.data(), ip->QB.child.size(), s);
            ip->quad_prim = upload(ip->QB.prim.data(), ip->QB.prim.size(), s);
            ip->quad_cbox = upload(ip->QB.cbox.data(), ip->QB.cbox.size(), s);
        }
        IPCHECK(hipStreamSynchronize(s));
        return FGPU_OK;
    });
}

// the face (or quadrangle) chain, the cell picks and the interpolation on the device, out and
// (located) cells copied back; false, with nothing written, when the host has to resolve it
static bool device_surface(fgpu_interp *ip, int n, int *cells, double *out, int rank, bool locate) {
    hipStream_t st = ip->ctx->stream;
    ensure_chain(ip, n);
    IPCHECK(hipMemsetAsync(ip->chain_flag, 0, sizeof(int), st));
    tet_candidates(ip, n, ip->cand);
    if (locate) {
        const int *perm = point_order(ip, n, ip->TB.n_leaves, ip->tri_box);
        hipLaunchKernelGGL(k_candidates<TriCells>, dim3(grid(n)), dim3(256), 0, st, n, ip->pts, ip->TB.n_leaves,
                           ip->tri_child, ip->tri_esc, ip->tri_box, ip->tri_cbox, ip->tri_leafbox, ip->tri_prim,
                           TriCells{ip->tri_tab}, ip->tri_cen, ip->cand2, ip->chain_lst, perm);
        // the first guess: |-1| as a face, or |-1| / 3 from the quadrangle numbering
        run_chain(ip, n, ip->cand2, ip->tri_nbr_ptr, ip->tri_nbr, ip->T.n, rank == 3 ? 0 : 1, ip->cells2);
    } else {
        IPCHECK(hipMemcpyAsync(ip->cells2, cells, n * sizeof(int), hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_surface_pick, dim3(grid(n)), dim3(256), 0, st, n, ip->pts, rank, locate ? 1 : 0, ip->tri_tab,
                       ip->tri2tet, ip->quad2hex, ip->cand, ip->det0, ip->det, ip->d_nbr_ptr, ip->d_nbr, ip->n_tets,
                       ip->cells2, ip->tet, ip->chain_flag);
    hipLaunchKernelGGL(k_interp, dim3(grid(n)), dim3(256), 0, st, n, ip->pts, rank, 0, ip->tet, ip->det0, ip->det,
                       ip->tets, ip->c10, ip->hexes, ip->xyz, ip->sol, ip->cells, ip->out);
    IPCHECK(hipGetLastError());
    // (when the flag is set the host path overwrites cells and out again)
    int flag = 0;
    IPCHECK(hipMemcpyAsync(&flag, ip->chain_flag, sizeof(int), hipMemcpyDeviceToHost, st));
    if (locate) IPCHECK(hipMemcpyAsync(cells, ip->cells2, n * sizeof(int), hipMemcpyDeviceToHost, st));
    IPCHECK(hipMemcpyAsync(out, ip->out, 5 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
    IPCHECK(hipStreamSynchronize(st));
    return flag == 0;
}

// SolutionReader::calc_full_interpolation / calc_interpolation for dim 2 (src/SolutionReader.cpp:
// 43-123, 136-190): faces (rank 1, 2) or quads (rank 3) located with the chained guess and the
// values interpolated in the tet or hex the face picks; with centroids, the exact centroid
// match and the cell's node average
static int surface_run(fgpu_interp *ip, int n, const double *xyz, int rank, int centroids, int *cells, double *out,
                       bool locate) {
    if (!ip || n < 0 || (n > 0 && (!xyz || !cells || !out)) || rank < 1 || rank > 3) return FGPU_EINVAL;
    if (n == 0) return FGPU_OK;
    auto fail = [&](const char *msg) {
        if (ip->ctx) ip->ctx->err = msg;
        return FGPU_EINVAL;
    };
    if (ip->T.n == 0) return fail("no surface triangles set");
    if ((rank == 3 || (rank == 2 && centroids)) && ip->T.quads.empty())
        return fail("rank 3 and the rank 2 centroids need the quadrangles of the faces");
    if (rank >= 2 && !centroids && ip->n_hexes == 0) return fail("ranks 2 and 3 need the hexes of the mesh");
    const int n_cells = rank == 3 ? 3 * ip->T.n : ip->T.n;
    return guard(ip, [&] {
        if (!locate)
            for (int i = 0; i < n; ++i)
                if (cells[i] <= -n_cells || cells[i] >= n_cells || (centroids && cells[i] < 0))
                    throw std::invalid_argument("cell index out of range");
        IPCHECK(hipSetDevice(ip->ctx->device));
        hipStream_t st = ip->ctx->stream;
        ensure_points(ip, n);
        IPCHECK(hipMemcpyAsync(ip->pts, xyz, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
        std::vector<int> use(n);
        if (centroids) {
            if (locate) {
                const bool q = rank == 3;
                hipLaunchKernelGGL(k_centroid_match, dim3(grid(n)), dim3(256), 0, st, n, ip->pts,
                                   q ? ip->QB.n_leaves : ip->TB.n_leaves, q ? ip->quad_child : ip->tri_child,
                                   q ? ip->quad_cbox : ip->tri_cbox, q ? ip->quad_prim : ip->tri_prim,
                                   q ? ip->quad_cen : ip->tri_cen, ip->cand2);
                IPCHECK(hipGetLastError());
                std::vector<int4> mc(n);
                IPCHECK(hipMemcpyAsync(mc.data(), ip->cand2, n * sizeof(int4), hipMemcpyDeviceToHost, st));
                IPCHECK(hipStreamSynchronize(st));
                for (int i = 0; i < n; ++i)
                    if (mc[i].w) throw std::runtime_error("point search stack overflow");
                resolve_centroid_chain(q ? ip->T.qcen : ip->T.centroid, n, xyz, mc.data(), cells, use.data());
            } else {
                use.assign(cells, cells + n);
            }
            IPCHECK(hipMemcpyAsync(ip->tet, use.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
            const int nn = rank == 3 ? 4 : (rank == 2 ? 6 : 3);
            const int *nodes = rank == 3 ? ip->quad_nodes : (rank == 2 ? ip->tri6 : ip->tri_nodes);
            hipLaunchKernelGGL(k_interp_nodes, dim3(grid(n)), dim3(256), 0, st, n, ip->tet, nodes, nn, 1.0 / nn,
                               ip->sol, ip->out);
        } else if (!ip->host_chain && device_surface(ip, n, cells, out, rank, locate)) {
            return FGPU_OK;
        } else {
            // the chain and the cell picks on the host (FGPU_INTERP_HOST_CHAIN, or a point in
            // more than kChainW - 1 faces, or a quadrangle pick with NaN weights)
            tet_candidates(ip, n, ip->cand);
            if (locate)
                hipLaunchKernelGGL(k_candidates<TriCells>, dim3(grid(n)), dim3(256), 0, st, n, ip->pts,
                                   ip->TB.n_leaves, ip->tri_child, ip->tri_esc, ip->tri_box, ip->tri_cbox,
                                   ip->tri_leafbox, ip->tri_prim, TriCells{ip->tri_tab}, ip->tri_cen, ip->cand2,
                                   nullptr, nullptr);
            IPCHECK(hipGetLastError());
            std::vector<int4> tc(n), fc(locate ? n : 0);
            IPCHECK(hipMemcpyAsync(tc.data(), ip->cand, n * sizeof(int4), hipMemcpyDeviceToHost, st));
            if (locate) IPCHECK(hipMemcpyAsync(fc.data(), ip->cand2, n * sizeof(int4), hipMemcpyDeviceToHost, st));
            IPCHECK(hipStreamSynchronize(st));
            for (int i = 0; i < n; ++i)
                if (tc[i].w || (locate && fc[i].w)) throw std::runtime_error("point search stack overflow");
            if (locate) resolve_tri_chain(ip->T, n, xyz, fc.data(), rank == 3, cells);
            ensure_host_tets(ip);
            for (int i = 0; i < n; ++i) {
                const double *p = xyz + 3 * (size_t)i;
                use[i] = rank == 3 ? surface_hex(ip->T, ip->L, p, std::abs(cells[i]), tc[i])
                                   : surface_tet(ip->T, ip->L, p, std::abs(cells[i]), tc[i]);
            }
            IPCHECK(hipMemcpyAsync(ip->tet, use.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_interp, dim3(grid(n)), dim3(256), 0, st, n, ip->pts, rank, 0, ip->tet, ip->det0,
                               ip->det, ip->tets, ip->c10, ip->hexes, ip->xyz, ip->sol, ip->cells, ip->out);
        }
        IPCHECK(hipGetLastError());
        IPCHECK(hipMemcpyAsync(out, ip->out, 5 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
        IPCHECK(hipStreamSynchronize(st));
        return FGPU_OK;
    });
}

extern "C" int fgpu_interp_surface_points(fgpu_interp *ip, int n, const double *xyz, int rank, int centroids,
                                          int *cells, double *out) {
    return surface_run(ip, n, xyz, rank, centroids, cells, out, true);
}

extern "C" int fgpu_interp_surface_cells(fgpu_interp *ip, int n, const double *xyz, int rank, int centroids,
                                         const int *cells, double *out) {
    return surface_run(ip, n, xyz, rank, centroids, const_cast<int *>(cells), out, false);
}

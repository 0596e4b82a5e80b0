// Mesh import on the GPU: the same tables as the host import (fgpu_mesh.cpp), bit for bit, built
// with device sorts and scans.  The deal.II steps it stands for: DealSolver::import_mesh
// (src/DealSolver.cpp:192-209: unused vertices dropped, UCD -> lexicographic order, negatively
// oriented cells inverted), mark_boundary (:461-518), calc_vertex2dof (:318-341); then the
// internal order the element kernels use (Morton-sorted cells, first-touch dofs, patches, the
// dof-major slot layout, boundary-face groups, sub-hex layouts).
//
// Sequential steps of the host import and their parallel forms:
//   faces of exactly one cell      two stable radix sorts of the sorted face vertices, runs
//   first-touch dof order          atomicMin of the first entry per dof, radix sort on it
//   greedy patches (add a cell while the patch's distinct nodes stay within the cap)
//                                  prev[e] = the previous entry with the same dof; from start a
//                                  the window's distinct count is #{e : prev[e] < 8a}, so every
//                                  start's greedy length comes out in parallel and the host walks
//                                  the chain of patch starts
//   per-patch node order (by decreasing entry count, ties by first appearance)
//                                  one workgroup per patch: block scan, block radix sort
//   slot positions (patches ascending per dof)   stable radix sort of the patch nodes by dof
// Compiled without floating-point contraction: centre Jacobians, centroids, Morton keys and
// the layout test round as the host's loops do.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "femocs_gpu.h"
#include "fgpu_internal.h"

namespace fgpu {
namespace {

struct ImpHipErr : std::runtime_error {
    hipError_t e;
    ImpHipErr(const std::string &w, hipError_t e_) : std::runtime_error(w), e(e_) {}
};
#define IMCHECK(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) throw ImpHipErr(std::string(#call) + ": " + hipGetErrorString(e_), e_); \
    } while (0)

// temporaries of one import (allocation failures raise ImpHipErr)
struct ImpFail {
    [[noreturn]] static void raise(hipError_t e, const char *what) {
        throw ImpHipErr(std::string(what) + ": " + hipGetErrorString(e), e);
    }
};
using Bufs = DeviceArena<ImpFail>;
// device arrays the mesh keeps (freed by the mesh)
template <class T> T *keep(size_t n) {
    void *q = nullptr;
    IMCHECK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    return static_cast<T *>(q);
}
template <class T> void d2h(std::vector<T> &dst, const T *src, size_t n, hipStream_t s) {
    dst.resize(n);
    if (n) IMCHECK(hipMemcpyAsync(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

inline int grid(size_t n) { return (int)std::max<size_t>(1, (n + 255) / 256); }
inline int bits_for(long long n) {   // bits holding 0 .. n-1 (at least 1)
    int b = 1;
    while (b < 62 && (1LL << b) < n) ++b;
    return b;
}

// --- topology ---------------------------------------------------------------------------------

__global__ void k_mark_used(long long n8, const int *__restrict__ ucd, int nv, int *__restrict__ used,
                            int *__restrict__ bad) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n8) return;
    const int v = ucd[i];
    if (v < 0 || v >= nv) bad[0] = 1;
    else used[v] = 1;
}

__global__ void k_compact(int nv, const int *__restrict__ used, const int *__restrict__ pos, const double *__restrict__ xyz,
                          int *__restrict__ dof_of_vertex, int *__restrict__ vertex_of_dof, double *__restrict__ x) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    if (!used[v]) {
        dof_of_vertex[v] = -1;
        return;
    }
    const int d = pos[v];
    dof_of_vertex[v] = d;
    vertex_of_dof[d] = v;
    for (int a = 0; a < 3; ++a) x[3 * (size_t)d + a] = xyz[3 * (size_t)v + a];
}

// trilinear Jacobian determinant at the cell centre (lexicographic vertices), summed in the
// host import's order
__device__ double center_det_d(const double *__restrict__ x, const int *l) {
    double J[3][3];
    for (int a = 0; a < 3; ++a) {
        double p[8];
        for (int n = 0; n < 8; ++n) p[n] = x[3 * (size_t)l[n] + a];
        J[a][0] = 0.25 * ((p[1] - p[0]) + (p[3] - p[2]) + (p[5] - p[4]) + (p[7] - p[6]));
        J[a][1] = 0.25 * ((p[2] - p[0]) + (p[3] - p[1]) + (p[6] - p[4]) + (p[7] - p[5]));
        J[a][2] = 0.25 * ((p[4] - p[0]) + (p[5] - p[1]) + (p[6] - p[2]) + (p[7] - p[3]));
    }
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

// UCD -> lexicographic (deal.II's ucd_to_deal inverted), eta flip of inverted cells
__global__ void k_lex(int nc, const int *__restrict__ ucd, const int *__restrict__ dov, const double *__restrict__ x,
                      int *__restrict__ lex) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int u2l[8] = {0, 1, 4, 5, 3, 2, 7, 6};
    int l[8];
    for (int k = 0; k < 8; ++k) l[k] = dov[ucd[8 * (size_t)c + u2l[k]]];
    if (center_det_d(x, l) < 0) {
        int t;
        t = l[0]; l[0] = l[2]; l[2] = t;
        t = l[1]; l[1] = l[3]; l[3] = t;
        t = l[4]; l[4] = l[6]; l[6] = t;
        t = l[5]; l[5] = l[7]; l[7] = t;
    }
    for (int k = 0; k < 8; ++k) lex[8 * (size_t)c + k] = l[k];
}

__device__ inline void face_verts(int f, int fv[4]) {
    const int t[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
    for (int i = 0; i < 4; ++i) fv[i] = t[f][i];
}

// per (cell, face): its sorted vertices v0 < v1 < v2 < v3 as two keys, (v2, v3) and (v0, v1)
__global__ void k_face_keys(int nc, const int *__restrict__ lex, int nbits, unsigned long long *__restrict__ ka,
                            unsigned long long *__restrict__ kb, int *__restrict__ slot) {
    const long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (s >= 6LL * nc) return;
    const int c = (int)(s / 6), f = (int)(s % 6);
    int fv[4], v[4];
    face_verts(f, fv);
    for (int i = 0; i < 4; ++i) v[i] = lex[8 * (size_t)c + fv[i]];
    // sorting network of four
    int t;
#define FG_CSWAP(i, j) if (v[i] > v[j]) { t = v[i]; v[i] = v[j]; v[j] = t; }
    FG_CSWAP(0, 1) FG_CSWAP(2, 3) FG_CSWAP(0, 2) FG_CSWAP(1, 3) FG_CSWAP(1, 2)
#undef FG_CSWAP
    ka[s] = ((unsigned long long)v[2] << nbits) | (unsigned long long)v[3];
    kb[s] = ((unsigned long long)v[0] << nbits) | (unsigned long long)v[1];
    slot[s] = (int)s;
}

__global__ void k_gather_u64(int n, const unsigned long long *__restrict__ src, const int *__restrict__ idx,
                             unsigned long long *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// eq[i] = 1 when sorted face i has the same four vertices as face i-1
__global__ void k_face_eq(int n, const unsigned long long *__restrict__ kbs, const unsigned long long *__restrict__ ka,
                          const int *__restrict__ slot, int *__restrict__ eq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    eq[i] = i > 0 && kbs[i] == kbs[i - 1] && ka[slot[i]] == ka[slot[i - 1]];
}

// a face is on the boundary when it matches neither neighbour; a run of three is an error
__global__ void k_face_bd(int n, const int *__restrict__ eq, const int *__restrict__ slot, int *__restrict__ is_bd,
                          int *__restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int e0 = eq[i], e1 = i + 1 < n ? eq[i + 1] : 0;
    is_bd[slot[i]] = !e0 && !e1;
    if (e0 && i > 0 && eq[i - 1]) bad[0] = 1;
}

// boundary faces in cell x face order, with the centres of their four vertices
__global__ void k_bfaces(int n6, const int *__restrict__ is_bd, const int *__restrict__ pos, const int *__restrict__ lex,
                         const double *__restrict__ x, int *__restrict__ bf_cell, int *__restrict__ bf_face,
                         double *__restrict__ ctr) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n6 || !is_bd[s]) return;
    const int b = pos[s], c = s / 6, f = s % 6;
    bf_cell[b] = c;
    bf_face[b] = f;
    int fv[4];
    face_verts(f, fv);
    for (int a = 0; a < 3; ++a) {
        double t = 0;
        for (int i = 0; i < 4; ++i) t += x[3 * (size_t)lex[8 * (size_t)c + fv[i]] + a];
        ctr[3 * (size_t)b + a] = t / 4.0;
    }
}

// bounding box of n points (3 doubles each): per-block minima and maxima, then one block
constexpr int kBoxBlocks = 256;
__global__ __launch_bounds__(256) void k_box_part(int n, const double *__restrict__ p, double *__restrict__ part) {
    __shared__ double s[6][256];
    double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        for (int a = 0; a < 3; ++a) {
            const double q = p[3 * (size_t)i + a];
            v[a] = fmin(v[a], q);
            v[3 + a] = fmax(v[3 + a], q);
        }
    for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; ++k)
                s[k][threadIdx.x] = k < 3 ? fmin(s[k][threadIdx.x], s[k][threadIdx.x + w]) : fmax(s[k][threadIdx.x], s[k][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}
__global__ __launch_bounds__(256) void k_box_final(int nparts, const double *__restrict__ part, double *__restrict__ box) {
    if (threadIdx.x >= 6) return;
    const int k = threadIdx.x;
    double v = k < 3 ? 1e300 : -1e300;
    for (int b = 0; b < nparts; ++b) v = k < 3 ? fmin(v, part[6 * b + k]) : fmax(v, part[6 * b + k]);
    box[k] = v;
}

__global__ void k_bf_ids(int nb, const double *__restrict__ ctr, const double *__restrict__ box, int top, int bottom,
                         int sides, int other, int *__restrict__ bf_id) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const double eps = 1e-6;
    const double *p = ctr + 3 * (size_t)i;
    const double *lo = box, *hi = box + 3;
    const bool side = fabs(p[0] - lo[0]) <= eps || fabs(p[0] - hi[0]) <= eps || fabs(p[1] - lo[1]) <= eps ||
                      fabs(p[1] - hi[1]) <= eps;
    int id;
    if (side) id = sides;
    else if (fabs(p[2] - hi[2]) <= eps) id = top;
    else if (fabs(p[2] - lo[2]) <= eps) id = bottom;
    else id = other;
    bf_id[i] = id;
}

__global__ void k_fill(long long n, int v, int *__restrict__ a) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// the last (cell, vertex) entry in cell order touching each dof: the largest 8c + k
__global__ void k_v2c(int nc, const int *__restrict__ lex, int *__restrict__ key) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= 8LL * nc) return;
    atomicMax(&key[lex[e]], (int)e);
}

// --- internal order ---------------------------------------------------------------------------

__device__ inline unsigned long long spread3_d(unsigned long long v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

// Morton key of each cell's vertex centroid in the dofs' bounding cube
__global__ void k_morton(int nc, const int *__restrict__ lex, const double *__restrict__ x, const double *__restrict__ box,
                         unsigned long long *__restrict__ key, int *__restrict__ cell) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double span = 0;
    for (int a = 0; a < 3; ++a) span = fmax(span, box[3 + a] - box[a]);
    if (span <= 0) span = 1;
    unsigned long long k = 0;
    for (int a = 0; a < 3; ++a) {
        double s = 0;
        for (int n = 0; n < 8; ++n) s += x[3 * (size_t)lex[8 * (size_t)c + n] + a];
        s = (s / 8.0 - box[a]) / span;
        const unsigned long long q = (unsigned long long)fmin(2097151.0, fmax(0.0, s * 2097152.0));
        k |= spread3_d(q) << a;
    }
    key[c] = k;
    cell[c] = c;
}

__global__ void k_inverse(int n, const int *__restrict__ perm, int *__restrict__ inv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[perm[i]] = i;
}

__global__ void k_iota(int n, int *__restrict__ a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// first entry (8 * internal cell + vertex) touching each dof
__global__ void k_first_touch(int nc, const int *__restrict__ cell_of_icell, const int *__restrict__ lex, int *__restrict__ ft) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= 8LL * nc) return;
    const int i = (int)(e >> 3), k = (int)(e & 7);
    atomicMin(&ft[lex[8 * (size_t)cell_of_icell[i] + k]], (int)e);
}

__global__ void k_icells(int nc, const int *__restrict__ cell_of_icell, const int *__restrict__ lex,
                         const int *__restrict__ idof_of_dof, int *__restrict__ icells) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= 8LL * nc) return;
    const int i = (int)(e >> 3), k = (int)(e & 7);
    icells[e] = idof_of_dof[lex[8 * (size_t)cell_of_icell[i] + k]];
}

__global__ void k_coords(int nd, const int *__restrict__ dof_of_idof, const double *__restrict__ x, double *__restrict__ X64,
                         float *__restrict__ X32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const int d = dof_of_idof[i];
    for (int a = 0; a < 3; ++a) {
        const double v = x[3 * (size_t)d + a];
        X64[4 * (size_t)i + a] = v;
        X32[4 * (size_t)i + a] = (float)v;
    }
    X64[4 * (size_t)i + 3] = 0.0;
    X32[4 * (size_t)i + 3] = 0.0f;
}

// layout of each internal cell: 1 when its eta flip is closer to the tet-split sub-hex
__device__ double subhex_dev(const double *__restrict__ X, const int *l, bool flip) {
    const double sub[8][3] = {{0, 0, 0}, {0.5, 0, 0}, {0, 0, 0.5}, {1.0 / 3, 0, 1.0 / 3},
                              {0, 0.5, 0}, {1.0 / 3, 1.0 / 3, 0}, {0, 1.0 / 3, 1.0 / 3}, {0.25, 0.25, 0.25}};
    const int fl[8] = {2, 3, 0, 1, 6, 7, 4, 5};
    double x[8][3];
    for (int n = 0; n < 8; ++n)
        for (int a = 0; a < 3; ++a) x[n][a] = X[4 * (size_t)l[flip ? fl[n] : n] + a];
    double u[3][3];
    for (int a = 0; a < 3; ++a) {
        u[0][a] = 2 * (x[1][a] - x[0][a]);
        u[1][a] = 2 * (x[4][a] - x[0][a]);
        u[2][a] = 2 * (x[2][a] - x[0][a]);
    }
    double size = 0, d = 0;
    for (int i = 0; i < 3; ++i) size += sqrt(u[i][0] * u[i][0] + u[i][1] * u[i][1] + u[i][2] * u[i][2]);
    for (int n = 0; n < 8; ++n)
        for (int a = 0; a < 3; ++a)
            d = fmax(d, fabs(x[0][a] + sub[n][0] * u[0][a] + sub[n][1] * u[1][a] + sub[n][2] * u[2][a] - x[n][a]));
    return size > 0 ? d / size : 1e300;
}
__global__ void k_layout(int nc, const int *__restrict__ icells, const double *__restrict__ X, unsigned char *__restrict__ layout) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int l[8];
    for (int k = 0; k < 8; ++k) l[k] = icells[8 * (size_t)c + k];
    layout[c] = subhex_dev(X, l, true) < subhex_dev(X, l, false) ? 1 : 0;
}

// --- patches ----------------------------------------------------------------------------------

// prev/next entry with the same dof (entries sorted by dof, ties by entry)
__global__ void k_prev_next(int n, const int *__restrict__ sk, const int *__restrict__ se, int *__restrict__ prev,
                            int *__restrict__ next) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int e = se[t];
    prev[e] = (t > 0 && sk[t - 1] == sk[t]) ? se[t - 1] : -1;
    next[e] = (t + 1 < n && sk[t + 1] == sk[t]) ? se[t + 1] : INT_MAX;
}

// Greedy patch from every start cell a: the longest run of at most cpp cells whose distinct
// nodes stay within cap, and that count.  A workgroup takes 64 consecutive starts and stages
// the prev entries of their windows in LDS; a wave handles one start at a time (each lane
// cpp / 64 consecutive cells).
constexpr int kStarts = 64;
__global__ __launch_bounds__(256) void k_patch_len(int nc, int cpp, int cap, const int *__restrict__ prev,
                                                   int2 *__restrict__ len) {
    __shared__ int sp[8 * (kStarts + kPatchCells)];
    const int a0 = blockIdx.x * kStarts;
    const int wcells = min(kStarts + cpp - 1, nc - a0);
    for (int i = threadIdx.x; i < 8 * wcells; i += 256) sp[i] = prev[8 * (size_t)a0 + i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, per = cpp / 64;
    for (int s = wave; s < kStarts; s += 4) {
        const int a = a0 + s;
        if (a >= nc) break;   // uniform over the wave
        const int lim = 8 * a;
        int cum[4] = {0, 0, 0, 0}, tot = 0;
        for (int q = 0; q < per; ++q) {
            const int idx = lane * per + q, c = a + idx;
            int n = cap + 1;   // past the mesh: never within the cap
            if (c < nc) {
                n = 0;
                const int *pc = sp + 8 * (c - a0);
                for (int j = 0; j < 8; ++j) n += pc[j] < lim;
            }
            tot += n;
            cum[q] = tot;
        }
        int incl = tot;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        const int excl = incl - tot;
        int L = 0, np = 0;
        for (int q = 0; q < per; ++q)
            if (excl + cum[q] <= cap) {
                L = lane * per + q + 1;
                np = excl + cum[q];
            }
        for (int o = 32; o > 0; o >>= 1) {
            L = max(L, __shfl_xor(L, o));
            np = max(np, __shfl_xor(np, o));
        }
        if (lane == 0) len[a] = make_int2(L, np);
    }
}

// One workgroup per patch: local node ids by first appearance (block scan over the cells),
// entry counts by walking each node's next links within the patch, the final order by
// decreasing count then first appearance (block radix sort of (2048 - count, id)),
// then the node list, the local index of every entry, and per node its entries ascending.
constexpr int kSortItems = (kPatchNodes + 255) / 256;
__global__ __launch_bounds__(256) void k_patch_build(const int *__restrict__ cell0, const int *__restrict__ node0,
                                                     const int *__restrict__ icells, const int *__restrict__ prev,
                                                     const int *__restrict__ next, int *__restrict__ nodes,
                                                     uint16_t *__restrict__ lidx, uint16_t *__restrict__ lptr,
                                                     uint16_t *__restrict__ lent, int *__restrict__ bad) {
    using BScan = hipcub::BlockScan<int, 256>;
    using BSort = hipcub::BlockRadixSort<unsigned int, 256, kSortItems>;
    __shared__ union {
        typename BScan::TempStorage scan;
        typename BSort::TempStorage sort;
    } tmp;
    __shared__ int snx[8 * kPatchCells];
    __shared__ int sfirst[kPatchNodes], sdeg[kPatchNodes], snew[kPatchNodes], soff[kPatchNodes];
    const int p = blockIdx.x, t = threadIdx.x;
    const int a = cell0[p], b = cell0[p + 1], ncp = b - a, ne = 8 * ncp;
    const int E0 = 8 * a, E1 = 8 * b;
    const int nb0 = node0[p], np = node0[p + 1] - nb0;
    if (ncp > kPatchCells || np > kPatchNodes) {   // uniform; the host checks this before the launch
        if (t == 0) bad[0] = 1;
        return;
    }
    bool first[8];
    int cnt = 0;
    for (int j = 0; j < 8; ++j) {
        first[j] = false;
        if (t < ncp) {
            const int g = E0 + 8 * t + j;
            first[j] = prev[g] < E0;
            cnt += first[j];
            const int nx = next[g];
            snx[8 * t + j] = nx < E1 ? nx - E0 : -1;
        }
    }
    int off = 0, total = 0;
    BScan(tmp.scan).ExclusiveSum(cnt, off, total);
    if (total != np) {   // uniform
        if (t == 0) bad[0] = 1;
        return;
    }
    for (int j = 0; j < 8; ++j)
        if (first[j]) sfirst[off++] = 8 * t + j;
    __syncthreads();
    for (int i = t; i < np; i += 256) {
        int e = sfirst[i], d = 1;
        while ((e = snx[e]) >= 0) ++d;
        sdeg[i] = d;
    }
    __syncthreads();
    unsigned int key[kSortItems];
    for (int q = 0; q < kSortItems; ++q) {
        const int i = kSortItems * t + q;
        key[q] = i < np ? ((unsigned int)(2048 - sdeg[i]) << 10) | (unsigned int)i : 0xffffffffu;
    }
    BSort(tmp.sort).Sort(key, 0, 21);
    // blocked arrangement: thread t holds ranks kSortItems * t + q
    int dl[kSortItems], sum = 0;
    for (int q = 0; q < kSortItems; ++q) {
        const int r = kSortItems * t + q;
        dl[q] = 0;
        if (r < np) {
            const int i = (int)(key[q] & 1023u);
            snew[i] = r;
            dl[q] = sdeg[i];
        }
        sum += dl[q];
    }
    __syncthreads();
    int o = 0;
    BScan(tmp.scan).ExclusiveSum(sum, o);
    uint16_t *ptr = lptr + (size_t)nb0 + p;
    for (int q = 0; q < kSortItems; ++q) {
        const int r = kSortItems * t + q;
        if (r < np) {
            soff[r] = o;
            ptr[r] = (uint16_t)o;
        }
        o += dl[q];
    }
    if (t == 0) ptr[np] = (uint16_t)ne;
    __syncthreads();
    for (int i = t; i < np; i += 256) {
        const int r = snew[i];
        int e = sfirst[i], k = soff[r];
        nodes[nb0 + r] = icells[E0 + e];
        do {
            lidx[E0 + e] = (uint16_t)r;
            lent[E0 + k] = (uint16_t)e;
            ++k;
            e = snx[e];
        } while (e >= 0);
    }
}

// slot positions: patch nodes sorted by dof (stable), slot[j] = rank; gptr[d] = first rank of d
__global__ void k_slots(int n, int nd, const int *__restrict__ sd, const int *__restrict__ sj, int *__restrict__ slot,
                        int *__restrict__ gptr) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    slot[sj[t]] = t;
    if (t == 0 || sd[t] != sd[t - 1]) gptr[sd[t]] = t;
    if (t == n - 1) gptr[nd] = n;
}

// boundary-face slots (4 per face, in face order) -> dofs; faces' internal cells
__global__ void k_face_dofs(int nb, const int *__restrict__ bf_cell, const int *__restrict__ bf_face,
                            const int *__restrict__ lex, const int *__restrict__ idof_of_dof,
                            const int *__restrict__ icell_of_cell, int *__restrict__ fdof, int *__restrict__ cnt,
                            int *__restrict__ bf_icell) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int c = bf_cell[b];
    bf_icell[b] = icell_of_cell[c];
    int fv[4];
    face_verts(bf_face[b], fv);
    for (int j = 0; j < 4; ++j) {
        const int d = idof_of_dof[lex[8 * (size_t)c + fv[j]]];
        fdof[4 * (size_t)b + j] = d;
        atomicAdd(&cnt[d], 1);
    }
}

// gradient export tables in internal numbering
__global__ void k_v2c_internal(int nd, const int *__restrict__ dof_of_idof, const int *__restrict__ key,
                               const int *__restrict__ icell_of_cell, int *__restrict__ v2c, int *__restrict__ v2n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const int k = key[dof_of_idof[i]];
    v2c[i] = icell_of_cell[k >> 3];
    v2n[i] = k & 7;
}

// bit 15 of a cell's first local index carries its sub-hex layout
__global__ void k_layout_bit(int nc, const unsigned char *__restrict__ layout, uint16_t *__restrict__ lidx) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc && layout[c]) lidx[8 * (size_t)c] |= 0x8000;
}

// hipcub radix sort of pairs into caller buffers (temporary storage from B)
template <class K, class V>
void sort_pairs(Bufs &B, const K *kin, K *kout, const V *vin, V *vout, int n, int end_bit, hipStream_t st) {
    size_t tb = 0;
    IMCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, vin, vout, n, 0, end_bit, st));
    void *tmp = B.scratch(tb);
    IMCHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, kin, kout, vin, vout, n, 0, end_bit, st));
}
template <class T> void exclusive_sum(Bufs &B, const T *in, T *out, int n, hipStream_t st) {
    size_t tb = 0;
    IMCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, st));
    void *tmp = B.scratch(tb);
    IMCHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, in, out, n, st));
}
int read_int(const int *d, hipStream_t st) {
    int h = 0;
    IMCHECK(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, st));
    IMCHECK(hipStreamSynchronize(st));
    return h;
}

}  // namespace

int device_import(Mesh &m, hipStream_t st, int nv, const double *xyz, int nc, const int *ucd, int top, int bottom,
                  int sides, int other, std::string &err) {
    try {
        if (nv <= 0 || nc <= 0 || !xyz || !ucd || nc > INT_MAX / 8) {
            err = "mesh import: invalid sizes or null arrays";
            return FGPU_EINVAL;
        }
        // about 0.9 KB per cell and 0.1 KB per vertex of temporaries (sort buffers included)
        Bufs B(960 * (size_t)nc + 112 * (size_t)nv + (1 << 20));
        const long long n8 = 8LL * nc;
        const int n6 = 6 * nc;
        // --- topology
        int *d_ucd = B.get<int>(n8);
        double *d_xyz = B.get<double>(3 * (size_t)nv);
        IMCHECK(hipMemcpyAsync(d_ucd, ucd, n8 * sizeof(int), hipMemcpyHostToDevice, st));
        IMCHECK(hipMemcpyAsync(d_xyz, xyz, 3 * (size_t)nv * sizeof(double), hipMemcpyHostToDevice, st));
        int *used = B.get<int>(nv), *vpos = B.get<int>(nv), *bad = B.get<int>(4);
        hipLaunchKernelGGL(k_fill, dim3(grid(nv)), dim3(256), 0, st, (long long)nv, 0, used);
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, st, 4LL, 0, bad);
        hipLaunchKernelGGL(k_mark_used, dim3(grid(n8)), dim3(256), 0, st, n8, d_ucd, nv, used, bad);
        IMCHECK(hipGetLastError());
        if (read_int(bad, st)) {
            err = "mesh import: cell vertex index out of range";
            return FGPU_EINVAL;
        }
        exclusive_sum(B, used, vpos, nv, st);
        const int nd = read_int(vpos + nv - 1, st) + read_int(used + nv - 1, st);
        m.n_vertices = nv;
        m.n_cells = nc;
        m.n_dofs = nd;
        int *dov = B.get<int>(nv), *vod = B.get<int>(nd);
        double *x = B.get<double>(3 * (size_t)nd);
        hipLaunchKernelGGL(k_compact, dim3(grid(nv)), dim3(256), 0, st, nv, used, vpos, d_xyz, dov, vod, x);
        int *lex = B.get<int>(n8);
        hipLaunchKernelGGL(k_lex, dim3(grid(nc)), dim3(256), 0, st, nc, d_ucd, dov, x, lex);
        // faces: sort by (v0, v1, v2, v3), boundary = unmatched
        const int nbits = bits_for(nd);
        auto *ka = B.get<unsigned long long>(n6), *kb = B.get<unsigned long long>(n6);
        auto *kas = B.get<unsigned long long>(n6), *kbg = B.get<unsigned long long>(n6), *kbs = B.get<unsigned long long>(n6);
        int *fslot = B.get<int>(n6), *fslot1 = B.get<int>(n6), *fslot2 = B.get<int>(n6);
        hipLaunchKernelGGL(k_face_keys, dim3(grid(n6)), dim3(256), 0, st, nc, lex, nbits, ka, kb, fslot);
        sort_pairs(B, ka, kas, fslot, fslot1, n6, 2 * nbits, st);
        hipLaunchKernelGGL(k_gather_u64, dim3(grid(n6)), dim3(256), 0, st, n6, kb, fslot1, kbg);
        sort_pairs(B, kbg, kbs, fslot1, fslot2, n6, 2 * nbits, st);
        int *eq = B.get<int>(n6), *is_bd = B.get<int>(n6), *bpos = B.get<int>(n6);
        hipLaunchKernelGGL(k_face_eq, dim3(grid(n6)), dim3(256), 0, st, n6, kbs, ka, fslot2, eq);
        hipLaunchKernelGGL(k_face_bd, dim3(grid(n6)), dim3(256), 0, st, n6, eq, fslot2, is_bd, bad);
        IMCHECK(hipGetLastError());
        if (read_int(bad, st)) {
            err = "mesh import: a face is shared by more than two cells";
            return FGPU_EINVAL;
        }
        exclusive_sum(B, is_bd, bpos, n6, st);
        const int nb = read_int(bpos + n6 - 1, st) + read_int(is_bd + n6 - 1, st);
        int *bf_cell = B.get<int>(nb), *bf_face = keep<int>(nb);
        m.d_bf_face = bf_face;
        double *ctr = B.get<double>(3 * (size_t)nb), *part = B.get<double>(6 * kBoxBlocks), *fbox = B.get<double>(6);
        double *xbox = B.get<double>(6);
        hipLaunchKernelGGL(k_bfaces, dim3(grid(n6)), dim3(256), 0, st, n6, is_bd, bpos, lex, x, bf_cell, bf_face, ctr);
        m.d_bf_id = keep<int>(nb);
        if (nb > 0) {
            hipLaunchKernelGGL(k_box_part, dim3(kBoxBlocks), dim3(256), 0, st, nb, ctr, part);
            hipLaunchKernelGGL(k_box_final, dim3(1), dim3(256), 0, st, kBoxBlocks, part, fbox);
            hipLaunchKernelGGL(k_bf_ids, dim3(grid(nb)), dim3(256), 0, st, nb, ctr, fbox, top, bottom, sides, other, m.d_bf_id);
        }
        int *v2key = B.get<int>(nd);
        hipLaunchKernelGGL(k_fill, dim3(grid(nd)), dim3(256), 0, st, (long long)nd, -1, v2key);
        hipLaunchKernelGGL(k_v2c, dim3(grid(n8)), dim3(256), 0, st, nc, lex, v2key);
        // --- internal order: cells along the Morton curve, dofs by first touch
        hipLaunchKernelGGL(k_box_part, dim3(kBoxBlocks), dim3(256), 0, st, nd, x, part);
        hipLaunchKernelGGL(k_box_final, dim3(1), dim3(256), 0, st, kBoxBlocks, part, xbox);
        auto *mkey = B.get<unsigned long long>(nc), *mkeys = B.get<unsigned long long>(nc);
        int *mcell = B.get<int>(nc), *coi = B.get<int>(nc), *ioc = B.get<int>(nc);
        hipLaunchKernelGGL(k_morton, dim3(grid(nc)), dim3(256), 0, st, nc, lex, x, xbox, mkey, mcell);
        sort_pairs(B, mkey, mkeys, mcell, coi, nc, 63, st);
        hipLaunchKernelGGL(k_inverse, dim3(grid(nc)), dim3(256), 0, st, nc, coi, ioc);
        int *ft = B.get<int>(nd), *fts = B.get<int>(nd), *dseq = B.get<int>(nd), *doi = B.get<int>(nd), *iod = B.get<int>(nd);
        hipLaunchKernelGGL(k_fill, dim3(grid(nd)), dim3(256), 0, st, (long long)nd, INT_MAX, ft);
        hipLaunchKernelGGL(k_first_touch, dim3(grid(n8)), dim3(256), 0, st, nc, coi, lex, ft);
        hipLaunchKernelGGL(k_iota, dim3(grid(nd)), dim3(256), 0, st, nd, dseq);
        sort_pairs(B, ft, fts, dseq, doi, nd, bits_for(n8), st);
        hipLaunchKernelGGL(k_inverse, dim3(grid(nd)), dim3(256), 0, st, nd, doi, iod);
        m.d_cells = keep<int>(n8);
        hipLaunchKernelGGL(k_icells, dim3(grid(n8)), dim3(256), 0, st, nc, coi, lex, iod, m.d_cells);
        m.d_X64 = keep<double>(4 * (size_t)nd);
        m.d_X32 = keep<float>(4 * (size_t)nd);
        hipLaunchKernelGGL(k_coords, dim3(grid(nd)), dim3(256), 0, st, nd, doi, x, m.d_X64, m.d_X32);
        m.d_layout = keep<unsigned char>(nc);
        hipLaunchKernelGGL(k_layout, dim3(grid(nc)), dim3(256), 0, st, nc, m.d_cells, m.d_X64, m.d_layout);
        IMCHECK(hipGetLastError());
        // --- patches: greedy lengths from every start, the chain of starts on the host
        const int cpp = default_patch_cells(nc);
        const int cap = std::min(kPatchNodes, 3 * cpp + 64);
        if (cpp % 64 != 0 || cpp > kPatchCells) throw std::logic_error("mesh import: unexpected patch size");
        int *ekey = B.get<int>(n8), *eseq = B.get<int>(n8), *eks = B.get<int>(n8), *es = B.get<int>(n8);
        int *prev = B.get<int>(n8), *next = B.get<int>(n8);
        IMCHECK(hipMemcpyAsync(ekey, m.d_cells, n8 * sizeof(int), hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_iota, dim3(grid(n8)), dim3(256), 0, st, (int)n8, eseq);
        sort_pairs(B, ekey, eks, eseq, es, (int)n8, nbits, st);
        hipLaunchKernelGGL(k_prev_next, dim3(grid(n8)), dim3(256), 0, st, (int)n8, eks, es, prev, next);
        int2 *len = B.get<int2>(nc);
        hipLaunchKernelGGL(k_patch_len, dim3((nc + kStarts - 1) / kStarts), dim3(256), 0, st, nc, cpp, cap, prev, len);
        IMCHECK(hipGetLastError());
        std::vector<int2> hlen;
        d2h(hlen, len, nc, st);
        IMCHECK(hipStreamSynchronize(st));
        Patches &P = m.patches;
        P = Patches();
        P.cells = cpp;
        P.cell0.assign(1, 0);
        P.node0.assign(1, 0);
        for (int c = 0; c < nc;) {
            const int2 l = hlen[c];
            if (l.x < 1 || l.x > nc - c || l.y < 1 || l.y > cap) throw std::logic_error("mesh import: patch length out of range");
            c += l.x;
            P.cell0.push_back(c);
            P.node0.push_back(P.node0.back() + l.y);
            P.node_cap = std::max(P.node_cap, l.y);
            P.n++;
        }
        P.node_cap = (P.node_cap + 7) / 8 * 8;
        const int S = P.node0.back();
        m.n_slots = S;
        m.d_pcell0 = keep<int>(P.n + 1);
        m.d_pnode0 = keep<int>(P.n + 1);
        IMCHECK(hipMemcpyAsync(m.d_pcell0, P.cell0.data(), (P.n + 1) * sizeof(int), hipMemcpyHostToDevice, st));
        IMCHECK(hipMemcpyAsync(m.d_pnode0, P.node0.data(), (P.n + 1) * sizeof(int), hipMemcpyHostToDevice, st));
        m.d_pnodes = keep<int>(S);
        m.d_plidx = keep<uint16_t>(n8);
        m.d_plptr = keep<uint16_t>((size_t)S + P.n);
        m.d_plent = keep<uint16_t>(n8);
        hipLaunchKernelGGL(k_patch_build, dim3(P.n), dim3(256), 0, st, m.d_pcell0, m.d_pnode0, m.d_cells, prev, next,
                           m.d_pnodes, m.d_plidx, m.d_plptr, m.d_plent, bad);
        IMCHECK(hipGetLastError());
        // slots
        int *sseq = B.get<int>(S), *sd = B.get<int>(S), *sj = B.get<int>(S);
        hipLaunchKernelGGL(k_iota, dim3(grid(S)), dim3(256), 0, st, S, sseq);
        sort_pairs(B, m.d_pnodes, sd, sseq, sj, S, nbits, st);
        m.d_pslot = keep<int>(S);
        m.d_gptr = keep<int>((size_t)nd + 1);
        hipLaunchKernelGGL(k_slots, dim3(grid(S)), dim3(256), 0, st, S, nd, sd, sj, m.d_pslot, m.d_gptr);
        // boundary-face groups
        m.d_bf_icell = keep<int>(nb);
        m.d_fgptr = keep<int>((size_t)nd + 1);
        m.d_fgidx = keep<int>(4 * (size_t)nb);
        int *fdof = B.get<int>(4 * (size_t)nb), *fcnt = B.get<int>((size_t)nd + 1), *fseq = B.get<int>(4 * (size_t)nb);
        int *fdofs = B.get<int>(4 * (size_t)nb);
        hipLaunchKernelGGL(k_fill, dim3(grid(nd + 1)), dim3(256), 0, st, (long long)nd + 1, 0, fcnt);
        if (nb > 0) {
            hipLaunchKernelGGL(k_face_dofs, dim3(grid(nb)), dim3(256), 0, st, nb, bf_cell, bf_face, lex, iod, ioc, fdof, fcnt,
                               m.d_bf_icell);
            hipLaunchKernelGGL(k_iota, dim3(grid(4 * (size_t)nb)), dim3(256), 0, st, 4 * nb, fseq);
            sort_pairs(B, fdof, fdofs, fseq, m.d_fgidx, 4 * nb, nbits, st);
        }
        exclusive_sum(B, fcnt, m.d_fgptr, nd + 1, st);
        m.d_v2c = keep<int>(nd);
        m.d_v2n = keep<int>(nd);
        hipLaunchKernelGGL(k_v2c_internal, dim3(grid(nd)), dim3(256), 0, st, nd, doi, v2key, ioc, m.d_v2c, m.d_v2n);
        IMCHECK(hipGetLastError());
        // --- host mirrors
        d2h(m.dof_of_vertex, dov, nv, st);
        d2h(m.vertex_of_dof, vod, nd, st);
        d2h(m.x, x, 3 * (size_t)nd, st);
        d2h(m.cells_lex, lex, n8, st);
        d2h(m.bf_cell, bf_cell, nb, st);
        d2h(m.bf_face, bf_face, nb, st);
        d2h(m.bf_id, m.d_bf_id, nb, st);
        d2h(m.cell_of_icell, coi, nc, st);
        d2h(m.icell_of_cell, ioc, nc, st);
        d2h(m.dof_of_idof, doi, nd, st);
        d2h(m.idof_of_dof, iod, nd, st);
        // (the patch lists, the vertex -> cell map and the internal cells stay on the device:
        // the host reads them back only to rebuild patches)
        d2h(m.layout, m.d_layout, nc, st);
        int hbad = 0;
        IMCHECK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
        IMCHECK(hipStreamSynchronize(st));
        if (hbad) throw std::logic_error("mesh import: patch tables inconsistent");
        hipLaunchKernelGGL(k_layout_bit, dim3(grid(nc)), dim3(256), 0, st, nc, m.d_layout, m.d_plidx);
        IMCHECK(hipGetLastError());
        IMCHECK(hipStreamSynchronize(st));
        m.patch_gen++;
        return FGPU_OK;
    } catch (const ImpHipErr &e) {
        err = e.what();
        return e.e == hipErrorOutOfMemory ? FGPU_ENOMEM : FGPU_EHIP;
    } catch (const std::bad_alloc &) {
        err = "host allocation failed";
        return FGPU_ENOMEM;
    } catch (const std::exception &e) {
        err = e.what();
        return FGPU_EINVAL;
    }
}

// 64-bit FNV-1a of every import table: the host mirrors, then the device arrays
namespace {
unsigned long long fnv(const void *p, size_t bytes, unsigned long long h = 1469598103934665603ULL) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < bytes; ++i) {
        h ^= b[i];
        h *= 1099511628211ULL;
    }
    return h;
}
template <class T> unsigned long long fnv_vec(const std::vector<T> &v) { return fnv(v.data(), v.size() * sizeof(T)); }
template <class T> unsigned long long fnv_dev(const T *d, size_t n, hipStream_t st) {
    std::vector<T> h;
    d2h(h, d, n, st);
    IMCHECK(hipStreamSynchronize(st));
    return fnv_vec(h);
}
}  // namespace
}  // namespace fgpu

extern "C" int fgpu_mesh_digest(const fgpu_mesh *m, int n, unsigned long long *out) {
    using namespace fgpu;
    if (!m || !out || n < FGPU_MESH_DIGEST_TABLES) return FGPU_EINVAL;
    fgpu_ctx *c = m->ctx;
    try {
        IMCHECK(hipSetDevice(c->device));
        hipStream_t st = c->stream;
        const Patches &P = m->patches;
        const size_t nd = m->n_dofs, nc = m->n_cells, nb = m->bf_cell.size(), S = m->n_slots, np = P.n;
        int k = 0;
        out[k++] = fnv_vec(m->vertex_of_dof);
        out[k++] = fnv_vec(m->dof_of_vertex);
        out[k++] = fnv_vec(m->cells_lex);
        out[k++] = fnv_vec(m->x);
        out[k++] = fnv_vec(m->bf_cell);
        out[k++] = fnv_vec(m->bf_face);
        out[k++] = fnv_vec(m->bf_id);
        out[k++] = fnv_vec(m->icell_of_cell);
        out[k++] = fnv_vec(m->cell_of_icell);
        out[k++] = fnv_vec(m->idof_of_dof);
        out[k++] = fnv_vec(m->dof_of_idof);
        {
            const int head[4] = {P.n, P.cells, P.node_cap, m->n_slots};
            unsigned long long h = fnv(head, sizeof(head));
            h = fnv(P.cell0.data(), P.cell0.size() * 4, h);
            h = fnv(P.node0.data(), P.node0.size() * 4, h);
            out[k++] = h;
        }
        out[k++] = fnv_vec(m->layout);
        out[k++] = fnv_dev(m->d_cells, 8 * nc, st);
        out[k++] = fnv_dev(m->d_X64, 4 * nd, st);
        out[k++] = fnv_dev(m->d_X32, 4 * nd, st);
        out[k++] = fnv_dev(m->d_gptr, nd + 1, st);
        out[k++] = fnv_dev(m->d_pcell0, np + 1, st);
        out[k++] = fnv_dev(m->d_pnode0, np + 1, st);
        out[k++] = fnv_dev(m->d_pnodes, S, st);
        out[k++] = fnv_dev(m->d_pslot, S, st);
        out[k++] = fnv_dev(m->d_plidx, 8 * nc, st);
        out[k++] = fnv_dev(m->d_plptr, S + np, st);
        out[k++] = fnv_dev(m->d_plent, 8 * nc, st);
        out[k++] = fnv_dev(m->d_bf_icell, nb, st);
        out[k++] = fnv_dev(m->d_bf_face, nb, st);
        out[k++] = fnv_dev(m->d_bf_id, nb, st);
        out[k++] = fnv_dev(m->d_fgptr, nd + 1, st);
        out[k++] = fnv_dev(m->d_fgidx, 4 * nb, st);
        out[k++] = fnv_dev(m->d_v2c, nd, st);
        out[k++] = fnv_dev(m->d_v2n, nd, st);
        out[k++] = fnv_dev(m->d_layout, nc, st);
        return FGPU_OK;
    } catch (const ImpHipErr &e) {
        c->err = e.what();
        return FGPU_EHIP;
    } catch (const std::bad_alloc &) {
        c->err = "host allocation failed";
        return FGPU_ENOMEM;
    }
}

// The reference's tet -> 4-hex split on the GPU (Tethex Mesh::convert_3D, src/Tethex.cpp:965-1169),
// with its numbering: edges by (larger, smaller) vertex (IncidenceMatrix, :374-425), faces by
// first encounter over (tet, local edge, opposite vertex) (Mesh::face_numeration, :1272-1340),
// new nodes after the vertices in the order edges, faces, tets (Mesh::set_new_vertices,
// :897-919), hex 4t + k around tet vertex k with the front/back swap of change_vertices_order
// (:1565-1590), and three quadrangles per boundary triangle (Mesh::convert_triangles,
// :1104-1169).  The sequential first-encounter numbering becomes two device sorts: faces keyed
// by their sorted vertices carry their earliest encounter rank, and are numbered by that rank.
// Compiled without floating-point contraction: the new coordinates and the quadrangle measure
// round as the reference's loops do.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "femocs_gpu.h"
#include "fgpu_internal.h"

// the split's results stay on the device until fgpu_tethex_export copies them out
struct fgpu_tethex {
    int device = 0;
    int nv = 0, ne = 0, nf = 0, nt = 0, ntri = 0;
    double *xyz = nullptr;        // 3 per node
    int *hexes = nullptr;         // 8 per hex
    int *quads = nullptr;         // 4 per quadrangle
    ~fgpu_tethex() {
        if (!xyz && !hexes && !quads) return;
        (void)hipSetDevice(device);
        if (xyz) (void)hipFree(xyz);
        if (hexes) (void)hipFree(hexes);
        if (quads) (void)hipFree(quads);
    }
};

namespace {

struct TxErr : std::runtime_error {
    using std::runtime_error::runtime_error;
};
#define TXCHECK(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) throw TxErr(std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

// device buffers of one split (allocation failures raise TxErr)
struct TxFail {
    [[noreturn]] static void raise(hipError_t e, const char *what) {
        throw TxErr(std::string(what) + ": " + hipGetErrorString(e));
    }
};
using Bufs = fgpu::DeviceArena<TxFail>;

constexpr int kLexOfUcd[8] = {0, 1, 4, 5, 3, 2, 7, 6};

inline int grid(size_t n) { return (int)((n + 255) / 256); }

// local edges of a tet in edge_numeration order: pairs (i, j) with v_i > v_j; lne -> (i, j)
__device__ inline int tet_local_edges(const int v[4], int ei[6], int ej[6]) {
    int n = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            if (v[i] > v[j]) {
                ei[n] = i;
                ej[n] = j;
                ++n;
            }
    return n;
}

__device__ inline unsigned long long edge_key(int hi, int lo) {
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

__device__ inline unsigned long long face_key(int a, int b, int c) {
    // sorted ascending; 21 bits each
    int t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
    return ((unsigned long long)a << 42) | ((unsigned long long)b << 21) | (unsigned long long)c;
}

template <class K> __device__ inline int lower_bound(const K *a, int n, K key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// per tet: its 6 edge keys in local order; per tet face (opposite local vertex q): the face
// key and the first encounter rank t*24 + 4*lne + p within the tet
__global__ void k_tet_keys(int nt, const int *__restrict__ tets, unsigned long long *__restrict__ ekeys,
                           unsigned long long *__restrict__ fkeys, unsigned long long *__restrict__ frank) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int v[4];
    for (int k = 0; k < 4; ++k) v[k] = tets[4 * (size_t)t + k];
    int ei[6], ej[6];
    tet_local_edges(v, ei, ej);
    unsigned long long best[4] = {ULLONG_MAX, ULLONG_MAX, ULLONG_MAX, ULLONG_MAX};
    for (int l = 0; l < 6; ++l) {
        ekeys[6 * (size_t)t + l] = edge_key(v[ei[l]], v[ej[l]]);
        // vertices off the edge, in tet order: (edge, w) is the face opposite the other one
        for (int w = 0; w < 4; ++w) {
            if (w == ei[l] || w == ej[l]) continue;
            const int q = 6 - ei[l] - ej[l] - w;
            const unsigned long long r = 24ULL * t + 4 * l + w;
            if (r < best[q]) best[q] = r;
        }
    }
    for (int q = 0; q < 4; ++q) {
        const int a = v[(q + 1) & 3], b = v[(q + 2) & 3], c = v[(q + 3) & 3];
        fkeys[4 * (size_t)t + q] = face_key(a, b, c);
        frank[4 * (size_t)t + q] = best[q];
    }
}

// 1 at the first entry of every run of equal sorted face keys
__global__ void k_face_first(int n, const unsigned long long *__restrict__ skey, int *__restrict__ head) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || skey[i] != skey[i - 1]) ? 1 : 0;
}

// unique sorted face keys with the earliest rank of each (the smaller rank of the run)
__global__ void k_face_compact(int n, const unsigned long long *__restrict__ skey, const unsigned long long *__restrict__ srank,
                               const int *__restrict__ pos, unsigned long long *__restrict__ ukey,
                               unsigned long long *__restrict__ urank, int *__restrict__ uidx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (i > 0 && skey[i] == skey[i - 1])) return;
    unsigned long long r = srank[i];
    for (int j = i + 1; j < n && skey[j] == skey[i]; ++j) r = srank[j] < r ? srank[j] : r;
    const int u = pos[i];
    ukey[u] = skey[i];
    urank[u] = r;
    uidx[u] = u;
}

// face number of each unique key (index in rank order)
__global__ void k_face_number(int nf, const int *__restrict__ order, int *__restrict__ fnum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nf) fnum[order[i]] = i;
}

__global__ void k_edge_nodes(int ne, const unsigned long long *__restrict__ ukey, const double *__restrict__ xyz,
                             double *__restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int hi = (int)(ukey[e] >> 32), lo = (int)(ukey[e] & 0xffffffffu);
    for (int d = 0; d < 3; ++d) {
        double c = 0.0;
        c += xyz[3 * (size_t)lo + d];
        c += xyz[3 * (size_t)hi + d];
        out[3 * (size_t)e + d] = c / 2;
    }
}

__global__ void k_face_nodes(int nf, const unsigned long long *__restrict__ ukey, const int *__restrict__ fnum,
                             const double *__restrict__ xyz, double *__restrict__ out) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nf) return;
    const unsigned long long k = ukey[u];
    const int a = (int)(k >> 42), b = (int)((k >> 21) & 0x1fffff), c3 = (int)(k & 0x1fffff);
    const int f = fnum[u];
    for (int d = 0; d < 3; ++d) {
        double c = 0.0;
        c += xyz[3 * (size_t)a + d];
        c += xyz[3 * (size_t)b + d];
        c += xyz[3 * (size_t)c3 + d];
        out[3 * (size_t)f + d] = c / 3;
    }
}

__device__ inline double center_det_lex(const double x[8][3]) {
    double d[3][3];
    for (int a = 0; a < 3; ++a) {
        d[0][a] = 0.25 * ((x[1][a] - x[0][a]) + (x[3][a] - x[2][a]) + (x[5][a] - x[4][a]) + (x[7][a] - x[6][a]));
        d[1][a] = 0.25 * ((x[2][a] - x[0][a]) + (x[3][a] - x[1][a]) + (x[6][a] - x[4][a]) + (x[7][a] - x[5][a]));
        d[2][a] = 0.25 * ((x[4][a] - x[0][a]) + (x[5][a] - x[1][a]) + (x[6][a] - x[2][a]) + (x[7][a] - x[3][a]));
    }
    return d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) - d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0]) +
           d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
}

// per tet: centroid node and its four hexes
__global__ void k_tet_hexes(int nt, int nv, int ne, int nf, const int *__restrict__ tets,
                            const unsigned long long *__restrict__ eukey, const unsigned long long *__restrict__ fukey,
                            const int *__restrict__ fnum, const double *__restrict__ all, double *__restrict__ cen,
                            int *__restrict__ hexes) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int v[4];
    for (int k = 0; k < 4; ++k) v[k] = tets[4 * (size_t)t + k];
    for (int d = 0; d < 3; ++d) {
        double c = 0.0;
        for (int k = 0; k < 4; ++k) c += all[3 * (size_t)v[k] + d];
        cen[3 * (size_t)t + d] = c / 4;
    }
    int ei[6], ej[6], eid[6];
    tet_local_edges(v, ei, ej);
    for (int l = 0; l < 6; ++l) eid[l] = lower_bound(eukey, ne, edge_key(v[ei[l]], v[ej[l]]));
    const int F0 = nv + ne, C0 = nv + ne + nf;
    for (int k = 0; k < 4; ++k) {
        int s[3], o[3], n = 0;
        for (int l = 0; l < 6 && n < 3; ++l)
            if (ei[l] == k || ej[l] == k) {
                s[n] = eid[l];
                o[n] = ei[l] == k ? ej[l] : ei[l];   // local index of the edge's other end
                ++n;
            }
        // the face through v and the other ends of two of its edges
        auto face = [&](int a, int b) {
            const int u = lower_bound(fukey, nf, face_key(v[k], v[o[a]], v[o[b]]));
            return F0 + fnum[u];
        };
        int h[8] = {v[k], nv + s[0], face(0, 1), nv + s[1], nv + s[2], face(0, 2), C0 + t, face(1, 2)};
        double x[8][3];
        for (int m = 0; m < 8; ++m) {
            const int node = h[kLexOfUcd[m]];
            const double *p = node < C0 ? all + 3 * (size_t)node : cen + 3 * (size_t)(node - C0);
            x[m][0] = p[0];
            x[m][1] = p[1];
            x[m][2] = p[2];
        }
        if (center_det_lex(x) < 0)
            for (int m = 0; m < 4; ++m) {
                const int tmp = h[m];
                h[m] = h[m + 4];
                h[m + 4] = tmp;
            }
        for (int m = 0; m < 8; ++m) hexes[8 * (4 * (size_t)t + k) + m] = h[m];
    }
}

// cell_measure_2D (src/Tethex.cpp:1592-1599) of 4 points in deal.II order, x and y only
__device__ inline double measure_2d(const double *p0, const double *p1, const double *p2, const double *p3) {
    const double x[4] = {p0[0], p1[0], p2[0], p3[0]}, y[4] = {p0[1], p1[1], p2[1], p3[1]};
    return (-x[1] * y[0] + x[1] * y[3] + y[0] * x[2] + x[0] * y[1] - x[0] * y[2] - y[1] * x[3] - x[2] * y[3] +
            x[3] * y[2]) /
           2.;
}

__global__ void k_tri_quads(int ntri, int nv, int ne, int nf, const int *__restrict__ tris,
                            const unsigned long long *__restrict__ eukey, const unsigned long long *__restrict__ fukey,
                            const int *__restrict__ fnum, const double *__restrict__ all, int *__restrict__ quads,
                            int *__restrict__ bad) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= ntri) return;
    int v[3];
    for (int k = 0; k < 3; ++k) v[k] = tris[3 * (size_t)f + k];
    int ei[3], ej[3], eid[3], n = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (v[i] > v[j] && n < 3) {
                ei[n] = i;
                ej[n] = j;
                ++n;
            }
    if (n != 3) {
        bad[0] = 1;
        return;
    }
    for (int l = 0; l < 3; ++l) {
        const unsigned long long key = edge_key(v[ei[l]], v[ej[l]]);
        eid[l] = lower_bound(eukey, ne, key);
        if (eid[l] >= ne || eukey[eid[l]] != key) bad[0] = 1;
    }
    const unsigned long long fk = face_key(v[0], v[1], v[2]);
    const int u = lower_bound(fukey, nf, fk);
    if (u >= nf || fukey[u] != fk) {
        bad[0] = 1;
        return;
    }
    const int fnode = nv + ne + fnum[u];
    for (int k = 0; k < 3; ++k) {
        int s[2], m = 0;
        for (int l = 0; l < 3 && m < 2; ++l)
            if (ei[l] == k || ej[l] == k) s[m++] = eid[l];
        int q[4] = {v[k], nv + s[0], fnode, nv + s[1]};
        if (measure_2d(all + 3 * (size_t)q[0], all + 3 * (size_t)q[1], all + 3 * (size_t)q[3], all + 3 * (size_t)q[2]) < 0) {
            const int tmp = q[1];
            q[1] = q[3];
            q[3] = tmp;
        }
        for (int m2 = 0; m2 < 4; ++m2) quads[4 * (3 * (size_t)f + k) + m2] = q[m2];
    }
}

template <class F> int tx_guard(fgpu_ctx *c, F &&f) {
    try {
        return f();
    } catch (const TxErr &e) {
        if (c) c->err = e.what();
        return FGPU_EHIP;
    } catch (const std::bad_alloc &) {
        if (c) c->err = "host allocation failed";
        return FGPU_ENOMEM;
    } catch (const std::exception &e) {
        if (c) c->err = e.what();
        return FGPU_EINVAL;
    }
}

}  // namespace

extern "C" int fgpu_tethex_split(fgpu_ctx *c, int nv, const double *xyz, int nt, const int *tets, int ntri, const int *tris,
                                 fgpu_tethex **out) {
    if (!c || !out || nv <= 0 || nt <= 0 || ntri < 0 || !xyz || !tets || (ntri > 0 && !tris)) return FGPU_EINVAL;
    *out = nullptr;
    return tx_guard(c, [&] {
        if (nv >= (1 << 21)) throw std::invalid_argument("tet split: at most 2097151 vertices");
        if (nt > INT_MAX / 24) throw std::invalid_argument("tet split: too many tets");
        for (size_t i = 0; i < 4 * (size_t)nt; ++i)
            if (tets[i] < 0 || tets[i] >= nv) throw std::invalid_argument("tet split: tet vertex out of range");
        for (size_t i = 0; i < 3 * (size_t)ntri; ++i)
            if (tris[i] < 0 || tris[i] >= nv) throw std::invalid_argument("tet split: triangle vertex out of range");
        for (int t = 0; t < nt; ++t) {
            const int *v = tets + 4 * (size_t)t;
            if (v[0] == v[1] || v[0] == v[2] || v[0] == v[3] || v[1] == v[2] || v[1] == v[3] || v[2] == v[3])
                throw std::invalid_argument("tet split: a tet repeats a vertex");
        }
        TXCHECK(hipSetDevice(c->device));
        hipStream_t st = c->stream;
        fgpu_tethex *h = new fgpu_tethex();
        std::unique_ptr<fgpu_tethex> guard_h(h);
        h->device = c->device;
        // about 0.9 KB per tet and 32 bytes per vertex (sort buffers included)
        Bufs B(960 * (size_t)nt + 32 * (size_t)nv + (1 << 20));
        const int n6 = 6 * nt, n4 = 4 * nt;
        int *d_tets = B.get<int>(n4);
        double *d_xyz = B.get<double>(3 * (size_t)nv);
        TXCHECK(hipMemcpyAsync(d_tets, tets, n4 * sizeof(int), hipMemcpyHostToDevice, st));
        TXCHECK(hipMemcpyAsync(d_xyz, xyz, 3 * (size_t)nv * sizeof(double), hipMemcpyHostToDevice, st));
        auto *ek = B.get<unsigned long long>(n6), *eks = B.get<unsigned long long>(n6);
        auto *fk = B.get<unsigned long long>(n4), *fks = B.get<unsigned long long>(n4);
        auto *fr = B.get<unsigned long long>(n4), *frs = B.get<unsigned long long>(n4);
        hipLaunchKernelGGL(k_tet_keys, dim3(grid(nt)), dim3(256), 0, st, nt, d_tets, ek, fk, fr);
        TXCHECK(hipGetLastError());
        // edges: sort the keys, keep one of each
        size_t tb = 0, tb2 = 0;
        TXCHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, ek, eks, n6, 0, 64, st));
        TXCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, fk, fks, fr, frs, n4, 0, 63, st));
        size_t tb3 = 0;
        auto *euk = B.get<unsigned long long>(n6);
        int *d_count = B.get<int>(2);
        TXCHECK(hipcub::DeviceSelect::Unique(nullptr, tb3, eks, euk, d_count, n6, st));
        int *pos = B.get<int>(n4);
        int *head = B.get<int>(n4);
        size_t tb4 = 0;
        TXCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb4, head, pos, n4, st));
        const size_t tbmax = std::max(std::max(tb, tb2), std::max(tb3, tb4));
        void *tmp = B.get<unsigned char>(tbmax);
        tb = tbmax;
        TXCHECK(hipcub::DeviceRadixSort::SortKeys(tmp, tb, ek, eks, n6, 0, 64, st));
        tb = tbmax;
        TXCHECK(hipcub::DeviceSelect::Unique(tmp, tb, eks, euk, d_count, n6, st));
        // faces: sort by key, earliest rank per key, number by that rank
        tb = tbmax;
        TXCHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, fk, fks, fr, frs, n4, 0, 63, st));
        hipLaunchKernelGGL(k_face_first, dim3(grid(n4)), dim3(256), 0, st, n4, fks, head);
        tb = tbmax;
        TXCHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, head, pos, n4, st));
        int h_cnt[2] = {0, 0}, last_head = 0, last_pos = 0;
        TXCHECK(hipMemcpyAsync(&h_cnt[0], d_count, sizeof(int), hipMemcpyDeviceToHost, st));
        TXCHECK(hipMemcpyAsync(&last_head, head + n4 - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        TXCHECK(hipMemcpyAsync(&last_pos, pos + n4 - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        TXCHECK(hipStreamSynchronize(st));
        const int ne = h_cnt[0], nf = last_pos + last_head;
        auto *fuk = B.get<unsigned long long>(nf), *fur = B.get<unsigned long long>(nf);
        auto *furs = B.get<unsigned long long>(nf);
        int *fidx = B.get<int>(nf), *forder = B.get<int>(nf), *fnum = B.get<int>(nf);
        hipLaunchKernelGGL(k_face_compact, dim3(grid(n4)), dim3(256), 0, st, n4, fks, frs, pos, fuk, fur, fidx);
        size_t tb5 = 0;
        TXCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb5, fur, furs, fidx, forder, nf, 0, 64, st));
        void *tmp2 = B.get<unsigned char>(tb5);
        TXCHECK(hipcub::DeviceRadixSort::SortPairs(tmp2, tb5, fur, furs, fidx, forder, nf, 0, 64, st));
        hipLaunchKernelGGL(k_face_number, dim3(grid(nf)), dim3(256), 0, st, nf, forder, fnum);
        // new nodes and the cells
        const size_t n_nodes = (size_t)nv + ne + nf + nt;
        TXCHECK(hipMalloc(&h->xyz, 3 * n_nodes * sizeof(double)));
        TXCHECK(hipMalloc(&h->hexes, 8 * (size_t)n4 * sizeof(int)));
        TXCHECK(hipMalloc(&h->quads, 12 * (size_t)std::max(1, ntri) * sizeof(int)));
        double *all = h->xyz;
        TXCHECK(hipMemcpyAsync(all, d_xyz, 3 * (size_t)nv * sizeof(double), hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_edge_nodes, dim3(grid(ne)), dim3(256), 0, st, ne, euk, d_xyz, all + 3 * (size_t)nv);
        hipLaunchKernelGGL(k_face_nodes, dim3(grid(nf)), dim3(256), 0, st, nf, fuk, fnum, d_xyz,
                           all + 3 * ((size_t)nv + ne));
        int *d_hex = h->hexes;
        hipLaunchKernelGGL(k_tet_hexes, dim3(grid(nt)), dim3(256), 0, st, nt, nv, ne, nf, d_tets, euk, fuk, fnum, all,
                           all + 3 * ((size_t)nv + ne + nf), d_hex);
        int *d_quad = h->quads;
        int *d_bad = B.get<int>(1);
        TXCHECK(hipMemsetAsync(d_bad, 0, sizeof(int), st));
        if (ntri > 0) {
            int *d_tris = B.get<int>(3 * (size_t)ntri);
            TXCHECK(hipMemcpyAsync(d_tris, tris, 3 * (size_t)ntri * sizeof(int), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_tri_quads, dim3(grid(ntri)), dim3(256), 0, st, ntri, nv, ne, nf, d_tris, euk, fuk, fnum,
                               all, d_quad, d_bad);
        }
        TXCHECK(hipGetLastError());
        h->nv = nv;
        h->ne = ne;
        h->nf = nf;
        h->nt = nt;
        h->ntri = ntri;
        int bad = 0;
        TXCHECK(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st));
        TXCHECK(hipStreamSynchronize(st));
        if (bad) throw std::invalid_argument("tet split: a boundary triangle is not a face of the tets");
        *out = guard_h.release();
        return FGPU_OK;
    });
}

extern "C" int fgpu_tethex_sizes(const fgpu_tethex *h, int *n_nodes, int *n_edges, int *n_faces) {
    if (!h) return FGPU_EINVAL;
    if (n_nodes) *n_nodes = h->nv + h->ne + h->nf + h->nt;
    if (n_edges) *n_edges = h->ne;
    if (n_faces) *n_faces = h->nf;
    return FGPU_OK;
}

extern "C" int fgpu_tethex_export(const fgpu_tethex *h, double *xyz, int *hexes, int *quads) {
    if (!h) return FGPU_EINVAL;
    const size_t n_nodes = (size_t)h->nv + h->ne + h->nf + h->nt;
    if (hipSetDevice(h->device) != hipSuccess) return FGPU_EHIP;
    if (xyz && hipMemcpy(xyz, h->xyz, 3 * n_nodes * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return FGPU_EHIP;
    if (hexes && hipMemcpy(hexes, h->hexes, 32 * (size_t)h->nt * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return FGPU_EHIP;
    if (quads && h->ntri > 0 &&
        hipMemcpy(quads, h->quads, 12 * (size_t)h->ntri * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return FGPU_EHIP;
    return FGPU_OK;
}

extern "C" void fgpu_tethex_destroy(fgpu_tethex *h) { delete h; }

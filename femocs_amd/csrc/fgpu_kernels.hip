// gfx950 kernels of the field-solve path.
//
// Q1 hexahedra, QGauss(2) quadrature, trilinear geometry recomputed in registers.  The
// operator is applied matrix-free in two deterministic steps per product:
//   k_patch   per patch of consecutive cells: node data staged in LDS, one thread per cell
//             (sum-factorised reference gradients, per-point metric, transposed sum
//             factorisation), per-node sums in LDS (FP64: in a fixed order; FP32: exact
//             fixed-point sums), one store per (node, patch) into the dof-major slot buffer
//   k_gather  one thread per dof: add the dof's slots in ascending order, fused with the
//             CG vector work of that step and a block reduction into per-block partials,
//             which a one-block-per-system scalar kernel sums in block order
// No floating-point atomics anywhere, so a system gives bit-identical results whatever
// other systems share the launch.  Every kernel walks a block table (system, local block),
// which is how one launch serves a batch of meshes of any sizes.
#include <hip/hip_runtime.h>

#include <cmath>

#include "fgpu_kernels.h"

// a * b + c fuses only within one expression, as written: two kernels that evaluate the same
// element code round identically, whatever the compiler's scheduling around it
#pragma clang fp contract(on)

namespace fgpu {

#define FG_G0 0.21132486540518711775   // (1 - 1/sqrt(3)) / 2: QGauss(2) point on [0,1]
#define FG_G1 0.78867513459481288225   // (1 + 1/sqrt(3)) / 2

template <class T> struct Vec4T;
template <> struct Vec4T<double> { using type = double4; };
template <> struct Vec4T<float> { using type = float4; };

// 1D Q1 basis j at Gauss point a: 1 - g_a for j = 0, g_a for j = 1
template <class T> __device__ __forceinline__ T Wt(int a, int j) { return a == j ? (T)FG_G1 : (T)FG_G0; }

// values D[i][m] at the two nodes of both directions -> values at the 2x2 Gauss points
template <class T> __device__ __forceinline__ void tp2(const T D[2][2], T F[2][2]) {
    T E[2][2];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int m = 0; m < 2; ++m) E[p][m] = Wt<T>(p, 0) * D[0][m] + Wt<T>(p, 1) * D[1][m];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int s = 0; s < 2; ++s) F[p][s] = Wt<T>(s, 0) * E[p][0] + Wt<T>(s, 1) * E[p][1];
}

// transpose of tp2
template <class T> __device__ __forceinline__ void tp2t(const T S[2][2], T R[2][2]) {
    T E[2][2];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int m = 0; m < 2; ++m) E[p][m] = Wt<T>(0, m) * S[p][0] + Wt<T>(1, m) * S[p][1];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int m = 0; m < 2; ++m) R[j][m] = Wt<T>(0, j) * E[0][m] + Wt<T>(1, j) * E[1][m];
}

// reference derivative along direction dir of the trilinear interpolant of f (lexicographic
// nodes i + 2j + 4k), at the 4 distinct Gauss-point pairs of the two other directions
template <class T> __device__ __forceinline__ void dref(const T f[8], int dir, T F[2][2]) {
    T D[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (dir == 0) D[s][t] = f[1 + 2 * s + 4 * t] - f[2 * s + 4 * t];
            else if (dir == 1) D[s][t] = f[s + 2 + 4 * t] - f[s + 4 * t];
            else D[s][t] = f[s + 2 * t + 4] - f[s + 2 * t];
        }
    tp2<T>(D, F);
}

// scatter of a per-direction sum R back to nodes: +R on the upper node, -R on the lower
template <class T> __device__ __forceinline__ void dref_t(const T S[2][2], int dir, T y[8]) {
    T R[2][2];
    tp2t<T>(S, R);
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (dir == 0) { y[1 + 2 * s + 4 * t] += R[s][t]; y[2 * s + 4 * t] -= R[s][t]; }
            else if (dir == 1) { y[s + 2 + 4 * t] += R[s][t]; y[s + 4 * t] -= R[s][t]; }
            else { y[s + 2 * t + 4] += R[s][t]; y[s + 2 * t] -= R[s][t]; }
        }
}

// y = K_e u for one cell: K_e(i,j) = sum_q JxW (kappa_q grad phi_i . grad phi_j + gamma_q phi_i phi_j)
template <class T, bool COEF>
__device__ __forceinline__ void q1_cell(const T X[8][3], const T u[8], const T *kq, const T *gq, T y[8]) {
    // columns of the Jacobian: dx/dxi at (b,c), dx/deta at (a,c), dx/dzeta at (a,b)
    T J0[3][2][2], J1[3][2][2], J2[3][2][2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T f[8];
#pragma unroll
        for (int n = 0; n < 8; ++n) f[n] = X[n][c];
        dref<T>(f, 0, J0[c]);
        dref<T>(f, 1, J1[c]);
        dref<T>(f, 2, J2[c]);
    }
    T du0[2][2], du1[2][2], du2[2][2];
    dref<T>(u, 0, du0);
    dref<T>(u, 1, du1);
    dref<T>(u, 2, du2);
    T uq[8];
    if (COEF) {
        // values at the Gauss points (three 1D interpolation sweeps)
        T a1[8], a2[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int a = q & 1, r = q & 6;
            a1[q] = Wt<T>(a, 0) * u[r] + Wt<T>(a, 1) * u[r + 1];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int b = (q >> 1) & 1, r = q & 5;
            a2[q] = Wt<T>(b, 0) * a1[r] + Wt<T>(b, 1) * a1[r + 2];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = (q >> 2) & 1, r = q & 3;
            uq[q] = Wt<T>(c, 0) * a2[r] + Wt<T>(c, 1) * a2[r + 4];
        }
    }
    T S0[2][2] = {{0, 0}, {0, 0}}, S1[2][2] = {{0, 0}, {0, 0}}, S2[2][2] = {{0, 0}, {0, 0}};
    T mq[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int a = q & 1, b = (q >> 1) & 1, c = (q >> 2) & 1;
        const T c0x = J0[0][b][c], c0y = J0[1][b][c], c0z = J0[2][b][c];
        const T c1x = J1[0][a][c], c1y = J1[1][a][c], c1z = J1[2][a][c];
        const T c2x = J2[0][a][b], c2y = J2[1][a][b], c2z = J2[2][a][b];
        // rows of det * J^{-1}
        const T a0x = c1y * c2z - c1z * c2y, a0y = c1z * c2x - c1x * c2z, a0z = c1x * c2y - c1y * c2x;
        const T a1x = c2y * c0z - c2z * c0y, a1y = c2z * c0x - c2x * c0z, a1z = c2x * c0y - c2y * c0x;
        const T a2x = c0y * c1z - c0z * c1y, a2y = c0z * c1x - c0x * c1z, a2z = c0x * c1y - c0y * c1x;
        const T det = c0x * a0x + c0y * a0y + c0z * a0z;
        const T g0 = du0[b][c], g1 = du1[a][c], g2 = du2[a][b];
        // det * grad u
        const T vx = g0 * a0x + g1 * a1x + g2 * a2x;
        const T vy = g0 * a0y + g1 * a1y + g2 * a2y;
        const T vz = g0 * a0z + g1 * a1z + g2 * a2z;
        T s = (T)0.125 / det;
        if (COEF) s *= kq[q];
        S0[b][c] += s * (a0x * vx + a0y * vy + a0z * vz);
        S1[a][c] += s * (a1x * vx + a1y * vy + a1z * vz);
        S2[a][b] += s * (a2x * vx + a2y * vy + a2z * vz);
        if (COEF) mq[q] = (T)0.125 * det * gq[q] * uq[q];
    }
#pragma unroll
    for (int n = 0; n < 8; ++n) y[n] = 0;
    dref_t<T>(S0, 0, y);
    dref_t<T>(S1, 1, y);
    dref_t<T>(S2, 2, y);
    if (COEF) {
        // y_n += sum_q phi_n(q) m_q: transposed interpolation sweeps
        T b2[8], b1[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int k = (r >> 2) & 1, base = r & 3;
            b2[r] = Wt<T>(0, k) * mq[base] + Wt<T>(1, k) * mq[base + 4];
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int j = (r >> 1) & 1, base = r & 5;
            b1[r] = Wt<T>(0, j) * b2[base] + Wt<T>(1, j) * b2[base + 2];
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int i = r & 1, base = r & 6;
            y[r] += Wt<T>(0, i) * b1[base] + Wt<T>(1, i) * b1[base + 1];
        }
    }
}

// y_n = K_e(n,n) for one cell in one pass over the Gauss points:
// sum_q JxW (kappa_q |grad phi_n|^2 + gamma_q phi_n^2), with det J^{-1} rows as in q1_cell and
// the reference gradient of phi_n (n = i + 2j + 4k) at (a,b,c): (+-1) Wt(b,j) Wt(c,k) along xi,
// (+-1) Wt(a,i) Wt(c,k) along eta, (+-1) Wt(a,i) Wt(b,j) along zeta
template <class T, bool COEF>
__device__ __forceinline__ void q1_cell_diag(const T X[8][3], const T *kq, const T *gq, T y[8]) {
    T J0[3][2][2], J1[3][2][2], J2[3][2][2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T f[8];
#pragma unroll
        for (int n = 0; n < 8; ++n) f[n] = X[n][c];
        dref<T>(f, 0, J0[c]);
        dref<T>(f, 1, J1[c]);
        dref<T>(f, 2, J2[c]);
    }
#pragma unroll
    for (int n = 0; n < 8; ++n) y[n] = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int a = q & 1, b = (q >> 1) & 1, c = (q >> 2) & 1;
        const T c0x = J0[0][b][c], c0y = J0[1][b][c], c0z = J0[2][b][c];
        const T c1x = J1[0][a][c], c1y = J1[1][a][c], c1z = J1[2][a][c];
        const T c2x = J2[0][a][b], c2y = J2[1][a][b], c2z = J2[2][a][b];
        const T a0x = c1y * c2z - c1z * c2y, a0y = c1z * c2x - c1x * c2z, a0z = c1x * c2y - c1y * c2x;
        const T a1x = c2y * c0z - c2z * c0y, a1y = c2z * c0x - c2x * c0z, a1z = c2x * c0y - c2y * c0x;
        const T a2x = c0y * c1z - c0z * c1y, a2y = c0z * c1x - c0x * c1z, a2z = c0x * c1y - c0y * c1x;
        const T det = c0x * a0x + c0y * a0y + c0z * a0z;
        T s = (T)0.125 / det;
        if (COEF) s *= kq[q];
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            const int i = n & 1, j = (n >> 1) & 1, k = (n >> 2) & 1;
            const T w0 = (i ? (T)1 : (T)-1) * Wt<T>(b, j) * Wt<T>(c, k);
            const T w1 = (j ? (T)1 : (T)-1) * Wt<T>(a, i) * Wt<T>(c, k);
            const T w2 = (k ? (T)1 : (T)-1) * Wt<T>(a, i) * Wt<T>(b, j);
            const T vx = w0 * a0x + w1 * a1x + w2 * a2x;
            const T vy = w0 * a0y + w1 * a1y + w2 * a2y;
            const T vz = w0 * a0z + w1 * a1z + w2 * a2z;
            y[n] += s * (vx * vx + vy * vy + vz * vz);
            if (COEF) {
                const T phi = Wt<T>(a, i) * Wt<T>(b, j) * Wt<T>(c, k);
                y[n] += (T)0.125 * det * gq[q] * phi * phi;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// block sum of NV values; result valid in thread 0
template <int NV> __device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double sh[NV][kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        v[i] = wave_sum(v[i]);
        if (lane == 0) sh[i][w] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < kBlock / 64; ++k) s += sh[i][k];
            v[i] = s;
        }
    }
}

// the dof's element outputs sit contiguously in the slot buffer, in ascending cell order
// Global-memory views of the descriptor's pointers: loads and stores through them are
// global_load/global_store (scalar base + vector offset) instead of flat accesses, which also
// count against the LDS counter and would make every LDS wait wait for them too.
template <class T> using gcp = const __attribute__((address_space(1))) T *;
template <class T> using gwp = __attribute__((address_space(1))) T *;
template <class T> __device__ __forceinline__ gcp<T> G(const T *p) { return (gcp<T>)p; }
template <class T> __device__ __forceinline__ gwp<T> GW(T *p) { return (gwp<T>)p; }
// HIP's vector structs cannot be copied out of an address-space-qualified reference: vector
// loads go through clang vector types
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
template <class T> struct ExtV4;
template <> struct ExtV4<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct ExtV4<double> { typedef double type __attribute__((ext_vector_type(4))); };
__device__ __forceinline__ uint4 ldg(const uint4 *p) {
    const u32x4 v = G(reinterpret_cast<const u32x4 *>(p))[0];
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 ldg(const float2 *p) {
    const f32x2 v = G(reinterpret_cast<const f32x2 *>(p))[0];
    return make_float2(v.x, v.y);
}
template <class V4, class T> __device__ __forceinline__ V4 ldg4(const V4 *p) {
    const typename ExtV4<T>::type v = G(reinterpret_cast<const typename ExtV4<T>::type *>(p))[0];
    V4 r;
    r.x = v.x;
    r.y = v.y;
    r.z = v.z;
    r.w = v.w;
    return r;
}

// coarse part of the preconditioned residual at dof i under the two-level FP32 CG: the
// coarse correction of the dof's block, 0 on constrained dofs (Desc::cc must be set)
template <class T> __device__ __forceinline__ T coarse_term(const Desc &d, int i) {
    // both loads unconditional and independent (a select on the mask would make the second
    // wait for the first); c * 1 and c * 0 are exact
    const T c = G(reinterpret_cast<const T *>(d.cc))[i >> d.agg_shift];
    const T keep = (T)(1 - (int)G(d.dmask)[i]);
    return c * keep;
}

// Slot sums of the kDofIter dofs a thread of a dof workgroup handles (base + threadIdx.x +
// m * kBlock), each in ascending slot order.  The slot offsets of all of them are loaded
// first, then up to four slots each in one wave (indices clamped into the dof's own range, the
// extra values masked to 0), so a thread waits twice instead of once per slot; dofs with more
// than four slots (rare) finish in a loop.
// Slot sums for dofs with long slot lists: for each round of the thread's dofs, the wave's 64
// consecutive dofs own one contiguous slot range, which the wave loads coalesced into its LDS
// row (kWideChunk slots per pass); each lane then adds its own dof's slots from LDS in ascending
// order.  Called by every lane of the wave; a dof past the end has the empty range at gptr[n].
constexpr int kWideChunk = 256;
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <class T> __device__ __forceinline__ void gather_slots_wide(const Desc &d, int base, T (&q)[kDofIter]) {
    __shared__ double wide_lds[kBlock / 64][kWideChunk];
    T *buf = reinterpret_cast<T *>(wide_lds[threadIdx.x >> 6]);
    const auto Ye = G(reinterpret_cast<const T *>(d.Ye));
    const auto gp = G(d.gptr);
    const int lane = threadIdx.x & 63, n = d.n_dofs;
#pragma unroll
    for (int m = 0; m < kDofIter; ++m) {
        const int i = base + threadIdx.x + m * kBlock;
        const int ic = i < n ? i : n;
        const int b = gp[ic];
        const int e = i < n ? gp[ic + 1] : b;
        const int g0 = __shfl(b, 0, 64), g1 = __shfl(e, 63, 64);
        T s = 0;
        for (int c0 = g0; c0 < g1; c0 += kWideChunk) {
            const int cn = min(kWideChunk, g1 - c0);
            for (int k = lane; k < cn; k += 64) buf[k] = Ye[c0 + k];
            wave_lds_sync();
            const int lo = max(b, c0), hi = min(e, c0 + cn);
            for (int k = lo; k < hi; ++k) s += buf[k - c0];
            wave_lds_sync();
        }
        q[m] = s;
    }
}

template <class T> __device__ __forceinline__ void gather_slots4(const Desc &d, int base, T (&q)[kDofIter]) {
    if (d.wide_slots) {
        gather_slots_wide<T>(d, base, q);
        return;
    }
    const auto Ye = G(reinterpret_cast<const T *>(d.Ye));
    const auto gp = G(d.gptr);
    int b[kDofIter], e[kDofIter];
#pragma unroll
    for (int m = 0; m < kDofIter; ++m) {
        const int i = base + threadIdx.x + m * kBlock;
        const int ii = i < d.n_dofs ? i : d.n_dofs - 1;
        b[m] = gp[ii];
        e[m] = gp[ii + 1];
    }
    T v[kDofIter][4];
#pragma unroll
    for (int m = 0; m < kDofIter; ++m)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int idx = b[m] + k < e[m] ? b[m] + k : b[m];
            const T y = Ye[idx];
            v[m][k] = b[m] + k < e[m] ? y : (T)0;
        }
#pragma unroll
    for (int m = 0; m < kDofIter; ++m) {
        T s = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) s += v[m][k];
        // long lists (Q2 vertex dofs touch about two dozen cells): four loads in flight at a
        // time, summed in the same order
        int k = b[m] + 4;
        for (; k + 4 <= e[m]; k += 4) {
            const T y0 = Ye[k], y1 = Ye[k + 1], y2 = Ye[k + 2], y3 = Ye[k + 3];
            s += y0;
            s += y1;
            s += y2;
            s += y3;
        }
        for (; k < e[m]; ++k) s += Ye[k];
        q[m] = s;
    }
}


// ---------------------------------------------------------------------------------------
// element kernel: one workgroup per patch of up to kPatchCells consecutive cells
// ---------------------------------------------------------------------------------------
// The patch's distinct nodes (at most kPatchNodes) are staged in LDS: coordinates and the
// input value u = ua + beta * ub (MASKIN: zero on constrained dofs, the lifted columns).
// Each thread applies one cell reading its 8 nodes from LDS and leaves its 8 outputs in LDS;
// each node's contributions are then summed in a fixed order and written once to the
// node's slot for this patch.
//   MODE 0: y = K_e u          MODE 1: y_n = K_e(n,n)          MODE 2: y_n = int phi_n f
enum { EM_APPLY = 0, EM_DIAG = 1, EM_RHS = 2, EM_VOL = 3 };

extern __shared__ __attribute__((aligned(16))) unsigned char fg_smem[];



// ---------------------------------------------------------------------------------------
// tet-split meshes: every hex of the Tethex split of a tet is the image of one reference
// sub-hex under the affine map of its own frame u1 = 2(x1-x0), u2 = 2(x4-x0), u3 = 2(x2-x0)
// (lexicographic nodes: vertex, three edge centres, three face centres, tet centre).  With
// J = A J_ref, w det(J) J^-1 J^-T = P_q^T Mh P_q, P_q = sqrt(|c_q|) J_ref,q^-T, c_q = w det J_ref,q
// (negative for every q) and Mh = -det(A) (A^T A)^-1: six numbers per cell replace the
// coordinates, and the per-point metric costs 27 multiply-adds.
// ---------------------------------------------------------------------------------------
// lambda coordinates of the reference sub-hex nodes in the frame (u1, u2, u3)
static const double kSubHex[8][3] = {{0, 0, 0}, {0.5, 0, 0}, {0, 0, 0.5}, {1.0 / 3, 0, 1.0 / 3},
                                     {0, 0.5, 0}, {1.0 / 3, 1.0 / 3, 0}, {0, 1.0 / 3, 1.0 / 3}, {0.25, 0.25, 0.25}};
// The sub-hex is symmetric under the permutations of the tet's three edge directions (with the
// matching permutation of xi: node 1 lies on u1, node 4 on u2, node 2 on u3), which map Gauss
// points onto Gauss points.  So P_q of the points (1,0,0), (0,1,0), (0,0,1) are row and column
// permutations of one table, likewise for (1,1,0), (1,0,1), (0,1,1): four tables (points 0, 1,
// 3, 7) hold all eight, P_q[i][j] = P_rep(q)[kAffRow[q][i]][kAffCol[q][j]] (checked when the
// tables are built).  Fewer constants keep them in scalar registers.
__constant__ float kAffP[4][9];
// mass weights of the reference sub-hex: w_q |det J_ref,q| (the cell's |det A| scales them),
// per representative point
__constant__ float kAffW[4];
constexpr int kAffRep[8] = {0, 1, 1, 2, 1, 2, 2, 3};    // table of point q (0, 1, 3, 7 -> 0..3)
constexpr int kAffRow[8][3] = {{0, 1, 2}, {0, 1, 2}, {1, 2, 0}, {0, 1, 2}, {1, 0, 2}, {0, 2, 1}, {1, 0, 2}, {0, 1, 2}};
constexpr int kAffCol[8][3] = {{0, 1, 2}, {0, 1, 2}, {2, 0, 1}, {0, 1, 2}, {2, 1, 0}, {0, 2, 1}, {2, 1, 0}, {0, 1, 2}};

// P_q[i][j] for a compile-time q (the loops over q are unrolled)
__device__ __forceinline__ float aff_p(int q, int i, int j) { return kAffP[kAffRep[q]][3 * kAffRow[q][i] + kAffCol[q][j]]; }
__constant__ double kSubHexD[8][3] = {{0, 0, 0}, {0.5, 0, 0}, {0, 0, 0.5}, {1.0 / 3, 0, 1.0 / 3},
                                      {0, 0.5, 0}, {1.0 / 3, 1.0 / 3, 0}, {0, 1.0 / 3, 1.0 / 3}, {0.25, 0.25, 0.25}};

hipError_t upload_affine_table(hipStream_t s) {
    const double g[2] = {FG_G0, FG_G1};
    float P[8][9], Wm[8];
    static_assert(sizeof(kAffP) == 4 * 9 * sizeof(float), "four representative tables");
    for (int q = 0; q < 8; ++q) {
        const double xi[3] = {g[q & 1], g[(q >> 1) & 1], g[(q >> 2) & 1]};
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int n = 0; n < 8; ++n) {
            double N[3], dN[3];
            for (int d = 0; d < 3; ++d) {
                const int bit = (n >> d) & 1;
                N[d] = bit ? xi[d] : 1 - xi[d];
                dN[d] = bit ? 1 : -1;
            }
            const double gr[3] = {dN[0] * N[1] * N[2], N[0] * dN[1] * N[2], N[0] * N[1] * dN[2]};
            for (int i = 0; i < 3; ++i)
                for (int d = 0; d < 3; ++d) J[i][d] += kSubHex[n][i] * gr[d];
        }
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        // J^-T = cofactor(J) / det
        double C[3][3];
        C[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        C[0][1] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        C[0][2] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        C[1][0] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
        C[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
        C[1][2] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
        C[2][0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
        C[2][1] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
        C[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        const double sc = std::sqrt(std::fabs(0.125 * det)) / det;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) P[q][3 * i + j] = (float)(C[i][j] * sc);
        Wm[q] = (float)std::fabs(0.125 * det);
    }
    (void)s;
    // the representatives, and the symmetry every other point must satisfy exactly
    const int rep_q[4] = {0, 1, 3, 7};
    float R[4][9], Rw[4];
    for (int r = 0; r < 4; ++r) {
        for (int k = 0; k < 9; ++k) R[r][k] = P[rep_q[r]][k];
        Rw[r] = Wm[rep_q[r]];
    }
    for (int q = 0; q < 8; ++q) {
        if (Wm[q] != Rw[kAffRep[q]]) return hipErrorInvalidValue;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (P[q][3 * i + j] != R[kAffRep[q]][3 * kAffRow[q][i] + kAffCol[q][j]]) return hipErrorInvalidValue;
    }
    const hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(kAffP), R, sizeof(R), 0, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(kAffW), Rw, sizeof(Rw), 0, hipMemcpyHostToDevice);
}

// y = K_e u for a tet-split hex with metric M (m00 m01 m02 m11 m12 m22)
__device__ __forceinline__ void q1_cell_affine(const float M[6], const float u[8], float y[8]) {
    float du0[2][2], du1[2][2], du2[2][2];
    dref<float>(u, 0, du0);
    dref<float>(u, 1, du1);
    dref<float>(u, 2, du2);
    float S0[2][2] = {{0, 0}, {0, 0}}, S1[2][2] = {{0, 0}, {0, 0}}, S2[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int a = q & 1, b = (q >> 1) & 1, c = (q >> 2) & 1;
        const float g0 = du0[b][c], g1 = du1[a][c], g2 = du2[a][b];
        const float s0 = aff_p(q, 0, 0) * g0 + aff_p(q, 0, 1) * g1 + aff_p(q, 0, 2) * g2;
        const float s1 = aff_p(q, 1, 0) * g0 + aff_p(q, 1, 1) * g1 + aff_p(q, 1, 2) * g2;
        const float s2 = aff_p(q, 2, 0) * g0 + aff_p(q, 2, 1) * g1 + aff_p(q, 2, 2) * g2;
        const float t0 = M[0] * s0 + M[1] * s1 + M[2] * s2;
        const float t1 = M[1] * s0 + M[3] * s1 + M[4] * s2;
        const float t2 = M[2] * s0 + M[4] * s1 + M[5] * s2;
        S0[b][c] += aff_p(q, 0, 0) * t0 + aff_p(q, 1, 0) * t1 + aff_p(q, 2, 0) * t2;
        S1[a][c] += aff_p(q, 0, 1) * t0 + aff_p(q, 1, 1) * t1 + aff_p(q, 2, 1) * t2;
        S2[a][b] += aff_p(q, 0, 2) * t0 + aff_p(q, 1, 2) * t1 + aff_p(q, 2, 2) * t2;
    }
#pragma unroll
    for (int n = 0; n < 8; ++n) y[n] = 0;
    dref_t<float>(S0, 0, y);
    dref_t<float>(S1, 1, y);
    dref_t<float>(S2, 2, y);
}

// y = K_e u for a tet-split hex with metric M and per-point stiffness and mass coefficients
// kq, gq (lexicographic Gauss points): sum_q kq (P_q g)^T M (P_q g) + gq |det A| w_q |det J_ref,q|
// phi phi^T.  |det A| is passed in: det M equals it, but the FP32 metric of a flat cell does not
// fix its determinant (cancellation)
__device__ __forceinline__ void q1_cell_affine_coef(const float M[6], const float detA, const float kq[8], const float gq[8],
                                                    const float u[8], float y[8]) {
    float du0[2][2], du1[2][2], du2[2][2];
    dref<float>(u, 0, du0);
    dref<float>(u, 1, du1);
    dref<float>(u, 2, du2);
    // values at the Gauss points (three 1D interpolation sweeps)
    float a1[8], a2[8], uq[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int a = q & 1, r = q & 6;
        a1[q] = Wt<float>(a, 0) * u[r] + Wt<float>(a, 1) * u[r + 1];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int b = (q >> 1) & 1, r = q & 5;
        a2[q] = Wt<float>(b, 0) * a1[r] + Wt<float>(b, 1) * a1[r + 2];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int c = (q >> 2) & 1, r = q & 3;
        uq[q] = Wt<float>(c, 0) * a2[r] + Wt<float>(c, 1) * a2[r + 4];
    }
    float S0[2][2] = {{0, 0}, {0, 0}}, S1[2][2] = {{0, 0}, {0, 0}}, S2[2][2] = {{0, 0}, {0, 0}};
    float mq[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int a = q & 1, b = (q >> 1) & 1, c = (q >> 2) & 1;
        const float g0 = du0[b][c], g1 = du1[a][c], g2 = du2[a][b];
        const float s0 = aff_p(q, 0, 0) * g0 + aff_p(q, 0, 1) * g1 + aff_p(q, 0, 2) * g2;
        const float s1 = aff_p(q, 1, 0) * g0 + aff_p(q, 1, 1) * g1 + aff_p(q, 1, 2) * g2;
        const float s2 = aff_p(q, 2, 0) * g0 + aff_p(q, 2, 1) * g1 + aff_p(q, 2, 2) * g2;
        const float k = kq[q];
        const float t0 = k * (M[0] * s0 + M[1] * s1 + M[2] * s2);
        const float t1 = k * (M[1] * s0 + M[3] * s1 + M[4] * s2);
        const float t2 = k * (M[2] * s0 + M[4] * s1 + M[5] * s2);
        S0[b][c] += aff_p(q, 0, 0) * t0 + aff_p(q, 1, 0) * t1 + aff_p(q, 2, 0) * t2;
        S1[a][c] += aff_p(q, 0, 1) * t0 + aff_p(q, 1, 1) * t1 + aff_p(q, 2, 1) * t2;
        S2[a][b] += aff_p(q, 0, 2) * t0 + aff_p(q, 1, 2) * t1 + aff_p(q, 2, 2) * t2;
        mq[q] = gq[q] * detA * kAffW[kAffRep[q]] * uq[q];
    }
#pragma unroll
    for (int n = 0; n < 8; ++n) y[n] = 0;
    dref_t<float>(S0, 0, y);
    dref_t<float>(S1, 1, y);
    dref_t<float>(S2, 2, y);
    float b2[8], b1[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int k = (r >> 2) & 1, base = r & 3;
        b2[r] = Wt<float>(0, k) * mq[base] + Wt<float>(1, k) * mq[base + 4];
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int j = (r >> 1) & 1, base = r & 5;
        b1[r] = Wt<float>(0, j) * b2[base] + Wt<float>(1, j) * b2[base + 2];
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int i = r & 1, base = r & 6;
        y[r] += Wt<float>(0, i) * b1[base] + Wt<float>(1, i) * b1[base + 1];
    }
}

// per cell: the metric of the tet-split form (metric[6 c ..]), |det A| (metric[6 nc + c]) and,
// per block, the largest deviation of the nodes from that form relative to the frame size
__global__ __launch_bounds__(kBlock) void k_metric(int nc, const int *__restrict__ cells,
                                                   const unsigned char *__restrict__ layout, const double4 *__restrict__ X,
                                                   float *__restrict__ metric, double *__restrict__ devmax) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double dev = 0;
    if (c < nc) {
        // an eta-flipped cell is read in the flipped order: the sub-hex form then holds with
        // the opposite orientation of the frame
        const bool flip = layout[c] != 0;
        double x[8][3];
        for (int n = 0; n < 8; ++n) {
            const double4 p = X[cells[8 * (size_t)c + (flip ? kEtaFlip[n] : n)]];
            x[n][0] = p.x;
            x[n][1] = p.y;
            x[n][2] = p.z;
        }
        double u[3][3];
        for (int a = 0; a < 3; ++a) {
            u[0][a] = 2 * (x[1][a] - x[0][a]);
            u[1][a] = 2 * (x[4][a] - x[0][a]);
            u[2][a] = 2 * (x[2][a] - x[0][a]);
        }
        const double size = sqrt(u[0][0] * u[0][0] + u[0][1] * u[0][1] + u[0][2] * u[0][2]) +
                            sqrt(u[1][0] * u[1][0] + u[1][1] * u[1][1] + u[1][2] * u[1][2]) +
                            sqrt(u[2][0] * u[2][0] + u[2][1] * u[2][1] + u[2][2] * u[2][2]);
        for (int n = 0; n < 8; ++n)
            for (int a = 0; a < 3; ++a) {
                const double pr = x[0][a] + kSubHexD[n][0] * u[0][a] + kSubHexD[n][1] * u[1][a] + kSubHexD[n][2] * u[2][a];
                dev = fmax(dev, fabs(pr - x[n][a]) / size);
            }
        // rows of adj(A), A = [u1 u2 u3] as columns
        double r[3][3];
        for (int i = 0; i < 3; ++i) {
            const double *p = u[(i + 1) % 3], *q = u[(i + 2) % 3];
            r[i][0] = p[1] * q[2] - p[2] * q[1];
            r[i][1] = p[2] * q[0] - p[0] * q[2];
            r[i][2] = p[0] * q[1] - p[1] * q[0];
        }
        const double det = u[0][0] * r[0][0] + u[0][1] * r[0][1] + u[0][2] * r[0][2];
        const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
        for (int k = 0; k < 6; ++k) {
            const double *a = r[ij[k][0]], *b = r[ij[k][1]];
            metric[6 * (size_t)c + k] = (float)((a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / fabs(det));
        }
        metric[6 * (size_t)nc + c] = (float)fabs(det);
        // a positively oriented cell has det(A) < 0 as laid out, det(A) > 0 when eta-flipped
        if (!(flip ? det > 0 : det < 0)) dev = 1e300;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dev = fmax(dev, __shfl_xor(dev, o, 64));
    __shared__ double s[kBlock / 64];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = dev;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; ++k) dev = fmax(dev, s[k]);
        dev = fmax(dev, s[0]);
        devmax[blockIdx.x] = dev;
    }
}

int metric_blocks(int nc) { return (nc + kBlock - 1) / kBlock; }
void launch_metric(hipStream_t s, int nc, const int *cells, const unsigned char *layout, const double *X, float *metric,
                   double *devmax) {
    if (nc > 0)
        hipLaunchKernelGGL(k_metric, dim3(metric_blocks(nc)), dim3(kBlock), 0, s, nc, cells, layout,
                           reinterpret_cast<const double4 *>(X), metric, devmax);
}

// AFF: every system of the launch has the tet-split FP32 form (the general path is not
// compiled in, which frees registers for more resident waves)
template <class T, bool COEF, bool MASKIN, bool ACTIVE_ONLY, int MODE, bool AFF = false>
__global__ __launch_bounds__(kBlock, sizeof(T) == 8 ? 3 : 4) void k_patch(const Desc *__restrict__ D, const ETab *__restrict__ tab, int scal_slot,
                                                     const double *__restrict__ fq) {
    constexpr int kRounds = (kPatchNodes + kBlock - 1) / kBlock;
    const ETab t = tab[blockIdx.x];
    if (t.nc == 0) return;
    const Desc &d = D[t.sys];
    // read now, tested once the first patch's lists are in flight (one latency fewer)
    const double active = ACTIVE_ONLY ? G(d.scal)[S_ACTIVE] : 1.0;
    using V4 = typename Vec4T<T>::type;
    // FP32 apply of a tet-split mesh: per-cell metric, no coordinates (uniform per workgroup)
    // (with COEF: per-point coefficients on top of the metric)
    const bool affine = AFF || (sizeof(T) == 4 && MODE == EM_APPLY && d.metric != nullptr);
    // FP32: the per-node sums are exact integer sums of fixed-point contributions (LDS 64-bit
    // atomics, scale per patch), so no entry lists are read and the order does not matter;
    // FP64: the cell outputs are summed per node along the entry lists in a fixed order
    constexpr bool FX = sizeof(T) == 4;
    // dynamic LDS: node coordinates (not for the tet-split form) and inputs (node_cap each);
    // FP32: fixed-point node sums (node_cap) and the patch's largest output per wave;
    // FP64: cell outputs, entry lists
    T *sx = reinterpret_cast<T *>(fg_smem);
    T *sy = sx + d.node_cap, *sz = sy + d.node_cap;
    T *su = affine ? sx : sz + d.node_cap, *sout = su + d.node_cap;
    uint4 *slent = reinterpret_cast<uint4 *>(sout + 8 * d.pcells);
    unsigned long long *sacc = reinterpret_cast<unsigned long long *>(su + d.node_cap);
    float *smax = reinterpret_cast<float *>(sacc + d.node_cap);
    const V4 *X = reinterpret_cast<const V4 *>(d.X);
    const auto ua = G(reinterpret_cast<const T *>(d.ua));
    const T *ub = reinterpret_cast<const T *>(d.ub);
    // converted where it is used, after the first loads are in flight
    const double beta64 = MODE == EM_APPLY ? G(d.scal)[scal_slot] : 0.0;   // used only with ub
    const auto Ye = GW(reinterpret_cast<T *>(d.Ye));
    const int tid = threadIdx.x;
    // A patch's first wave of loads: its tables (cell local ids, entry lists, metric and
    // coefficients, node slots and ids).  Indices past the patch are clamped to its first node
    // or cell, so every load is unconditional and none waits for another of its wave.
    struct Tables {
        uint4 packed, lent;
        float2 mq[3];
        float4 kc[2], gc[2];
        float vdet;
        int eb[kRounds], ee[kRounds], os[kRounds], nd[kRounds];
    };
    auto load_tables = [&](Tables &a, const int p, const int c0, const int nc, const int n0, const int np) {
        const int ct = tid < nc ? tid : 0;
        a.packed = ldg(reinterpret_cast<const uint4 *>(d.plidx) + c0 + ct);
        a.lent = make_uint4(0, 0, 0, 0);
        if constexpr (!FX) a.lent = ldg(reinterpret_cast<const uint4 *>(d.plent) + c0 + ct);
        a.mq[0] = a.mq[1] = a.mq[2] = make_float2(0.f, 0.f);
        a.kc[0] = a.kc[1] = a.gc[0] = a.gc[1] = make_float4(0.f, 0.f, 0.f, 0.f);
        a.vdet = 0.f;
        if constexpr (sizeof(T) == 4 && MODE == EM_APPLY) {
            if (affine) {
                const float2 *mp = reinterpret_cast<const float2 *>(d.metric + 6 * (size_t)(c0 + ct));
                a.mq[0] = ldg(mp);
                a.mq[1] = ldg(mp + 1);
                a.mq[2] = ldg(mp + 2);
                if constexpr (COEF) {
                    const float4 *kp = reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(d.kq) + 8 * (size_t)(c0 + ct));
                    const float4 *gp = reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(d.gq) + 8 * (size_t)(c0 + ct));
                    a.kc[0] = ldg4<float4, float>(kp);
                    a.kc[1] = ldg4<float4, float>(kp + 1);
                    a.gc[0] = ldg4<float4, float>(gp);
                    a.gc[1] = ldg4<float4, float>(gp + 1);
                    a.vdet = G(d.metric)[6 * (size_t)d.n_cells + c0 + ct];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const int j = tid + r * kBlock;
            const int jj = j < np ? j : 0;
            if constexpr (!FX) {
                a.eb[r] = G(d.plptr)[n0 + p + jj];
                a.ee[r] = G(d.plptr)[n0 + p + jj + 1];
            } else {
                a.eb[r] = a.ee[r] = 0;
            }
            a.os[r] = G(d.pslot)[n0 + jj];
            a.nd[r] = G(d.pnodes)[n0 + jj];
        }
    };
    // one patch from its loaded tables; the first one (check) leaves the workgroup early when
    // the system is idle, tested only once that patch's lists are in flight.  Returns false to
    // leave.  (Loading the next patch's tables ahead, in workgroups of several patches, was
    // measured: the extra registers cost more resident waves than the hidden latency gained.
    // Loading only its node ids ahead keeps 8 waves resident and changed nothing: the scattered
    // node accesses cost memory requests, not latency.)
    auto patch = [&](Tables &a, const int p, const int c0, const int nc, const int n0, const int np, const bool check) -> bool {
        const uint4 packed = a.packed;
        const uint4 lent = a.lent;
        const float2 *mq = a.mq;
        const float4 *kc = a.kc, *gc = a.gc;
        const float vdet = a.vdet;
        int *eb = a.eb, *ee = a.ee, *os = a.os, *nd = a.nd;
        (void)p;
        // the node lists had to arrive before the second wave is addressed: store the cell
        // entry lists meanwhile (keeps their load in the first wave)
        if constexpr (!FX) {
            if (tid < nc) slent[tid] = lent;
        }
        if (ACTIVE_ONLY && check && active == 0.0) return false;   // uniform: the workgroup leaves
        const T beta = (T)beta64;
        T va[kRounds], vb[kRounds], vc[kRounds];
        unsigned char mk[kRounds];
        V4 xs[kRounds];
        const auto ubp = G(ub ? ub : reinterpret_cast<const T *>(d.ua));   // never read through null
        const bool cin = d.cc != nullptr && d.cc_in != 0;
        const auto ccp = G(cin ? reinterpret_cast<const T *>(d.cc) : reinterpret_cast<const T *>(d.ua));
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            if (!affine) xs[r] = ldg4<V4, T>(X + nd[r]);
            if (MODE == EM_APPLY) {
                va[r] = ua[nd[r]];
                vb[r] = ubp[nd[r]];
                vc[r] = ccp[cin ? nd[r] >> d.agg_shift : 0];
                mk[r] = G(d.dmask)[nd[r]];
            }
        }

#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const int j = tid + r * kBlock;
            if (j < np) {
                if constexpr (FX) sacc[j] = 0ull;
                if (!affine) {
                    sx[j] = xs[r].x;
                    sy[j] = xs[r].y;
                    sz[j] = xs[r].z;
                }
                if (MODE == EM_APPLY) {
                    T v = va[r];
                    // two-level FP32 CG: z carries its coarse part where it is read
                    if (cin) v += vc[r] * (T)(1 - (int)mk[r]);
                    if (ub) v += beta * vb[r];
                    if (MASKIN && mk[r]) v = 0;
                    su[j] = v;
                }
            } else {
                eb[r] = ee[r] = 0;
            }
        }
        __syncthreads();
        // phase 2: one cell per thread, from LDS (FP32: outputs yc for local nodes al)
        T yc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int al[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if constexpr (sizeof(T) == 4 && MODE == EM_APPLY) {
            if (affine && tid < nc) {
                // an eta-flipped cell (bit 15 of its first index) is computed in flipped order
                const bool flip = (packed.x & 0x8000u) != 0;
                const int l0[8] = {(int)(packed.x & 0x7fff), (int)(packed.x >> 16), (int)(packed.y & 0xffff),
                                   (int)(packed.y >> 16), (int)(packed.z & 0xffff), (int)(packed.z >> 16),
                                   (int)(packed.w & 0xffff), (int)(packed.w >> 16)};
                int lid[8];
#pragma unroll
                for (int n = 0; n < 8; ++n) lid[n] = flip ? l0[kEtaFlip[n]] : l0[n];
                const float M[6] = {mq[0].x, mq[0].y, mq[1].x, mq[1].y, mq[2].x, mq[2].y};
                float u[8], y[8];
#pragma unroll
                for (int n = 0; n < 8; ++n) u[n] = su[lid[n]];
                if constexpr (COEF) {
                    // the Gauss points of a flipped cell flip with its vertices
                    const float k0[8] = {kc[0].x, kc[0].y, kc[0].z, kc[0].w, kc[1].x, kc[1].y, kc[1].z, kc[1].w};
                    const float g0[8] = {gc[0].x, gc[0].y, gc[0].z, gc[0].w, gc[1].x, gc[1].y, gc[1].z, gc[1].w};
                    float kf[8], gf[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        kf[q] = flip ? k0[kEtaFlip[q]] : k0[q];
                        gf[q] = flip ? g0[kEtaFlip[q]] : g0[q];
                    }
                    q1_cell_affine_coef(M, vdet, kf, gf, u, y);
                } else {
                    q1_cell_affine(M, u, y);
                }
#pragma unroll
                for (int n = 0; n < 8; ++n) {
                    yc[n] = y[n];
                    al[n] = lid[n];
                }
            }
        }
        if (!AFF && tid < nc && !affine) {
            const int c = c0 + tid;
            const int lid[8] = {(int)(packed.x & 0x7fff), (int)(packed.x >> 16), (int)(packed.y & 0xffff), (int)(packed.y >> 16),
                                (int)(packed.z & 0xffff), (int)(packed.z >> 16), (int)(packed.w & 0xffff), (int)(packed.w >> 16)};
            T Xe[8][3];
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                Xe[n][0] = sx[lid[n]];
                Xe[n][1] = sy[lid[n]];
                Xe[n][2] = sz[lid[n]];
            }
            T kq[8], gq[8];
            if (COEF) {
                const auto k = G(reinterpret_cast<const T *>(d.kq) + 8 * (size_t)c);
                const auto g = G(reinterpret_cast<const T *>(d.gq) + 8 * (size_t)c);
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    kq[q] = k[q];
                    gq[q] = g[q];
                }
            }
            T y[8];
            if (MODE == EM_APPLY) {
                T u[8];
#pragma unroll
                for (int n = 0; n < 8; ++n) u[n] = su[lid[n]];
                q1_cell<T, COEF>(Xe, u, kq, gq, y);
            } else if (MODE == EM_DIAG) {
                q1_cell_diag<T, COEF>(Xe, kq, gq, y);
            } else {
                for (int n = 0; n < 8; ++n) y[n] = 0;
                for (int q = 0; q < 8; ++q) {
                    const int a = q & 1, b = (q >> 1) & 1, cc = (q >> 2) & 1;
                    T J[3][3];
                    for (int co = 0; co < 3; ++co) {
                        T f[8], F[2][2];
                        for (int n = 0; n < 8; ++n) f[n] = Xe[n][co];
                        dref<T>(f, 0, F); J[co][0] = F[b][cc];
                        dref<T>(f, 1, F); J[co][1] = F[a][cc];
                        dref<T>(f, 2, F); J[co][2] = F[a][b];
                    }
                    const T det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                  J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                  J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                    if (MODE == EM_VOL) {
                        // JxW(q) to every node of the cell, not shape-weighted
                        const T jxw = (T)0.125 * det;
                        for (int n = 0; n < 8; ++n) y[n] += jxw;
                    } else {
                        const T jxw = (T)0.125 * det * (T)G(fq)[8 * (size_t)c + q];
                        for (int n = 0; n < 8; ++n)
                            y[n] += Wt<T>(a, n & 1) * Wt<T>(b, (n >> 1) & 1) * Wt<T>(cc, (n >> 2) & 1) * jxw;
                    }
                }
            }
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                if constexpr (FX) {
                    yc[n] = y[n];
                    al[n] = lid[n];
                } else {
                    sout[8 * tid + n] = y[n];
                }
            }
        }
        if constexpr (FX) {
            // fixed point: the patch's largest |output| m < 2^e gives the scale 2^(50 - e), so up
            // to 2^13 contributions of a node add without overflow, each exact to m 2^-50 (far
            // below FP32 rounding).  A non-finite output makes the patch's outputs NaN.
            // fmaxf skips NaN, so a non-finite output is caught through the sum of the eight
            float m = 0.f;
            if (tid < nc) {
                float sum = 0.f;
#pragma unroll
                for (int n = 0; n < 8; ++n) {
                    m = fmaxf(m, fabsf((float)yc[n]));
                    sum += (float)yc[n];
                }
                if (!(fabsf(sum) <= 3.0e38f)) m = __builtin_inff();
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
            if ((tid & 63) == 0) smax[tid >> 6] = m;
            __syncthreads();
            m = fmaxf(fmaxf(smax[0], smax[1]), fmaxf(smax[2], smax[3]));
            const bool finite = m < __builtin_inff();
            int e = 0;
            (void)frexpf(finite ? m : 0.f, &e);
            // scaled outputs |y| 2^(50 - e) < 2^50: rounded to integers by adding 1.5 * 2^52 in
            // FP64 (round to nearest even, ulp 1 in that binade); the integer is the sum's bit
            // pattern minus the constant's
            const double sc = (finite && m > 0.f) ? ldexp(1.0, 50 - e) : 1.0;
            if (tid < nc && finite) {
                constexpr double kMagic = 6755399441055744.0;                  // 1.5 * 2^52
                constexpr long long kMagicBits = 0x4338000000000000ll;
#pragma unroll
                for (int n = 0; n < 8; ++n) {
                    // y sc is exact (a power of two), so the fused add rounds once
                    const double t = __fma_rn((double)yc[n], sc, kMagic);
                    atomicAdd(&sacc[al[n]], (unsigned long long)(__double_as_longlong(t) - kMagicBits));
                }
            }
            __syncthreads();
            const double isc = 1.0 / sc;            // a power of two: exact
#pragma unroll
            for (int r = 0; r < kRounds; ++r) {
                const int j = tid + r * kBlock;
                if (j < np) Ye[os[r]] = finite ? (T)((double)(long long)sacc[j] * isc) : (T)__builtin_nanf("");
            }
        } else {
            __syncthreads();
            // phase 3: per patch node, its cell contributions in ascending order
            const uint16_t *le = reinterpret_cast<const uint16_t *>(slent);
#pragma unroll
            for (int r = 0; r < kRounds; ++r) {
                const int j = tid + r * kBlock;
                if (j < np) {
                    T s = 0;
                    for (int e = eb[r]; e < ee[r]; ++e) s += sout[le[e]];
                    Ye[os[r]] = s;
                }
            }
        }
        __syncthreads();
        return true;
    };
    Tables a;
    load_tables(a, t.p, t.c0, t.nc, t.n0, t.np);
    if (!patch(a, t.p, t.c0, t.nc, t.n0, t.np, true)) return;
    const int p_end = min(d.n_patch, t.p + d.ppb);
    for (int p = t.p + 1; p < p_end; ++p) {
        const int c0 = G(d.pcell0)[p], n0 = G(d.pnode0)[p];
        const int nc = G(d.pcell0)[p + 1] - c0, np = G(d.pnode0)[p + 1] - n0;
        load_tables(a, p, c0, nc, n0, np);
        patch(a, p, c0, nc, n0, np, false);
    }
}

#include "fgpu_cg.inc"

// ---------------------------------------------------------------------------------------
// setup kernels (FP64)
// ---------------------------------------------------------------------------------------
__global__ void k_enforce_dirichlet(int n, const unsigned char *__restrict__ mask, const double *__restrict__ g, double *__restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && mask[i]) x[i] = g[i];
}

// b_I -= (A g)_I, b_D = diag_D g_D, x_D = g_D
__global__ void k_lift(int n, const unsigned char *__restrict__ mask, const double *__restrict__ g, const double *__restrict__ Ag,
                       const double *__restrict__ diag, double *__restrict__ b, double *__restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mask[i]) {
        b[i] = g[i] * diag[i];
        x[i] = g[i];
    } else if (Ag) {       // null: g is zero, so is A g
        b[i] -= Ag[i];
    }
}

__global__ void k_invdiag(int n, const unsigned char *__restrict__ mask, const double *__restrict__ diag, double *__restrict__ inv64,
                          float *__restrict__ inv32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = mask[i] ? 0.0 : 1.0 / diag[i];
    inv64[i] = v;
    inv32[i] = (float)v;
}

__global__ void k_set_const(int n, double v, double *__restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// Neumann load on the boundary faces with id bid: 4 contributions per face into fslot
// (DealSolver::assemble_rhs, QGauss(2) on the face, bilinear face map)
__global__ void k_face_load(int nb, const int *__restrict__ bf_icell, const int *__restrict__ bf_face, const int *__restrict__ bf_id,
                            int bid, const double *__restrict__ fval, const int *__restrict__ cells, const double4 *__restrict__ X,
                            double *__restrict__ fslot) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nb) return;
    double out[4] = {0, 0, 0, 0};
    if (bf_id[f] == bid) {
        const int face = bf_face[f], dd = face / 2, side = face % 2;
        const int a0 = dd == 0 ? 1 : 0, a1 = dd == 2 ? 1 : 2;
        const int *cl = cells + 8 * (size_t)bf_icell[f];
        // face vertices in lexicographic order of (a0, a1)
        int fv[4];
        for (int j = 0; j < 4; ++j) {
            int idx[3];
            idx[dd] = side;
            idx[a0] = j & 1;
            idx[a1] = j >> 1;
            fv[j] = cl[idx[0] + 2 * idx[1] + 4 * idx[2]];
        }
        double P[4][3];
        for (int j = 0; j < 4; ++j) {
            const double4 p = X[fv[j]];
            P[j][0] = p.x;
            P[j][1] = p.y;
            P[j][2] = p.z;
        }
        const double g = fval[f];
        for (int q = 0; q < 4; ++q) {
            const int qa = q & 1, qb = q >> 1;
            double t0[3], t1[3];
            for (int c = 0; c < 3; ++c) {
                t0[c] = Wt<double>(qb, 0) * (P[1][c] - P[0][c]) + Wt<double>(qb, 1) * (P[3][c] - P[2][c]);
                t1[c] = Wt<double>(qa, 0) * (P[2][c] - P[0][c]) + Wt<double>(qa, 1) * (P[3][c] - P[1][c]);
            }
            const double cx = t0[1] * t1[2] - t0[2] * t1[1];
            const double cy = t0[2] * t1[0] - t0[0] * t1[2];
            const double cz = t0[0] * t1[1] - t0[1] * t1[0];
            const double jxw = sqrt(cx * cx + cy * cy + cz * cz) * 0.25;
            for (int j = 0; j < 4; ++j) out[j] += Wt<double>(qa, j & 1) * Wt<double>(qb, j >> 1) * g * jxw;
        }
    }
    for (int j = 0; j < 4; ++j) fslot[4 * (size_t)f + j] = out[j];
}

__global__ void k_face_gather(int n, const int *__restrict__ fgptr, const int *__restrict__ fgidx, const double *__restrict__ fslot,
                              double *__restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0;
    for (int k = fgptr[i]; k < fgptr[i + 1]; ++k) s += fslot[fgidx[k]];
    b[i] += s;
}

// min / max of a vector (single block per call, n moderate)
__global__ __launch_bounds__(kBlock) void k_minmax(int n, const double *__restrict__ x, double *__restrict__ out) {
    double mn = 1e100, mx = -1e100;
    for (int i = threadIdx.x; i < n; i += kBlock) {
        mn = fmin(mn, x[i]);
        mx = fmax(mx, x[i]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, o, 64));
        mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    __shared__ double s[2][kBlock / 64];
    if ((threadIdx.x & 63) == 0) {
        s[0][threadIdx.x >> 6] = mn;
        s[1][threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; ++k) {
            mn = fmin(mn, s[0][k]);
            mx = fmax(mx, s[1][k]);
        }
        mn = fmin(mn, s[0][0]);
        mx = fmax(mx, s[1][0]);
        out[0] = mn;
        out[1] = mx;
    }
}

// new positions of a subset of dofs (internal numbering) in both coordinate copies
__global__ void k_scatter_coords(int n, const int *__restrict__ idx, const double *__restrict__ xyz, double4 *__restrict__ X64,
                                 float4 *__restrict__ X32) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = idx[k];
    const double x = xyz[3 * (size_t)k], y = xyz[3 * (size_t)k + 1], z = xyz[3 * (size_t)k + 2];
    X64[i] = make_double4(x, y, z, 0.0);
    X32[i] = make_float4((float)x, (float)y, (float)z, 0.f);
}

// linear mesh motion: row r of the map moves dof idof[r] to base[r] + sum_k w[k] disp[ctrl[k]]
__global__ void k_motion(int rows, const int *__restrict__ idof, const int *__restrict__ ptr, const int *__restrict__ ctrl,
                         const double *__restrict__ w, const double *__restrict__ base, const double *__restrict__ disp,
                         double4 *__restrict__ X64, float4 *__restrict__ X32) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double x = 0, y = 0, z = 0;
    for (int k = ptr[r]; k < ptr[r + 1]; ++k) {
        const double wk = w[k];
        const double *d = disp + 3 * (size_t)ctrl[k];
        x += wk * d[0];
        y += wk * d[1];
        z += wk * d[2];
    }
    x += base[3 * (size_t)r];
    y += base[3 * (size_t)r + 1];
    z += base[3 * (size_t)r + 2];
    const int i = idof[r];
    X64[i] = make_double4(x, y, z, 0.0);
    X32[i] = make_float4((float)x, (float)y, (float)z, 0.f);
}

// per block: the smallest Jacobian determinant at the cell centres of its cells, in the
// orientation fixed at import (same formula as the import's orientation test)
__global__ __launch_bounds__(kBlock) void k_center_det_min(int n, const int *__restrict__ cells, const double4 *__restrict__ X,
                                                           double *__restrict__ out) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double v = 1e300;
    if (c < n) {
        double p[8][3];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double4 q = X[cells[8 * (size_t)c + k]];
            p[k][0] = q.x;
            p[k][1] = q.y;
            p[k][2] = q.z;
        }
        double J[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            J[a][0] = 0.25 * ((p[1][a] - p[0][a]) + (p[3][a] - p[2][a]) + (p[5][a] - p[4][a]) + (p[7][a] - p[6][a]));
            J[a][1] = 0.25 * ((p[2][a] - p[0][a]) + (p[3][a] - p[1][a]) + (p[6][a] - p[4][a]) + (p[7][a] - p[5][a]));
            J[a][2] = 0.25 * ((p[4][a] - p[0][a]) + (p[5][a] - p[1][a]) + (p[6][a] - p[2][a]) + (p[7][a] - p[3][a]));
        }
        v = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    __shared__ double s[kBlock / 64];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; ++k) v = fmin(v, s[k]);
        v = fmin(v, s[0]);
        out[blockIdx.x] = v;
    }
}

// out[0] = min of in[0..n) (one workgroup; min is exact, so the order does not matter)
__global__ __launch_bounds__(kBlock) void k_min_reduce(int n, const double *__restrict__ in, double *__restrict__ out) {
    double v = 1e300;
    for (int i = threadIdx.x; i < n; i += kBlock) v = fmin(v, in[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    __shared__ double s[kBlock / 64];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; ++k) v = fmin(v, s[k]);
        v = fmin(v, s[0]);
        out[0] = v;
    }
}
void launch_min_reduce(hipStream_t s, int n, const double *in, double *out) {
    hipLaunchKernelGGL(k_min_reduce, dim3(1), dim3(kBlock), 0, s, n, in, out);
}

// resistivity of bulk Cu, (T [K], rho [Ohm nm]) pairs of the reference's table
// (src/PhysicalQuantityData.cpp:47-76); rho(T) = 10 * linear interpolation [Ohm A]
__constant__ double kRhoT[26] = {200, 250, 273.15, 300, 350, 400, 450, 500, 550, 600, 650, 700, 750,
                                 800, 850, 900, 950, 1000, 1050, 1100, 1150, 1200, 1250, 1300, 1350, 1400};
__constant__ double kRhoV[26] = {
    1.04880891725670e1, 1.38746906313026e1, 1.54302576609579e1, 1.72303253767393e1, 2.05822342557969e1,
    2.39476190263668e1, 2.73376276457821e1, 3.07598544167200e1, 3.42197627890272e1, 3.77214997429762e1,
    4.12683756079926e1, 4.48631564739703e1, 4.85082484331790e1, 5.22058173711257e1, 5.59578691681847e1,
    5.97663048918425e1, 6.36329597840890e1, 6.75596315079091e1, 7.15481011300350e1, 7.56001491045029e1,
    7.97175677635730e1, 8.39021713382574e1, 8.81558042149690e1, 9.24803479251429e1, 9.68777272230961e1,
    1.01349915510292e2};

// PhysicalQuantities::sigma / kappa (src/PhysicalQuantities.cpp:31-64, 173-185): T clamped to
// the table, rho = 10 * lin_interp, sigma = 1 / rho, kappa = L T sigma
__device__ inline double pq_sigma(double T) {
    T = T < kRhoT[0] ? kRhoT[0] : (T > kRhoT[25] ? kRhoT[25] : T);
    double r;
    if (T <= kRhoT[0]) r = kRhoV[0];
    else if (T >= kRhoT[25]) r = kRhoV[25];
    else {
        int i1 = 0;
        while (kRhoT[i1] < T) ++i1;
        const int i2 = i1 - 1;
        r = kRhoV[i2] + (kRhoV[i1] - kRhoV[i2]) * (T - kRhoT[i2]) / (kRhoT[i1] - kRhoT[i2]);
    }
    return 1.0 / (10. * r);
}

// HeatSolver::assemble_local_cell (src/CurrentHeatSolver.cpp:347-393) per Gauss point q of
// cell c (internal order): T_old and grad phi_c interpolated at q; stiffness coefficient
// theta kappa(T_old), mass coefficient gamma, source gamma T_old + sigma(T_old) |grad phi_c|^2.
// theta < 1 (the theta scheme, Crank-Nicolson at 1/2): kx = -(1 - theta) kappa(T_old), the
// coefficient of the explicit stiffness term applied to T_old
__global__ __launch_bounds__(kBlock) void k_heat_terms(int nc, const int *__restrict__ cells, const double4 *__restrict__ X,
                             const double *__restrict__ T, const double *__restrict__ phic, double gamma, double lorentz,
                             double theta, double *__restrict__ kq, double *__restrict__ gq, float *__restrict__ kq32,
                             float *__restrict__ gq32, double *__restrict__ f, double *__restrict__ kx) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int *cl = cells + 8 * (size_t)c;
    double xs[3][8], t[8], u[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const double4 p = X[cl[n]];
        xs[0][n] = p.x;
        xs[1][n] = p.y;
        xs[2][n] = p.z;
        t[n] = T[cl[n]];
        u[n] = phic[cl[n]];
    }
    double J0[3][2][2], J1[3][2][2], J2[3][2][2], d0[2][2], d1[2][2], d2[2][2];
#pragma unroll
    for (int co = 0; co < 3; ++co) {
        dref<double>(xs[co], 0, J0[co]);
        dref<double>(xs[co], 1, J1[co]);
        dref<double>(xs[co], 2, J2[co]);
    }
    dref<double>(u, 0, d0);
    dref<double>(u, 1, d1);
    dref<double>(u, 2, d2);
    // unrolled: the Gauss point's corner indices are constants, so the tables stay in registers
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int a = q & 1, b = (q >> 1) & 1, cc = (q >> 2) & 1;
        // T at the Gauss point: trilinear weights
        double Tq = 0;
#pragma unroll
        for (int n = 0; n < 8; ++n)
            Tq += Wt<double>(a, n & 1) * Wt<double>(b, (n >> 1) & 1) * Wt<double>(cc, (n >> 2) & 1) * t[n];
        const double c0x = J0[0][b][cc], c0y = J0[1][b][cc], c0z = J0[2][b][cc];
        const double c1x = J1[0][a][cc], c1y = J1[1][a][cc], c1z = J1[2][a][cc];
        const double c2x = J2[0][a][b], c2y = J2[1][a][b], c2z = J2[2][a][b];
        const double a0x = c1y * c2z - c1z * c2y, a0y = c1z * c2x - c1x * c2z, a0z = c1x * c2y - c1y * c2x;
        const double a1x = c2y * c0z - c2z * c0y, a1y = c2z * c0x - c2x * c0z, a1z = c2x * c0y - c2y * c0x;
        const double a2x = c0y * c1z - c0z * c1y, a2y = c0z * c1x - c0x * c1z, a2z = c0x * c1y - c0y * c1x;
        const double det = c0x * a0x + c0y * a0y + c0z * a0z;
        const double g0 = d0[b][cc], g1 = d1[a][cc], g2 = d2[a][b];
        const double vx = (g0 * a0x + g1 * a1x + g2 * a2x) / det;
        const double vy = (g0 * a0y + g1 * a1y + g2 * a2y) / det;
        const double vz = (g0 * a0z + g1 * a1z + g2 * a2z) / det;
        const double sg = pq_sigma(Tq);
        const double Tc = Tq < kRhoT[0] ? kRhoT[0] : (Tq > kRhoT[25] ? kRhoT[25] : Tq);
        const double kap = lorentz * Tc * pq_sigma(Tc);
        const size_t o = 8 * (size_t)c + q;
        kq[o] = theta * kap;
        gq[o] = gamma;
        kq32[o] = (float)(theta * kap);
        gq32[o] = (float)gamma;
        f[o] = gamma * Tq + sg * (vx * vx + vy * vy + vz * vz);
        if (kx) kx[o] = -(1.0 - theta) * kap;
    }
}

void launch_heat_terms(hipStream_t s, int nc, const int *cells, const double *X, const double *T, const double *phic,
                       double gamma, double lorentz, double theta, double *kq, double *gq, float *kq32, float *gq32,
                       double *f, double *kx) {
    if (nc > 0)
        hipLaunchKernelGGL(k_heat_terms, dim3((nc + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nc, cells,
                           reinterpret_cast<const double4 *>(X), T, phic, gamma, lorentz, theta, kq, gq, kq32, gq32, f,
                           kx);
}

// temperature-dependent conductivity of the current solve, per Gauss point of cell c
// (internal order): sigma = sigma0 / (1 + alpha (T - 300 K)) with T interpolated at q;
// no mass term
__global__ __launch_bounds__(kBlock) void k_sigma_terms(int nc, const int *__restrict__ cells, const double *__restrict__ T, double sigma0,
                              double alpha, double *__restrict__ kq, double *__restrict__ gq, float *__restrict__ kq32,
                              float *__restrict__ gq32) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int *cl = cells + 8 * (size_t)c;
    double t[8];
    for (int n = 0; n < 8; ++n) t[n] = T[cl[n]];
    for (int q = 0; q < 8; ++q) {
        const int a = q & 1, b = (q >> 1) & 1, cc = (q >> 2) & 1;
        double Tq = 0;
        for (int n = 0; n < 8; ++n)
            Tq += Wt<double>(a, n & 1) * Wt<double>(b, (n >> 1) & 1) * Wt<double>(cc, (n >> 2) & 1) * t[n];
        const double sg = sigma0 / (1.0 + alpha * (Tq - 300.0));
        const size_t o = 8 * (size_t)c + q;
        kq[o] = sg;
        gq[o] = 0.0;
        kq32[o] = (float)sg;
        gq32[o] = 0.0f;
    }
}

void launch_sigma_terms(hipStream_t s, int nc, const int *cells, const double *T, double sigma0, double alpha, double *kq,
                        double *gq, float *kq32, float *gq32) {
    if (nc > 0)
        hipLaunchKernelGGL(k_sigma_terms, dim3((nc + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nc, cells, T, sigma0,
                           alpha, kq, gq, kq32, gq32);
}

// -grad u at the Gauss point v2n of cell v2c (DealSolver::export_solution_grad)
__global__ __launch_bounds__(kBlock) void k_grad_export(int n, const int *__restrict__ v2c, const int *__restrict__ v2n, const int *__restrict__ cells,
                              const double4 *__restrict__ X, const double *__restrict__ u, double *__restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = v2c[i], q = v2n[i];
    const int a = q & 1, b = (q >> 1) & 1, cc = (q >> 2) & 1;
    const int *cl = cells + 8 * (size_t)c;
    double J[3][3], uv[8], du[3];
    for (int co = 0; co < 3; ++co) {
        double f[8], F[2][2];
        for (int n2 = 0; n2 < 8; ++n2) {
            const double4 p = X[cl[n2]];
            f[n2] = co == 0 ? p.x : (co == 1 ? p.y : p.z);
        }
        dref<double>(f, 0, F); J[co][0] = F[b][cc];
        dref<double>(f, 1, F); J[co][1] = F[a][cc];
        dref<double>(f, 2, F); J[co][2] = F[a][b];
    }
    for (int n2 = 0; n2 < 8; ++n2) uv[n2] = u[cl[n2]];
    {
        double F[2][2];
        dref<double>(uv, 0, F); du[0] = F[b][cc];
        dref<double>(uv, 1, F); du[1] = F[a][cc];
        dref<double>(uv, 2, F); du[2] = F[a][b];
    }
    // grad u = J^{-T} du: solve J^T g = du
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double inv[3][3];
    inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    for (int k = 0; k < 3; ++k) {
        // dxi_j/dx_k = inv[j][k]
        const double g = inv[0][k] * du[0] + inv[1][k] * du[1] + inv[2][k] * du[2];
        grad[3 * (size_t)i + k] = -g;
    }
}

// ---------------------------------------------------------------------------------------
// explicit instantiations through launch helpers
// ---------------------------------------------------------------------------------------
template <class T>
void launch_elem(hipStream_t s, const Desc *D, const ETab *tab, int nblk, bool coef, bool maskin, bool active_only, int scal_slot,
                 size_t lds, bool affine_all) {
    if (nblk <= 0) return;
    const dim3 g(nblk), b(kBlock);
    if (sizeof(T) == 4 && affine_all && !maskin) {
#define FG_AFF(C, A) hipLaunchKernelGGL((k_patch<T, C, false, A, EM_APPLY, true>), g, b, lds, s, D, tab, scal_slot, nullptr)
        if (coef) { if (active_only) FG_AFF(true, true); else FG_AFF(true, false); }
        else { if (active_only) FG_AFF(false, true); else FG_AFF(false, false); }
#undef FG_AFF
        return;
    }
#define FG_ELEM(C, M, A) hipLaunchKernelGGL((k_patch<T, C, M, A, EM_APPLY>), g, b, lds, s, D, tab, scal_slot, nullptr)
    if (coef) {
        if (maskin) { if (active_only) FG_ELEM(true, true, true); else FG_ELEM(true, true, false); }
        else { if (active_only) FG_ELEM(true, false, true); else FG_ELEM(true, false, false); }
    } else {
        if (maskin) { if (active_only) FG_ELEM(false, true, true); else FG_ELEM(false, true, false); }
        else { if (active_only) FG_ELEM(false, false, true); else FG_ELEM(false, false, false); }
    }
#undef FG_ELEM
}
template void launch_elem<double>(hipStream_t, const Desc *, const ETab *, int, bool, bool, bool, int, size_t, bool);
template void launch_elem<float>(hipStream_t, const Desc *, const ETab *, int, bool, bool, bool, int, size_t, bool);

template <class T> void launch_elem_diag(hipStream_t s, const Desc *D, const ETab *tab, int nblk, bool coef, size_t lds) {
    if (nblk <= 0) return;
    if (coef) hipLaunchKernelGGL((k_patch<T, true, false, false, EM_DIAG>), dim3(nblk), dim3(kBlock), lds, s, D, tab, 0, nullptr);
    else hipLaunchKernelGGL((k_patch<T, false, false, false, EM_DIAG>), dim3(nblk), dim3(kBlock), lds, s, D, tab, 0, nullptr);
}
template void launch_elem_diag<double>(hipStream_t, const Desc *, const ETab *, int, bool, size_t);

void launch_elem_rhs(hipStream_t s, const Desc *D, const ETab *tab, int nblk, const double *fq, size_t lds) {
    if (nblk > 0) hipLaunchKernelGGL((k_patch<double, false, false, false, EM_RHS>), dim3(nblk), dim3(kBlock), lds, s, D, tab, 0, fq);
}

// per slot the volume of its cell (the dof volumes once gathered)
void launch_elem_vol(hipStream_t s, const Desc *D, const ETab *tab, int nblk, size_t lds) {
    if (nblk > 0) hipLaunchKernelGGL((k_patch<double, false, false, false, EM_VOL>), dim3(nblk), dim3(kBlock), lds, s, D, tab, 0, nullptr);
}

static inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

// rho = b / vol (PoissonSolver::assemble, src/PoissonSolver.cpp:196-207)
__global__ void k_charge(int n, const double *__restrict__ b, const double *__restrict__ vol, double *__restrict__ rho) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rho[i] = b[i] / vol[i];
}
void launch_charge(hipStream_t s, int n, const double *b, const double *vol, double *rho) {
    if (n > 0) hipLaunchKernelGGL(k_charge, dim3(grid_for(n)), dim3(kBlock), 0, s, n, b, vol, rho);
}

void launch_enforce_dirichlet(hipStream_t s, int n, const unsigned char *mask, const double *g, double *x) {
    if (n > 0) hipLaunchKernelGGL(k_enforce_dirichlet, dim3(grid_for(n)), dim3(kBlock), 0, s, n, mask, g, x);
}
void launch_lift(hipStream_t s, int n, const unsigned char *mask, const double *g, const double *Ag, const double *diag,
                 double *b, double *x) {
    if (n > 0) hipLaunchKernelGGL(k_lift, dim3(grid_for(n)), dim3(kBlock), 0, s, n, mask, g, Ag, diag, b, x);
}
// ------------------------------------------------------------------------------------------
// coarse operator of the two-level preconditioner: E = Z^T A Z from A applied to the
// indicators of one colour of dof blocks at a time, then E^-1 by Gauss-Jordan sweeps
// ------------------------------------------------------------------------------------------
__global__ void k_coarse_indicator(int n, const unsigned char *__restrict__ mask, const int *__restrict__ color,
                                   int shift, int c, double *__restrict__ u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = (!mask[i] && color[i >> shift] == c) ? 1.0 : 0.0;
}

// out[q] = sum of q over the free dofs of the 256-dof round q (one workgroup per dof block of
// kAggPerBlock rounds, fixed order)
__global__ __launch_bounds__(kBlock) void k_coarse_sums(int n, const unsigned char *__restrict__ mask,
                                                         const double *__restrict__ q, double *__restrict__ out) {
    double v[kAggPerBlock];
#pragma unroll
    for (int a = 0; a < kAggPerBlock; ++a) v[a] = 0.0;
#pragma unroll
    for (int m = 0; m < kDofIter; ++m) {
        const int i = blockIdx.x * kDofBlock + threadIdx.x + m * kBlock;
        if (i < n && !mask[i]) v[m / (kDofIter / kAggPerBlock)] += q[i];
    }
    block_sum<kAggPerBlock>(v);
    if (threadIdx.x == 0) {
        const int na = (n + kAggMin - 1) / kAggMin;
#pragma unroll
        for (int a = 0; a < kAggPerBlock; ++a)
            if (blockIdx.x * kAggPerBlock + a < na) out[blockIdx.x * kAggPerBlock + a] = v[a];
    }
}

// Gauss-Jordan exchange on a block K of kGjBlock pivots (no pivoting: E is SPD), out of place,
// with P = A_KK: B_KK = P^-1, B_Kj = P^-1 A_Kj, B_iK = -A_iK P^-1, B_ij = A_ij - A_iK P^-1 A_Kj.
// One block is the same as its scalar exchanges in turn (b_kk = 1/a_kk, b_kj = a_kj/a_kk,
// b_ik = -a_ik/a_kk, b_ij = a_ij - a_ik a_kj/a_kk); after every block the matrix holds the
// inverse.  Three launches per block: P^-1 (one workgroup, scalar exchanges in LDS), the panels
// V = A_iK P^-1 and W = P^-1 A_Kj, then the update.
constexpr int kGjBlock = 32;

// P^-1 of the bb x bb block at (k, k), bb <= kGjBlock, into Pi (kGjBlock x kGjBlock, row-major)
__global__ __launch_bounds__(256) void k_gj_pivot(int n, int k, int bb, const double *__restrict__ A,
                                                  double *__restrict__ Pi) {
    __shared__ double P[kGjBlock][kGjBlock + 1];
    constexpr int kPer = kGjBlock * kGjBlock / 256;
    for (int e = threadIdx.x; e < kGjBlock * kGjBlock; e += 256) {
        const int r = e / kGjBlock, c = e % kGjBlock;
        P[r][c] = (r < bb && c < bb) ? A[(size_t)(k + r) * n + k + c] : 0.0;
    }
    __syncthreads();
    for (int p = 0; p < bb; ++p) {
        double v[kPer];
        const double piv = P[p][p];
#pragma unroll
        for (int m = 0; m < kPer; ++m) {
            const int e = threadIdx.x + m * 256, r = e / kGjBlock, c = e % kGjBlock;
            if (r >= bb || c >= bb) v[m] = 0.0;
            else if (r == p && c == p) v[m] = 1.0 / piv;
            else if (r == p) v[m] = P[r][c] / piv;
            else if (c == p) v[m] = -P[r][c] / piv;
            else v[m] = P[r][c] - P[r][p] * P[p][c] / piv;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < kPer; ++m) {
            const int e = threadIdx.x + m * 256;
            P[e / kGjBlock][e % kGjBlock] = v[m];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < kGjBlock * kGjBlock; e += 256) Pi[e] = P[e / kGjBlock][e % kGjBlock];
}

// V[i][q] = sum_p A[i][k+p] Pi[p][q] (n x kGjBlock) and W[q][j] = sum_p Pi[q][p] A[k+p][j]
// (kGjBlock x n), q < bb: threads [0, n bb) for V, [n bb, 2 n bb) for W
__global__ __launch_bounds__(256) void k_gj_panel(int n, int k, int bb, const double *__restrict__ A,
                                                  const double *__restrict__ Pi, double *__restrict__ V,
                                                  double *__restrict__ W) {
    __shared__ double sP[kGjBlock * kGjBlock];
    for (int e = threadIdx.x; e < kGjBlock * kGjBlock; e += 256) sP[e] = Pi[e];
    __syncthreads();
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x, nb = (size_t)n * bb;
    if (t < nb) {
        const int i = (int)(t / bb), q = (int)(t % bb);
        double s = 0.0;
        for (int p = 0; p < bb; ++p) s += A[(size_t)i * n + k + p] * sP[p * kGjBlock + q];
        V[(size_t)i * kGjBlock + q] = s;
    } else if (t < 2 * nb) {
        const int q = (int)((t - nb) / n), j = (int)((t - nb) % n);
        double s = 0.0;
        for (int p = 0; p < bb; ++p) s += sP[q * kGjBlock + p] * A[(size_t)(k + p) * n + j];
        W[(size_t)q * n + j] = s;
    }
}

__global__ __launch_bounds__(256) void k_gj_update(int n, int k, int bb, const double *__restrict__ A,
                                                   const double *__restrict__ Pi, const double *__restrict__ V,
                                                   const double *__restrict__ W, double *__restrict__ B) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (size_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    const bool ri = i >= k && i < k + bb, cj = j >= k && j < k + bb;
    double v;
    if (ri && cj) v = Pi[(i - k) * kGjBlock + (j - k)];
    else if (ri) v = W[(size_t)(i - k) * n + j];
    else if (cj) v = -V[(size_t)i * kGjBlock + (j - k)];
    else {
        v = A[e];
        for (int q = 0; q < bb; ++q) v -= V[(size_t)i * kGjBlock + q] * A[(size_t)(k + q) * n + j];
    }
    B[e] = v;
}

// out = a / scale in FP16 (scale a power of two: the division is exact before the rounding);
// a: n x n, out: n rows of ld >= n entries, zero past n
__global__ void k_to_half(int n, int ld, const double *__restrict__ a, double inv_scale, _Float16 *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * ld) return;
    const int r = (int)(i / ld), c = (int)(i % ld);
    out[i] = c < n ? (_Float16)(a[(size_t)r * n + c] * inv_scale) : (_Float16)0.0f;
}

void launch_coarse_indicator(hipStream_t s, int n, const unsigned char *mask, const int *color, int shift, int c,
                             double *u) {
    if (n > 0) hipLaunchKernelGGL(k_coarse_indicator, dim3((n + 255) / 256), dim3(256), 0, s, n, mask, color, shift, c, u);
}
void launch_coarse_sums(hipStream_t s, int n, const unsigned char *mask, const double *q, double *out) {
    const int nb = (n + kDofBlock - 1) / kDofBlock;
    if (nb > 0) hipLaunchKernelGGL(k_coarse_sums, dim3(nb), dim3(kBlock), 0, s, n, mask, q, out);
}
int gj_block() { return kGjBlock; }
// one block exchange on pivots [k, k + bb): Pi holds kGjBlock^2 doubles, V and W n kGjBlock each
void launch_gj_block(hipStream_t s, int n, int k, int bb, const double *A, double *B, double *Pi, double *V,
                     double *W) {
    const size_t e = (size_t)n * n, nb = 2 * (size_t)n * bb;
    if (!e || bb <= 0 || bb > kGjBlock) return;
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(256), 0, s, n, k, bb, A, Pi);
    hipLaunchKernelGGL(k_gj_panel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, n, k, bb, A, Pi, V, W);
    hipLaunchKernelGGL(k_gj_update, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, s, n, k, bb, A, Pi, V, W, B);
}
void launch_to_half(hipStream_t s, int n, int ld, const double *a, double scale, _Float16 *out) {
    const size_t e = (size_t)n * ld;
    if (e) hipLaunchKernelGGL(k_to_half, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, s, n, ld, a, 1.0 / scale, out);
}

void launch_invdiag(hipStream_t s, int n, const unsigned char *mask, const double *diag, double *inv64, float *inv32) {
    if (n > 0) hipLaunchKernelGGL(k_invdiag, dim3(grid_for(n)), dim3(kBlock), 0, s, n, mask, diag, inv64, inv32);
}
void launch_set_const(hipStream_t s, int n, double v, double *x) {
    if (n > 0) hipLaunchKernelGGL(k_set_const, dim3(grid_for(n)), dim3(kBlock), 0, s, n, v, x);
}
void launch_face_load(hipStream_t s, int nb, const int *bf_icell, const int *bf_face, const int *bf_id, int bid,
                      const double *fval, const int *cells, const double *X, double *fslot) {
    if (nb > 0)
        hipLaunchKernelGGL(k_face_load, dim3(grid_for(nb)), dim3(kBlock), 0, s, nb, bf_icell, bf_face, bf_id, bid, fval, cells,
                           reinterpret_cast<const double4 *>(X), fslot);
}
void launch_face_gather(hipStream_t s, int n, const int *fgptr, const int *fgidx, const double *fslot, double *b) {
    if (n > 0) hipLaunchKernelGGL(k_face_gather, dim3(grid_for(n)), dim3(kBlock), 0, s, n, fgptr, fgidx, fslot, b);
}
void launch_minmax(hipStream_t s, int n, const double *x, double *out) {
    hipLaunchKernelGGL(k_minmax, dim3(1), dim3(kBlock), 0, s, n, x, out);
}
int center_det_blocks(int n) { return grid_for(n); }
void launch_motion(hipStream_t s, int rows, const int *idof, const int *ptr, const int *ctrl, const double *w,
                   const double *base, const double *disp, double *X64, float *X32) {
    if (rows > 0)
        hipLaunchKernelGGL(k_motion, dim3(grid_for(rows)), dim3(kBlock), 0, s, rows, idof, ptr, ctrl, w, base, disp,
                           reinterpret_cast<double4 *>(X64), reinterpret_cast<float4 *>(X32));
}
void launch_scatter_coords(hipStream_t s, int n, const int *idx, const double *xyz, double *X64, float *X32) {
    if (n > 0)
        hipLaunchKernelGGL(k_scatter_coords, dim3(grid_for(n)), dim3(kBlock), 0, s, n, idx, xyz,
                           reinterpret_cast<double4 *>(X64), reinterpret_cast<float4 *>(X32));
}
void launch_center_det_min(hipStream_t s, int n, const int *cells, const double *X, double *out) {
    if (n > 0)
        hipLaunchKernelGGL(k_center_det_min, dim3(grid_for(n)), dim3(kBlock), 0, s, n, cells,
                           reinterpret_cast<const double4 *>(X), out);
}
void launch_grad_export(hipStream_t s, int n, const int *v2c, const int *v2n, const int *cells, const double *X, const double *u,
                        double *grad) {
    if (n > 0)
        hipLaunchKernelGGL(k_grad_export, dim3(grid_for(n)), dim3(kBlock), 0, s, n, v2c, v2n, cells,
                           reinterpret_cast<const double4 *>(X), u, grad);
}

}  // namespace fgpu

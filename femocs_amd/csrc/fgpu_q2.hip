// Q2 Laplace solve: the reference's Poisson path with the element degree raised to 2, i.e. a
// recompile of include/DealSolver.h:130-131 with shape_degree = 2 (FE_Q(2), QGauss(3) on cells
// and QGauss<2>(3) on faces, trilinear MappingQ1 geometry), on the hexes of an imported mesh.
//
// Dofs: the mesh's vertices (the Q1 dofs, same numbering) first, then edge midpoints, face
// centres and cell centres in order of first appearance over the internal (space-filling
// curve) cell order.  Matrix-free apply, sum-factorised: 9 threads per cell (one per column of
// the 3 x 3 x 3 node grid), 28 cells per workgroup, three 1D contractions into the 27 Gauss
// points through LDS, the metric of the trilinear map at each point, three transposed
// contractions back; each cell's 27 outputs go to dof-major slots, which the CG gather of the
// Q1 path sums in a fixed order (deterministic, no atomics).  The CG, the Jacobi and
// Chebyshev-Jacobi preconditioners and the scalar kernels are the Q1 path's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "fgpu_cheb.h"
#include "fgpu_internal.h"
#include "fgpu_kernels.h"

using namespace fgpu;

namespace {

struct HipErr {
    hipError_t e;
    const char *what;
};
#define Q2CHECK(call)                                     \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) throw HipErr{e_, #call};    \
    } while (0)

template <class T> T *q2alloc(size_t n) {
    void *p = nullptr;
    Q2CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return static_cast<T *>(p);
}
template <class T> void q2free(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

constexpr int kQ2CellsPerWave = 7;           // 9 threads each: 63 of a wave's 64 lanes
constexpr int kQ2Cells = 4 * kQ2CellsPerWave; // cells per workgroup of 4 waves

// 1D Q2 basis on the nodes 0, 1/2, 1 at the QGauss(3) points of [0, 1]: B[q][n], derivative
// D[q][n], weights W[q]
__constant__ double kQB[3][3], kQD[3][3], kQW[3];

void q2_tables(double B[3][3], double D[3][3], double W[3]) {
    const double g[3] = {0.5 - 0.5 * std::sqrt(0.6), 0.5, 0.5 + 0.5 * std::sqrt(0.6)};
    const double w[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
    for (int q = 0; q < 3; ++q) {
        const double x = g[q];
        B[q][0] = (1 - x) * (1 - 2 * x);
        B[q][1] = 4 * x * (1 - x);
        B[q][2] = x * (2 * x - 1);
        D[q][0] = 4 * x - 3;
        D[q][1] = 4 - 8 * x;
        D[q][2] = 4 * x - 1;
        W[q] = w[q];
    }
}

__device__ __constant__ double kQG[3] = {0.1127016653792583, 0.5, 0.8872983346207417};

// Tet-split meshes: every hex is the image of one reference sub-hex (frame coordinates of its
// lexicographic nodes below: tet vertex, three edge midpoints, three face centroids, centroid)
// under the affine map x = x0 + A lambda, A = [2(x1 - x0), 2(x4 - x0), 2(x2 - x0)].  With
// J = A J_ref(xi), w |det J| J^-1 J^-T = P_q^T M P_q for the per-cell metric
// M = |det A| (A^T A)^-1 and the constants P_q = sqrt(w_q |det J_ref,q|) J_ref,q^-T (sign of
// det J_ref,q included, it cancels): 6 numbers per cell replace the Jacobian at each of the 27
// Gauss points.  An eta-flipped cell (layout 1) has the form in the order n ^ 2 of its
// vertices, i.e. with the node grid's middle index reversed.
static const double kQ2SubHex[8][3] = {{0, 0, 0}, {0.5, 0, 0}, {0, 0, 0.5}, {1.0 / 3, 0, 1.0 / 3},
                                       {0, 0.5, 0}, {1.0 / 3, 1.0 / 3, 0}, {0, 1.0 / 3, 1.0 / 3}, {0.25, 0.25, 0.25}};
__constant__ double kQ2P[27][9];

void q2_affine_table(double P[27][9]) {
    const double g[3] = {0.5 - 0.5 * std::sqrt(0.6), 0.5, 0.5 + 0.5 * std::sqrt(0.6)};
    const double w[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
    for (int q = 0; q < 27; ++q) {
        const int qi = q % 3, qj = (q / 3) % 3, qk = q / 9;
        const double xi[3] = {g[qi], g[qj], g[qk]};
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
                for (int d = 0; d < 3; ++d) J[i][d] += kQ2SubHex[n][i] * gr[d];
        }
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        // cofactor(J) = det J^-T
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
        const double sc = std::sqrt(std::fabs(w[qi] * w[qj] * w[qk] * det)) / det;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) P[q][3 * i + j] = C[i][j] * sc;
    }
}

// per cell (internal order): the FP64 metric of the tet-split form and, per block, the largest
// deviation of the vertices from that form relative to the frame size (1e300 for a cell whose
// frame has the wrong orientation for its layout)
__global__ __launch_bounds__(256) void k_q2_metric(int nc, const int *__restrict__ c27, const unsigned char *__restrict__ layout,
                                                   const double4 *__restrict__ X, double *__restrict__ qm,
                                                   double *__restrict__ devmax) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    double dev = 0;
    if (c < nc) {
        const bool flip = layout[c] != 0;
        double x[8][3];
        for (int n = 0; n < 8; ++n) {
            const int v = flip ? (n ^ 2) : n;     // vertex (a, b, c) is grid node 2a + 6b + 18c
            const double4 p = X[c27[27 * (size_t)c + 2 * (v & 1) + 6 * ((v >> 1) & 1) + 18 * (v >> 2)]];
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
        for (int n = 0; n < 8; ++n) {
            // the sub-hex node coordinates of the frame (kQ2SubHex, exact in binary but 1/3 and 1/4)
            const double l0 = (n == 1) ? 0.5 : (n == 3 || n == 5) ? 1.0 / 3 : (n == 7) ? 0.25 : 0.0;
            const double l1 = (n == 4) ? 0.5 : (n == 5 || n == 6) ? 1.0 / 3 : (n == 7) ? 0.25 : 0.0;
            const double l2 = (n == 2) ? 0.5 : (n == 3 || n == 6) ? 1.0 / 3 : (n == 7) ? 0.25 : 0.0;
            for (int a = 0; a < 3; ++a) {
                const double pr = x[0][a] + l0 * u[0][a] + l1 * u[1][a] + l2 * u[2][a];
                dev = fmax(dev, fabs(pr - x[n][a]) / size);
            }
        }
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
            qm[6 * (size_t)c + k] = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / fabs(det);
        }
        if (!(flip ? det > 0 : det < 0)) dev = 1e300;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dev = fmax(dev, __shfl_xor(dev, o, 64));
    __shared__ double sm[4];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = dev;
    __syncthreads();
    if (threadIdx.x == 0) devmax[blockIdx.x] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}

// columns c0, c1, c2 (d x / d xi_a) of the trilinear map of vertices X[8][3] (lexicographic) at
// the reference point (x, y, z)
__device__ __forceinline__ void q2_jacobian(const double (*X)[3], double x, double y, double z, double c[3][3]) {
    const double mx = 1 - x, my = 1 - y, mz = 1 - z;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        c[0][d] = my * mz * (X[1][d] - X[0][d]) + y * mz * (X[3][d] - X[2][d]) + my * z * (X[5][d] - X[4][d]) +
                  y * z * (X[7][d] - X[6][d]);
        c[1][d] = mx * mz * (X[2][d] - X[0][d]) + x * mz * (X[3][d] - X[1][d]) + mx * z * (X[6][d] - X[4][d]) +
                  x * z * (X[7][d] - X[5][d]);
        c[2][d] = mx * my * (X[4][d] - X[0][d]) + x * my * (X[5][d] - X[1][d]) + mx * y * (X[6][d] - X[2][d]) +
                  x * y * (X[7][d] - X[3][d]);
    }
}

// rows a_k of det * J^-1 and det
__device__ __forceinline__ double q2_adj(const double c[3][3], double a[3][3]) {
    a[0][0] = c[1][1] * c[2][2] - c[1][2] * c[2][1];
    a[0][1] = c[1][2] * c[2][0] - c[1][0] * c[2][2];
    a[0][2] = c[1][0] * c[2][1] - c[1][1] * c[2][0];
    a[1][0] = c[2][1] * c[0][2] - c[2][2] * c[0][1];
    a[1][1] = c[2][2] * c[0][0] - c[2][0] * c[0][2];
    a[1][2] = c[2][0] * c[0][1] - c[2][1] * c[0][0];
    a[2][0] = c[0][1] * c[1][2] - c[0][2] * c[1][1];
    a[2][1] = c[0][2] * c[1][0] - c[0][0] * c[1][2];
    a[2][2] = c[0][0] * c[1][1] - c[0][1] * c[1][0];
    return c[0][0] * a[0][0] + c[0][1] * a[0][1] + c[0][2] * a[0][2];
}

// lexicographic node of the vertex v = a + 2b + 4c of the Q1 cell
__host__ __device__ __forceinline__ int q2_vnode(int v) { return 2 * (v & 1) + 6 * ((v >> 1) & 1) + 18 * (v >> 2); }

struct Q2Args {
    int nc;
    const int *c27, *spos;
    const double4 *X;            // vertex coordinates (Q2 vertex dofs = the mesh's internal dofs)
    const double *ua, *ub;       // input u = ua + beta ub (ub may be null)
    const double *scal;          // S_BETA, S_ACTIVE
    const unsigned char *dmask;  // masked inputs (maskin)
    double *Ye;                  // dof-major slots
    int maskin, active_only;
    const double *qm;            // tet-split form: 6 per cell (AFF kernels only)
    const unsigned char *layout; // tet-split form: 1 for an eta-flipped cell
};

// orders a wave's LDS writes before its lanes' later reads (no other wave takes part)
__device__ __forceinline__ void q2_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// y = K_c u for every cell, K_c(i, j) = sum_q JxW grad phi_i . grad phi_j (QGauss(3)).
// AFF: every cell has the tet-split form (per-cell metric, no coordinates); a flipped cell is
// computed with its node grid's middle index reversed (thread column tb reads and writes row
// 2 - tb), which relabels the nodes of the same element matrix.
template <bool AFF>
__global__ __launch_bounds__(256) void k_q2_apply(Q2Args a) {
    __shared__ double sX[AFF ? 1 : kQ2Cells][8][3];
    // one exchange row per cell, overwritten stage after stage: each stage reads what it needs
    // into registers, the wave syncs, then it writes (a cell's row belongs to one wave)
    __shared__ double sA[kQ2Cells][81];
    if (a.active_only && a.scal[S_ACTIVE] == 0.0) return;
    const int tid = threadIdx.x;
    // a cell's 9 threads lie in one wave, so the exchanges between them need no workgroup
    // barrier: a wave's own LDS writes are visible to its lanes after a wave-level fence
    const int lane = tid & 63, wv = tid >> 6;
    const int lc = wv * kQ2CellsPerWave + lane / 9, t = lane % 9, ta = t % 3, tb = t / 3;
    const bool inwg = lane < 9 * kQ2CellsPerWave;
    // the grid is a multiple of 8 and workgroups are dispatched round-robin over the 8 XCDs:
    // workgroup w runs cell block (w mod 8) * (grid / 8) + w / 8, so each XCD (its own L2) gets
    // one contiguous range of cells and the dofs and slot lines they share
    const int per = gridDim.x >> 3;
    const int blk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int c = blk * kQ2Cells + lc;
    const bool live = inwg && c < a.nc;
    const int cc = live ? c : 0;
    const int *cd = a.c27 + 27 * (size_t)cc;
    const double beta = a.ub ? a.scal[S_BETA] : 0.0;
    // the row of the grid this thread's column reads and writes
    const int tr = (AFF && a.layout[cc]) ? 2 - tb : tb;
    double M[6] = {0, 0, 0, 0, 0, 0};
    if (AFF) {
#pragma unroll
        for (int k = 0; k < 6; ++k) M[k] = a.qm[6 * (size_t)cc + k];
    }
    // the column's output slots, loaded with its inputs (they are stored at the very end)
    int slot[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) slot[k] = a.spos[27 * (size_t)cc + ta + 3 * tr + 9 * k];
    // column (i = ta, j = tb) of the nodal values; vertex coordinates of the cell
    double u[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int dof = cd[ta + 3 * tr + 9 * k];
        double v = a.ua[dof];
        if (a.ub) v += beta * a.ub[dof];
        if (a.maskin && a.dmask[dof]) v = 0.0;
        u[k] = v;
    }
    if (!AFF && inwg && t < 8) {
        const double4 p = a.X[cd[q2_vnode(t)]];
        sX[lc][t][0] = p.x;
        sX[lc][t][1] = p.y;
        sX[lc][t][2] = p.z;
    }
    // 1: along z (thread (i, j)): T1[i,j,qk] and its derivative T1d
    if (inwg) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            double s = 0, sd = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                s += kQB[q][k] * u[k];
                sd += kQD[q][k] * u[k];
            }
            sA[lc][ta + 3 * tb + 9 * q] = s;
            sA[lc][27 + ta + 3 * tb + 9 * q] = sd;
        }
    }
    q2_wave_sync();
    // 2: along y (thread (i, qk)): T2 (B), T2y (D) from T1; T2z (B) from T1d
    if (inwg) {
        const int i = ta, qk = tb;
        double t1[3], t1d[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            t1[j] = sA[lc][i + 3 * j + 9 * qk];
            t1d[j] = sA[lc][27 + i + 3 * j + 9 * qk];
        }
        q2_wave_sync();
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            double s = 0, sy = 0, sz = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                s += kQB[q][j] * t1[j];
                sy += kQD[q][j] * t1[j];
                sz += kQB[q][j] * t1d[j];
            }
            sA[lc][i + 3 * q + 9 * qk] = s;
            sA[lc][27 + i + 3 * q + 9 * qk] = sy;
            sA[lc][54 + i + 3 * q + 9 * qk] = sz;
        }
    }
    q2_wave_sync();
    // 3: along x (thread (qj, qk)): reference gradient at (qi, qj, qk), metric, and back along x
    if (inwg) {
        const int qj = ta, qk = tb;
        double t2[3], t2y[3], t2z[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            t2[i] = sA[lc][i + 3 * qj + 9 * qk];
            t2y[i] = sA[lc][27 + i + 3 * qj + 9 * qk];
            t2z[i] = sA[lc][54 + i + 3 * qj + 9 * qk];
        }
        q2_wave_sync();
        double hx[3], hy[3], hz[3];
#pragma unroll
        for (int qi = 0; qi < 3; ++qi) {
            double g0 = 0, g1 = 0, g2 = 0;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                g0 += kQD[qi][i] * t2[i];
                g1 += kQB[qi][i] * t2y[i];
                g2 += kQB[qi][i] * t2z[i];
            }
            if (AFF) {
                // P_q^T M P_q g
                const double *P = kQ2P[qi + 3 * qj + 9 * qk];
                const double s0 = P[0] * g0 + P[1] * g1 + P[2] * g2;
                const double s1 = P[3] * g0 + P[4] * g1 + P[5] * g2;
                const double s2 = P[6] * g0 + P[7] * g1 + P[8] * g2;
                const double t0 = M[0] * s0 + M[1] * s1 + M[2] * s2;
                const double t1 = M[1] * s0 + M[3] * s1 + M[4] * s2;
                const double t2 = M[2] * s0 + M[4] * s1 + M[5] * s2;
                hx[qi] = P[0] * t0 + P[3] * t1 + P[6] * t2;
                hy[qi] = P[1] * t0 + P[4] * t1 + P[7] * t2;
                hz[qi] = P[2] * t0 + P[5] * t1 + P[8] * t2;
            } else {
                double cj[3][3], ad[3][3];
                q2_jacobian(sX[lc], kQG[qi], kQG[qj], kQG[qk], cj);
                const double det = q2_adj(cj, ad);
                // det * grad u, then w / det * (a_k . det grad u)
                const double vx = g0 * ad[0][0] + g1 * ad[1][0] + g2 * ad[2][0];
                const double vy = g0 * ad[0][1] + g1 * ad[1][1] + g2 * ad[2][1];
                const double vz = g0 * ad[0][2] + g1 * ad[1][2] + g2 * ad[2][2];
                const double s = kQW[qi] * kQW[qj] * kQW[qk] / det;
                hx[qi] = s * (ad[0][0] * vx + ad[0][1] * vy + ad[0][2] * vz);
                hy[qi] = s * (ad[1][0] * vx + ad[1][1] * vy + ad[1][2] * vz);
                hz[qi] = s * (ad[2][0] * vx + ad[2][1] * vy + ad[2][2] * vz);
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double s = 0, sy = 0, sz = 0;
#pragma unroll
            for (int qi = 0; qi < 3; ++qi) {
                s += kQD[qi][i] * hx[qi];
                sy += kQB[qi][i] * hy[qi];
                sz += kQB[qi][i] * hz[qi];
            }
            sA[lc][i + 3 * qj + 9 * qk] = s;
            sA[lc][27 + i + 3 * qj + 9 * qk] = sy;
            sA[lc][54 + i + 3 * qj + 9 * qk] = sz;
        }
    }
    q2_wave_sync();
    // 4: back along y (thread (i, qk))
    if (inwg) {
        const int i = ta, qk = tb;
        double s1[3], s1y[3], s1z[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            s1[q] = sA[lc][i + 3 * q + 9 * qk];
            s1y[q] = sA[lc][27 + i + 3 * q + 9 * qk];
            s1z[q] = sA[lc][54 + i + 3 * q + 9 * qk];
        }
        q2_wave_sync();
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = 0, sz = 0;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                s += kQB[q][j] * s1[q] + kQD[q][j] * s1y[q];
                sz += kQB[q][j] * s1z[q];
            }
            sA[lc][i + 3 * j + 9 * qk] = s;
            sA[lc][27 + i + 3 * j + 9 * qk] = sz;
        }
    }
    q2_wave_sync();
    // 5: back along z (thread (i, j)), one store per node into its slot
    if (live) {
        const int i = ta, j = tb;
        double s2[3], s2z[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            s2[q] = sA[lc][i + 3 * j + 9 * q];
            s2z[q] = sA[lc][27 + i + 3 * j + 9 * q];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double y = 0;
#pragma unroll
            for (int q = 0; q < 3; ++q) y += kQB[q][k] * s2[q] + kQD[q][k] * s2z[q];
            a.Ye[slot[k]] = y;
        }
    }
}

// diagonal: one thread per (cell, node), sum over the 27 Gauss points of JxW |grad phi_n|^2
__global__ __launch_bounds__(256) void k_q2_diag(int nc, const int *__restrict__ c27, const int *__restrict__ spos,
                                                 const double4 *__restrict__ X, double *__restrict__ Ye) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 27 * (size_t)nc) return;
    const size_t c = e / 27;
    const int n = (int)(e % 27), ni = n % 3, nj = (n / 3) % 3, nk = n / 9;
    double x[8][3];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const double4 p = X[c27[27 * c + q2_vnode(v)]];
        x[v][0] = p.x;
        x[v][1] = p.y;
        x[v][2] = p.z;
    }
    double sum = 0;
    for (int qk = 0; qk < 3; ++qk)
        for (int qj = 0; qj < 3; ++qj)
            for (int qi = 0; qi < 3; ++qi) {
                double cj[3][3], ad[3][3];
                q2_jacobian(x, kQG[qi], kQG[qj], kQG[qk], cj);
                const double det = q2_adj(cj, ad);
                const double g0 = kQD[qi][ni] * kQB[qj][nj] * kQB[qk][nk];
                const double g1 = kQB[qi][ni] * kQD[qj][nj] * kQB[qk][nk];
                const double g2 = kQB[qi][ni] * kQB[qj][nj] * kQD[qk][nk];
                const double vx = g0 * ad[0][0] + g1 * ad[1][0] + g2 * ad[2][0];
                const double vy = g0 * ad[0][1] + g1 * ad[1][1] + g2 * ad[2][1];
                const double vz = g0 * ad[0][2] + g1 * ad[1][2] + g2 * ad[2][2];
                sum += kQW[qi] * kQW[qj] * kQW[qk] / det * (vx * vx + vy * vy + vz * vz);
            }
    Ye[spos[e]] = sum;
}

// Neumann load of the boundary faces with id bid: 9 contributions per face into fslot
// (DealSolver::assemble_rhs, src/DealSolver.cpp:390-430, with QGauss<2>(3) and the 9 face dofs)
__global__ void k_q2_face(int nb, const int *__restrict__ icell, const int *__restrict__ face, const int *__restrict__ id,
                          int bid, double value, const int *__restrict__ c27, const double4 *__restrict__ X,
                          double *__restrict__ fslot) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nb) return;
    double out[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (id[f] == bid) {
        const int fc = face[f], d = fc / 2, side = fc % 2;
        const int ax0 = d == 0 ? 1 : 0, ax1 = d == 2 ? 1 : 2;
        double x[8][3];
        for (int v = 0; v < 8; ++v) {
            const double4 p = X[c27[27 * (size_t)icell[f] + q2_vnode(v)]];
            x[v][0] = p.x;
            x[v][1] = p.y;
            x[v][2] = p.z;
        }
        for (int qt = 0; qt < 3; ++qt)
            for (int qs = 0; qs < 3; ++qs) {
                double xi[3];
                xi[d] = side;
                xi[ax0] = kQG[qs];
                xi[ax1] = kQG[qt];
                double cj[3][3];
                q2_jacobian(x, xi[0], xi[1], xi[2], cj);
                const double *u = cj[ax0], *v = cj[ax1];
                const double nx = u[1] * v[2] - u[2] * v[1], ny = u[2] * v[0] - u[0] * v[2], nz = u[0] * v[1] - u[1] * v[0];
                const double jxw = kQW[qs] * kQW[qt] * sqrt(nx * nx + ny * ny + nz * nz) * value;
                // face nodes in ascending lexicographic order: (ax0 index, ax1 index) = (m % 3, m / 3)
                for (int m = 0; m < 9; ++m) out[m] += kQB[qs][m % 3] * kQB[qt][m / 3] * jxw;
            }
    }
    for (int m = 0; m < 9; ++m) fslot[9 * (size_t)f + m] = out[m];
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------------------------
struct fgpu_q2 {
    fgpu_ctx *ctx = nullptr;
    fgpu_mesh *mesh = nullptr;
    int nv = 0, n_dofs = 0, n_cells = 0, nb = 0, nb_dof = 0;
    std::vector<int> c27;                 // internal cells x 27, internal q2 dofs
    std::vector<int> api_of_q2, q2_of_api;
    std::vector<unsigned char> face_dofs_mask;
    int *d_c27 = nullptr, *d_spos = nullptr, *d_gptr = nullptr, *d_fgptr = nullptr, *d_fgidx = nullptr;
    double *fslot = nullptr;
    // system (internal dof order)
    double *b = nullptr, *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *Ye = nullptr;
    double *invd = nullptr, *diag = nullptr, *g = nullptr, *tmp = nullptr, *cd = nullptr, *cres = nullptr;
    double *ccoef = nullptr, *hist = nullptr, *scal = nullptr, *part = nullptr, *dflag = nullptr;
    float *invd32 = nullptr;
    unsigned char *d_dmask = nullptr;
    std::vector<unsigned char> dmask;
    std::vector<double> gval;
    bool lifted = false, diag_valid = false;
    int diag_gen = -1;                    // coordinate generation of the diagonal
    double cheb_lmax = 0;
    // tet-split form: per-cell FP64 metric, valid for the mesh's coordinate generation metric_gen
    double *d_qm = nullptr;
    bool affine = false;
    int metric_gen = -1;
    Desc *d_desc = nullptr;
    int2 *d_dtab = nullptr;
    ~fgpu_q2() {
        for (double **v : {&b, &x, &r, &z, &p, &q, &Ye, &invd, &diag, &g, &tmp, &cd, &cres, &ccoef, &hist, &scal, &part,
                           &dflag, &fslot})
            q2free(*v);
        q2free(invd32);
        q2free(d_qm);
        q2free(d_dmask);
        q2free(d_c27);
        q2free(d_spos);
        q2free(d_gptr);
        q2free(d_fgptr);
        q2free(d_fgidx);
        q2free(d_desc);
        q2free(d_dtab);
    }
};

namespace {

template <class F> int q2_guarded(fgpu_ctx *ctx, F &&f) {
    try {
        return f();
    } catch (const HipErr &h) {
        if (ctx) ctx->err = std::string(h.what) + ": " + hipGetErrorString(h.e);
        return h.e == hipErrorOutOfMemory ? FGPU_ENOMEM : FGPU_EHIP;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->err = "host allocation failed";
        return FGPU_ENOMEM;
    } catch (const std::exception &ex) {
        if (ctx) ctx->err = ex.what();
        return FGPU_EINVAL;
    }
}

// node tuple of lexicographic node n
inline void tup(int n, int t[3]) {
    t[0] = n % 3;
    t[1] = (n / 3) % 3;
    t[2] = n / 9;
}

// Q2 dof numbering of the internal cells (lexicographic Q1 vertices `icells`, nv vertices):
// vertices keep their numbers; every edge, face and cell interior gets the next number at its
// first appearance in (cell, node) order.  Returns the dof count.
int number_dofs(const std::vector<int> &icells, int nv, int nc, std::vector<int> &c27) {
    c27.assign(27 * (size_t)nc, -1);
    for (int c = 0; c < nc; ++c)
        for (int v = 0; v < 8; ++v) c27[27 * (size_t)c + q2_vnode(v)] = icells[8 * (size_t)c + v];
    int next = nv;
    for (int kind = 1; kind <= 3; ++kind) {
        std::vector<int> nodes;
        for (int n = 0; n < 27; ++n) {
            int t[3];
            tup(n, t);
            if ((t[0] == 1) + (t[1] == 1) + (t[2] == 1) == kind) nodes.push_back(n);
        }
        const int nv_ent = 1 << kind;       // vertices spanned: 2 (edge), 4 (face), 8 (cell)
        const size_t ne = nodes.size() * (size_t)nc;
        std::vector<int> key(ne * nv_ent);
        for (int c = 0; c < nc; ++c)
            for (size_t m = 0; m < nodes.size(); ++m) {
                int t[3];
                tup(nodes[m], t);
                int *k = &key[((size_t)c * nodes.size() + m) * nv_ent];
                int cnt = 0;
                for (int cz = 0; cz < 2; ++cz)
                    for (int cy = 0; cy < 2; ++cy)
                        for (int cx = 0; cx < 2; ++cx) {
                            const int want[3] = {cx, cy, cz};
                            bool ok = true;
                            for (int d = 0; d < 3; ++d)
                                if (t[d] != 1 && want[d] != t[d] / 2) ok = false;
                            if (ok) k[cnt++] = icells[8 * (size_t)c + cx + 2 * cy + 4 * cz];
                        }
                std::sort(k, k + nv_ent);
            }
        std::vector<size_t> order(ne);
        std::iota(order.begin(), order.end(), 0);
        auto less = [&](size_t a, size_t b) {
            const int *ka = &key[a * nv_ent], *kb = &key[b * nv_ent];
            for (int i = 0; i < nv_ent; ++i)
                if (ka[i] != kb[i]) return ka[i] < kb[i];
            return a < b;
        };
        std::sort(order.begin(), order.end(), less);
        // runs of equal keys: the first entry of a run is its first appearance
        std::vector<size_t> first(ne);
        for (size_t i = 0; i < ne;) {
            size_t j = i + 1;
            while (j < ne && std::equal(&key[order[i] * nv_ent], &key[order[i] * nv_ent] + nv_ent, &key[order[j] * nv_ent]))
                ++j;
            for (size_t k = i; k < j; ++k) first[order[k]] = order[i];
            i = j;
        }
        std::vector<int> id(ne, -1);
        for (size_t e = 0; e < ne; ++e) {
            if (first[e] == e) id[e] = next++;
            const int c = (int)(e / nodes.size()), m = (int)(e % nodes.size());
            c27[27 * (size_t)c + nodes[m]] = id[first[e]];
        }
    }
    return next;
}

void q2_upload_tables() {
    static bool done[64] = {};
    int dev = 0;
    Q2CHECK(hipGetDevice(&dev));
    if (dev < 64 && done[dev]) return;
    double B[3][3], D[3][3], W[3], P[27][9];
    q2_tables(B, D, W);
    q2_affine_table(P);
    Q2CHECK(hipMemcpyToSymbol(HIP_SYMBOL(kQB), B, sizeof(B), 0, hipMemcpyHostToDevice));
    Q2CHECK(hipMemcpyToSymbol(HIP_SYMBOL(kQD), D, sizeof(D), 0, hipMemcpyHostToDevice));
    Q2CHECK(hipMemcpyToSymbol(HIP_SYMBOL(kQW), W, sizeof(W), 0, hipMemcpyHostToDevice));
    Q2CHECK(hipMemcpyToSymbol(HIP_SYMBOL(kQ2P), P, sizeof(P), 0, hipMemcpyHostToDevice));
    if (dev < 64) done[dev] = true;
}

Desc q2_desc(fgpu_q2 *s) {
    Desc d{};
    d.n_dofs = s->n_dofs;
    d.n_cells = s->n_cells;
    d.nb_dof = s->nb_dof;
    d.gptr = s->d_gptr;
    d.wide_slots = 1;             // vertex dofs own about two dozen slots each
    d.dmask = s->d_dmask;
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
    d.cd = s->cd;
    d.cres = s->cres;
    d.ccoef = s->ccoef;
    return d;
}

void q2_upload_desc(fgpu_q2 *s, const Desc &d) {
    Q2CHECK(hipMemcpyAsync(s->d_desc, &d, sizeof(Desc), hipMemcpyHostToDevice, s->ctx->stream));
    Q2CHECK(hipStreamSynchronize(s->ctx->stream));
}

// the tet-split metric for the mesh's current coordinates, and whether every cell has the form
// (the mesh's use-affine switch applies to Q2 too); synchronises when the coordinates changed
void q2_ensure_metric(fgpu_q2 *s) {
    fgpu_mesh *m = s->mesh;
    if (s->metric_gen == m->coord_gen) return;
    hipStream_t st = s->ctx->stream;
    const int nc = s->n_cells, nb = (nc + 255) / 256;
    if (!s->d_qm) s->d_qm = q2alloc<double>(6 * (size_t)std::max(nc, 1));
    double *dev = q2alloc<double>(std::max(nb, 1));
    double mx = 0;
    try {
        if (nc > 0 && m->d_layout) {
            hipLaunchKernelGGL(k_q2_metric, dim3(nb), dim3(256), 0, st, nc, s->d_c27, m->d_layout,
                               reinterpret_cast<const double4 *>(m->d_X64), s->d_qm, dev);
            Q2CHECK(hipGetLastError());
            std::vector<double> h(nb);
            Q2CHECK(hipMemcpyAsync(h.data(), dev, nb * sizeof(double), hipMemcpyDeviceToHost, st));
            Q2CHECK(hipStreamSynchronize(st));
            for (double v : h) mx = std::max(mx, v);
        } else {
            mx = 1e300;
        }
    } catch (...) {
        q2free(dev);
        throw;
    }
    q2free(dev);
    s->affine = nc > 0 && mx <= 1e-9;
    s->metric_gen = m->coord_gen;
}

void q2_apply(fgpu_q2 *s, const double *ua, const double *ub, bool maskin, bool active_only) {
    q2_ensure_metric(s);
    const bool aff = s->affine && s->mesh->affine_allowed;
    Q2Args a{s->n_cells, s->d_c27, s->d_spos, reinterpret_cast<const double4 *>(s->mesh->d_X64), ua, ub, s->scal,
             s->d_dmask, s->Ye, maskin ? 1 : 0, active_only ? 1 : 0, s->d_qm, s->mesh->d_layout};
    const int grid = (s->n_cells + 8 * kQ2Cells - 1) / (8 * kQ2Cells) * 8;   // cell blocks, a multiple of 8
    if (grid <= 0) return;
    if (aff) hipLaunchKernelGGL(k_q2_apply<true>, dim3(grid), dim3(256), 0, s->ctx->stream, a);
    else hipLaunchKernelGGL(k_q2_apply<false>, dim3(grid), dim3(256), 0, s->ctx->stream, a);
}

void q2_ensure_diag(fgpu_q2 *s) {
    if (s->diag_valid && s->diag_gen == s->mesh->coord_gen) return;
    hipStream_t st = s->ctx->stream;
    const size_t ne = 27 * (size_t)s->n_cells;
    if (ne > 0)
        hipLaunchKernelGGL(k_q2_diag, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, s->n_cells, s->d_c27, s->d_spos,
                           reinterpret_cast<const double4 *>(s->mesh->d_X64), s->Ye);
    Desc d = q2_desc(s);
    d.q = s->diag;
    q2_upload_desc(s, d);
    launch_gather_plain<double>(st, s->d_desc, s->d_dtab, s->nb_dof, false);
    launch_invdiag(st, s->n_dofs, s->d_dmask, s->diag, s->invd, s->invd32);
    Q2CHECK(hipGetLastError());
    s->diag_valid = true;
    s->diag_gen = s->mesh->coord_gen;
}

}  // namespace

extern "C" int fgpu_q2_create(fgpu_mesh *m, fgpu_q2 **out) {
    if (!m || !out) return FGPU_EINVAL;
    *out = nullptr;
    fgpu_ctx *c = m->ctx;
    fgpu_q2 *s = new (std::nothrow) fgpu_q2();
    if (!s) return FGPU_ENOMEM;
    const int rc = q2_guarded(c, [&] {
        Q2CHECK(hipSetDevice(c->device));
        hipStream_t st = c->stream;
        q2_upload_tables();
        s->ctx = c;
        s->mesh = m;
        const int nc = m->n_cells, nv = m->n_dofs;
        if ((int)m->cell_of_icell.size() != nc || (int)m->cells_lex.size() != 8 * nc)
            throw std::invalid_argument("mesh tables missing");
        std::vector<int> icells(8 * (size_t)nc);
        for (int i = 0; i < nc; ++i)
            for (int k = 0; k < 8; ++k)
                icells[8 * (size_t)i + k] = m->idof_of_dof[m->cells_lex[8 * (size_t)m->cell_of_icell[i] + k]];
        s->n_cells = nc;
        s->nv = nv;
        s->n_dofs = number_dofs(icells, nv, nc, s->c27);
        const int n = s->n_dofs;
        // dof-major slots: each dof's (cell, node) entries in ascending order
        std::vector<int> gptr(n + 1, 0), spos(27 * (size_t)nc);
        for (int d : s->c27) gptr[d + 1]++;
        for (int i = 0; i < n; ++i) gptr[i + 1] += gptr[i];
        {
            std::vector<int> fill(gptr.begin(), gptr.end() - 1);
            for (size_t e = 0; e < s->c27.size(); ++e) spos[e] = fill[s->c27[e]]++;
        }
        // boundary faces: 9 slots per face, dof -> face slots in ascending order
        const int nb = (int)m->bf_cell.size();
        s->nb = nb;
        std::vector<int> fgptr(n + 1, 0), fgidx(9 * (size_t)nb), ficell(nb);
        std::vector<int> fnodes[6];
        for (int f = 0; f < 6; ++f)
            for (int k = 0; k < 27; ++k) {
                int t[3];
                tup(k, t);
                if (t[f / 2] == 2 * (f % 2)) fnodes[f].push_back(k);
            }
        for (int f = 0; f < nb; ++f) {
            ficell[f] = m->icell_of_cell[m->bf_cell[f]];
            for (int k : fnodes[m->bf_face[f]]) fgptr[s->c27[27 * (size_t)ficell[f] + k] + 1]++;
        }
        for (int i = 0; i < n; ++i) fgptr[i + 1] += fgptr[i];
        {
            std::vector<int> fill(fgptr.begin(), fgptr.end() - 1);
            for (int f = 0; f < nb; ++f)
                for (int mm = 0; mm < 9; ++mm) fgidx[fill[s->c27[27 * (size_t)ficell[f] + fnodes[m->bf_face[f]][mm]]]++] = 9 * f + mm;
        }
        // API numbering: the mesh's dof order for vertices, internal order for the rest
        s->api_of_q2.resize(n);
        s->q2_of_api.resize(n);
        for (int i = 0; i < n; ++i) {
            const int api = i < nv ? m->dof_of_idof[i] : i;
            s->api_of_q2[i] = api;
            s->q2_of_api[api] = i;
        }
        s->dmask.assign(n, 0);
        s->gval.assign(n, 0.0);
        s->nb_dof = ((std::max((n + kDofBlock - 1) / kDofBlock, 1) + 7) / 8) * 8;
        // device
        s->d_c27 = q2alloc<int>(s->c27.size());
        s->d_spos = q2alloc<int>(spos.size());
        s->d_gptr = q2alloc<int>(gptr.size());
        s->d_fgptr = q2alloc<int>(fgptr.size());
        s->d_fgidx = q2alloc<int>(fgidx.size());
        Q2CHECK(hipMemcpyAsync(s->d_c27, s->c27.data(), s->c27.size() * 4, hipMemcpyHostToDevice, st));
        Q2CHECK(hipMemcpyAsync(s->d_spos, spos.data(), spos.size() * 4, hipMemcpyHostToDevice, st));
        Q2CHECK(hipMemcpyAsync(s->d_gptr, gptr.data(), gptr.size() * 4, hipMemcpyHostToDevice, st));
        Q2CHECK(hipMemcpyAsync(s->d_fgptr, fgptr.data(), fgptr.size() * 4, hipMemcpyHostToDevice, st));
        if (!fgidx.empty()) Q2CHECK(hipMemcpyAsync(s->d_fgidx, fgidx.data(), fgidx.size() * 4, hipMemcpyHostToDevice, st));
        for (double **v : {&s->b, &s->x, &s->r, &s->z, &s->p, &s->q, &s->invd, &s->diag, &s->g, &s->tmp, &s->cd, &s->cres})
            *v = q2alloc<double>(n);
        s->invd32 = q2alloc<float>(n);
        s->Ye = q2alloc<double>(27 * (size_t)nc);
        s->fslot = q2alloc<double>(9 * (size_t)std::max(nb, 1));
        s->ccoef = q2alloc<double>(2 * kChebMax + 1);
        s->hist = q2alloc<double>(2 * kHist);
        s->scal = q2alloc<double>(kScal);
        s->part = q2alloc<double>(kPartRegions * (size_t)s->nb_dof);
        s->dflag = q2alloc<double>(1);
        s->d_dmask = q2alloc<unsigned char>(n);
        s->d_desc = q2alloc<Desc>(1);
        s->d_dtab = q2alloc<int2>(s->nb_dof);
        std::vector<int2> tab(s->nb_dof);
        const int per = s->nb_dof / 8;
        for (int i = 0; i < s->nb_dof; ++i) tab[i] = make_int2(0, (i % 8) * per + i / 8);   // one slice per XCD
        Q2CHECK(hipMemcpyAsync(s->d_dtab, tab.data(), tab.size() * sizeof(int2), hipMemcpyHostToDevice, st));
        Q2CHECK(hipMemsetAsync(s->d_dmask, 0, n, st));
        launch_set_const(st, n, 0.0, s->b);
        launch_set_const(st, n, 0.0, s->x);
        launch_set_const(st, n, 0.0, s->g);
        Q2CHECK(hipGetLastError());
        Q2CHECK(hipStreamSynchronize(st));
        m->refs++;
        c->refs++;
        return FGPU_OK;
    });
    if (rc != FGPU_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return FGPU_OK;
}

void fgpu_ctx_release_ref(fgpu_ctx *c);
void fgpu_mesh_release_ref(fgpu_mesh *m);

extern "C" void fgpu_q2_destroy(fgpu_q2 *s) {
    if (!s) return;
    fgpu_ctx *c = s->ctx;
    fgpu_mesh *m = s->mesh;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete s;
    fgpu_mesh_release_ref(m);
    fgpu_ctx_release_ref(c);
}

extern "C" int fgpu_q2_sizes(const fgpu_q2 *s, int *n_dofs, int *n_cells) {
    if (!s) return FGPU_EINVAL;
    if (n_dofs) *n_dofs = s->n_dofs;
    if (n_cells) *n_cells = s->n_cells;
    return FGPU_OK;
}

extern "C" int fgpu_q2_export_dofs(const fgpu_q2 *s, double *xyz, int *cells27) {
    if (!s) return FGPU_EINVAL;
    const fgpu_mesh *m = s->mesh;
    if (xyz) {
        // support points: vertices, then the mean of the entity's vertices
        std::vector<int> done(s->n_dofs, 0);
        for (int i = 0; i < s->n_cells; ++i) {
            const int *cd = &s->c27[27 * (size_t)i];
            for (int n = 0; n < 27; ++n) {
                const int d = cd[n];
                if (done[d]) continue;
                done[d] = 1;
                int t[3];
                tup(n, t);
                double acc[3] = {0, 0, 0};
                int cnt = 0;
                for (int cz = 0; cz < 2; ++cz)
                    for (int cy = 0; cy < 2; ++cy)
                        for (int cx = 0; cx < 2; ++cx) {
                            const int w[3] = {cx, cy, cz};
                            bool ok = true;
                            for (int k = 0; k < 3; ++k)
                                if (t[k] != 1 && w[k] != t[k] / 2) ok = false;
                            if (!ok) continue;
                            const int v = m->dof_of_idof[cd[q2_vnode(cx + 2 * cy + 4 * cz)]];
                            for (int k = 0; k < 3; ++k) acc[k] += m->x[3 * (size_t)v + k];
                            ++cnt;
                        }
                const int api = s->api_of_q2[d];
                for (int k = 0; k < 3; ++k) xyz[3 * (size_t)api + k] = acc[k] / cnt;
            }
        }
    }
    if (cells27)
        for (int cl = 0; cl < s->n_cells; ++cl) {
            const int i = m->icell_of_cell[cl];
            for (int n = 0; n < 27; ++n) cells27[27 * (size_t)cl + n] = s->api_of_q2[s->c27[27 * (size_t)i + n]];
        }
    return FGPU_OK;
}

extern "C" int fgpu_q2_reset(fgpu_q2 *s) {
    if (!s) return FGPU_EINVAL;
    return q2_guarded(s->ctx, [&] {
        Q2CHECK(hipSetDevice(s->ctx->device));
        launch_set_const(s->ctx->stream, s->n_dofs, 0.0, s->b);
        std::fill(s->dmask.begin(), s->dmask.end(), 0);
        std::fill(s->gval.begin(), s->gval.end(), 0.0);
        s->lifted = false;
        Q2CHECK(hipStreamSynchronize(s->ctx->stream));
        return FGPU_OK;
    });
}

extern "C" int fgpu_q2_add_dirichlet(fgpu_q2 *s, int bid, double value) {
    if (!s) return FGPU_EINVAL;
    const fgpu_mesh *m = s->mesh;
    for (size_t f = 0; f < m->bf_cell.size(); ++f) {
        if (m->bf_id[f] != bid) continue;
        const int i = m->icell_of_cell[m->bf_cell[f]], fc = m->bf_face[f];
        for (int k = 0; k < 27; ++k) {
            int t[3];
            tup(k, t);
            if (t[fc / 2] != 2 * (fc % 2)) continue;
            const int d = s->c27[27 * (size_t)i + k];
            s->dmask[d] = 1;
            s->gval[d] = value;
        }
    }
    return FGPU_OK;
}

extern "C" int fgpu_q2_add_neumann(fgpu_q2 *s, int bid, double value) {
    if (!s) return FGPU_EINVAL;
    return q2_guarded(s->ctx, [&] {
        Q2CHECK(hipSetDevice(s->ctx->device));
        if (s->lifted) throw std::invalid_argument("Neumann loads are added before fgpu_q2_apply_dirichlet");
        hipStream_t st = s->ctx->stream;
        const fgpu_mesh *m = s->mesh;
        if (s->nb > 0) {
            hipLaunchKernelGGL(k_q2_face, dim3((s->nb + 255) / 256), dim3(256), 0, st, s->nb, m->d_bf_icell, m->d_bf_face,
                               m->d_bf_id, bid, value, s->d_c27, reinterpret_cast<const double4 *>(m->d_X64), s->fslot);
            launch_face_gather(st, s->n_dofs, s->d_fgptr, s->d_fgidx, s->fslot, s->b);
        }
        Q2CHECK(hipGetLastError());
        Q2CHECK(hipStreamSynchronize(st));
        return FGPU_OK;
    });
}

// the Dirichlet lifting as the Q1 path: b_I -= (A g)_I, b_D = diag_D g_D, x_D = g_D
extern "C" int fgpu_q2_apply_dirichlet(fgpu_q2 *s) {
    if (!s) return FGPU_EINVAL;
    return q2_guarded(s->ctx, [&] {
        Q2CHECK(hipSetDevice(s->ctx->device));
        hipStream_t st = s->ctx->stream;
        const int n = s->n_dofs;
        bool g_zero = true;
        for (int i = 0; i < n; ++i) g_zero = g_zero && s->gval[i] == 0.0 && !std::signbit(s->gval[i]);
        Q2CHECK(hipMemcpyAsync(s->d_dmask, s->dmask.data(), n, hipMemcpyHostToDevice, st));
        Q2CHECK(hipMemcpyAsync(s->g, s->gval.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
        s->diag_valid = false;
        s->cheb_lmax = 0;
        q2_ensure_diag(s);
        if (g_zero) {
            launch_lift(st, n, s->d_dmask, s->g, nullptr, s->diag, s->b, s->x);
        } else {
            q2_apply(s, s->g, nullptr, false, false);
            Desc d = q2_desc(s);
            d.q = s->tmp;
            q2_upload_desc(s, d);
            launch_gather_plain<double>(st, s->d_desc, s->d_dtab, s->nb_dof, false);
            launch_lift(st, n, s->d_dmask, s->g, s->tmp, s->diag, s->b, s->x);
        }
        Q2CHECK(hipGetLastError());
        Q2CHECK(hipStreamSynchronize(st));
        s->lifted = true;
        return FGPU_OK;
    });
}

extern "C" int fgpu_q2_fill_solution(fgpu_q2 *s, double v) {
    if (!s) return FGPU_EINVAL;
    return q2_guarded(s->ctx, [&] {
        Q2CHECK(hipSetDevice(s->ctx->device));
        launch_set_const(s->ctx->stream, s->n_dofs, v, s->x);
        Q2CHECK(hipStreamSynchronize(s->ctx->stream));
        return FGPU_OK;
    });
}

static int q2_get(fgpu_q2 *s, const double *dev, double *out) {
    return q2_guarded(s->ctx, [&] {
        Q2CHECK(hipSetDevice(s->ctx->device));
        std::vector<double> h(s->n_dofs);
        Q2CHECK(hipMemcpyAsync(h.data(), dev, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream));
        Q2CHECK(hipStreamSynchronize(s->ctx->stream));
        for (int i = 0; i < s->n_dofs; ++i) out[s->api_of_q2[i]] = h[i];
        return FGPU_OK;
    });
}

extern "C" int fgpu_q2_get_solution(fgpu_q2 *s, double *x) { return (!s || !x) ? FGPU_EINVAL : q2_get(s, s->x, x); }
extern "C" int fgpu_q2_get_rhs(fgpu_q2 *s, double *b) { return (!s || !b) ? FGPU_EINVAL : q2_get(s, s->b, b); }

// y = A x with the un-lifted operator (API order)
extern "C" int fgpu_q2_apply(fgpu_q2 *s, const double *x, double *y) {
    if (!s || !x || !y) return FGPU_EINVAL;
    return q2_guarded(s->ctx, [&] {
        Q2CHECK(hipSetDevice(s->ctx->device));
        hipStream_t st = s->ctx->stream;
        std::vector<double> h(s->n_dofs);
        for (int i = 0; i < s->n_dofs; ++i) h[i] = x[s->api_of_q2[i]];
        Q2CHECK(hipMemcpyAsync(s->tmp, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st));
        q2_apply(s, s->tmp, nullptr, false, false);
        Desc d = q2_desc(s);
        d.q = s->q;
        q2_upload_desc(s, d);
        launch_gather_plain<double>(st, s->d_desc, s->d_dtab, s->nb_dof, false);
        Q2CHECK(hipGetLastError());
        return q2_get(s, s->q, y);
    });
}

// average duration of `reps` element applies of the CG (input z + beta p), HIP events on the
// context's stream
extern "C" int fgpu_q2_time_apply(fgpu_q2 *s, int reps, double *avg_ms) {
    if (!s || reps <= 0 || !avg_ms) return FGPU_EINVAL;
    return q2_guarded(s->ctx, [&] {
        Q2CHECK(hipSetDevice(s->ctx->device));
        hipStream_t st = s->ctx->stream;
        q2_apply(s, s->z, s->p, false, false);
        hipEvent_t e0, e1;
        Q2CHECK(hipEventCreate(&e0));
        Q2CHECK(hipEventCreate(&e1));
        Q2CHECK(hipEventRecord(e0, st));
        for (int i = 0; i < reps; ++i) q2_apply(s, s->z, s->p, false, false);
        Q2CHECK(hipEventRecord(e1, st));
        Q2CHECK(hipEventSynchronize(e1));
        float ms = 0;
        Q2CHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *avg_ms = ms / reps;
        return FGPU_OK;
    });
}

namespace {

// FP64 CG with the given preconditioner (Jacobi, none, or Chebyshev of degree deg), from the
// present solution; hist records the first coefficients (Lanczos)
void q2_cg(fgpu_q2 *s, const fgpu_solve_opts *o, int pc, int deg, bool hist, int max_it, fgpu_solve_stats *stats) {
    hipStream_t st = s->ctx->stream;
    const int n = s->n_dofs;
    Desc d = q2_desc(s);
    d.pc = pc;
    d.cheb_deg = deg;
    d.hist = hist ? s->hist : nullptr;
    q2_upload_desc(s, d);
    const Desc *D = s->d_desc;
    std::vector<double> h(kScal, 0.0);
    h[S_ACTIVE] = 1.0;
    h[S_MAXIT] = max_it;
    h[S_TOL] = o->tol;
    h[S_REL] = o->relative ? 1.0 : 0.0;
    Q2CHECK(hipMemcpyAsync(s->scal, h.data(), kScal * sizeof(double), hipMemcpyHostToDevice, st));
    launch_enforce_dirichlet(st, n, s->d_dmask, s->g, s->x);
    auto cheb = [&] {
        for (int j = 1; j < deg; ++j) {
            q2_apply(s, s->cd, nullptr, false, true);
            launch_gather_cheb(st, D, s->d_dtab, s->nb_dof, j);
        }
    };
    // residual of the lifted system and the first preconditioned residual
    q2_apply(s, s->x, nullptr, true, true);
    launch_gather_res<double>(st, D, s->d_dtab, s->nb_dof);
    if (pc == FGPU_PC_CHEBYSHEV) cheb();
    launch_scal_res(st, D, D, 1, 0);
    const int k = std::max(2, o->check_every > 0 ? o->check_every : 32);
    double done = 0;
    for (int ch = 0; ch < max_it / k + 2; ++ch) {
        launch_collect(st, D, 1, S_DONE, s->dflag);
        Q2CHECK(hipMemcpyAsync(&done, s->dflag, sizeof(double), hipMemcpyDeviceToHost, st));
        Q2CHECK(hipStreamSynchronize(st));
        if (done != 0.0) break;
        for (int it = 0; it < k; ++it) {
            q2_apply(s, s->z, s->p, false, true);
            launch_gather_cg<double>(st, D, s->d_dtab, s->nb_dof);
            launch_scal_alpha(st, D, 1);
            launch_update_cg<double>(st, D, s->d_dtab, s->nb_dof);
            if (pc == FGPU_PC_CHEBYSHEV) cheb();
            launch_scal_beta(st, D, 1, 0);
        }
        Q2CHECK(hipGetLastError());
    }
    double hs[kScal];
    Q2CHECK(hipMemcpy(hs, s->scal, sizeof(hs), hipMemcpyDeviceToHost));
    const int it = (int)hs[S_IT];
    stats->iterations = hs[S_CONV] != 0.0 ? it : -std::max(it, 1);
    stats->outer = 1;
    stats->residual = std::sqrt(std::max(0.0, hs[S_RR]));
    stats->rhs_norm = std::sqrt(std::max(0.0, hs[S_BB]));
}

}  // namespace

// DealSolver::solve_cg (src/DealSolver.cpp:443-458) for the Q2 system: FP64 CG, preconditioner
// FGPU_PC_NONE, FGPU_PC_JACOBI or FGPU_PC_CHEBYSHEV (Lanczos estimate from the first Jacobi
// steps of the solve, kept while the system is unchanged)
extern "C" int fgpu_q2_solve(fgpu_q2 *s, const fgpu_solve_opts *o, fgpu_solve_stats *stats) {
    if (!s || !o) return FGPU_EINVAL;
    return q2_guarded(s->ctx, [&] {
        Q2CHECK(hipSetDevice(s->ctx->device));
        if (o->precision != FGPU_FP64) throw std::invalid_argument("the Q2 solve runs in FGPU_FP64");
        if (o->precond < FGPU_PC_NONE || o->precond > FGPU_PC_CHEBYSHEV) throw std::invalid_argument("unknown preconditioner");
        if (!s->lifted) throw std::invalid_argument("apply the Dirichlet conditions before solving");
        q2_ensure_diag(s);
        fgpu_solve_stats st{}, pre{};
        const int max_it = std::max(0, o->max_it);
        if (o->precond != FGPU_PC_CHEBYSHEV) {
            q2_cg(s, o, o->precond, 0, false, max_it, &st);
        } else {
            int used = 0;
            if (!(s->cheb_lmax > 0)) {
                q2_cg(s, o, FGPU_PC_JACOBI, 0, true, std::min(max_it, kLanczosSteps), &pre);
                used = std::abs(pre.iterations);
                if (pre.iterations > 0 || used >= max_it) {
                    if (stats) *stats = pre;
                    return FGPU_OK;
                }
                double h[2 * kHist];
                Q2CHECK(hipMemcpy(h, s->hist, sizeof(h), hipMemcpyDeviceToHost));
                const double lm = lanczos_lmax(h, used);
                if (!(lm > 0) || !std::isfinite(lm)) throw std::runtime_error("Lanczos estimate of lambda_max failed");
                s->cheb_lmax = lm;
            }
            const int deg = o->cheb_degree > 0 ? std::min(o->cheb_degree, kChebMax) : cheb_auto_degree(s->n_dofs);
            std::vector<double> cf;
            cheb_coefficients(s->cheb_lmax, deg, cf);
            Q2CHECK(hipMemcpy(s->ccoef, cf.data(), cf.size() * sizeof(double), hipMemcpyHostToDevice));
            q2_cg(s, o, FGPU_PC_CHEBYSHEV, deg, false, std::max(1, max_it - used), &st);
            st.iterations = st.iterations > 0 ? st.iterations + used : st.iterations - used;
        }
        if (stats) *stats = st;
        return FGPU_OK;
    });
}

extern "C" int fgpu_q2_uses_tet_split(fgpu_q2 *s, int *on) {
    if (!s || !on) return FGPU_EINVAL;
    return q2_guarded(s->ctx, [&] {
        Q2CHECK(hipSetDevice(s->ctx->device));
        q2_ensure_metric(s);
        *on = (s->affine && s->mesh->affine_allowed) ? 1 : 0;
        return FGPU_OK;
    });
}

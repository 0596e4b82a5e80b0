/* CPU reference timing (test and benchmark infrastructure, never part of the product path):
 * the oracle's preconditioned CG (orc_pcg in oracle_fem.c, which follows SolverCG with
 * PreconditionSSOR, src/DealSolver.cpp:443-458) with the sparse matrix-vector product and the
 * vector operations spread over OpenMP threads.  The SSOR sweeps stay sequential, as deal.II's
 * PreconditionSSOR runs them.  Only the summation order of the dot products differs from
 * orc_pcg, so iteration counts may differ by rounding; the tests use orc_pcg. */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void spmv_omp(const orc_csr *A, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < A->n; ++i) {
        double s = 0;
        for (int64_t p = A->rowptr[i]; p < A->rowptr[i + 1]; ++p) s += A->val[p] * x[A->col[p]];
        y[i] = s;
    }
}

static double dot_omp(int n, const double *a, const double *b) {
    double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* z = M^{-1} r: identity, Jacobi, or SSOR(w) with sequential forward and backward sweeps */
static void precondition_omp(const orc_csr *A, int kind, double w, const double *r, double *z) {
    const int n = A->n;
    if (kind == 0) {
        memcpy(z, r, sizeof(double) * n);
    } else if (kind == 1) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) z[i] = r[i] / A->val[A->diag[i]];
    } else {
        for (int i = 0; i < n; ++i) {
            double s = r[i];
            for (int64_t p = A->rowptr[i]; p < A->diag[i]; ++p) s -= w * A->val[p] * z[A->col[p]];
            z[i] = s / A->val[A->diag[i]];
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) z[i] *= A->val[A->diag[i]];
        for (int i = n - 1; i >= 0; --i) {
            double s = z[i];
            for (int64_t p = A->diag[i] + 1; p < A->rowptr[i + 1]; ++p) s -= w * A->val[p] * z[A->col[p]];
            z[i] = s / A->val[A->diag[i]];
        }
        const double sc = w * (2.0 - w);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) z[i] *= sc;
    }
}

int orc_pcg_omp(const orc_csr *A, const double *b, double *x, int max_it, double tol, int relative, int precond,
                double omega, int nthreads, double *res_out) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const int n = A->n;
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    double *p = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * n);
    spmv_omp(A, x, q);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
    const double thr = relative ? tol * sqrt(dot_omp(n, b, b)) : tol;
    double res = sqrt(dot_omp(n, r, r));
    int it = 0, ok = res <= thr;
    if (!ok) {
        precondition_omp(A, precond, omega, r, z);
        memcpy(p, z, sizeof(double) * n);
        double rz = dot_omp(n, r, z);
        while (it < max_it) {
            ++it;
            spmv_omp(A, p, q);
            const double alpha = rz / dot_omp(n, p, q);
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) {
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
            }
            res = sqrt(dot_omp(n, r, r));
            if (res <= thr) {
                ok = 1;
                break;
            }
            precondition_omp(A, precond, omega, r, z);
            const double rz1 = dot_omp(n, r, z);
            const double beta = rz1 / rz;
            rz = rz1;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        }
    }
    free(r);
    free(z);
    free(p);
    free(q);
    if (res_out) *res_out = res;
    return ok ? it : -it;
}

int orc_omp_max_threads(void) { return omp_get_max_threads(); }

/*
 * CPU oracle, finite-element part.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates the reference's deal.II-based path:
 *   mesh import + boundary marking   src/DealSolver.cpp:192-209, 461-518
 *   Q1 / QGauss(2) cell matrices     src/PoissonSolver.cpp:239-263, src/CurrentHeatSolver.cpp:347-393
 *   Neumann face load                src/DealSolver.cpp:390-430
 *   Dirichlet lifting                src/DealSolver.cpp:433-440 (deal.II apply_boundary_values)
 *   CG + SSOR / Jacobi               src/DealSolver.cpp:443-458 (deal.II SolverCG, PreconditionSSOR)
 *   nodal gradient export            src/DealSolver.cpp:281-301
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* UCD (femocs) -> deal.II lexicographic: lex[k] = ucd[UCD2LEX[k]].  deal.II's ucd_to_deal
 * table {0,1,5,4,2,3,7,6} maps the other way; LinearHexahedra::shape_funs_dealii
 * (src/InterpolatorCells.cpp:1355-1358) uses this direction. */
static const int UCD2LEX[8] = {0, 1, 4, 5, 3, 2, 7, 6};
/* vertices of deal.II face f (xi_{f/2} = f%2), lexicographic */
static const int FACE_VERTS[6][4] = {
    {0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};

/* ------------------------------------------------------------------------------------- */
/* Q1 element kernel                                                                      */
/* ------------------------------------------------------------------------------------- */
static double gauss_pt(int i) {
    /* QGauss<1>(2) on [0,1] */
    const double s = 0.5 / sqrt(3.0);
    return i ? 0.5 + s : 0.5 - s;
}

/* shape values and reference gradients of lexicographic Q1 node n at xi */
static void q1_shape(const double xi[3], double phi[8], double dphi[8][3]) {
    for (int n = 0; n < 8; ++n) {
        const int b[3] = {n & 1, (n >> 1) & 1, (n >> 2) & 1};
        double l[3], dl[3];
        for (int d = 0; d < 3; ++d) {
            l[d] = b[d] ? xi[d] : 1.0 - xi[d];
            dl[d] = b[d] ? 1.0 : -1.0;
        }
        phi[n] = l[0] * l[1] * l[2];
        dphi[n][0] = dl[0] * l[1] * l[2];
        dphi[n][1] = l[0] * dl[1] * l[2];
        dphi[n][2] = l[0] * l[1] * dl[2];
    }
}

/* J[a][b] = dx_a / dxi_b; returns det, fills inverse Jinv[b][a] = dxi_b/dx_a */
static double jacobian(const double xe[8][3], double dphi[8][3], double J[3][3], double Ji[3][3]) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double s = 0;
            for (int n = 0; n < 8; ++n) s += xe[n][a] * dphi[n][b];
            J[a][b] = s;
        }
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
                     - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                     + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const double id = 1.0 / det;
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    return det;
}

/* physical gradients at quadrature point q (lexicographic, x fastest) */
static double qpoint(const double xe[8][3], int q, double phi[8], double grad[8][3]) {
    const double xi[3] = {gauss_pt(q & 1), gauss_pt((q >> 1) & 1), gauss_pt((q >> 2) & 1)};
    double dphi[8][3], J[3][3], Ji[3][3];
    q1_shape(xi, phi, dphi);
    const double det = jacobian(xe, dphi, J, Ji);
    for (int n = 0; n < 8; ++n)
        for (int a = 0; a < 3; ++a)
            grad[n][a] = Ji[0][a] * dphi[n][0] + Ji[1][a] * dphi[n][1] + Ji[2][a] * dphi[n][2];
    return det * 0.125; /* JxW, QGauss<3>(2) weight 1/8 */
}

void orc_element_matrix(const double xe[8][3], const double kappa_q[8], const double gamma_q[8],
                        double Ke[8][8]) {
    memset(Ke, 0, sizeof(double) * 64);
    for (int q = 0; q < 8; ++q) {
        double phi[8], g[8][3];
        const double jxw = qpoint(xe, q, phi, g);
        const double k = kappa_q ? kappa_q[q] : 1.0;
        const double m = gamma_q ? gamma_q[q] : 0.0;
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j)
                Ke[i][j] += jxw * (k * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2])
                                   + m * phi[i] * phi[j]);
    }
}

static void cell_coords(const orc_mesh *m, int c, double xe[8][3]) {
    for (int n = 0; n < 8; ++n) {
        const int v = m->cells[8 * c + n];
        xe[n][0] = m->x[3 * v];
        xe[n][1] = m->x[3 * v + 1];
        xe[n][2] = m->x[3 * v + 2];
    }
}

/* ------------------------------------------------------------------------------------- */
/* mesh import                                                                            */
/* ------------------------------------------------------------------------------------- */
typedef struct {
    int v[4];
    int cell, face;
} face_rec;

static int cmp_face(const void *a, const void *b) {
    const face_rec *x = a, *y = b;
    for (int i = 0; i < 4; ++i)
        if (x->v[i] != y->v[i]) return x->v[i] < y->v[i] ? -1 : 1;
    if (x->cell != y->cell) return x->cell < y->cell ? -1 : 1;
    return x->face - y->face;
}

static void sort4(int v[4]) {
    for (int i = 1; i < 4; ++i)
        for (int j = i; j > 0 && v[j - 1] > v[j]; --j) {
            int t = v[j];
            v[j] = v[j - 1];
            v[j - 1] = t;
        }
}

static int on_bd2(double v, double b1, double b2, double eps) {
    return fabs(v - b1) <= eps || fabs(v - b2) <= eps;
}

int orc_mesh_import(orc_mesh *m, int nv, const double *xyz, int nc, const int *ucd, int top,
                    int bottom, int sides, int other) {
    memset(m, 0, sizeof(*m));
    int *dof_of = malloc(sizeof(int) * (nv > 0 ? nv : 1));
    for (int i = 0; i < nv; ++i) dof_of[i] = -1;
    for (int64_t i = 0; i < 8LL * nc; ++i) {
        if (ucd[i] < 0 || ucd[i] >= nv) {
            free(dof_of);
            return -1;
        }
        dof_of[ucd[i]] = 1;
    }
    int nd = 0;
    for (int i = 0; i < nv; ++i)
        if (dof_of[i] > 0) dof_of[i] = nd++;
    m->n_dofs = nd;
    m->n_cells = nc;
    m->x = malloc(sizeof(double) * 3 * (nd > 0 ? nd : 1));
    m->vertex_of_dof = malloc(sizeof(int) * (nd > 0 ? nd : 1));
    for (int i = 0; i < nv; ++i)
        if (dof_of[i] >= 0) {
            const int d = dof_of[i];
            m->vertex_of_dof[d] = i;
            m->x[3 * d] = xyz[3 * i];
            m->x[3 * d + 1] = xyz[3 * i + 1];
            m->x[3 * d + 2] = xyz[3 * i + 2];
        }
    m->cells = malloc(sizeof(int) * 8 * (nc > 0 ? nc : 1));
    for (int c = 0; c < nc; ++c) {
        int u[8];
        for (int k = 0; k < 8; ++k) u[k] = dof_of[ucd[8 * c + k]];
        int *l = m->cells + 8 * c;
        for (int k = 0; k < 8; ++k) l[k] = u[UCD2LEX[k]];
        /* GridReordering::invert_all_cells_of_negative_grid: swap UCD i <-> i+4 of cells with
         * negative measure (eta flip in lexicographic order) */
        double xe[8][3], phi[8], dphi[8][3], J[3][3], Ji[3][3];
        cell_coords(m, c, xe);
        const double xi[3] = {0.5, 0.5, 0.5};
        q1_shape(xi, phi, dphi);
        if (jacobian(xe, dphi, J, Ji) < 0) {
            int t;
            t = l[0]; l[0] = l[2]; l[2] = t;
            t = l[1]; l[1] = l[3]; l[3] = t;
            t = l[4]; l[4] = l[6]; l[6] = t;
            t = l[5]; l[5] = l[7]; l[7] = t;
        }
    }
    free(dof_of);

    /* boundary faces: faces owned by exactly one cell */
    const int64_t nf = 6LL * nc;
    face_rec *f = malloc(sizeof(face_rec) * (nf > 0 ? nf : 1));
    for (int c = 0; c < nc; ++c)
        for (int k = 0; k < 6; ++k) {
            face_rec *r = f + 6LL * c + k;
            for (int i = 0; i < 4; ++i) r->v[i] = m->cells[8 * c + FACE_VERTS[k][i]];
            sort4(r->v);
            r->cell = c;
            r->face = k;
        }
    qsort(f, nf, sizeof(face_rec), cmp_face);
    unsigned char *is_bd = calloc(nf > 0 ? nf : 1, 1);
    for (int64_t i = 0; i < nf;) {
        int64_t j = i + 1;
        while (j < nf && memcmp(f[j].v, f[i].v, sizeof(int) * 4) == 0) ++j;
        if (j - i == 1) is_bd[6LL * f[i].cell + f[i].face] = 1;
        i = j;
    }
    free(f);
    int nb = 0;
    for (int64_t i = 0; i < nf; ++i) nb += is_bd[i];
    m->n_bfaces = nb;
    m->bface_cell = malloc(sizeof(int) * (nb > 0 ? nb : 1));
    m->bface_face = malloc(sizeof(int) * (nb > 0 ? nb : 1));
    m->bface_id = malloc(sizeof(int) * (nb > 0 ? nb : 1));
    double *ctr = malloc(sizeof(double) * 3 * (nb > 0 ? nb : 1));
    double lo[3] = {1e16, 1e16, 1e16}, hi[3] = {-1e16, -1e16, -1e16};
    int k = 0;
    for (int c = 0; c < nc; ++c)
        for (int fc = 0; fc < 6; ++fc)
            if (is_bd[6LL * c + fc]) {
                m->bface_cell[k] = c;
                m->bface_face[k] = fc;
                for (int d = 0; d < 3; ++d) {
                    double s = 0;
                    for (int i = 0; i < 4; ++i) s += m->x[3 * m->cells[8 * c + FACE_VERTS[fc][i]] + d];
                    ctr[3 * k + d] = s / 4.0;
                    if (ctr[3 * k + d] < lo[d]) lo[d] = ctr[3 * k + d];
                    if (ctr[3 * k + d] > hi[d]) hi[d] = ctr[3 * k + d];
                }
                ++k;
            }
    const double eps = 1e-6;
    for (int i = 0; i < nb; ++i) {
        const double *p = ctr + 3 * i;
        int id;
        if (on_bd2(p[0], lo[0], hi[0], eps) || on_bd2(p[1], lo[1], hi[1], eps)) id = sides;
        else if (fabs(p[2] - hi[2]) <= eps) id = top;
        else if (fabs(p[2] - lo[2]) <= eps) id = bottom;
        else id = other;
        m->bface_id[i] = id;
    }
    free(ctr);
    free(is_bd);
    return 0;
}

void orc_mesh_free(orc_mesh *m) {
    free(m->x);
    free(m->vertex_of_dof);
    free(m->cells);
    free(m->bface_cell);
    free(m->bface_face);
    free(m->bface_id);
    memset(m, 0, sizeof(*m));
}

/* ------------------------------------------------------------------------------------- */
/* CSR                                                                                    */
/* ------------------------------------------------------------------------------------- */
static int cmp_int(const void *a, const void *b) {
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

int orc_csr_pattern(orc_csr *A, const orc_mesh *m) {
    const int n = m->n_dofs, nc = m->n_cells;
    memset(A, 0, sizeof(*A));
    int *cnt = calloc(n + 1, sizeof(int));
    for (int64_t i = 0; i < 8LL * nc; ++i) cnt[m->cells[i] + 1]++;
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    int *v2c = malloc(sizeof(int) * 8 * (nc > 0 ? nc : 1));
    int *pos = malloc(sizeof(int) * (n + 1));
    memcpy(pos, cnt, sizeof(int) * (n + 1));
    for (int c = 0; c < nc; ++c)
        for (int k = 0; k < 8; ++k) v2c[pos[m->cells[8 * c + k]]++] = c;
    free(pos);
    A->n = n;
    A->rowptr = malloc(sizeof(int64_t) * (n + 1));
    A->diag = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    int64_t cap = 32LL * n + 64, nnz = 0;
    A->col = malloc(sizeof(int) * cap);
    int *tmp = malloc(sizeof(int) * 8 * 512);
    A->rowptr[0] = 0;
    for (int i = 0; i < n; ++i) {
        int t = 0;
        for (int e = cnt[i]; e < cnt[i + 1]; ++e) {
            const int c = v2c[e];
            for (int k = 0; k < 8; ++k) tmp[t++] = m->cells[8 * c + k];
            if (t > 8 * 504) break;
        }
        qsort(tmp, t, sizeof(int), cmp_int);
        int u = 0;
        for (int j = 0; j < t; ++j)
            if (j == 0 || tmp[j] != tmp[j - 1]) tmp[u++] = tmp[j];
        if (nnz + u > cap) {
            cap = 2 * (nnz + u);
            A->col = realloc(A->col, sizeof(int) * cap);
        }
        for (int j = 0; j < u; ++j) {
            if (tmp[j] == i) A->diag[i] = nnz + j;
            A->col[nnz + j] = tmp[j];
        }
        nnz += u;
        A->rowptr[i + 1] = nnz;
    }
    free(tmp);
    free(v2c);
    free(cnt);
    A->nnz = nnz;
    A->val = calloc(nnz > 0 ? nnz : 1, sizeof(double));
    return 0;
}

void orc_csr_free(orc_csr *A) {
    free(A->rowptr);
    free(A->col);
    free(A->val);
    free(A->diag);
    memset(A, 0, sizeof(*A));
}

static int64_t csr_find(const orc_csr *A, int i, int j) {
    int64_t lo = A->rowptr[i], hi = A->rowptr[i + 1] - 1;
    while (lo <= hi) {
        const int64_t mid = (lo + hi) / 2;
        if (A->col[mid] == j) return mid;
        if (A->col[mid] < j) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

void orc_csr_spmv(const orc_csr *A, const double *x, double *y) {
    for (int i = 0; i < A->n; ++i) {
        double s = 0;
        for (int64_t p = A->rowptr[i]; p < A->rowptr[i + 1]; ++p) s += A->val[p] * x[A->col[p]];
        y[i] = s;
    }
}

void orc_assemble(orc_csr *A, const orc_mesh *m, const double *kappa_q, const double *gamma_q) {
    memset(A->val, 0, sizeof(double) * A->nnz);
    for (int c = 0; c < m->n_cells; ++c) {
        double xe[8][3], Ke[8][8];
        cell_coords(m, c, xe);
        orc_element_matrix(xe, kappa_q ? kappa_q + 8 * c : NULL, gamma_q ? gamma_q + 8 * c : NULL, Ke);
        const int *l = m->cells + 8 * c;
        /* DealSolver::copy_global_cell, src/DealSolver.cpp:64-72 */
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j) A->val[csr_find(A, l[i], l[j])] += Ke[i][j];
    }
}

void orc_assemble_cell_rhs(double *b, const orc_mesh *m, const double *f_q) {
    for (int c = 0; c < m->n_cells; ++c) {
        double xe[8][3], be[8] = {0};
        cell_coords(m, c, xe);
        for (int q = 0; q < 8; ++q) {
            double phi[8], g[8][3];
            const double jxw = qpoint(xe, q, phi, g);
            for (int i = 0; i < 8; ++i) be[i] += jxw * phi[i] * f_q[8 * c + q];
        }
        for (int i = 0; i < 8; ++i) b[m->cells[8 * c + i]] += be[i];
    }
}

/* DealSolver::calc_dof_volumes (src/DealSolver.cpp:344-366): every quadrature weight JxW(q)
 * of a cell is added to each of the cell's eight dofs, unweighted by the shape functions */
void orc_dof_volumes(const orc_mesh *m, double *vol) {
    for (int i = 0; i < m->n_dofs; ++i) vol[i] = 0.0;
    for (int c = 0; c < m->n_cells; ++c) {
        double xe[8][3];
        cell_coords(m, c, xe);
        for (int q = 0; q < 8; ++q) {
            double phi[8], g[8][3];
            const double jxw = qpoint(xe, q, phi, g);
            for (int i = 0; i < 8; ++i) vol[m->cells[8 * c + i]] += jxw;
        }
    }
}

/* PoissonSolver::assemble, write-step branch (src/PoissonSolver.cpp:196-207): the right-hand
 * side before the Dirichlet conditions, divided by the dof volumes */
void orc_charge_density(const orc_mesh *m, const double *b, double *rho) {
    orc_dof_volumes(m, rho);
    for (int i = 0; i < m->n_dofs; ++i) rho[i] = b[i] / rho[i];
}

void orc_assemble_face_rhs(double *b, const orc_mesh *m, int bid, const double *fv, double cst) {
    int k = 0;
    for (int i = 0; i < m->n_bfaces; ++i) {
        if (m->bface_id[i] != bid) continue;
        const double g = fv ? fv[k] : cst;
        ++k;
        const int c = m->bface_cell[i], f = m->bface_face[i];
        const int d = f / 2, s = f % 2;
        const int a0 = d == 0 ? 1 : 0, a1 = d == 2 ? 1 : 2;
        double xe[8][3], be[8] = {0};
        cell_coords(m, c, xe);
        for (int q = 0; q < 4; ++q) {
            double xi[3];
            xi[d] = (double)s;
            xi[a0] = gauss_pt(q & 1);
            xi[a1] = gauss_pt(q >> 1);
            double phi[8], dphi[8][3], t0[3] = {0, 0, 0}, t1[3] = {0, 0, 0};
            q1_shape(xi, phi, dphi);
            for (int n = 0; n < 8; ++n)
                for (int a = 0; a < 3; ++a) {
                    t0[a] += xe[n][a] * dphi[n][a0];
                    t1[a] += xe[n][a] * dphi[n][a1];
                }
            const double cx = t0[1] * t1[2] - t0[2] * t1[1];
            const double cy = t0[2] * t1[0] - t0[0] * t1[2];
            const double cz = t0[0] * t1[1] - t0[1] * t1[0];
            const double jxw = sqrt(cx * cx + cy * cy + cz * cz) * 0.25;
            for (int n = 0; n < 8; ++n) be[n] += phi[n] * g * jxw;
        }
        for (int n = 0; n < 8; ++n) b[m->cells[8 * c + n]] += be[n];
    }
}

int orc_dirichlet_mask(const orc_mesh *m, int bid, unsigned char *mask) {
    int cnt = 0;
    for (int i = 0; i < m->n_bfaces; ++i)
        if (m->bface_id[i] == bid)
            for (int k = 0; k < 4; ++k) {
                const int v = m->cells[8 * m->bface_cell[i] + FACE_VERTS[m->bface_face[i]][k]];
                if (!mask[v]) {
                    mask[v] = 1;
                    ++cnt;
                }
            }
    return cnt;
}

void orc_apply_dirichlet(orc_csr *A, double *b, double *x, const unsigned char *mask,
                         const double *values) {
    double first_diag = 1.0;
    for (int i = 0; i < A->n; ++i)
        if (A->val[A->diag[i]] != 0) {
            first_diag = A->val[A->diag[i]];
            break;
        }
    for (int i = 0; i < A->n; ++i) {
        if (!mask[i]) continue;
        const double g = values ? values[i] : 0.0;
        for (int64_t p = A->rowptr[i]; p < A->rowptr[i + 1]; ++p)
            if (p != A->diag[i]) A->val[p] = 0;
        if (A->val[A->diag[i]] == 0) A->val[A->diag[i]] = first_diag;
        b[i] = g * A->val[A->diag[i]];
        x[i] = g;
        for (int64_t p = A->rowptr[i]; p < A->rowptr[i + 1]; ++p) {
            const int j = A->col[p];
            if (j == i) continue;
            const int64_t q = csr_find(A, j, i);
            if (q >= 0) {
                b[j] -= A->val[q] * g;
                A->val[q] = 0;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------- */
/* preconditioned CG                                                                      */
/* ------------------------------------------------------------------------------------- */
static double dot(int n, const double *a, const double *b) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* z = M^{-1} r, M_SSOR = (D + wL) D^{-1} (D + wU) / (w (2 - w)) */
static void precondition(const orc_csr *A, int kind, double w, const double *r, double *z) {
    const int n = A->n;
    if (kind == 0) {
        memcpy(z, r, sizeof(double) * n);
    } else if (kind == 1) {
        for (int i = 0; i < n; ++i) z[i] = r[i] / A->val[A->diag[i]];
    } else {
        for (int i = 0; i < n; ++i) {
            double s = r[i];
            for (int64_t p = A->rowptr[i]; p < A->diag[i]; ++p) s -= w * A->val[p] * z[A->col[p]];
            z[i] = s / A->val[A->diag[i]];
        }
        for (int i = 0; i < n; ++i) z[i] *= A->val[A->diag[i]];
        for (int i = n - 1; i >= 0; --i) {
            double s = z[i];
            for (int64_t p = A->diag[i] + 1; p < A->rowptr[i + 1]; ++p) s -= w * A->val[p] * z[A->col[p]];
            z[i] = s / A->val[A->diag[i]];
        }
        const double sc = w * (2.0 - w);
        for (int i = 0; i < n; ++i) z[i] *= sc;
    }
}

int orc_pcg(const orc_csr *A, const double *b, double *x, int max_it, double tol, int relative,
            int precond, double omega, double *res_out) {
    const int n = A->n;
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    double *p = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * n);
    orc_csr_spmv(A, x, q);
    for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
    const double thr = relative ? tol * sqrt(dot(n, b, b)) : tol;
    double res = sqrt(dot(n, r, r));
    int it = 0, ok = res <= thr;
    if (!ok) {
        precondition(A, precond, omega, r, z);
        memcpy(p, z, sizeof(double) * n);
        double rz = dot(n, r, z);
        while (it < max_it) {
            ++it;
            orc_csr_spmv(A, p, q);
            const double alpha = rz / dot(n, p, q);
            for (int i = 0; i < n; ++i) {
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
            }
            res = sqrt(dot(n, r, r));
            if (res <= thr) {
                ok = 1;
                break;
            }
            precondition(A, precond, omega, r, z);
            const double rz1 = dot(n, r, z);
            const double beta = rz1 / rz;
            rz = rz1;
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

/* ------------------------------------------------------------------------------------- */
/* gradient export                                                                         */
/* ------------------------------------------------------------------------------------- */
void orc_qp_fields(const orc_mesh *m, const double *u, double *val, double *grad) {
    for (int c = 0; c < m->n_cells; ++c) {
        double xe[8][3];
        cell_coords(m, c, xe);
        for (int q = 0; q < 8; ++q) {
            double phi[8], g[8][3];
            qpoint(xe, q, phi, g);
            double v = 0, gr[3] = {0, 0, 0};
            for (int k = 0; k < 8; ++k) {
                const double uk = u[m->cells[8 * c + k]];
                v += uk * phi[k];
                for (int a = 0; a < 3; ++a) gr[a] += uk * g[k][a];
            }
            if (val) val[8 * (size_t)c + q] = v;
            if (grad)
                for (int a = 0; a < 3; ++a) grad[24 * (size_t)c + 3 * q + a] = gr[a];
        }
    }
}

void orc_export_solution_grad(const orc_mesh *m, const double *u, double *grad) {
    const int n = m->n_dofs;
    int *v2c = malloc(sizeof(int) * (n > 0 ? n : 1)), *v2n = malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int c = 0; c < m->n_cells; ++c)
        for (int k = 0; k < 8; ++k) {
            v2c[m->cells[8 * c + k]] = c;
            v2n[m->cells[8 * c + k]] = k;
        }
    for (int v = 0; v < n; ++v) {
        double xe[8][3], phi[8], g[8][3];
        cell_coords(m, v2c[v], xe);
        qpoint(xe, v2n[v], phi, g);
        double s[3] = {0, 0, 0};
        for (int k = 0; k < 8; ++k)
            for (int a = 0; a < 3; ++a) s[a] += u[m->cells[8 * v2c[v] + k]] * g[k][a];
        for (int a = 0; a < 3; ++a) grad[3 * v + a] = -1.0 * s[a];
    }
    free(v2c);
    free(v2n);
}

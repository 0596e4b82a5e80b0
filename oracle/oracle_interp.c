/*
 * CPU oracle, transfer-to-atoms part and coefficient tables.  TEST INFRASTRUCTURE ONLY
 * (see oracle.h).  Every routine keeps the reference's operation order so that, compiled
 * without FP contraction, it rounds exactly as the reference does:
 *   InterpolatorCells::locate_cell       src/InterpolatorCells.cpp:268-307
 *   LinearTetrahedra::precompute etc.    src/InterpolatorCells.cpp:523-661, 720-742
 *   LinearHexahedra                      src/InterpolatorCells.cpp:1205-1562
 *   Interpolator::store_elfield          src/Interpolator.cpp:125-140, 172-190
 *   PhysicalQuantities                   src/PhysicalQuantities.cpp:31-64, 173-185
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ZERO 1e-15 /* InterpolatorCells::zero, include/InterpolatorCells.h:241 */

typedef struct { double x, y, z; } v3;
static v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 scl(v3 a, double r) { return mk(a.x * r, a.y * r, a.z * r); }
static v3 dvd(v3 a, double r) { return mk(a.x / r, a.y / r, a.z / r); }
static v3 ld(const double *xyz, int i) { return mk(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]); }

/* InterpolatorCells::determinant overloads, src/InterpolatorCells.cpp:471-500 */
static double det2(v3 v1, v3 v2) {
    return v1.x * (v2.y - v2.z) - v1.y * (v2.x - v2.z) + v1.z * (v2.x - v2.y);
}
static double det3(v3 v1, v3 v2, v3 v3_) {
    return v1.x * (v2.y * v3_.z - v3_.y * v2.z) - v2.x * (v1.y * v3_.z - v3_.y * v1.z)
         + v3_.x * (v1.y * v2.z - v2.y * v1.z);
}
static double det4(v3 v1, v3 v2, v3 v3_, v3 v4) {
    const double d1 = det3(v2, v3_, v4);
    const double d2 = det3(v1, v3_, v4);
    const double d3 = det3(v1, v2, v4);
    const double d4 = det3(v1, v2, v3_);
    return d4 - d3 + d2 - d1;
}

/* ------------------------------------------------------------------------------------- */
/* linear tetrahedra                                                                      */
/* ------------------------------------------------------------------------------------- */
int orc_lintet_build(orc_lintet *L, int n_nodes, const double *xyz, int nt, const int *tets,
                     const int *tet_nbr, const int *tet_marker, int search_region) {
    memset(L, 0, sizeof(*L));
    L->n_tets = nt;
    L->det0 = malloc(sizeof(double) * (nt ? nt : 1));
    L->det = malloc(sizeof(double) * 16 * (nt ? nt : 1));
    L->centroid = malloc(sizeof(double) * 3 * (nt ? nt : 1));
    L->marker = malloc(sizeof(int) * (nt ? nt : 1));
    /* node2tets in increasing tet order */
    int *cnt = calloc(n_nodes + 1, sizeof(int));
    for (int i = 0; i < 4 * nt; ++i) cnt[tets[i] + 1]++;
    for (int i = 0; i < n_nodes; ++i) cnt[i + 1] += cnt[i];
    int *n2t = malloc(sizeof(int) * 4 * (nt ? nt : 1)), *pos = malloc(sizeof(int) * (n_nodes + 1));
    memcpy(pos, cnt, sizeof(int) * (n_nodes + 1));
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < 4; ++k) n2t[pos[tets[4 * t + k]]++] = t;
    free(pos);
    /* neighbours: face neighbours in stored order, then node-sharing tets per node */
    int64_t cap = 64LL * nt + 16, nn = 0;
    L->nbr = malloc(sizeof(int) * cap);
    L->nbr_ptr = malloc(sizeof(int) * (nt + 1));
    L->nbr_ptr[0] = 0;
    for (int t = 0; t < nt; ++t) {
        const int *nb = tet_nbr + 4 * t;
        for (int k = 0; k < 4; ++k)
            if (nb[k] >= 0) {
                if (nn >= cap) L->nbr = realloc(L->nbr, sizeof(int) * (cap *= 2));
                L->nbr[nn++] = nb[k];
            }
        for (int k = 0; k < 4; ++k) {
            const int node = tets[4 * t + k];
            for (int e = cnt[node]; e < cnt[node + 1]; ++e) {
                const int o = n2t[e];
                if (o != t && o != nb[0] && o != nb[1] && o != nb[2] && o != nb[3]) {
                    if (nn >= cap) L->nbr = realloc(L->nbr, sizeof(int) * (cap *= 2));
                    L->nbr[nn++] = o;
                }
            }
        }
        L->nbr_ptr[t + 1] = (int)nn;
    }
    free(n2t);
    free(cnt);
    for (int t = 0; t < nt; ++t) {
        const int *s = tets + 4 * t;
        v3 v1 = ld(xyz, s[0]), v2 = ld(xyz, s[1]), v3_ = ld(xyz, s[2]), v4 = ld(xyz, s[3]);
        /* TetgenCells::get_centroid: sum in order, times 1/dim */
        v3 c = mk(0, 0, 0);
        c = add(c, v1); c = add(c, v2); c = add(c, v3_); c = add(c, v4);
        c = scl(c, 1.0 / 4);
        L->centroid[3 * t] = c.x; L->centroid[3 * t + 1] = c.y; L->centroid[3 * t + 2] = c.z;
        const double d0 = det4(v1, v2, v3_, v4);
        L->det0[t] = 1.0 / d0;
        double *D = L->det + 16 * t;
        double d1, dd2, d3, dd4;
        d1 = det2(mk(v2.y, v3_.y, v4.y), mk(v2.z, v3_.z, v4.z));
        dd2 = det2(mk(v2.x, v3_.x, v4.x), mk(v2.z, v3_.z, v4.z));
        d3 = det2(mk(v2.x, v3_.x, v4.x), mk(v2.y, v3_.y, v4.y));
        dd4 = det3(mk(v2.x, v3_.x, v4.x), mk(v2.y, v3_.y, v4.y), mk(v2.z, v3_.z, v4.z));
        D[0] = d1; D[1] = -dd2; D[2] = d3; D[3] = -dd4;
        d1 = det2(mk(v1.y, v3_.y, v4.y), mk(v1.z, v3_.z, v4.z));
        dd2 = det2(mk(v1.x, v3_.x, v4.x), mk(v1.z, v3_.z, v4.z));
        d3 = det2(mk(v1.x, v3_.x, v4.x), mk(v1.y, v3_.y, v4.y));
        dd4 = det3(mk(v1.x, v3_.x, v4.x), mk(v1.y, v3_.y, v4.y), mk(v1.z, v3_.z, v4.z));
        D[4] = -d1; D[5] = dd2; D[6] = -d3; D[7] = dd4;
        d1 = det2(mk(v1.y, v2.y, v4.y), mk(v1.z, v2.z, v4.z));
        dd2 = det2(mk(v1.x, v2.x, v4.x), mk(v1.z, v2.z, v4.z));
        d3 = det2(mk(v1.x, v2.x, v4.x), mk(v1.y, v2.y, v4.y));
        dd4 = det3(mk(v1.x, v2.x, v4.x), mk(v1.y, v2.y, v4.y), mk(v1.z, v2.z, v4.z));
        D[8] = d1; D[9] = -dd2; D[10] = d3; D[11] = -dd4;
        d1 = det2(mk(v1.y, v2.y, v3_.y), mk(v1.z, v2.z, v3_.z));
        dd2 = det2(mk(v1.x, v2.x, v3_.x), mk(v1.z, v2.z, v3_.z));
        d3 = det2(mk(v1.x, v2.x, v3_.x), mk(v1.y, v2.y, v3_.y));
        dd4 = det3(v1, v2, v3_);
        D[12] = -d1; D[13] = dd2; D[14] = -d3; D[15] = dd4;
        /* LinearTetrahedra::narrow_search_to (src/InterpolatorCells.cpp:720-742) */
        if (search_region == 3) L->marker[t] = tet_marker[t] != 3;
        else if (search_region == 1) L->marker[t] = tet_marker[t] == 3;
        else L->marker[t] = 0;
    }
    return 0;
}

void orc_lintet_free(orc_lintet *L) {
    free(L->det0); free(L->det); free(L->centroid); free(L->nbr_ptr); free(L->nbr); free(L->marker);
    memset(L, 0, sizeof(*L));
}

static double dot4(const double p[3], const double *d) {
    /* Vec4(point, 1).dotProduct(det_k) */
    return p[0] * d[0] + p[1] * d[1] + p[2] * d[2] + 1.0 * d[3];
}

int orc_lintet_point_in_cell(const orc_lintet *L, const double p[3], int i) {
    const double *D = L->det + 16 * i;
    const double d0 = L->det0[i];
    if (d0 * dot4(p, D) < -ZERO) return 0;
    if (d0 * dot4(p, D + 4) < -ZERO) return 0;
    if (d0 * dot4(p, D + 8) < -ZERO) return 0;
    if (d0 * dot4(p, D + 12) < -ZERO) return 0;
    return 1;
}

void orc_lintet_bcc(const orc_lintet *L, const double p[3], int t, double b[4]) {
    const double *D = L->det + 16 * t;
    for (int k = 0; k < 4; ++k) b[k] = ZERO + L->det0[t] * dot4(p, D + 4 * k);
}

int orc_lintet_locate(const orc_lintet *L, const double p[3], int guess) {
    const int n = L->n_tets;
    if (guess >= 0 && guess < n) {
        if (orc_lintet_point_in_cell(L, p, guess)) return guess;
        for (int e = L->nbr_ptr[guess]; e < L->nbr_ptr[guess + 1]; ++e)
            if (orc_lintet_point_in_cell(L, p, L->nbr[e])) return L->nbr[e];
    }
    double best = 1e100;
    int best_i = 0;
    for (int c = 0; c < n; ++c) {
        if (L->marker[c] == 0 && orc_lintet_point_in_cell(L, p, c)) return c;
        const double xx = p[0] - L->centroid[3 * c];
        const double yy = p[1] - L->centroid[3 * c + 1];
        const double zz = p[2] - L->centroid[3 * c + 2];
        const double d2 = xx * xx + yy * yy + zz * zz;
        if (d2 < best) {
            best = d2;
            best_i = c;
        }
    }
    return -best_i;
}

int orc_linhex_locate(const orc_lintet *L, const double p[3], int guess) {
    int tet = guess / 4;
    tet = orc_lintet_locate(L, p, tet);
    const int sign = tet < 0 ? -1 : 1;
    tet = abs(tet);
    double b[4];
    orc_lintet_bcc(L, p, tet, b);
    if (b[0] >= b[1] && b[0] >= b[2] && b[0] >= b[3]) return sign * (4 * tet + 0);
    if (b[1] >= b[0] && b[1] >= b[2] && b[1] >= b[3]) return sign * (4 * tet + 1);
    if (b[2] >= b[0] && b[2] >= b[1] && b[2] >= b[3]) return sign * (4 * tet + 2);
    if (b[3] >= b[0] && b[3] >= b[1] && b[3] >= b[2]) return sign * (4 * tet + 3);
    return -1;
}

void orc_locate_chain(const orc_lintet *L, int np, const double *xyz, int rank, int *cells) {
    int cell = -1;
    for (int i = 0; i < np; ++i) {
        const double *p = xyz + 3 * i;
        const int g = abs(cell);
        cell = rank == 3 ? orc_linhex_locate(L, p, g) : orc_lintet_locate(L, p, g);
        cells[i] = cell;
    }
}

/* ------------------------------------------------------------------------------------- */
/* linear hexahedra                                                                       */
/* ------------------------------------------------------------------------------------- */
static void linhex_uvw(const v3 x[8], const double p[3], double *u_out, double *v_out, double *w_out) {
    /* LinearHexahedra::precompute, src/InterpolatorCells.cpp:1237-1244 */
    v3 f0s = dvd(add(add(add(add(add(add(add(x[0], x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    v3 f1 = dvd(sub(add(add(sub(sub(add(add(scl(x[0], -1), x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    v3 f2 = dvd(add(add(sub(sub(add(add(sub(scl(x[0], -1), x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    v3 f3 = dvd(add(add(add(add(sub(sub(sub(scl(x[0], -1), x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    v3 f4 = dvd(sub(add(sub(add(sub(add(sub(x[0], x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    v3 f5 = dvd(sub(add(add(sub(add(sub(sub(x[0], x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    v3 f6 = dvd(add(add(sub(sub(sub(sub(add(x[0], x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    v3 f7 = dvd(sub(add(sub(add(add(sub(add(scl(x[0], -1), x[1]), x[2]), x[3]), x[4]), x[5]), x[6]), x[7]), 8.0);
    /* project_to_nat_coords, src/InterpolatorCells.cpp:1272-1317 */
    v3 f0 = sub(mk(p[0], p[1], p[2]), f0s);
    double u = 0, v = 0, w = 0;
    for (int i = 0; i < 20; ++i) {
        v3 f = sub(sub(sub(sub(sub(sub(sub(f0, scl(f1, u)), scl(f2, v)), scl(f3, w)), scl(f4, u * v)),
                           scl(f5, u * w)), scl(f6, v * w)), scl(f7, u * v * w));
        v3 fu = add(add(add(f1, scl(f4, v)), scl(f5, w)), scl(f7, v * w));
        v3 fv = add(add(add(f2, scl(f4, u)), scl(f6, w)), scl(f7, u * w));
        v3 fw = add(add(add(f3, scl(f5, u)), scl(f6, v)), scl(f7, u * v));
        double D = det3(fu, fv, fw);
        D = 1.0 / D;
        const double du = det3(f, fv, fw) * D;
        const double dv = det3(fu, f, fw) * D;
        const double dw = det3(fu, fv, f) * D;
        u += du;
        v += dv;
        w += dw;
        if (du * du + dv * dv + dw * dw < ZERO) break;
    }
    *u_out = u;
    *v_out = v;
    *w_out = w;
}

void orc_linhex_shape_functions(const double xe[8][3], const double p[3], double sf[8]) {
    v3 x[8];
    double u, v, w;
    for (int i = 0; i < 8; ++i) x[i] = mk(xe[i][0], xe[i][1], xe[i][2]);
    linhex_uvw(x, p, &u, &v, &w);
    /* shape_functions, src/InterpolatorCells.cpp:1334-1353 */
    sf[0] = (1 - u) * (1 - v) * (1 - w) / 8.0;
    sf[1] = (1 + u) * (1 - v) * (1 - w) / 8.0;
    sf[2] = (1 + u) * (1 + v) * (1 - w) / 8.0;
    sf[3] = (1 - u) * (1 + v) * (1 - w) / 8.0;
    sf[4] = (1 - u) * (1 - v) * (1 + w) / 8.0;
    sf[5] = (1 + u) * (1 - v) * (1 + w) / 8.0;
    sf[6] = (1 + u) * (1 + v) * (1 + w) / 8.0;
    sf[7] = (1 - u) * (1 + v) * (1 + w) / 8.0;
}

/* LinearHexahedra::shape_fun_grads(point, hex) (src/InterpolatorCells.cpp:1363-1416) and
 * interp_gradient(point, cell) (:407-421): E = -sum_i grad N_i(p) scalar2_i */
void orc_linhex_gradient(const double *nodes, const int *hexes, const double *sol5, int n, const double *pts,
                         const int *hex, double *out) {
    for (int q = 0; q < n; ++q) {
        const int *h = hexes + 8 * hex[q];
        v3 x[8], dN[8], J[3], Ji[3], e = mk(0, 0, 0);
        double u, v, w;
        for (int k = 0; k < 8; ++k) x[k] = mk(nodes[3 * h[k]], nodes[3 * h[k] + 1], nodes[3 * h[k] + 2]);
        linhex_uvw(x, pts + 3 * q, &u, &v, &w);
        dN[0] = mk(-(1 - v) * (1 - w), -(1 - u) * (1 - w), -(1 - u) * (1 - v));
        dN[1] = mk((1 - v) * (1 - w), -(1 + u) * (1 - w), -(1 + u) * (1 - v));
        dN[2] = mk((1 + v) * (1 - w), (1 + u) * (1 - w), -(1 + u) * (1 + v));
        dN[3] = mk(-(1 + v) * (1 - w), (1 - u) * (1 - w), -(1 - u) * (1 + v));
        dN[4] = mk(-(1 - v) * (1 + w), -(1 - u) * (1 + w), (1 - u) * (1 - v));
        dN[5] = mk((1 - v) * (1 + w), -(1 + u) * (1 + w), (1 + u) * (1 - v));
        dN[6] = mk((1 + v) * (1 + w), (1 + u) * (1 + w), (1 + u) * (1 + v));
        dN[7] = mk(-(1 + v) * (1 + w), (1 - u) * (1 + w), (1 - u) * (1 + v));
        for (int k = 0; k < 8; ++k) dN[k] = scl(dN[k], 0.125);
        J[0] = J[1] = J[2] = mk(0, 0, 0);
        for (int k = 0; k < 8; ++k) {
            J[0] = add(J[0], scl(x[k], dN[k].x));
            J[1] = add(J[1], scl(x[k], dN[k].y));
            J[2] = add(J[2], scl(x[k], dN[k].z));
        }
        const double Jdet = 1.0 / det3(J[0], J[1], J[2]);
        Ji[0] = scl(mk(J[1].y * J[2].z - J[2].y * J[1].z, J[2].x * J[1].z - J[1].x * J[2].z,
                       J[1].x * J[2].y - J[2].x * J[1].y), Jdet);
        Ji[1] = scl(mk(J[2].y * J[0].z - J[0].y * J[2].z, J[0].x * J[2].z - J[2].x * J[0].z,
                       J[2].x * J[0].y - J[0].x * J[2].y), Jdet);
        Ji[2] = scl(mk(J[0].y * J[1].z - J[1].y * J[0].z, J[1].x * J[0].z - J[0].x * J[1].z,
                       J[0].x * J[1].y - J[1].x * J[0].y), Jdet);
        for (int k = 0; k < 8; ++k) {
            v3 g = mk(0, 0, 0);
            g = add(g, scl(Ji[0], dN[k].x));
            g = add(g, scl(Ji[1], dN[k].y));
            g = add(g, scl(Ji[2], dN[k].z));
            e = sub(e, scl(g, sol5[5 * h[k] + 4]));
        }
        out[3 * q] = e.x;
        out[3 * q + 1] = e.y;
        out[3 * q + 2] = e.z;
    }
}

/* shape_fun_grads(hex, node), src/InterpolatorCells.cpp:1319-1332, 1462-1505 */
static void corner_grads(const double *xyz, const int *shex, int node, v3 sfg[8]) {
    static const int UVW[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                  {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    static const int NN[8][3] = {{1, 3, 4}, {0, 2, 5}, {3, 1, 6}, {2, 0, 7},
                                 {5, 7, 0}, {4, 6, 1}, {7, 5, 2}, {6, 4, 3}};
    const double uvw[3] = {UVW[node][0], UVW[node][1], UVW[node][2]};
    const int *nnn = NN[node];
    v3 vec0 = ld(xyz, shex[node]);
    double J[3][3];
    for (int i = 0; i < 3; ++i) {
        v3 r = scl(sub(vec0, ld(xyz, shex[nnn[i]])), uvw[i] * 0.5);
        J[i][0] = r.x; J[i][1] = r.y; J[i][2] = r.z;
    }
    double Jdet = det3(mk(J[0][0], J[0][1], J[0][2]), mk(J[1][0], J[1][1], J[1][2]),
                       mk(J[2][0], J[2][1], J[2][2]));
    Jdet = 1.0 / Jdet;
    double Ji[3][3] = {
        {J[1][1] * J[2][2] - J[2][1] * J[1][2], J[2][1] * J[0][2] - J[0][1] * J[2][2],
         J[0][1] * J[1][2] - J[1][1] * J[0][2]},
        {J[2][0] * J[1][2] - J[1][0] * J[2][2], J[0][0] * J[2][2] - J[2][0] * J[0][2],
         J[1][0] * J[0][2] - J[0][0] * J[1][2]},
        {J[1][0] * J[2][1] - J[2][0] * J[1][1], J[2][0] * J[0][1] - J[0][0] * J[2][1],
         J[0][0] * J[1][1] - J[1][0] * J[0][1]}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ji[i][j] *= Jdet;
    double g[8][3];
    memset(g, 0, sizeof(g));
    for (int i = 0; i < 3; ++i) {
        g[node][i] = 0.5 * (Ji[i][0] * uvw[0] + Ji[i][1] * uvw[1] + Ji[i][2] * uvw[2]);
        for (int j = 0; j < 3; ++j) g[nnn[j]][i] = -0.5 * Ji[i][j] * uvw[j];
    }
    for (int k = 0; k < 8; ++k) sfg[k] = mk(g[k][0], g[k][1], g[k][2]);
}

void orc_nodal_field(int n_nodes, const double *xyz, int nh, const int *hexes, const int *hex_marker,
                     int region_sign, const double *phi, double *out) {
    /* node2cells in hex order (src/Interpolator.cpp:61-76) */
    int *cnt = calloc(n_nodes + 1, sizeof(int));
    for (int h = 0; h < nh; ++h)
        if ((region_sign > 0 && hex_marker[h] > 0) || (region_sign < 0 && hex_marker[h] < 0))
            for (int k = 0; k < 8; ++k) cnt[hexes[8 * h + k] + 1]++;
    for (int i = 0; i < n_nodes; ++i) cnt[i + 1] += cnt[i];
    int *lst = malloc(sizeof(int) * (cnt[n_nodes] ? cnt[n_nodes] : 1));
    int *pos = malloc(sizeof(int) * (n_nodes + 1));
    memcpy(pos, cnt, sizeof(int) * (n_nodes + 1));
    for (int h = 0; h < nh; ++h)
        if ((region_sign > 0 && hex_marker[h] > 0) || (region_sign < 0 && hex_marker[h] < 0))
            for (int k = 0; k < 8; ++k) lst[pos[hexes[8 * h + k]]++] = 8 * h + k;
    for (int n = 0; n < n_nodes; ++n) {
        v3 mean = mk(0, 0, 0);
        const int nf = cnt[n + 1] - cnt[n];
        if (nf > 0) {
            for (int e = cnt[n]; e < cnt[n + 1]; ++e) {
                const int h = lst[e] / 8, node = lst[e] % 8;
                v3 sfg[8];
                corner_grads(xyz, hexes + 8 * h, node, sfg);
                v3 vi = mk(0, 0, 0);
                for (int i = 0; i < 8; ++i) vi = sub(vi, scl(sfg[i], phi[hexes[8 * h + i]]));
                mean = add(mean, vi);
            }
            mean = scl(mean, 1.0 / nf);
        }
        out[3 * n] = mean.x; out[3 * n + 1] = mean.y; out[3 * n + 2] = mean.z;
    }
    free(pos);
    free(lst);
    free(cnt);
}

/* ------------------------------------------------------------------------------------- */
/* quadratic tetrahedra, nodal solution, smoothing, per-point interpolation               */
/* ------------------------------------------------------------------------------------- */
static int common_entry(const int *a, int na, const int *b, int nb) {
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j)
            if (a[i] == b[j]) return a[i];
    return -1;
}

void orc_quadtet_cells(int nt, const int *tets, const int *hexes, const int *node_kind, int *cells10) {
    for (int t = 0; t < nt; ++t) {
        int e[4][8], ne[4] = {0, 0, 0, 0};
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 8; ++k) {
                const int node = hexes[8 * (4 * t + i) + k];
                if (node_kind[node] == 2) e[i][ne[i]++] = node;
            }
        int *c = cells10 + 10 * t;
        for (int k = 0; k < 4; ++k) c[k] = tets[4 * t + k];
        c[4] = common_entry(e[0], ne[0], e[1], ne[1]);
        c[5] = common_entry(e[1], ne[1], e[2], ne[2]);
        c[6] = common_entry(e[2], ne[2], e[0], ne[0]);
        c[7] = common_entry(e[0], ne[0], e[3], ne[3]);
        c[8] = common_entry(e[1], ne[1], e[3], ne[3]);
        c[9] = common_entry(e[2], ne[2], e[3], ne[3]);
    }
}

void orc_extract_nodal(int n_nodes, const double *xyz, int nh, const int *hexes, const int *hex_marker,
                       int region_sign, const unsigned char *in_region, const double *phi, const double *rho,
                       double empty, double *sol5) {
    double *E = malloc(sizeof(double) * 3 * (n_nodes ? n_nodes : 1));
    orc_nodal_field(n_nodes, xyz, nh, hexes, hex_marker, region_sign, phi, E);
    for (int n = 0; n < n_nodes; ++n) {
        double *s = sol5 + 5 * n;
        if (in_region[n]) {
            s[0] = E[3 * n]; s[1] = E[3 * n + 1]; s[2] = E[3 * n + 2];
            s[3] = rho ? rho[n] : 0.0;
            s[4] = phi[n];
        } else {
            for (int k = 0; k < 5; ++k) s[k] = empty;
        }
    }
    free(E);
}

static double dist2(const double *a, const double *b) {
    const double xx = a[0] - b[0], yy = a[1] - b[1], zz = a[2] - b[2];
    return xx * xx + yy * yy + zz * zz;
}

double orc_decay_factor(const double *xyz, int nt, const int *tets) {
    double emax = -1e308;
    for (int t = 0; t < nt; ++t) {
        const double *n1 = xyz + 3 * tets[4 * t], *n2 = xyz + 3 * tets[4 * t + 1];
        const double *n3 = xyz + 3 * tets[4 * t + 2], *n4 = xyz + 3 * tets[4 * t + 3];
        const double e[6] = {dist2(n1, n2), dist2(n1, n3), dist2(n1, n4), dist2(n2, n3), dist2(n2, n4), dist2(n3, n4)};
        for (int k = 0; k < 6; ++k) emax = e[k] > emax ? e[k] : emax;
    }
    return -1.0 / sqrt(emax);
}

void orc_smooth_nodal(int n_nodes, int n_tetnodes, const double *xyz, const int *node_kind, int nh,
                      const int *hexes, const int *hex_marker, int vacuum, double decay, double *sol5) {
    (void)n_nodes;
    const int mult = vacuum ? 1 : -1;
    /* neighbour lists in hex order */
    int *cnt = calloc(n_tetnodes + 1, sizeof(int));
    for (int h = 0; h < nh; ++h) {
        int tn = mult * hex_marker[h];
        if (tn <= 0) continue;
        tn -= 1;
        for (int k = 0; k < 8; ++k) {
            const int node = hexes[8 * h + k];
            if (node != tn && node_kind[node] >= 2) cnt[tn + 1]++;
        }
    }
    for (int i = 0; i < n_tetnodes; ++i) cnt[i + 1] += cnt[i];
    int *lst = malloc(sizeof(int) * (cnt[n_tetnodes] ? cnt[n_tetnodes] : 1));
    int *pos = malloc(sizeof(int) * (n_tetnodes + 1));
    memcpy(pos, cnt, sizeof(int) * (n_tetnodes + 1));
    for (int h = 0; h < nh; ++h) {
        int tn = mult * hex_marker[h];
        if (tn <= 0) continue;
        tn -= 1;
        for (int k = 0; k < 8; ++k) {
            const int node = hexes[8 * h + k];
            if (node != tn && node_kind[node] >= 2) lst[pos[tn]++] = node;
        }
    }
    for (int i = 0; i < n_tetnodes; ++i) {
        if (cnt[i + 1] == cnt[i]) continue;
        v3 vec = mk(0, 0, 0);
        double w_sum = 0;
        for (int e = cnt[i]; e < cnt[i + 1]; ++e) {
            const int nb = lst[e];
            const double w = exp(decay * sqrt(dist2(xyz + 3 * i, xyz + 3 * nb)));
            w_sum += w;
            vec = add(vec, scl(mk(sol5[5 * nb], sol5[5 * nb + 1], sol5[5 * nb + 2]), w));
        }
        if (w_sum > 0) {
            vec = scl(vec, 1.0 / w_sum);
            sol5[5 * i] = vec.x; sol5[5 * i + 1] = vec.y; sol5[5 * i + 2] = vec.z;
        }
    }
    free(pos);
    free(lst);
    free(cnt);
}

/* weights of point p in cell c (already |c|) and their nodes */
static int cell_weights(const orc_lintet *L, const double *xyz, const int *tets, const int *cells10,
                        const int *hexes, const double p[3], int rank, int c, double w[10], int nodes[10]) {
    if (rank == 3) {
        double xe[8][3];
        for (int k = 0; k < 8; ++k) {
            nodes[k] = hexes[8 * c + k];
            for (int a = 0; a < 3; ++a) xe[k][a] = xyz[3 * nodes[k] + a];
        }
        orc_linhex_shape_functions(xe, p, w);
        return 8;
    }
    double b[4];
    orc_lintet_bcc(L, p, c, b);
    if (rank == 1) {
        for (int k = 0; k < 4; ++k) {
            w[k] = b[k];
            nodes[k] = tets[4 * c + k];
        }
        return 4;
    }
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
    for (int k = 0; k < 10; ++k) nodes[k] = cells10[10 * c + k];
    return 10;
}

static void interp_one(const orc_lintet *L, const double *xyz, const int *tets, const int *cells10,
                       const int *hexes, const double *sol5, const double p[3], int rank, int cell, double *out) {
    double w[10];
    int nodes[10];
    const int n = cell_weights(L, xyz, tets, cells10, hexes, p, rank, abs(cell), w, nodes);
    double vx = 0, vy = 0, vz = 0, s1 = 0, s2 = 0;
    for (int i = 0; i < n; ++i) {
        const double *s = sol5 + 5 * nodes[i];
        vx += s[0] * w[i];
        vy += s[1] * w[i];
        vz += s[2] * w[i];
        s1 += s[3] * w[i];
        s2 += s[4] * w[i];
    }
    out[0] = vx; out[1] = vy; out[2] = vz; out[3] = s1; out[4] = s2;
}

void orc_interp_points(const orc_lintet *L, const double *xyz, const int *tets, const int *cells10,
                       const int *hexes, const double *sol5, int np, const double *pts, int rank,
                       int *cells, double *out) {
    int cell = -1;
    for (int i = 0; i < np; ++i) {
        const double *p = pts + 3 * i;
        const int g = abs(cell);
        cell = rank == 3 ? orc_linhex_locate(L, p, g) : orc_lintet_locate(L, p, g);
        cells[i] = cell;
        interp_one(L, xyz, tets, cells10, hexes, sol5, p, rank, cell, out + 5 * i);
    }
}

void orc_interp_cells(const orc_lintet *L, const double *xyz, const int *tets, const int *cells10,
                      const int *hexes, const double *sol5, int np, const double *pts, int rank,
                      const int *cells, double *out) {
    for (int i = 0; i < np; ++i)
        interp_one(L, xyz, tets, cells10, hexes, sol5, pts + 3 * i, rank, cells[i], out + 5 * i);
}

/* ------------------------------------------------------------------------------------- */
/* linear triangles (surface atoms, dim 2)                                               */
/* ------------------------------------------------------------------------------------- */
static v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static void st3(double *d, int i, v3 a) { d[3 * i] = a.x; d[3 * i + 1] = a.y; d[3 * i + 2] = a.z; }

int orc_lintri_build(orc_lintri *T, int n_nodes, const double *xyz, int nt, const int *tris, const int *tri2tet) {
    memset(T, 0, sizeof(*T));
    T->n = nt;
    const size_t m = nt ? nt : 1;
    T->v0 = malloc(sizeof(double) * 3 * m); T->e1 = malloc(sizeof(double) * 3 * m);
    T->e2 = malloc(sizeof(double) * 3 * m); T->pv = malloc(sizeof(double) * 3 * m);
    T->nrm = malloc(sizeof(double) * 3 * m); T->cen = malloc(sizeof(double) * 3 * m);
    T->maxd = malloc(sizeof(double) * m); T->tri2tet = malloc(sizeof(int) * 2 * m);
    memcpy(T->tri2tet, tri2tet, sizeof(int) * 2 * nt);
    T->tris = malloc(sizeof(int) * 3 * m);
    memcpy(T->tris, tris, sizeof(int) * 3 * nt);
    /* node -> triangles, increasing */
    int *cnt = calloc(n_nodes + 1, sizeof(int));
    for (int i = 0; i < 3 * nt; ++i) cnt[tris[i] + 1]++;
    for (int i = 0; i < n_nodes; ++i) cnt[i + 1] += cnt[i];
    int *n2t = malloc(sizeof(int) * 3 * m), *pos = malloc(sizeof(int) * (n_nodes + 1));
    memcpy(pos, cnt, sizeof(int) * (n_nodes + 1));
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) n2t[pos[tris[3 * t + k]]++] = t;
    free(pos);
    /* neighbours: for each vertex, the other triangles touching it (repeats kept) */
    T->nbr_ptr = malloc(sizeof(int) * (nt + 1));
    int64_t total = 0;
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) total += cnt[tris[3 * t + k] + 1] - cnt[tris[3 * t + k]] - 1;
    T->nbr = malloc(sizeof(int) * (total ? total : 1));
    int64_t nn = 0;
    T->nbr_ptr[0] = 0;
    for (int t = 0; t < nt; ++t) {
        for (int k = 0; k < 3; ++k) {
            const int node = tris[3 * t + k];
            for (int e = cnt[node]; e < cnt[node + 1]; ++e)
                if (n2t[e] != t) T->nbr[nn++] = n2t[e];
        }
        T->nbr_ptr[t + 1] = (int)nn;
    }
    free(n2t);
    free(cnt);
    for (int t = 0; t < nt; ++t) {
        v3 a = ld(xyz, tris[3 * t]), b = ld(xyz, tris[3 * t + 1]), c = ld(xyz, tris[3 * t + 2]);
        v3 e1 = sub(b, a), e2 = sub(c, a);
        /* TetgenFaces::calc_appendices: normalized e1 x e2; centroid (v0 + v1 + v2) / 3 */
        v3 n = cross(e1, e2);
        const double nl = sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
        if (nl > 0) {
            const double f = 1 / nl;
            n = mk(n.x * f, n.y * f, n.z * f);
        }
        v3 cen = dvd(add(add(a, b), c), 3.0);
        /* LinearTriangles::precompute */
        v3 pv = cross(n, e2);
        const double idet = 1.0 / dot(e1, pv);
        st3(T->v0, t, a);
        st3(T->e1, t, scl(e1, idet));
        st3(T->e2, t, e2);
        st3(T->pv, t, scl(pv, idet));
        st3(T->nrm, t, n);
        T->maxd[t] = sqrt(e2.x * e2.x + e2.y * e2.y + e2.z * e2.z);
        st3(T->cen, t, cen);
    }
    return 0;
}

void orc_lintri_free(orc_lintri *T) {
    free(T->v0); free(T->e1); free(T->e2); free(T->pv); free(T->nrm); free(T->maxd); free(T->cen);
    free(T->nbr_ptr); free(T->nbr); free(T->tri2tet); free(T->tris);
    free(T->quads); free(T->quad2hex); free(T->tri6); free(T->qcen);
    memset(T, 0, sizeof(*T));
}

int orc_lintri_point_in_cell(const orc_lintri *T, const double p[3], int f) {
    v3 tv = sub(mk(p[0], p[1], p[2]), ld(T->v0, f));
    const double u = dot(tv, ld(T->pv, f));
    if (u < -ZERO || u > 1 + ZERO) return 0;
    v3 q = cross(tv, ld(T->e1, f));
    const double v = dot(q, ld(T->nrm, f));
    if (v < -ZERO || u + v > 1 + ZERO) return 0;
    return fabs(dot(q, ld(T->e2, f))) < T->maxd[f];
}

void orc_lintri_shape_functions(const orc_lintri *T, const double p[3], int f, double w[3]) {
    v3 tv = sub(mk(p[0], p[1], p[2]), ld(T->v0, f));
    v3 q = cross(tv, ld(T->e1, f));
    const double v = dot(tv, ld(T->pv, f));
    const double ww = dot(q, ld(T->nrm, f));
    const double u = 1.0 - v - ww;
    w[0] = ZERO + u; w[1] = ZERO + v; w[2] = ZERO + ww;
}

double orc_lintri_fast_distance(const orc_lintri *T, const double p[3], int f) {
    v3 tv = sub(mk(p[0], p[1], p[2]), ld(T->v0, f));
    v3 q = cross(tv, ld(T->e1, f));
    return dot(ld(T->e2, f), q);
}

int orc_lintri_locate(const orc_lintri *T, const double p[3], int guess) {
    const int n = T->n;
    if (guess >= 0 && guess < n) {
        if (orc_lintri_point_in_cell(T, p, guess)) return guess;
        for (int e = T->nbr_ptr[guess]; e < T->nbr_ptr[guess + 1]; ++e)
            if (orc_lintri_point_in_cell(T, p, T->nbr[e])) return T->nbr[e];
    }
    double best = 1e100;
    int best_i = 0;
    for (int c = 0; c < n; ++c) {
        if (orc_lintri_point_in_cell(T, p, c)) return c;
        const double xx = p[0] - T->cen[3 * c], yy = p[1] - T->cen[3 * c + 1], zz = p[2] - T->cen[3 * c + 2];
        const double d2 = xx * xx + yy * yy + zz * zz;
        if (d2 < best) {
            best = d2;
            best_i = c;
        }
    }
    return -best_i;
}

/* quadrangles of the surface faces; hex of tet t sharing the quad = 4t + position of the face
 * vertex in t (the only hex of t holding that vertex) */
int orc_lintri_set_quads(orc_lintri *T, int n_nodes, const double *xyz, const int *node_kind, const int *tets,
                         const int *tris, const int *quads) {
    (void)n_nodes;
    const int nq = 3 * T->n;
    const size_t m = nq ? nq : 1;
    T->quads = malloc(sizeof(int) * 4 * m);
    T->quad2hex = malloc(sizeof(int) * 2 * m);
    T->tri6 = malloc(sizeof(int) * 6 * (T->n ? T->n : 1));
    T->qcen = malloc(sizeof(double) * 3 * m);
    memcpy(T->quads, quads, sizeof(int) * 4 * nq);
    for (int q = 0; q < nq; ++q) {
        /* TetgenCells<4>::get_centroid: sum from zero, times 1/4 */
        double c[3] = {0, 0, 0};
        for (int k = 0; k < 4; ++k)
            for (int a = 0; a < 3; ++a) c[a] += xyz[3 * quads[4 * q + k] + a];
        for (int a = 0; a < 3; ++a) T->qcen[3 * q + a] = c[a] * (1.0 / 4);
        const int f = q / 3, v = tris[3 * f + q % 3];
        for (int r = 0; r < 2; ++r) {
            const int t = T->tri2tet[2 * f + r];
            int j = 0;
            while (j < 4 && tets[4 * t + j] != v) ++j;
            if (j == 4) return -1;
            T->quad2hex[2 * q + r] = 4 * t + j;
        }
    }
    /* QuadraticTriangles::calc_cell: edge-centroid nodes shared by quads (0,1), (1,2), (2,0) */
    for (int f = 0; f < T->n; ++f) {
        int e[3][4], ne[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 4; ++k) {
                const int node = quads[4 * (3 * f + i) + k];
                if (node_kind[node] == 2) e[i][ne[i]++] = node;
            }
        for (int k = 0; k < 3; ++k) T->tri6[6 * f + k] = tris[3 * f + k];
        for (int i = 0; i < 3; ++i) T->tri6[6 * f + 3 + i] = common_entry(e[i], ne[i], e[(i + 1) % 3], ne[(i + 1) % 3]);
    }
    return 0;
}

/* LinearTriangles/QuadraticTriangles::interp_solution: the tet the point is interpolated in
 * (the quadratic tets share the linear tets' predicates and search) */
static int surface_tet(const orc_lintri *T, const orc_lintet *L, const double p[3], int tri) {
    const int t0 = T->tri2tet[2 * tri], t1 = T->tri2tet[2 * tri + 1];
    if (fabs(orc_lintri_fast_distance(T, p, tri)) <= 100.0 * ZERO) return t0;
    if (orc_lintet_point_in_cell(L, p, t0)) return t0;
    return orc_lintet_locate(L, p, t1);
}

/* LinearHexahedra::point_in_cell (src/InterpolatorCells.cpp:1507-1527) */
static int linhex_point_in_cell(const orc_lintet *L, const double p[3], int hex) {
    double b[4];
    orc_lintet_bcc(L, p, hex / 4, b);
    if (!(b[0] >= 0 && b[1] >= 0 && b[2] >= 0 && b[3] >= 0)) return 0;
    switch (hex % 4) {
        case 0: return b[0] >= b[1] && b[0] >= b[2] && b[0] >= b[3];
        case 1: return b[1] >= b[0] && b[1] >= b[2] && b[1] >= b[3];
        case 2: return b[2] >= b[0] && b[2] >= b[1] && b[2] >= b[3];
        default: return b[3] >= b[0] && b[3] >= b[1] && b[3] >= b[2];
    }
}

/* LinearQuadrangles::interp_solution: the hex the point is interpolated in */
static int surface_hex(const orc_lintri *T, const orc_lintet *L, const double p[3], int quad) {
    const int h0 = T->quad2hex[2 * quad], h1 = T->quad2hex[2 * quad + 1];
    if (fabs(orc_lintri_fast_distance(T, p, quad / 3)) <= 100.0 * ZERO) return h0;
    if (linhex_point_in_cell(L, p, h0)) return h0;
    return orc_linhex_locate(L, p, h1);
}

/* LinearQuadrangles::locate_cell (src/InterpolatorCells.cpp:1963-1992) */
static int linquad_locate(const orc_lintri *T, const double p[3], int guess) {
    int tri = orc_lintri_locate(T, p, guess / 3);
    const int sign = tri < 0 ? -1 : 1;
    tri = abs(tri);
    double b[3];
    orc_lintri_shape_functions(T, p, tri, b);
    if (b[0] >= b[1] && b[0] >= b[2]) return sign * (3 * tri);
    if (b[1] >= b[0] && b[1] >= b[2]) return sign * (3 * tri + 1);
    if (b[2] >= b[0] && b[2] >= b[1]) return sign * (3 * tri + 2);
    return -1;
}

/* InterpolatorCells::locate_centroid (src/InterpolatorCells.cpp:309-333) */
static int locate_centroid(const double *cen, int n, const double p[3], int guess) {
    if (guess >= 0 && dist2(p, cen + 3 * guess) < ZERO) return guess;
    for (int c = 0; c < n; ++c)
        if (dist2(p, cen + 3 * c) < ZERO) return c;
    return -1;
}

/* InterpolatorCells::interp_centroid (:336-356): nodes weighted 1/n */
static void interp_centroid(const int *nodes, int nn, const double *sol5, double *out) {
    const double w = 1.0 / nn;
    double vx = 0, vy = 0, vz = 0, s1 = 0, s2 = 0;
    for (int i = 0; i < nn; ++i) {
        const double *s = sol5 + 5 * nodes[i];
        vx += s[0] * w;
        vy += s[1] * w;
        vz += s[2] * w;
        s1 += s[3] * w;
        s2 += s[4] * w;
    }
    out[0] = vx; out[1] = vy; out[2] = vz; out[3] = s1; out[4] = s2;
}

static void surface_value(const orc_lintri *T, const orc_lintet *L, const double *xyz, const int *tets,
                          const int *cells10, const int *hexes, const double *sol5, const double p[3], int rank,
                          int cell, double *out) {
    const int c = abs(cell);
    if (rank == 3) interp_one(L, xyz, tets, cells10, hexes, sol5, p, 3, surface_hex(T, L, p, c), out);
    else interp_one(L, xyz, tets, cells10, hexes, sol5, p, rank, surface_tet(T, L, p, c), out);
}

static void centroid_value(const orc_lintri *T, const double *sol5, int rank, int cell, double *out) {
    if (rank == 3) interp_centroid(T->quads + 4 * cell, 4, sol5, out);
    else if (rank == 2) interp_centroid(T->tri6 + 6 * cell, 6, sol5, out);
    else interp_centroid(T->tris + 3 * cell, 3, sol5, out);
}

void orc_interp_surface(const orc_lintri *T, const orc_lintet *L, const double *xyz, const int *tets,
                        const int *cells10, const int *hexes, const double *sol5, int np, const double *pts,
                        int rank, int centroids, int *cells, double *out) {
    int cell = -1;
    for (int i = 0; i < np; ++i) {
        const double *p = pts + 3 * i;
        double *o = out + 5 * i;
        if (centroids) {
            const double *cen = rank == 3 ? T->qcen : T->cen;
            const int r = locate_centroid(cen, rank == 3 ? 3 * T->n : T->n, p, cell);
            if (r >= 0) {
                cell = r;
                centroid_value(T, sol5, rank, cell, o);
            } else {
                for (int k = 0; k < 5; ++k) o[k] = -1.0;
            }
        } else {
            cell = rank == 3 ? linquad_locate(T, p, abs(cell)) : orc_lintri_locate(T, p, abs(cell));
            surface_value(T, L, xyz, tets, cells10, hexes, sol5, p, rank, cell, o);
        }
        cells[i] = cell;
    }
}

void orc_interp_surface_cells(const orc_lintri *T, const orc_lintet *L, const double *xyz, const int *tets,
                              const int *cells10, const int *hexes, const double *sol5, int np,
                              const double *pts, int rank, int centroids, const int *cells, double *out) {
    for (int i = 0; i < np; ++i) {
        if (centroids) centroid_value(T, sol5, rank, cells[i], out + 5 * i);
        else surface_value(T, L, xyz, tets, cells10, hexes, sol5, pts + 3 * i, rank, cells[i], out + 5 * i);
    }
}

/* ------------------------------------------------------------------------------------- */
/* physical quantities                                                                    */
/* ------------------------------------------------------------------------------------- */
/* resistivity of bulk Cu, (T [K], rho [Ohm nm]) pairs, src/PhysicalQuantityData.cpp:47-76 */
/* Pic::update_point_cell (src/Pic.cpp:185-195): deal cell -> femocs hex guess -> locate -> deal
 * cell, -1 outside */
void orc_update_cells(const orc_lintet *L, const int *deal2femocs, const int *femocs2deal, int n, const double *pts,
                      int *cells) {
    for (int i = 0; i < n; ++i) {
        const int h = orc_linhex_locate(L, pts + 3 * i, deal2femocs[cells[i]]);
        cells[i] = h < 0 ? -1 : femocs2deal[h];
    }
}

/* PoissonSolver::assemble_space_charge_fast (src/PoissonSolver.cpp:300-319) */
void orc_space_charge(const orc_mesh *m, const double *nodes, const int *hexes, const int *deal2femocs, int n,
                      const double *pts, const int *cells, double factor, double *b) {
    /* LinearHexahedra::shape_funs_dealii (src/InterpolatorCells.cpp:1355-1358): femocs (UCD)
     * order to the solver cell's lexicographic order */
    static const int ucd_of_lex[8] = {0, 1, 4, 5, 3, 2, 7, 6};
    for (int p = 0; p < n; ++p) {
        const int c = cells[p];
        const int *h = hexes + 8 * deal2femocs[c];
        double xe[8][3], sf[8];
        for (int k = 0; k < 8; ++k)
            for (int d = 0; d < 3; ++d) xe[k][d] = nodes[3 * h[k] + d];
        orc_linhex_shape_functions(xe, pts + 3 * p, sf);
        for (int k = 0; k < 8; ++k) b[m->cells[8 * c + k]] += sf[ucd_of_lex[k]] * factor;
    }
}

static const double RHO_T[26] = {200, 250, 273.15, 300, 350, 400, 450, 500, 550, 600, 650, 700, 750,
                                 800, 850, 900, 950, 1000, 1050, 1100, 1150, 1200, 1250, 1300, 1350, 1400};
static const double RHO_V[26] = {
    1.04880891725670e1, 1.38746906313026e1, 1.54302576609579e1, 1.72303253767393e1,
    2.05822342557969e1, 2.39476190263668e1, 2.73376276457821e1, 3.07598544167200e1,
    3.42197627890272e1, 3.77214997429762e1, 4.12683756079926e1, 4.48631564739703e1,
    4.85082484331790e1, 5.22058173711257e1, 5.59578691681847e1, 5.97663048918425e1,
    6.36329597840890e1, 6.75596315079091e1, 7.15481011300350e1, 7.56001491045029e1,
    7.97175677635730e1, 8.39021713382574e1, 8.81558042149690e1, 9.24803479251429e1,
    9.68777272230961e1, 1.01349915510292e2};

static double lin_interp(double x) {
    if (x <= RHO_T[0]) return RHO_V[0];
    if (x >= RHO_T[25]) return RHO_V[25];
    int i1 = 0;
    while (RHO_T[i1] < x) ++i1; /* std::lower_bound on first */
    const int i2 = i1 - 1;
    return RHO_V[i2] + (RHO_V[i1] - RHO_V[i2]) * (x - RHO_T[i2]) / (RHO_T[i1] - RHO_T[i2]);
}

static double clampT(double T) {
    if (T < RHO_T[0]) T = RHO_T[0];
    if (T > RHO_T[25]) T = RHO_T[25];
    return T;
}

double orc_resistivity(double T) { return 10. * lin_interp(clampT(T)); }
double orc_sigma(double T) { return 1.0 / orc_resistivity(T); }
double orc_kappa(double T, double lorentz) {
    T = clampT(T);
    return lorentz * T * orc_sigma(T);
}

/* element-wise orc_sigma / orc_kappa over n temperatures (same arithmetic per element) */
void orc_sigma_kappa_array(int n, const double *T, double lorentz, double *sigma, double *kappa) {
    for (int i = 0; i < n; ++i) {
        if (sigma) sigma[i] = orc_sigma(T[i]);
        if (kappa) kappa[i] = orc_kappa(T[i], lorentz);
    }
}

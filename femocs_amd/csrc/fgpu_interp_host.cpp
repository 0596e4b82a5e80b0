// Host-side tables of the transfer to atoms.  Compiled with -ffp-contract=off: the tet tables
// must round exactly as the reference computes them, since the located cells depend on the
// signs of these numbers to the last bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <stdexcept>

#include "fgpu_interp.h"

namespace fgpu {

namespace {

struct V3 {
    double x, y, z;
};
inline V3 ld(const double *xyz, int i) { return {xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2]}; }

// InterpolatorCells::determinant overloads, src/InterpolatorCells.cpp:471-500
inline double det2(V3 a, V3 b) { return a.x * (b.y - b.z) - a.y * (b.x - b.z) + a.z * (b.x - b.y); }
inline double det3(V3 a, V3 b, V3 c) {
    return a.x * (b.y * c.z - c.y * b.z) - b.x * (a.y * c.z - c.y * a.z) + c.x * (a.y * b.z - b.y * a.z);
}
inline double det4(V3 a, V3 b, V3 c, V3 d) {
    const double d1 = det3(b, c, d), d2 = det3(a, c, d), d3 = det3(a, b, d), d4 = det3(a, b, c);
    return d4 - d3 + d2 - d1;
}

// cofactor row of the three vertices p, q, r (the other one left out); the reference takes
// the 3x3 determinant of the last row from the untransposed vertices
inline void cof(V3 p, V3 q, V3 r, double sign, bool last, double *D) {
    const double d1 = det2({p.y, q.y, r.y}, {p.z, q.z, r.z});
    const double d2 = det2({p.x, q.x, r.x}, {p.z, q.z, r.z});
    const double d3 = det2({p.x, q.x, r.x}, {p.y, q.y, r.y});
    const double d4 = last ? det3(p, q, r) : det3({p.x, q.x, r.x}, {p.y, q.y, r.y}, {p.z, q.z, r.z});
    D[0] = sign * d1;
    D[1] = -sign * d2;
    D[2] = sign * d3;
    D[3] = -sign * d4;
}

inline float down(double v) {
    float f = (float)v;
    if ((double)f > v) f = std::nextafter(f, -INFINITY);
    return f;
}
inline float up(double v) {
    float f = (float)v;
    if ((double)f < v) f = std::nextafter(f, INFINITY);
    return f;
}

}  // namespace

void build_lintet(LinTetHost &L, int n_nodes, const double *xyz, int nt, const int *tets, const int *tet_nbr,
                  const int *tet_marker, int search_region) {
    L.n = nt;
    L.det0.resize(nt);
    L.det.resize(16 * (size_t)nt);
    L.centroid.resize(3 * (size_t)nt);
    L.marker.resize(nt);
    // node -> tets in increasing tet order
    std::vector<int> cnt(n_nodes + 1, 0), n2t(4 * (size_t)nt);
    for (size_t i = 0; i < 4 * (size_t)nt; ++i) cnt[tets[i] + 1]++;
    for (int i = 0; i < n_nodes; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> pos(cnt.begin(), cnt.end());
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < 4; ++k) n2t[pos[tets[4 * t + k]]++] = t;
    L.nbr_ptr.assign(nt + 1, 0);
    L.nbr.clear();
    L.nbr.reserve(40 * (size_t)nt);
    for (int t = 0; t < nt; ++t) {
        const int *nb = tet_nbr + 4 * (size_t)t;
        for (int k = 0; k < 4; ++k)
            if (nb[k] >= 0) L.nbr.push_back(nb[k]);
        for (int k = 0; k < 4; ++k) {
            const int node = tets[4 * t + k];
            for (int e = cnt[node]; e < cnt[node + 1]; ++e) {
                const int o = n2t[e];
                if (o != t && o != nb[0] && o != nb[1] && o != nb[2] && o != nb[3]) L.nbr.push_back(o);
            }
        }
        L.nbr_ptr[t + 1] = (int)L.nbr.size();
    }
    for (int t = 0; t < nt; ++t) {
        const int *s = tets + 4 * (size_t)t;
        const V3 v1 = ld(xyz, s[0]), v2 = ld(xyz, s[1]), v3 = ld(xyz, s[2]), v4 = ld(xyz, s[3]);
        // TetgenCells::get_centroid: running sum, then times 1/4
        V3 c{0, 0, 0};
        c = {c.x + v1.x, c.y + v1.y, c.z + v1.z};
        c = {c.x + v2.x, c.y + v2.y, c.z + v2.z};
        c = {c.x + v3.x, c.y + v3.y, c.z + v3.z};
        c = {c.x + v4.x, c.y + v4.y, c.z + v4.z};
        const double q = 1.0 / 4;
        L.centroid[3 * (size_t)t] = c.x * q;
        L.centroid[3 * (size_t)t + 1] = c.y * q;
        L.centroid[3 * (size_t)t + 2] = c.z * q;
        L.det0[t] = 1.0 / det4(v1, v2, v3, v4);
        double *D = &L.det[16 * (size_t)t];
        cof(v2, v3, v4, 1.0, false, D);
        cof(v1, v3, v4, -1.0, false, D + 4);
        cof(v1, v2, v4, 1.0, false, D + 8);
        cof(v1, v2, v3, -1.0, true, D + 12);
        if (search_region == 3) L.marker[t] = tet_marker[t] != 3;
        else if (search_region == 1) L.marker[t] = tet_marker[t] == 3;
        else L.marker[t] = 0;
    }
}

void LinTetHost::bcc(const double p[3], int t, double b[4]) const {
    const double *D = &det[16 * (size_t)t];
    for (int k = 0; k < 4; ++k) {
        const double *d = D + 4 * k;
        b[k] = kZero + det0[t] * (p[0] * d[0] + p[1] * d[1] + p[2] * d[2] + 1.0 * d[3]);
    }
}

bool LinTetHost::point_in_cell(const double p[3], int t) const {
    const double *D = &det[16 * (size_t)t];
    const double d0 = det0[t];
    for (int k = 0; k < 4; ++k) {
        const double *d = D + 4 * k;
        if (d0 * (p[0] * d[0] + p[1] * d[1] + p[2] * d[2] + 1.0 * d[3]) < -kZero) return false;
    }
    return true;
}

// ------------------------------------------------------------------------------------------
// linear triangles
// ------------------------------------------------------------------------------------------
namespace {
inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 vscl(V3 a, double r) { return {a.x * r, a.y * r, a.z * r}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline void put(double *d, V3 a) {
    d[0] = a.x;
    d[1] = a.y;
    d[2] = a.z;
}
inline V3 at(const double *d) { return {d[0], d[1], d[2]}; }
}  // namespace

void build_lintri(LinTriHost &T, int n_nodes, const double *xyz, int nf, const int *tris, const int *tri2tet) {
    T = LinTriHost{};
    T.n = nf;
    T.tab.assign(16 * (size_t)nf, 0.0);
    T.centroid.assign(3 * (size_t)nf, 0.0);
    T.tri2tet.assign(tri2tet, tri2tet + 2 * (size_t)nf);
    T.tris.assign(tris, tris + 3 * (size_t)nf);
    std::vector<int> ptr(n_nodes + 1, 0), n2f(3 * (size_t)nf);
    for (size_t i = 0; i < 3 * (size_t)nf; ++i) ptr[tris[i] + 1]++;
    for (int i = 0; i < n_nodes; ++i) ptr[i + 1] += ptr[i];
    {
        std::vector<int> pos(ptr.begin(), ptr.end() - 1);
        for (int f = 0; f < nf; ++f)
            for (int k = 0; k < 3; ++k) n2f[pos[tris[3 * (size_t)f + k]]++] = f;
    }
    T.nbr_ptr.assign(nf + 1, 0);
    for (int f = 0; f < nf; ++f) {
        for (int k = 0; k < 3; ++k) {
            const int node = tris[3 * (size_t)f + k];
            for (int e = ptr[node]; e < ptr[node + 1]; ++e)
                if (n2f[e] != f) T.nbr.push_back(n2f[e]);
        }
        T.nbr_ptr[f + 1] = (int)T.nbr.size();
    }
    for (int f = 0; f < nf; ++f) {
        const V3 a = ld(xyz, tris[3 * (size_t)f]), b = ld(xyz, tris[3 * (size_t)f + 1]), c = ld(xyz, tris[3 * (size_t)f + 2]);
        const V3 e1 = vsub(b, a), e2 = vsub(c, a);
        V3 nrm = cross(e1, e2);
        const double len = std::sqrt(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
        if (len > 0) nrm = vscl(nrm, 1 / len);
        const V3 pv = cross(nrm, e2);
        const double idet = 1.0 / dot(e1, pv);
        double *t = &T.tab[16 * (size_t)f];
        put(t, a);
        put(t + 3, vscl(e1, idet));
        put(t + 6, e2);
        put(t + 9, vscl(pv, idet));
        put(t + 12, nrm);
        t[15] = std::sqrt(e2.x * e2.x + e2.y * e2.y + e2.z * e2.z);
        const V3 s = {a.x + b.x + c.x, a.y + b.y + c.y, a.z + b.z + c.z};
        put(&T.centroid[3 * (size_t)f], {s.x / 3.0, s.y / 3.0, s.z / 3.0});
    }
}

bool LinTriHost::point_in_cell(const double p[3], int f) const {
    const double *t = &tab[16 * (size_t)f];
    const V3 tv = vsub({p[0], p[1], p[2]}, at(t));
    const double u = dot(tv, at(t + 9));
    if (u < -kZero || u > 1 + kZero) return false;
    const V3 q = cross(tv, at(t + 3));
    const double v = dot(q, at(t + 12));
    if (v < -kZero || u + v > 1 + kZero) return false;
    return std::fabs(dot(q, at(t + 6))) < t[15];
}

void LinTriHost::shape_functions(const double p[3], int f, double w[3]) const {
    const double *t = &tab[16 * (size_t)f];
    const V3 tv = vsub({p[0], p[1], p[2]}, at(t));
    const V3 q = cross(tv, at(t + 3));
    const double v = dot(tv, at(t + 9));
    const double ww = dot(q, at(t + 12));
    const double u = 1.0 - v - ww;
    w[0] = kZero + u;
    w[1] = kZero + v;
    w[2] = kZero + ww;
}

void set_quads(LinTriHost &T, const double *xyz, const int *node_kind, const int *tets, const int *quads) {
    const int nq = 3 * T.n;
    T.quads.assign(quads, quads + 4 * (size_t)nq);
    T.qcen.assign(3 * (size_t)nq, 0.0);
    T.quad2hex.assign(2 * (size_t)nq, 0);
    T.tri6.assign(6 * (size_t)T.n, 0);
    for (int q = 0; q < nq; ++q) {
        double c[3] = {0, 0, 0};
        for (int k = 0; k < 4; ++k)
            for (int a = 0; a < 3; ++a) c[a] += xyz[3 * (size_t)quads[4 * (size_t)q + k] + a];
        for (int a = 0; a < 3; ++a) T.qcen[3 * (size_t)q + a] = c[a] * (1.0 / 4);
        const int f = q / 3, v = T.tris[3 * (size_t)f + q % 3];
        if (quads[4 * (size_t)q] != v) throw std::invalid_argument("quad 3f+k must start at vertex k of face f");
        for (int r = 0; r < 2; ++r) {
            const int t = T.tri2tet[2 * (size_t)f + r];
            int j = 0;
            while (j < 4 && tets[4 * (size_t)t + j] != v) ++j;
            if (j == 4) throw std::invalid_argument("a face vertex is not a vertex of the face's tets");
            T.quad2hex[2 * (size_t)q + r] = 4 * t + j;
        }
    }
    for (int f = 0; f < T.n; ++f) {
        int e[3][4], ne[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 4; ++k) {
                const int node = quads[4 * (size_t)(3 * f + i) + k];
                if (node_kind[node] == 2) e[i][ne[i]++] = node;
            }
        for (int k = 0; k < 3; ++k) T.tri6[6 * (size_t)f + k] = T.tris[3 * (size_t)f + k];
        for (int i = 0; i < 3; ++i) {
            const int j = (i + 1) % 3;
            int c = -1;
            for (int a = 0; a < ne[i] && c < 0; ++a)
                for (int b = 0; b < ne[j]; ++b)
                    if (e[i][a] == e[j][b]) {
                        c = e[i][a];
                        break;
                    }
            T.tri6[6 * (size_t)f + 3 + i] = c;
        }
    }
}

double LinTriHost::fast_distance(const double p[3], int f) const {
    const double *t = &tab[16 * (size_t)f];
    const V3 tv = vsub({p[0], p[1], p[2]}, at(t));
    return dot(at(t + 6), cross(tv, at(t + 3)));
}

// ------------------------------------------------------------------------------------------
// linear BVH
// ------------------------------------------------------------------------------------------
namespace {

inline uint64_t spread21(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

struct Radix {
    const std::vector<uint64_t> &code;
    int n;
    int delta(int i, int j) const {
        if (j < 0 || j >= n) return -1;
        if (code[i] == code[j]) return 64 + __builtin_clz((unsigned)(i ^ j));
        return __builtin_clzll(code[i] ^ code[j]);
    }
};

}  // namespace

void build_bvh(int n, const std::vector<double> &lo, const std::vector<double> &hi, const std::vector<double> &cen, Bvh &B) {
    B = Bvh{};
    B.n_leaves = n;
    if (n <= 0) return;
    double cl[3] = {1e300, 1e300, 1e300}, ch[3] = {-1e300, -1e300, -1e300}, ext = 0;
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            cl[a] = std::min(cl[a], cen[3 * (size_t)i + a]);
            ch[a] = std::max(ch[a], cen[3 * (size_t)i + a]);
            ext = std::max(ext, hi[3 * (size_t)i + a] - lo[3 * (size_t)i + a]);
        }
    std::vector<uint64_t> key(n);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) {
        uint64_t m = 0;
        for (int a = 0; a < 3; ++a) {
            const double s = ch[a] > cl[a] ? (cen[3 * (size_t)i + a] - cl[a]) / (ch[a] - cl[a]) : 0.0;
            const uint64_t q = (uint64_t)std::min(2097151.0, std::max(0.0, s * 2097151.0));
            m |= spread21(q) << (2 - a);
        }
        key[i] = m;
    }
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return key[a] != key[b] ? key[a] < key[b] : a < b; });
    std::vector<uint64_t> code(n);
    B.prim.resize(n);
    B.leafbox.resize(6 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        const int t = order[i];
        code[i] = key[t];
        B.prim[i] = t;
        for (int a = 0; a < 3; ++a) {
            // widen by a margin far above the 1e-15 tolerance of the containment predicate
            const double m = 1e-9 * (std::fabs(lo[3 * (size_t)t + a]) + std::fabs(hi[3 * (size_t)t + a]) + ext);
            B.leafbox[6 * (size_t)i + a] = down(lo[3 * (size_t)t + a] - m);
            B.leafbox[6 * (size_t)i + 3 + a] = up(hi[3 * (size_t)t + a] + m);
        }
    }
    if (n == 1) return;
    const int ni = n - 1;
    B.child.assign(2 * (size_t)ni, 0);
    std::vector<int> parent_of_internal(ni, -1);
    Radix R{code, n};
    for (int i = 0; i < ni; ++i) {
        const int d = (R.delta(i, i + 1) - R.delta(i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = R.delta(i, i - d);
        int lmax = 2;
        while (R.delta(i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (R.delta(i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = R.delta(i, j);
        int s = 0;
        for (int div = 2;; div *= 2) {
            const int t = (l + div - 1) / div;
            if (R.delta(i, i + (s + t) * d) > dnode) s += t;
            if (t <= 1) break;
        }
        const int gamma = i + s * d + std::min(d, 0);
        const int left = std::min(i, j) == gamma ? ~gamma : gamma;
        const int right = std::max(i, j) == gamma + 1 ? ~(gamma + 1) : gamma + 1;
        B.child[2 * (size_t)i] = left;
        B.child[2 * (size_t)i + 1] = right;
        if (left >= 0) parent_of_internal[left] = i;
        if (right >= 0) parent_of_internal[right] = i;
    }
    // boxes bottom-up: post-order over the tree from the root (node 0)
    B.box.assign(6 * (size_t)ni, 0.f);
    B.cbox.assign(6 * (size_t)ni, 0.f);
    std::vector<int> stack{0}, post;
    post.reserve(ni);
    while (!stack.empty()) {
        const int v = stack.back();
        stack.pop_back();
        post.push_back(v);
        for (int k = 0; k < 2; ++k)
            if (B.child[2 * (size_t)v + k] >= 0) stack.push_back(B.child[2 * (size_t)v + k]);
    }
    if ((int)post.size() != ni) throw std::runtime_error("bvh: tree does not reach every internal node");
    for (auto it = post.rbegin(); it != post.rend(); ++it) {
        const int v = *it;
        float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        float cb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 2; ++k) {
            const int c = B.child[2 * (size_t)v + k];
            const float *cbx, *ccb;
            float lc[6];
            if (c >= 0) {
                cbx = &B.box[6 * (size_t)c];
                ccb = &B.cbox[6 * (size_t)c];
            } else {
                const int leaf = ~c, t = B.prim[leaf];
                cbx = &B.leafbox[6 * (size_t)leaf];
                for (int a = 0; a < 3; ++a) {
                    lc[a] = down(cen[3 * (size_t)t + a]);
                    lc[3 + a] = up(cen[3 * (size_t)t + a]);
                }
                ccb = lc;
            }
            for (int a = 0; a < 3; ++a) {
                b[a] = std::min(b[a], cbx[a]);
                b[3 + a] = std::max(b[3 + a], cbx[3 + a]);
                cb[a] = std::min(cb[a], ccb[a]);
                cb[3 + a] = std::max(cb[3 + a], ccb[3 + a]);
            }
        }
        std::copy(b, b + 6, &B.box[6 * (size_t)v]);
        std::copy(cb, cb + 6, &B.cbox[6 * (size_t)v]);
    }
}

void build_quadtet(int nt, const int *tets, const int *hexes, const int *node_kind, std::vector<int> &c10) {
    auto common = [](const int *a, int na, const int *b, int nb) {
        for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j)
                if (a[i] == b[j]) return a[i];
        return -1;
    };
    c10.assign(10 * (size_t)nt, -1);
    for (int t = 0; t < nt; ++t) {
        int e[4][8], ne[4] = {0, 0, 0, 0};
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 8; ++k) {
                const int node = hexes[8 * (4 * (size_t)t + i) + k];
                if (node_kind[node] == 2) e[i][ne[i]++] = node;
            }
        int *c = &c10[10 * (size_t)t];
        for (int k = 0; k < 4; ++k) c[k] = tets[4 * (size_t)t + k];
        c[4] = common(e[0], ne[0], e[1], ne[1]);
        c[5] = common(e[1], ne[1], e[2], ne[2]);
        c[6] = common(e[2], ne[2], e[0], ne[0]);
        c[7] = common(e[0], ne[0], e[3], ne[3]);
        c[8] = common(e[1], ne[1], e[3], ne[3]);
        c[9] = common(e[2], ne[2], e[3], ne[3]);
    }
}

// cand[i] = {lowest searchable tet containing p_i or -1, number of tets containing p_i,
//            nearest centroid (lowest index among equals) when no searchable tet contains p_i,
//            unused}
void resolve_chain(const LinTetHost &L, int n, const double *pts, const int4 *cand, int rank, int *tet) {
    int prev = -1;     // previous result in the rank's numbering (hex for rank 3)
    for (int i = 0; i < n; ++i) {
        const double *p = pts + 3 * (size_t)i;
        const int4 c = cand[i];
        const int g = rank == 3 ? std::abs(prev) / 4 : std::abs(prev);
        int r;
        if (c.y == 0) {
            r = -c.z;
        } else if (c.y == 1 && c.x >= 0) {
            r = c.x;
        } else {
            r = INT_MIN;
            if (g >= 0 && g < L.n) {
                if (L.point_in_cell(p, g)) {
                    r = g;
                } else {
                    for (int e = L.nbr_ptr[g]; e < L.nbr_ptr[g + 1]; ++e)
                        if (L.point_in_cell(p, L.nbr[e])) {
                            r = L.nbr[e];
                            break;
                        }
                }
            }
            if (r == INT_MIN) r = c.x >= 0 ? c.x : -c.z;
        }
        tet[i] = r;
        // rank 3: |hex| / 4 == |tet| for the next guess (hex = +-(4 |tet| + k), k < 4)
        prev = rank == 3 ? (r < 0 ? -(4 * -r) : 4 * r) : r;
    }
}

void resolve_tri_chain(const LinTriHost &T, int n, const double *pts, const int4 *cand, bool quads, int *cell) {
    int prev = -1;
    for (int i = 0; i < n; ++i) {
        const double *p = pts + 3 * (size_t)i;
        const int4 c = cand[i];
        const int g = quads ? std::abs(prev) / 3 : std::abs(prev);
        int r;
        if (c.y == 0) {
            r = -c.z;
        } else if (c.y == 1) {
            r = c.x;
        } else {
            r = INT_MIN;
            if (g < T.n) {
                if (T.point_in_cell(p, g)) {
                    r = g;
                } else {
                    for (int e = T.nbr_ptr[g]; e < T.nbr_ptr[g + 1]; ++e)
                        if (T.point_in_cell(p, T.nbr[e])) {
                            r = T.nbr[e];
                            break;
                        }
                }
            }
            if (r == INT_MIN) r = c.x;
        }
        if (quads) {
            const int sign = r < 0 ? -1 : 1, f = std::abs(r);
            double b[3];
            T.shape_functions(p, f, b);
            if (b[0] >= b[1] && b[0] >= b[2]) r = sign * (3 * f);
            else if (b[1] >= b[0] && b[1] >= b[2]) r = sign * (3 * f + 1);
            else if (b[2] >= b[0] && b[2] >= b[1]) r = sign * (3 * f + 2);
            else r = -1;
        }
        cell[i] = r;
        prev = r;
    }
}

void resolve_centroid_chain(const std::vector<double> &cen, int n, const double *pts, const int4 *cand, int *cell,
                            int *use) {
    int prev = -1;
    for (int i = 0; i < n; ++i) {
        const double *p = pts + 3 * (size_t)i;
        int r = -1;
        if (prev >= 0) {
            const double xx = p[0] - cen[3 * (size_t)prev], yy = p[1] - cen[3 * (size_t)prev + 1],
                         zz = p[2] - cen[3 * (size_t)prev + 2];
            if (xx * xx + yy * yy + zz * zz < kZero) r = prev;
        }
        if (r < 0) r = cand[i].x;
        if (r >= 0) prev = r;
        cell[i] = prev;
        use[i] = r;
    }
}

int locate_tet(const LinTetHost &L, const double p[3], int g, int4 c) {
    if (g >= 0 && g < L.n) {
        if (L.point_in_cell(p, g)) return g;
        for (int e = L.nbr_ptr[g]; e < L.nbr_ptr[g + 1]; ++e)
            if (L.point_in_cell(p, L.nbr[e])) return L.nbr[e];
    }
    return c.x >= 0 ? c.x : -c.z;
}

int surface_hex(const LinTriHost &T, const LinTetHost &L, const double p[3], int quad, int4 tet_cand) {
    const int h0 = T.quad2hex[2 * (size_t)quad], h1 = T.quad2hex[2 * (size_t)quad + 1];
    if (std::fabs(T.fast_distance(p, quad / 3)) <= 100.0 * kZero) return h0;
    // LinearHexahedra::point_in_cell (src/InterpolatorCells.cpp:1507-1527)
    double b[4];
    L.bcc(p, h0 / 4, b);
    if (b[0] >= 0 && b[1] >= 0 && b[2] >= 0 && b[3] >= 0) {
        const int k = h0 % 4;
        bool top = true;
        for (int j = 0; j < 4; ++j)
            if (j != k && !(b[k] >= b[j])) top = false;
        if (top) return h0;
    }
    // LinearHexahedra::locate_cell from h1 (src/InterpolatorCells.cpp:1536-1562)
    int t = locate_tet(L, p, h1 / 4, tet_cand);
    const int sign = t < 0 ? -1 : 1;
    t = std::abs(t);
    L.bcc(p, t, b);
    if (b[0] >= b[1] && b[0] >= b[2] && b[0] >= b[3]) return sign * (4 * t);
    if (b[1] >= b[0] && b[1] >= b[2] && b[1] >= b[3]) return sign * (4 * t + 1);
    if (b[2] >= b[0] && b[2] >= b[1] && b[2] >= b[3]) return sign * (4 * t + 2);
    if (b[3] >= b[0] && b[3] >= b[1] && b[3] >= b[2]) return sign * (4 * t + 3);
    return -1;
}

int surface_tet(const LinTriHost &T, const LinTetHost &L, const double p[3], int tri, int4 tet_cand) {
    const int t0 = T.tri2tet[2 * (size_t)tri], t1 = T.tri2tet[2 * (size_t)tri + 1];
    if (std::fabs(T.fast_distance(p, tri)) <= 100.0 * kZero) return t0;
    if (L.point_in_cell(p, t0)) return t0;
    return locate_tet(L, p, t1, tet_cand);
}

}  // namespace fgpu

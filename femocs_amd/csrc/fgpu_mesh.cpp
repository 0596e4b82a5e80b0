// Host-side mesh import: the deal.II triangulation steps the reference relies on
// (DealSolver::import_mesh, src/DealSolver.cpp:192-209; mark_boundary, :461-518), plus the
// internal ordering the GPU kernels use.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <utility>

#include "fgpu_internal.h"

namespace fgpu {

// lex[k] = ucd[kUcd2Lex[k]]; deal.II's ucd_to_deal {0,1,5,4,2,3,7,6} is the inverse map
static const int kUcd2Lex[8] = {0, 1, 4, 5, 3, 2, 7, 6};
// lexicographic vertices of deal.II face f (xi_{f/2} = f%2)
static const int kFaceVerts[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5},
                                     {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};

static double center_det(const double *x, const int *l) {
    // trilinear Jacobian at the cell centre, lexicographic vertices
    double J[3][3];
    for (int a = 0; a < 3; ++a) {
        const double *p[8];
        for (int n = 0; n < 8; ++n) p[n] = x + 3 * l[n];
        J[a][0] = 0.25 * ((p[1][a] - p[0][a]) + (p[3][a] - p[2][a]) + (p[5][a] - p[4][a]) + (p[7][a] - p[6][a]));
        J[a][1] = 0.25 * ((p[2][a] - p[0][a]) + (p[3][a] - p[1][a]) + (p[6][a] - p[4][a]) + (p[7][a] - p[5][a]));
        J[a][2] = 0.25 * ((p[4][a] - p[0][a]) + (p[5][a] - p[1][a]) + (p[6][a] - p[2][a]) + (p[7][a] - p[3][a]));
    }
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

int build_topology(Mesh &m, int nv, const double *xyz, int nc, const int *ucd, int top, int bottom,
                   int sides, int other, std::string &err) {
    if (nv < 0 || nc < 0 || (nc > 0 && (!xyz || !ucd))) {
        err = "mesh import: invalid sizes or null arrays";
        return FGPU_EINVAL;
    }
    m.n_vertices = nv;
    m.n_cells = nc;
    m.dof_of_vertex.assign(nv, -1);
    for (int64_t i = 0; i < 8LL * nc; ++i) {
        if (ucd[i] < 0 || ucd[i] >= nv) {
            err = "mesh import: cell vertex index out of range";
            return FGPU_EINVAL;
        }
        m.dof_of_vertex[ucd[i]] = 1;
    }
    // GridTools::delete_unused_vertices keeps the order of the used vertices
    int nd = 0;
    for (int i = 0; i < nv; ++i)
        if (m.dof_of_vertex[i] > 0) m.dof_of_vertex[i] = nd++;
    m.n_dofs = nd;
    m.vertex_of_dof.resize(nd);
    m.x.resize(3 * (size_t)nd);
    for (int i = 0; i < nv; ++i)
        if (m.dof_of_vertex[i] >= 0) {
            const int d = m.dof_of_vertex[i];
            m.vertex_of_dof[d] = i;
            for (int a = 0; a < 3; ++a) m.x[3 * (size_t)d + a] = xyz[3 * (size_t)i + a];
        }
    m.cells_lex.resize(8 * (size_t)nc);
    for (int c = 0; c < nc; ++c) {
        int *l = &m.cells_lex[8 * (size_t)c];
        for (int k = 0; k < 8; ++k) l[k] = m.dof_of_vertex[ucd[8 * (size_t)c + kUcd2Lex[k]]];
        // invert_all_cells_of_negative_grid swaps UCD vertex i and i+4, which is the eta
        // flip of the lexicographic cell
        if (center_det(m.x.data(), l) < 0) {
            std::swap(l[0], l[2]);
            std::swap(l[1], l[3]);
            std::swap(l[4], l[6]);
            std::swap(l[5], l[7]);
        }
    }
    // boundary faces = faces of exactly one cell.  Faces are bucketed by their smallest vertex
    // (a counting sort) and matched within each small bucket on their other three vertices
    struct FaceKey {
        int v1, v2, v3;
        int64_t slot;
    };
    std::vector<FaceKey> keys((size_t)6 * nc);
    std::vector<int64_t> bstart((size_t)nd + 1, 0);
    for (int c = 0; c < nc; ++c)
        for (int f = 0; f < 6; ++f) {
            int v[4];
            for (int i = 0; i < 4; ++i) v[i] = m.cells_lex[8 * (size_t)c + kFaceVerts[f][i]];
            std::sort(v, v + 4);
            keys[6 * (size_t)c + f] = {v[1], v[2], v[3], v[0]};   // slot holds v0 until bucketed
            bstart[v[0] + 1]++;
        }
    for (int d = 0; d < nd; ++d) bstart[d + 1] += bstart[d];
    std::vector<FaceKey> bucket(keys.size());
    {
        std::vector<int64_t> fill(bstart.begin(), bstart.end() - 1);
        for (size_t i = 0; i < keys.size(); ++i) {
            const int64_t v0 = keys[i].slot;
            bucket[fill[v0]++] = {keys[i].v1, keys[i].v2, keys[i].v3, (int64_t)i};
        }
    }
    std::vector<unsigned char> is_bd((size_t)6 * nc, 0);
    auto less = [](const FaceKey &a, const FaceKey &b) {
        if (a.v1 != b.v1) return a.v1 < b.v1;
        if (a.v2 != b.v2) return a.v2 < b.v2;
        if (a.v3 != b.v3) return a.v3 < b.v3;
        return a.slot < b.slot;
    };
    for (int d = 0; d < nd; ++d) {
        FaceKey *b0 = bucket.data() + bstart[d], *b1 = bucket.data() + bstart[d + 1];
        std::sort(b0, b1, less);
        for (FaceKey *i = b0; i < b1;) {
            FaceKey *j = i + 1;
            while (j < b1 && j->v1 == i->v1 && j->v2 == i->v2 && j->v3 == i->v3) ++j;
            if (j - i == 1) is_bd[i->slot] = 1;
            else if (j - i > 2) {
                err = "mesh import: a face is shared by more than two cells";
                return FGPU_EINVAL;
            }
            i = j;
        }
    }
    bucket.clear();
    bucket.shrink_to_fit();
    keys.clear();
    keys.shrink_to_fit();
    std::vector<double> ctr;
    double lo[3] = {1e16, 1e16, 1e16}, hi[3] = {-1e16, -1e16, -1e16};
    m.bf_cell.clear();
    m.bf_face.clear();
    for (int c = 0; c < nc; ++c)
        for (int f = 0; f < 6; ++f)
            if (is_bd[6 * (size_t)c + f]) {
                m.bf_cell.push_back(c);
                m.bf_face.push_back(f);
                for (int a = 0; a < 3; ++a) {
                    double s = 0;
                    for (int i = 0; i < 4; ++i) s += m.x[3 * (size_t)m.cells_lex[8 * (size_t)c + kFaceVerts[f][i]] + a];
                    s /= 4.0;
                    ctr.push_back(s);
                    lo[a] = std::min(lo[a], s);
                    hi[a] = std::max(hi[a], s);
                }
            }
    const size_t nb = m.bf_cell.size();
    m.bf_id.resize(nb);
    const double eps = 1e-6;
    for (size_t i = 0; i < nb; ++i) {
        const double *p = &ctr[3 * i];
        const bool side = std::fabs(p[0] - lo[0]) <= eps || std::fabs(p[0] - hi[0]) <= eps ||
                          std::fabs(p[1] - lo[1]) <= eps || std::fabs(p[1] - hi[1]) <= eps;
        if (side) m.bf_id[i] = sides;
        else if (std::fabs(p[2] - hi[2]) <= eps) m.bf_id[i] = top;
        else if (std::fabs(p[2] - lo[2]) <= eps) m.bf_id[i] = bottom;
        else m.bf_id[i] = other;
    }
    // vertex2cell / vertex2node (DealSolver::calc_vertex2dof, src/DealSolver.cpp:318-341): the
    // last cell in cell order that touches the vertex
    m.v2c_cell.assign(nd, 0);
    m.v2c_node.assign(nd, 0);
    for (int c = 0; c < nc; ++c)
        for (int k = 0; k < 8; ++k) {
            const int d = m.cells_lex[8 * (size_t)c + k];
            m.v2c_cell[d] = c;
            m.v2c_node[d] = k;
        }
    return FGPU_OK;
}

static uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

void build_patches(int nc, int nd, const std::vector<int> &icells, int cells_per_patch, Patches &P) {
    P = Patches();
    cells_per_patch = std::max(8, std::min(kPatchCells, cells_per_patch));
    P.cells = cells_per_patch;
    const int cap = std::min(kPatchNodes, 3 * cells_per_patch + 64);
    P.cell0.push_back(0);
    P.node0.push_back(0);
    std::vector<int> stamp(nd, -1), local(nd, 0);
    int c = 0;
    while (c < nc) {
        const int p = P.n, c_start = c;
        int np = 0;
        while (c < nc && c - c_start < cells_per_patch) {
            // distinct vertices of the cell not yet in the patch (a vertex repeated inside a
            // collapsed cell counts once, as in the device import)
            int fresh = 0;
            for (int k = 0; k < 8; ++k) {
                const int d = icells[8 * (size_t)c + k];
                bool repeat = false;
                for (int j = 0; j < k; ++j) repeat = repeat || icells[8 * (size_t)c + j] == d;
                fresh += stamp[d] != p && !repeat;
            }
            if (np + fresh > cap) break;
            for (int k = 0; k < 8; ++k) {
                const int d = icells[8 * (size_t)c + k];
                if (stamp[d] != p) {
                    stamp[d] = p;
                    local[d] = np++;
                    P.nodes.push_back(d);
                }
                P.lidx.push_back((uint16_t)local[d]);
            }
            ++c;
        }
        // the patch's nodes by decreasing number of cell entries (stable): the threads of a
        // wave then sum lists of similar length in the per-node reduction.  Only the thread
        // assignment changes; each node still adds its entries in ascending order.
        {
            const size_t nb = P.node0.back();
            std::vector<int> deg(np, 0), ord(np), newid(np);
            for (size_t e = 8 * (size_t)c_start; e < 8 * (size_t)c; ++e) deg[P.lidx[e]]++;
            for (int j = 0; j < np; ++j) ord[j] = j;
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return deg[a] > deg[b]; });
            std::vector<int> old(P.nodes.begin() + nb, P.nodes.end());
            for (int j = 0; j < np; ++j) {
                newid[ord[j]] = j;
                P.nodes[nb + j] = old[ord[j]];
            }
            for (size_t e = 8 * (size_t)c_start; e < 8 * (size_t)c; ++e) P.lidx[e] = (uint16_t)newid[P.lidx[e]];
        }
        P.cell0.push_back(c);
        P.node0.push_back((int)P.nodes.size());
        P.node_cap = std::max(P.node_cap, np);
        P.n++;
    }
    P.node_cap = (P.node_cap + 7) / 8 * 8;
    // per patch: local node -> its (cell, vertex) entries, ascending
    P.lptr.assign(P.nodes.size() + P.n, 0);
    P.lent.assign(8 * (size_t)nc, 0);
    for (int p = 0; p < P.n; ++p) {
        const int e0 = 8 * P.cell0[p], ne = 8 * (P.cell0[p + 1] - P.cell0[p]);
        const int np = P.node0[p + 1] - P.node0[p];
        uint16_t *ptr = &P.lptr[(size_t)P.node0[p] + p];
        std::vector<int> cnt(np + 1, 0);
        for (int e = 0; e < ne; ++e) cnt[P.lidx[e0 + e] + 1]++;
        for (int j = 0; j < np; ++j) cnt[j + 1] += cnt[j];
        for (int j = 0; j <= np; ++j) ptr[j] = (uint16_t)cnt[j];
        for (int e = 0; e < ne; ++e) P.lent[e0 + cnt[P.lidx[e0 + e]]++] = (uint16_t)e;
    }
}

// patch outputs are stored dof-major: dof i owns positions [gptr[i], gptr[i+1]) of the slot
// buffer, one per patch touching it, patches ascending (deterministic sum)
void build_slots(int nd, Patches &P, std::vector<int> &gptr) {
    gptr.assign(nd + 1, 0);
    for (int d : P.nodes) gptr[d + 1]++;
    for (int i = 0; i < nd; ++i) gptr[i + 1] += gptr[i];
    P.slot.resize(P.nodes.size());
    std::vector<int> pos(gptr.begin(), gptr.end() - 1);
    for (size_t j = 0; j < P.nodes.size(); ++j) P.slot[j] = pos[P.nodes[j]]++;
}

// Layout of each internal cell against the tet-split sub-hex (nodes of a tet vertex, three edge
// midpoints, three face centroids and the centroid, lexicographic): 0 as laid out, 1 when the
// eta flip of the cell's vertices has that form more closely
void cell_layouts(int nc, const std::vector<int> &icells, const std::vector<double> &X, std::vector<unsigned char> &layout) {
    static const double kSub[8][3] = {{0, 0, 0}, {0.5, 0, 0}, {0, 0, 0.5}, {1.0 / 3, 0, 1.0 / 3},
                                      {0, 0.5, 0}, {1.0 / 3, 1.0 / 3, 0}, {0, 1.0 / 3, 1.0 / 3}, {0.25, 0.25, 0.25}};
    static const int kId[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    auto dev = [&](const int *l, const int *perm) {
        double x[8][3];
        for (int n = 0; n < 8; ++n)
            for (int a = 0; a < 3; ++a) x[n][a] = X[4 * (size_t)l[perm[n]] + a];
        double u[3][3];
        for (int a = 0; a < 3; ++a) {
            u[0][a] = 2 * (x[1][a] - x[0][a]);
            u[1][a] = 2 * (x[4][a] - x[0][a]);
            u[2][a] = 2 * (x[2][a] - x[0][a]);
        }
        double size = 0, d = 0;
        for (int i = 0; i < 3; ++i) size += std::sqrt(u[i][0] * u[i][0] + u[i][1] * u[i][1] + u[i][2] * u[i][2]);
        for (int n = 0; n < 8; ++n)
            for (int a = 0; a < 3; ++a)
                d = std::max(d, std::fabs(x[0][a] + kSub[n][0] * u[0][a] + kSub[n][1] * u[1][a] + kSub[n][2] * u[2][a] -
                                          x[n][a]));
        return size > 0 ? d / size : 1e300;
    };
    layout.assign(nc, 0);
    for (int c = 0; c < nc; ++c) {
        const int *l = &icells[8 * (size_t)c];
        layout[c] = dev(l, kEtaFlip) < dev(l, kId) ? 1 : 0;
    }
}

void build_internal_order(Mesh &m, std::vector<int> &icells, std::vector<double> &X, std::vector<int> &gptr,
                          std::vector<int> &spos, std::vector<int> &fgptr, std::vector<int> &fgidx,
                          std::vector<int> &bf_icell) {
    const int nc = m.n_cells, nd = m.n_dofs;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int d = 0; d < nd; ++d)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], m.x[3 * (size_t)d + a]);
            hi[a] = std::max(hi[a], m.x[3 * (size_t)d + a]);
        }
    double span = 0;
    for (int a = 0; a < 3; ++a) span = std::max(span, hi[a] - lo[a]);
    if (span <= 0) span = 1;
    // cells along a Morton curve of their centroids: neighbouring cells share cache lines
    std::vector<uint64_t> key(nc);
    for (int c = 0; c < nc; ++c) {
        uint64_t k = 0;
        for (int a = 0; a < 3; ++a) {
            double s = 0;
            for (int n = 0; n < 8; ++n) s += m.x[3 * (size_t)m.cells_lex[8 * (size_t)c + n] + a];
            s = (s / 8.0 - lo[a]) / span;
            const uint64_t q = (uint64_t)std::min(2097151.0, std::max(0.0, s * 2097152.0));
            k |= spread3(q) << a;
        }
        key[c] = k;
    }
    // (key, cell) pairs sorted: equal keys keep the cell order
    std::vector<std::pair<uint64_t, int>> kc(nc);
    for (int c = 0; c < nc; ++c) kc[c] = {key[c], c};
    std::sort(kc.begin(), kc.end());
    m.cell_of_icell.resize(nc);
    for (int i = 0; i < nc; ++i) m.cell_of_icell[i] = kc[i].second;
    m.icell_of_cell.resize(nc);
    for (int i = 0; i < nc; ++i) m.icell_of_cell[m.cell_of_icell[i]] = i;
    // dofs by first touch along the sorted cells
    m.idof_of_dof.assign(nd, -1);
    m.dof_of_idof.resize(nd);
    int next = 0;
    for (int i = 0; i < nc; ++i)
        for (int k = 0; k < 8; ++k) {
            const int d = m.cells_lex[8 * (size_t)m.cell_of_icell[i] + k];
            if (m.idof_of_dof[d] < 0) {
                m.idof_of_dof[d] = next;
                m.dof_of_idof[next++] = d;
            }
        }
    icells.resize(8 * (size_t)nc);
    for (int i = 0; i < nc; ++i)
        for (int k = 0; k < 8; ++k)
            icells[8 * (size_t)i + k] = m.idof_of_dof[m.cells_lex[8 * (size_t)m.cell_of_icell[i] + k]];
    X.assign(4 * (size_t)nd, 0.0);
    for (int i = 0; i < nd; ++i)
        for (int a = 0; a < 3; ++a) X[4 * (size_t)i + a] = m.x[3 * (size_t)m.dof_of_idof[i] + a];
    build_patches(nc, nd, icells, default_patch_cells(nc), m.patches);
    build_slots(nd, m.patches, gptr);
    spos.clear();
    // dof -> boundary-face slots (4 per face, API face order)
    static const int kFV[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
    const size_t nb = m.bf_cell.size();
    bf_icell.resize(nb);
    fgptr.assign(nd + 1, 0);
    std::vector<int> fdof(4 * nb);
    for (size_t b = 0; b < nb; ++b) {
        bf_icell[b] = m.icell_of_cell[m.bf_cell[b]];
        for (int j = 0; j < 4; ++j) {
            const int d = m.idof_of_dof[m.cells_lex[8 * (size_t)m.bf_cell[b] + kFV[m.bf_face[b]][j]]];
            fdof[4 * b + j] = d;
            fgptr[d + 1]++;
        }
    }
    for (int i = 0; i < nd; ++i) fgptr[i + 1] += fgptr[i];
    fgidx.resize(4 * nb);
    {
        std::vector<int> pos(fgptr.begin(), fgptr.end() - 1);
        for (size_t s = 0; s < 4 * nb; ++s) fgidx[pos[fdof[s]]++] = (int)s;
    }
}

}  // namespace fgpu

// Transfer of the field solution to atoms: host-side tables of the tet/hex mesh (built once
// per mesh) shared by the HIP kernels in fgpu_interp.hip.
#pragma once

#include <cstdint>
#include <vector>

namespace fgpu {

constexpr double kZero = 1e-15;   // InterpolatorCells::zero (include/InterpolatorCells.h:241)
constexpr int kBvhStack = 96;     // traversal stack of the point queries (tree depth <= 64 + 32)
constexpr int kChainW = 64;       // possible results per point of the device chained guess

// Linear-tet tables exactly as LinearTetrahedra::precompute builds them
// (src/InterpolatorCells.cpp:523-629): 1/det0, four Vec4 cofactor rows, centroid, neighbour
// lists (face neighbours in stored order, then the other tets sharing each vertex) and the
// search marker of narrow_search_to (:720-742, 0 = searchable).
struct LinTetHost {
    int n = 0;
    std::vector<double> det0, det, centroid;
    std::vector<int> nbr_ptr, nbr, marker;
    bool point_in_cell(const double p[3], int t) const;
    void bcc(const double p[3], int t, double b[4]) const;     // shape_functions: bcc + 1e-15
};

void build_lintet(LinTetHost &L, int n_nodes, const double *xyz, int nt, const int *tets, const int *tet_nbr,
                  const int *tet_marker, int search_region);

// Linear-triangle tables of the surface faces as LinearTriangles::precompute builds them
// (src/InterpolatorCells.cpp:1585-1637, norms and centroids of TetgenFaces::calc_appendices,
// src/TetgenCells.cpp:387-409): per face 16 doubles {vert0, edge1/det, edge2, pvec/det, norm,
// |edge2|}, centroid, neighbour lists (for each face vertex, the other faces touching it in
// increasing order, repeats kept) and the two tets of the face.
struct LinTriHost {
    int n = 0;
    std::vector<double> tab, centroid;
    std::vector<int> nbr_ptr, nbr, tri2tet, tris;
    // quadrangles (3 per face, quad 3f+k around face vertex k), their centroids, their two
    // hexes (hex 4t+j around vertex j of tet t) and the 6-node faces; empty without quads
    std::vector<int> quads, quad2hex, tri6;
    std::vector<double> qcen;
    bool point_in_cell(const double p[3], int f) const;
    double fast_distance(const double p[3], int f) const;
    void shape_functions(const double p[3], int f, double w[3]) const;
};

void build_lintri(LinTriHost &T, int n_nodes, const double *xyz, int nf, const int *tris, const int *tri2tet);

// Quadrangles of the faces (Tethex convert_triangles, src/Tethex.cpp:1104-1170), their
// centroids (TetgenCells<4>::get_centroid, include/TetgenCells.h:142-151), the quad -> hex map
// of TetgenMesh::calc_quad2hex2quad_mapping (src/TetgenMesh.cpp:613-649) and the 6-node faces
// of QuadraticTriangles::calc_cell (src/InterpolatorCells.cpp:1866-1886).  Throws when a quad
// does not start at its face vertex or that vertex is not in the face's tets.
void set_quads(LinTriHost &T, const double *xyz, const int *node_kind, const int *tets, const int *quads);

// Linear BVH over the tets (Karras 2012 radix tree on 63-bit Morton codes of the centroids).
// Child c >= 0: internal node c; c < 0: leaf ~c.  Boxes are float, rounded outwards and
// widened by a margin, so a point the exact predicate accepts is never pruned.
struct Bvh {
    int n_leaves = 0;
    std::vector<int> child;        // 2 per internal node
    std::vector<float> box;        // 6 per internal node: tet bounding box
    std::vector<float> cbox;       // 6 per internal node: centroid bounding box
    std::vector<float> leafbox;    // 6 per leaf
    std::vector<int> prim;         // leaf -> tet
};

void build_bvh(int n, const std::vector<double> &lo, const std::vector<double> &hi, const std::vector<double> &centroid,
               Bvh &out);

// QuadraticTetrahedra::calc_cell (src/InterpolatorCells.cpp:1151-1174)
void build_quadtet(int nt, const int *tets, const int *hexes, const int *node_kind, std::vector<int> &cells10);

// Chain of SolutionReader::calc_full_interpolation resolved on the host from the per-point
// candidates found on the device (see fgpu_interp.hip): tet per point (negative outside).
void resolve_chain(const LinTetHost &L, int n, const double *pts, const int4 *cand, int rank, int *tet);

// The same chain over the surface faces (InterpolatorCells<3>::locate_cell, every face
// searchable); quads: LinearQuadrangles::locate_cell (src/InterpolatorCells.cpp:1963-1992), the
// face guess |previous quad| / 3 and the result the face's sub-quad of the largest weight.
void resolve_tri_chain(const LinTriHost &T, int n, const double *pts, const int4 *cand, bool quads, int *cell);

// SolutionReader::locate_interp_centroid chain (src/SolutionReader.cpp:67-90, 136-151;
// InterpolatorCells::locate_centroid, src/InterpolatorCells.cpp:309-333): cell per point (the
// previous one kept on a miss) and the cell to average over (-1 on a miss).
void resolve_centroid_chain(const std::vector<double> &cen, int n, const double *pts, const int4 *cand, int *cell,
                            int *use);

// LinearTetrahedra::locate_cell(p, guess) with the point's device candidates for the full scan.
int locate_tet(const LinTetHost &L, const double p[3], int guess, int4 cand);

// The tet LinearTriangles::interp_solution interpolates in (src/InterpolatorCells.cpp:1662-1686).
int surface_tet(const LinTriHost &T, const LinTetHost &L, const double p[3], int tri, int4 tet_cand);

// The hex LinearQuadrangles::interp_solution interpolates in (src/InterpolatorCells.cpp:1925-1942).
int surface_hex(const LinTriHost &T, const LinTetHost &L, const double p[3], int quad, int4 tet_cand);

}  // namespace fgpu

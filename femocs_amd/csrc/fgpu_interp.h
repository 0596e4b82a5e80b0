// Transfer of the field solution to atoms: host-side tables of the tet/hex mesh (built once
// per mesh) shared by the HIP kernels in fgpu_interp.hip.
#pragma once

#include <cstdint>
#include <vector>

namespace fgpu {

constexpr double kZero = 1e-15;   // InterpolatorCells::zero (include/InterpolatorCells.h:241)
constexpr int kBvhStack = 96;     // traversal stack of the point queries (tree depth <= 64 + 32)

// Linear-tet tables exactly as LinearTetrahedra::precompute builds them
// (src/InterpolatorCells.cpp:523-629): 1/det0, four Vec4 cofactor rows, centroid, neighbour
// lists (face neighbours in stored order, then the other tets sharing each vertex) and the
// search marker of narrow_search_to (:720-742, 0 = searchable).
struct LinTetHost {
    int n = 0;
    std::vector<double> det0, det, centroid;
    std::vector<int> nbr_ptr, nbr, marker;
    bool point_in_cell(const double p[3], int t) const;
};

void build_lintet(LinTetHost &L, int n_nodes, const double *xyz, int nt, const int *tets, const int *tet_nbr,
                  const int *tet_marker, int search_region);

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

}  // namespace fgpu

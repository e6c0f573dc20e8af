// partition.hpp -- geometric cluster tree and lattice-H leaf partition (host).
//
// Integer/compare work done once at setup, O(N log N): it stays on the host.
// Semantics follow HACApK at np=1 (one lattice block):
//   cluster tree: midpoint bisection of the longest bounding-box axis,
//                 leaf when size <= leaf       m_HACApK_base.f90:3272-3333
//   bounding boxes and diameters              m_HACApK_base.f90:3191-3229
//   leaf blocks: low rank iff admissible (min diameter <= eta * distance),
//                both sizes > leaf, and not index-adjacent (strong);
//                dense iff a cluster is a leaf or a size <= leaf+1
//                                             m_HACApK_base.f90:2937-2995
//   the root block is split into its 2x2 son blocks (generate_blrleaf at the
//   single lattice block, m_HACApK_base.f90:2372-2430)
//   leaves sorted by (row start, column start) m_HACApK_base.f90:3114-3166
#pragma once
#include <vector>
#include <cstdint>

namespace hbi {

struct Leaf {
    int ltmtx;    // 1 low rank, 2 dense
    int r0, nr;   // rows [r0, r0+nr) in cluster order
    int c0, nc;   // cols [c0, c0+nc)
    int kt;       // rank (low rank), 0 for dense
};

struct Partition {
    std::vector<int> lod;      // cluster position -> mesh element
    std::vector<Leaf> leaves;  // sorted by (r0, c0)
};

// coord: (n x 3) as three arrays x, y, z (y, z may be null -> 0)
Partition build_partition(int n, const double *x, const double *y, const double *z, int leaf, double eta);

}  // namespace hbi

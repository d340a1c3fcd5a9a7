// forest.h -- host description of the cells, leaves, particles and lists a
// plan uploads: one complete octree in the reference's numbering, or a
// batch of independent instances each on its own sparse octree of uniform
// depth (only the cells that hold particles), concatenated into one global
// numbering of cells, leaves and particles.
#pragma once

#include "common.h"

#include <sys/mman.h>

#include <cstdlib>
#include <memory>
#include <new>
#include <utility>
#include <vector>

namespace bb {

// std::allocator that leaves new elements uninitialised: for large tables
// that are filled right after they are sized.  Blocks of 64 MB or more are
// 2 MB aligned and marked for transparent huge pages, so that filling them
// (a few GB of M2L source tables for large batches) takes one page fault
// per 2 MB instead of one per 4 KB.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  static constexpr size_t kLarge = (size_t)64 << 20, kHuge = (size_t)2 << 20;
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
  T *allocate(size_t n)
  {
    const size_t bytes = n * sizeof(T);
    if (bytes < kLarge) return std::allocator<T>::allocate(n);
    void *p = nullptr;
    if (posix_memalign(&p, kHuge, (bytes + kHuge - 1) / kHuge * kHuge) != 0) throw std::bad_alloc();
    madvise(p, bytes, MADV_HUGEPAGE);  // advisory: ignored where unsupported
    return static_cast<T *>(p);
  }
  void deallocate(T *p, size_t n)
  {
    if (n * sizeof(T) < kLarge)
      std::allocator<T>::deallocate(p, n);
    else
      free(p);
  }
  template <class U>
  void construct(U *p) noexcept
  {
    ::new ((void *)p) U;
  }
  template <class U, class... A>
  void construct(U *p, A &&...a)
  {
    ::new ((void *)p) U(std::forward<A>(a)...);
  }
};

struct HostForest {
  int N = 0, ncell = 0, nleaf = 0, maxdepth = 0;
  bool complete = false;            // one complete octree (cells 0..ncell-1 as cell.h:27-28)
  std::vector<int> levsta;          // complete: first cell of each level, [maxdepth + 2]
  std::vector<real> celeng;         // cell length per level from the root, [maxdepth + 1]
  std::vector<real4> center;        // [ncell]
  std::vector<int> celllev;         // [ncell]
  std::vector<int> child;           // [ncell * 8], -1 when absent
  std::vector<int> plist, ploff;    // cells with children, grouped by level; ploff [maxdepth + 1]
  std::vector<int> leafcell, leafsta, leafcnt;  // [nleaf]
  std::vector<real> leafihalf;      // [nleaf] 1 / (half the leaf's length)
  std::vector<int> perm;            // [N] leaf order -> input index
  std::vector<int> leafof;          // [N] leaf of each leaf-ordered particle
  std::vector<int> nbr, nnbr;       // [nleaf * 27] neighbour leaves (self included), [nleaf]
  // interaction counts [ncell]; lists and operator indices [ncell * kPitchInteraction]
  // for a complete tree only (sparse forests read their sources from vsrc)
  std::vector<int> iinter, inter, kidx;
  double pairs = 0, ninter = 0;     // P2P pairs and M2L interactions (levels >= 2), per vector
  std::vector<int> inst_N, inst_depth, inst_p0, inst_leaf0;  // batch instances
  // sparse forests: field cells of levels >= 2 grouped by sibling class, and
  // the source cell of each of the class's 189 M2L offsets (ncell: empty)
  std::vector<int> cls_off, cls_cells;        // [9], [cells of levels >= 2]
  std::vector<int, NoInitAlloc<int>> vsrc;    // [ncell * 189]
  std::vector<int> cellcls;                   // sparse forests: sibling class of each cell [ncell]
};

// The reference's complete tree (anotherTree after build_tree).
int forest_from_tree(HostForest &f, const anotherTree *t);

// ninst instances: N[i] points each, concatenated in pos; depth[i] the
// uniform depth of instance i, or (depth == NULL or depth[i] <= 0) the
// smallest depth >= 2 whose leaves hold at most leafmax points (at most
// kMaxForestDepth).  Every instance sits in the root cell [-L/2, L/2]^3.
constexpr int kMaxForestDepth = 10;
int forest_from_points(HostForest &f, int ninst, const int *N, const real3 *pos, const int *depth, int leafmax,
                       real L, std::vector<int> *depth_out = nullptr);

}  // namespace bb

// Batched inversion of the diagonal tiles of a tiled triangular factor
// (potrs.hip), shared with the plan builders (tiled.hip).
#pragma once
#include <functional>
#include <vector>
#include "dm_internal.h"

namespace dm {

struct TrtriLeaf {
  const void* src;  // the block's origin in the factor tile
  void* dst;        // its origin in the inverse tile
  int lds, ldd, n, pad_;
};

// Plan of the batched inversion of every diagonal tile of a factor with
// square blocks, built once per (factor, right-hand sides) plan: device
// tables for the leaves and the per-level GEMM batches, and the inverse
// tiles (tile k: rows(k) x rows(k), column-major, at inv(k)).
template <typename T>
struct TrtriPlan {
  bool upper = true;
  int mt = 0, nb = 0;
  int64_t m = 0;
  void* mem = nullptr;       // [inverse tiles][level scratch][tables]
  size_t inv_bytes = 0;
  int nleaves = 0;
  const TrtriLeaf* leaves_dev = nullptr;
  struct Launch {
    const GemmJob<T>* jobs = nullptr;
    int njobs = 0;
    long total = 0;
    GemmCfgId cfg{1, 64, 64};
    bool vec = false;
  };
  std::vector<Launch> launches;  // two per level, in order

  std::vector<int> rowsv;  // rows of tile k

  T* inv(int k) const { return (T*)mem + (size_t)k * nb * nb; }
  int rows(int k) const { return rowsv[k]; }
  ~TrtriPlan();
  // tile(k): origin of the k-th diagonal tile of the factor (ld = rows(k))
  int build(bool up, int64_t m_, int nb_, const std::function<const T*(int)>& tile);
  // count tiles of at most nb_ rows: tile k at tile(k), rows(k) x rows(k)
  // with leading dimension ld(k) (the inverse tiles keep ld = rows(k))
  int build_list(bool up, int count, int nb_, const std::function<const T*(int)>& tile,
                 const std::function<int(int)>& rows_of, const std::function<int(int)>& ld_of);
  // Inverse tiles on stream s (a no-op once *abort is set).
  int run(hipStream_t s, const int* abort) const;
};

}  // namespace dm

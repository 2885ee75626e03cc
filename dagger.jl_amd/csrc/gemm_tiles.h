// Workgroup -> (job, C tile) mapping shared by the GEMM kernels.
#pragma once
#include "dm_internal.h"

namespace dm {

// Logical workgroup index such that each XCD receives a contiguous range.
__device__ __forceinline__ int xcd_remap(int bid, int nwg) {
  const int q = nwg >> 3, r = nwg & 7;
  const int xcd = bid & 7, slot = bid >> 3;
  const int base = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
  return base + slot;
}

// Decode the idx-th tile (column by column) of an upper triangle of tiles.
__device__ __forceinline__ void upper_tri_decode(int idx, int& tm, int& tn) {
  int c = (int)((sqrtf(8.0f * (float)idx + 1.0f) - 1.0f) * 0.5f);
  while ((c + 1) * (c + 2) / 2 <= idx) ++c;
  while (c * (c + 1) / 2 > idx) --c;
  tn = c;
  tm = idx - c * (c + 1) / 2;
}

// Finds the job and the C tile (tm, tn) of this workgroup in a batched
// launch (binary search over wg_begin) or a single-job launch (njobs == 0).
template <typename T>
__device__ __forceinline__ void locate_tile(const GemmJob<T>* __restrict__ jobs, int njobs,
                                            const GemmJob<T>& single, GemmJob<T>& job, int& tm,
                                            int& tn) {
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  int local;
  if (njobs == 0) {
    job = single;
    local = wg;
  } else {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].wg_begin <= wg) lo = mid; else hi = mid - 1;
    }
    job = jobs[lo];
    local = wg - job.wg_begin;
  }
  if (job.tri == 0) {
    tm = local % job.tiles_m;
    tn = local / job.tiles_m;
  } else if (job.tri == 1) {
    upper_tri_decode(local, tm, tn);
  } else {
    upper_tri_decode(local, tn, tm);
  }
}

}  // namespace dm

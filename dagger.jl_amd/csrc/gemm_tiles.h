// Workgroup -> (job, C tile) mapping shared by the GEMM kernels.
#pragma once
#include "dm_internal.h"

namespace dm {

// Decode the idx-th tile (column by column) of an upper triangle of tiles.
__device__ __forceinline__ void upper_tri_decode(int idx, int& tm, int& tn) {
  int c = (int)((sqrtf(8.0f * (float)idx + 1.0f) - 1.0f) * 0.5f);
  while ((c + 1) * (c + 2) / 2 <= idx) ++c;
  while (c * (c + 1) / 2 > idx) --c;
  tn = c;
  tm = idx - c * (c + 1) / 2;
}

// The logical tiles of this workgroup in a launch of `total` tiles over
// gridDim.x <= total workgroups: tiles t, t + step, ... below end.
// Workgroup b runs on XCD b % 8 (the dispatcher deals workgroups round-robin
// over the 8 XCDs); each XCD owns a contiguous range of tiles, in proportion
// to its workgroups, so neighbouring tiles of one job share an L2.  With one
// workgroup per tile every workgroup gets exactly one tile.
struct TileRange {
  int t, end, step;
};
__device__ __forceinline__ TileRange tile_range(int total) {
  const int G = gridDim.x, b = blockIdx.x;
  const int x = b & 7, slot = b >> 3;
  const int q = G >> 3, r = G & 7;
  const int nx = q + (x < r ? 1 : 0);                  // workgroups on XCD x
  const int base = x * q + (x < r ? x : r);            // workgroups on XCDs < x
  const int lo = (int)((long)total * base / G);
  const int hi = (int)((long)total * (base + nx) / G);
  return TileRange{lo + slot, hi, nx};
}

// Finds the job and the C tile (tm, tn) of logical tile `wg` in a batched
// launch (binary search over wg_begin) or a single-job launch (njobs == 0).
// packed: SYRK (tri) jobs enumerate only the tiles of their triangle (square
// tiles); otherwise every tile of a tri job is enumerated and the caller
// skips those outside the triangle.
template <typename T>
__device__ __forceinline__ void locate_tile(const GemmJob<T>* __restrict__ jobs, int njobs,
                                            const GemmJob<T>& single, int wg, GemmJob<T>& job,
                                            int& tm, int& tn, bool packed = true) {
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
  if (job.tri == 0 || !packed) {
    tm = local % job.tiles_m;
    tn = local / job.tiles_m;
  } else if (job.tri == 1) {
    upper_tri_decode(local, tm, tn);
  } else {
    upper_tri_decode(local, tn, tm);
  }
}

}  // namespace dm

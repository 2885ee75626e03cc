// LU with partial pivoting (RowMaximum) and without (NoPivot) on a tiled
// matrix: the panel and row-swap tile operations of Dagger's lu!
// (src/array/lu.jl:5-55 NoPivot, :221-371 RowMaximum); the TRSM and GEMM
// tile operations of the same loop nests run on the existing kernels.
//
// Panel (block column k, tiles k..mt-1, a tall Mp x W matrix): as the
// reference's CPU path does (factor_panel_lapack!, :160-204), the panel's
// tiles are gathered into one contiguous column-major buffer, factored by
// LAPACK's getrf semantics (partial pivoting by maximum magnitude, the
// first maximum on ties, a zero pivot sets info and skips the scaling) and
// scattered back.  The factorization is LAPACK getrf2's recursion:
//   left half (recursively); its row swaps applied to the right half;
//   A12 := L11^-1 A12 (unit lower TRSM); A22 -= A21 A12 (GEMM);
//   right half (recursively); its row swaps applied to the left half,
// with leaves of 32 columns factored column by column in one workgroup.
// Row swaps of the other block columns (swaprows_trail!, :239-262) apply
// the panel's pivots in order p = 1..npiv, one thread per column.
#include <algorithm>
#include <limits>
#include "dm_internal.h"

namespace dm {

constexpr int LU_LEAF = 32;       // columns of a leaf panel
constexpr int LU_THREADS = 1024;  // the leaf's one workgroup

// Global row r of a tiled block column starting at tile row 0 of `base`
// (tiles of mb rows, all `cols` wide, each with ld = its row count).
template <typename T>
__device__ __forceinline__ T* tiled_at(T* base, int64_t rows_total, int mb, int cols, int64_t r,
                                       int c) {
  const int64_t t = r / mb;
  const int64_t lr = r - t * mb;
  const int64_t ld = std::min<int64_t>(mb, rows_total - t * mb);
  return base + t * (int64_t)mb * cols + (int64_t)c * ld + lr;
}

// Tiled block-column region <-> contiguous Mp x W buffer (ld = Mp).
template <typename T>
__global__ void panel_pack_kernel(T* __restrict__ dst, T* __restrict__ tiles, int64_t Mp, int W,
                                  int mb, int unpack) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Mp * W) return;
  const int c = (int)(i / Mp);
  const int64_t r = i - (int64_t)c * Mp;
  T* t = tiled_at(tiles, Mp, mb, W, r, c);
  if (unpack) *t = dst[i];
  else dst[i] = *t;
}

// Factor columns [c0, c0 + nc) of the contiguous panel P (Mp x W, ld = Mp),
// rows [c0, Mp): for each column j, pick the pivot (maximum |P[r, j]|,
// r >= j, first on ties) when `pivot`, swap rows j and p within the leaf's
// columns, scale the column below the diagonal by the pivot's reciprocal
// (skipped for a zero pivot, which sets *info to j + 1 if unset) and update
// the rest of the leaf.  piv[j] receives p (0-based, panel rows).
template <typename T>
__global__ __launch_bounds__(LU_THREADS) void getf2_leaf_kernel(T* __restrict__ P, int64_t Mp,
                                                                int c0, int nc,
                                                                int* __restrict__ piv,
                                                                int* __restrict__ info,
                                                                const int* __restrict__ abort,
                                                                int pivot) {
  __shared__ T bval[LU_THREADS / 64];
  __shared__ int64_t bidx[LU_THREADS / 64];
  __shared__ int64_t pv;
  if (abort && *abort) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t ld = Mp;
  for (int j = c0; j < c0 + nc && j < Mp; ++j) {
    T* colj = P + (int64_t)j * ld;
    if (pivot) {
      // per-thread maximum over rows j + tid, j + tid + 1024, ... (first on ties)
      T best = T(-1);
      int64_t bi = Mp;
      for (int64_t r = j + tid; r < Mp; r += LU_THREADS) {
        const T v = fabs(colj[r]);
        if (v > best) { best = v; bi = r; }
      }
      // wave then workgroup reduction: larger value, then smaller row
      for (int off = 32; off > 0; off >>= 1) {
        const T ov = __shfl_down(best, off, 64);
        const int64_t oi = __shfl_down(bi, off, 64);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      if (lane == 0) { bval[wave] = best; bidx[wave] = bi; }
      __syncthreads();
      if (tid == 0) {
        T b = bval[0];
        int64_t x = bidx[0];
        for (int w = 1; w < LU_THREADS / 64; ++w)
          if (bval[w] > b || (bval[w] == b && bidx[w] < x)) { b = bval[w]; x = bidx[w]; }
        if (x >= Mp) x = j;  // an all-NaN column keeps its row
        pv = x;
        piv[j] = (int)x;
      }
      __syncthreads();
      const int64_t p = pv;
      if (p != j)
        for (int c = c0 + tid; c < c0 + nc; c += LU_THREADS) {
          T* a = P + (int64_t)c * ld;
          const T t = a[j];
          a[j] = a[p];
          a[p] = t;
        }
    } else if (tid == 0) {
      piv[j] = j;
    }
    __syncthreads();
    const T d = colj[j];
    if (d == T(0)) {
      if (tid == 0 && info && *info == 0) *info = j + 1;
    } else if (fabs(d) >= std::numeric_limits<T>::min()) {
      const T inv = T(1) / d;  // LAPACK getf2: scale by the reciprocal
      for (int64_t r = j + 1 + tid; r < Mp; r += LU_THREADS) colj[r] *= inv;
    } else {                   // ... or divide when it would overflow
      for (int64_t r = j + 1 + tid; r < Mp; r += LU_THREADS) colj[r] /= d;
    }
    __syncthreads();
    // rank-1 update of the leaf's remaining columns
    const int rest = c0 + nc - (j + 1);
    if (rest > 0) {
      for (int64_t r = j + 1 + tid; r < Mp; r += LU_THREADS) {
        const T l = colj[r];
        for (int c = j + 1; c < c0 + nc; ++c) {
          T* a = P + (int64_t)c * ld;
          a[r] -= l * a[j];
        }
      }
    }
    __syncthreads();
  }
}

// Row swaps of a contiguous panel: for p in [p0, p1) in order, swap rows p
// and piv[p] of columns [c0, c1).  One thread per column.
template <typename T>
__global__ void laswp_contig_kernel(T* __restrict__ P, int64_t ld, int c0, int c1, int p0, int p1,
                                    const int* __restrict__ piv, const int* __restrict__ abort) {
  if (abort && *abort) return;
  const int c = c0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= c1) return;
  T* a = P + (int64_t)c * ld;
  for (int p = p0; p < p1; ++p) {
    const int q = piv[p];
    if (q != p) {
      const T t = a[p];
      a[p] = a[q];
      a[q] = t;
    }
  }
}

// Panel-relative 0-based pivots -> the reference's ipiv (global 1-based
// rows); the panel's first zero pivot (local, 1-based) -> the global info.
__global__ void lu_ipiv_kernel(const int* __restrict__ piv, int npiv, int64_t row0,
                               int* __restrict__ ipiv, const int* __restrict__ local_info,
                               int* __restrict__ info, int* __restrict__ abort) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npiv) ipiv[i] = (int)(row0 + piv[i] + 1);
  if (i == 0 && *local_info != 0) {
    if (info && *info == 0) *info = (int)(row0 + *local_info);
    (void)abort;
  }
}

// Row swaps of a tiled block-column region (rows row0.. of the matrix): for
// p = 0..npiv-1 in order, swap region row p with region row ipiv[p] - 1 -
// row0.  One thread per column (the order of the swaps is kept per column).
template <typename T>
__global__ void laswp_tiled_kernel(T* __restrict__ base, int64_t Mr, int W, int mb,
                                   const int* __restrict__ ipiv, int npiv, int64_t row0,
                                   const int* __restrict__ abort) {
  if (abort && *abort) return;
  const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= W) return;
  for (int p = 0; p < npiv; ++p) {
    const int64_t q = (int64_t)ipiv[p] - 1 - row0;
    if (q != p && q >= 0 && q < Mr) {
      T* a = tiled_at(base, Mr, mb, W, p, c);
      T* b = tiled_at(base, Mr, mb, W, q, c);
      const T t = *a;
      *a = *b;
      *b = t;
    }
  }
}

namespace {
unsigned blocks_of(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// getrf2's recursion on columns [c0, c0 + nc) of the contiguous panel.
template <typename T>
int getrf_rec(T* P, int64_t Mp, int c0, int nc, int* piv, int* linfo, T* ws, int* abort,
              bool pivot, hipStream_t s) {
  if (nc <= LU_LEAF) {
    hipLaunchKernelGGL((getf2_leaf_kernel<T>), dim3(1), dim3(LU_THREADS), 0, s, P, Mp, c0, nc, piv,
                       linfo, abort, pivot ? 1 : 0);
    DM_LAUNCH_CHECK("getf2_leaf_kernel");
    return 0;
  }
  const int n1 = ((nc / 2 + LU_LEAF - 1) / LU_LEAF) * LU_LEAF, n2 = nc - n1;
  int rc = getrf_rec<T>(P, Mp, c0, n1, piv, linfo, ws, abort, pivot, s);
  if (rc) return rc;
  const int64_t ld = Mp;
  if (pivot) {  // left half's swaps on the right half
    hipLaunchKernelGGL((laswp_contig_kernel<T>), dim3(blocks_of(n2, 64)), dim3(64), 0, s, P, ld,
                       c0 + n1, c0 + nc, c0, c0 + n1, piv, abort);
    DM_LAUNCH_CHECK("laswp_contig_kernel");
  }
  T* A11 = P + (int64_t)c0 * ld + c0;
  T* A12 = P + (int64_t)(c0 + n1) * ld + c0;
  T* A21 = P + (int64_t)c0 * ld + c0 + n1;
  T* A22 = P + (int64_t)(c0 + n1) * ld + c0 + n1;
  rc = trsm_full<T>('L', 'L', 'N', 'U', n1, n2, T(1), A11, (int)ld, A12, (int)ld, ws, abort, s);
  if (rc) return rc;
  if (Mp > c0 + n1) {
    rc = gemm_single<T>(false, false, (int)(Mp - c0 - n1), n2, n1, T(-1), A21, (int)ld, A12,
                        (int)ld, T(1), A22, (int)ld, 0, abort, s);
    if (rc) return rc;
  }
  if (Mp > c0 + n1) {
    rc = getrf_rec<T>(P, Mp, c0 + n1, n2, piv, linfo, ws, abort, pivot, s);
    if (rc) return rc;
    if (pivot) {  // right half's swaps on the left half
      hipLaunchKernelGGL((laswp_contig_kernel<T>), dim3(blocks_of(n1, 64)), dim3(64), 0, s, P, ld,
                         c0, c0 + n1, c0 + n1, (int)std::min<int64_t>(c0 + nc, Mp), piv, abort);
      DM_LAUNCH_CHECK("laswp_contig_kernel");
    }
  }
  return 0;
}
}  // namespace

// Workspace elements of a panel factorization: the contiguous panel, the
// TRSM's inverted diagonal blocks, then W + 1 ints (pivots, local info).
template <typename T>
size_t getrf_panel_ws_bytes(int64_t Mp, int W) {
  const size_t panel = sizeof(T) * (size_t)Mp * W;
  const size_t trsm = sizeof(T) * trsm_ws_elems<T>();
  return ((panel + 255) / 256) * 256 + ((trsm + 255) / 256) * 256 + 4 * ((size_t)W + 64);
}

// LU of a tiled panel: the tiles (k..mt-1, k) starting at `tiles` (Mp rows
// in tiles of mb, W columns), pivots to ipiv (global 1-based, row0 = the
// panel's first global row), info to *info (global, first zero pivot).
template <typename T>
int getrf_panel(T* tiles, int64_t Mp, int W, int mb, int64_t row0, int* ipiv, int* info,
                int* abort, bool pivot, void* wsp, hipStream_t s) {
  if (Mp <= 0 || W <= 0) return 0;
  char* w = (char*)wsp;
  T* P = (T*)w;
  const size_t panel = ((sizeof(T) * (size_t)Mp * W + 255) / 256) * 256;
  T* tws = (T*)(w + panel);
  int* piv = (int*)(w + panel + ((sizeof(T) * trsm_ws_elems<T>() + 255) / 256) * 256);
  int* linfo = piv + W;
  const int npiv = (int)std::min<int64_t>(Mp, W);
  int rc = fill_i32_launch(linfo, 0, 1, s);
  if (rc) return rc;
  hipLaunchKernelGGL((panel_pack_kernel<T>), dim3(blocks_of(Mp * W, 256)), dim3(256), 0, s, P,
                     tiles, Mp, W, mb, 0);
  DM_LAUNCH_CHECK("panel_pack_kernel");
  rc = getrf_rec<T>(P, Mp, 0, npiv, piv, linfo, tws, abort, pivot, s);
  if (rc) return rc;
  if (W > npiv) {  // a panel wider than tall: the rest gets the swaps and L11^-1
    if (pivot) {
      hipLaunchKernelGGL((laswp_contig_kernel<T>), dim3(blocks_of(W - npiv, 64)), dim3(64), 0, s,
                         P, Mp, npiv, W, 0, npiv, piv, abort);
      DM_LAUNCH_CHECK("laswp_contig_kernel");
    }
    rc = trsm_full<T>('L', 'L', 'N', 'U', npiv, W - npiv, T(1), P, (int)Mp,
                      P + (int64_t)npiv * Mp, (int)Mp, tws, abort, s);
    if (rc) return rc;
  }
  hipLaunchKernelGGL((panel_pack_kernel<T>), dim3(blocks_of(Mp * W, 256)), dim3(256), 0, s, P,
                     tiles, Mp, W, mb, 1);
  DM_LAUNCH_CHECK("panel_pack_kernel");
  hipLaunchKernelGGL(lu_ipiv_kernel, dim3(blocks_of(std::max(npiv, 1), 256)), dim3(256), 0, s,
                     piv, npiv, row0, ipiv, linfo, info, abort);
  DM_LAUNCH_CHECK("lu_ipiv_kernel");
  return 0;
}

template <typename T>
int laswp_tiled(T* base, int64_t Mr, int W, int mb, const int* ipiv, int npiv, int64_t row0,
                const int* abort, hipStream_t s) {
  if (Mr <= 0 || W <= 0 || npiv <= 0) return 0;
  hipLaunchKernelGGL((laswp_tiled_kernel<T>), dim3(blocks_of(W, 64)), dim3(64), 0, s, base, Mr, W,
                     mb, ipiv, npiv, row0, abort);
  DM_LAUNCH_CHECK("laswp_tiled_kernel");
  return 0;
}

template size_t getrf_panel_ws_bytes<double>(int64_t, int);
template size_t getrf_panel_ws_bytes<float>(int64_t, int);
template int getrf_panel<double>(double*, int64_t, int, int, int64_t, int*, int*, int*, bool, void*,
                                 hipStream_t);
template int getrf_panel<float>(float*, int64_t, int, int, int64_t, int*, int*, int*, bool, void*,
                                hipStream_t);
template int laswp_tiled<double>(double*, int64_t, int, int, const int*, int, int64_t, const int*,
                                 hipStream_t);
template int laswp_tiled<float>(float*, int64_t, int, int, const int*, int, int64_t, const int*,
                                hipStream_t);

}  // namespace dm

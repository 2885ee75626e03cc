// Cholesky factorization of a diagonal tile (LAPACK potrf) on CDNA4.
//
// Replaces LAPACK.potrf! as called by potrf_checked! (src/array/cholesky.jl:
// 2-18, spawned at :39/:84).  info follows LAPACK: 0, or k > 0 when the
// leading minor of order k is not positive definite (pivot <= 0 or NaN).
//
// Structure: recursive 2x2 splitting on the host down to 64-column leaves.
//   U: potrf(A11); A12 := U11^-T A12 (trsm); A22 -= A12^T A12 (syrk); potrf(A22)
//   L: potrf(A11); A21 := A21 L11^-T (trsm); A22 -= A21 A21^T (syrk); potrf(A22)
// The off-diagonal work runs on the MFMA trsm/syrk kernels.  A leaf is one
// wave: the 64x64 block sits in LDS, lane c owns column c (U) or row c (L),
// and each step is a square root, a row scale and a rank-1 update with two
// single-wave barriers.  Every launch first checks the abort flag, so after
// a failed pivot the remaining launches are no-ops.
#include "dm_internal.h"

namespace dm {

constexpr int POTRF_LEAF = 64;

// Factor an n x n (n <= 64) block in place.  The unreferenced triangle is
// not touched.  On failure writes info = info_base + j + 1 (local) and, if
// abort != nullptr, *abort = global_base + j + 1.
template <typename T, bool UPPER>
__global__ __launch_bounds__(64) void potrf_leaf_kernel(T* __restrict__ A, int n, int lda,
                                                        int* __restrict__ info, int info_base,
                                                        int* __restrict__ abort, int global_base) {
  __shared__ T s[POTRF_LEAF][POTRF_LEAF + 1];  // s[r][c], symmetric view
  if (abort && *abort) return;
  if (info && *info) return;
  const int t = threadIdx.x;
  // Load the referenced triangle into the upper view s[r][c], r <= c.
  for (int c = 0; c < n; ++c) {
    if (t < n) {
      const int r = t;
      if (UPPER) {
        if (r <= c) s[r][c] = A[(size_t)c * lda + r];
      } else {
        // lower: element (c, r) with c >= r stored at s[r][c] (transpose)
        if (r <= c) s[r][c] = A[(size_t)r * lda + c];
      }
    }
  }
  __syncthreads();
  int fail = 0;
  for (int j = 0; j < n; ++j) {
    const T d = s[j][j];
    if (!(d > T(0))) { fail = j + 1; break; }
    const T sd = sqrt(d);
    const T inv = T(1) / sd;
    T u = T(0);
    if (t > j && t < n) {
      u = s[j][t] * inv;
      s[j][t] = u;
    }
    __syncthreads();
    if (t == j) s[j][j] = sd;
    if (t > j && t < n) {
      for (int r = j + 1; r <= t; ++r) s[r][t] -= s[j][r] * u;
    }
    __syncthreads();
  }
  // Write back the factored part (LAPACK leaves columns past a failure as
  // partially updated; the same holds here).
  for (int c = 0; c < n; ++c) {
    if (t < n) {
      const int r = t;
      if (r <= c) {
        if (UPPER) A[(size_t)c * lda + r] = s[r][c];
        else A[(size_t)r * lda + c] = s[r][c];
      }
    }
  }
  if (fail && t == 0) {
    if (info) *info = info_base + fail;
    if (abort) *abort = global_base + fail;
  }
}

template <typename T>
static int potrf_rec(bool upper, int n, T* A, int lda, int* info, int info_base, int* abort,
                     int global_base, const int* chk, T* ws, hipStream_t s) {
  if (n <= 0) return 0;
  if (n <= POTRF_LEAF) {
    if (upper)
      hipLaunchKernelGGL((potrf_leaf_kernel<T, true>), dim3(1), dim3(64), 0, s, A, n, lda, info,
                         info_base, abort, global_base);
    else
      hipLaunchKernelGGL((potrf_leaf_kernel<T, false>), dim3(1), dim3(64), 0, s, A, n, lda, info,
                         info_base, abort, global_base);
    DM_LAUNCH_CHECK("potrf_leaf_kernel");
    return 0;
  }
  // split point: half, rounded up to a multiple of the leaf size
  const int n1 = ((n / 2) + POTRF_LEAF - 1) / POTRF_LEAF * POTRF_LEAF;
  const int n2 = n - n1;
  T* A11 = A;
  T* A22 = A + (size_t)n1 * lda + n1;
  int rc = potrf_rec<T>(upper, n1, A11, lda, info, info_base, abort, global_base, chk, ws, s);
  if (rc) return rc;
  if (upper) {
    T* A12 = A + (size_t)n1 * lda;
    rc = trsm_full<T>('L', 'U', 'T', 'N', n1, n2, T(1), A11, lda, A12, lda, ws, chk, s);
    if (rc) return rc;
    rc = gemm_single<T>(true, false, n2, n2, n1, T(-1), A12, lda, A12, lda, T(1), A22, lda, 1, chk,
                        s);
  } else {
    T* A21 = A + n1;
    rc = trsm_full<T>('R', 'L', 'T', 'N', n2, n1, T(1), A11, lda, A21, lda, ws, chk, s);
    if (rc) return rc;
    rc = gemm_single<T>(false, true, n2, n2, n1, T(-1), A21, lda, A21, lda, T(1), A22, lda, 2, chk,
                        s);
  }
  if (rc) return rc;
  return potrf_rec<T>(upper, n2, A22, lda, info, info_base + n1, abort, global_base + n1, chk, ws,
                      s);
}

template <typename T>
int potrf_full(bool upper, int n, T* A, int lda, int* info, int* abort, int global_offset, T* ws,
               hipStream_t s) {
  // The trsm/syrk launches inside the recursion stop once a pivot failed:
  // they watch the plan-level abort flag when there is one, else info.
  const int* chk = abort ? abort : info;
  return potrf_rec<T>(upper, n, A, lda, info, 0, abort, global_offset, chk, ws, s);
}

template int potrf_full<double>(bool, int, double*, int, int*, int*, int, double*, hipStream_t);
template int potrf_full<float>(bool, int, float*, int, int*, int*, int, float*, hipStream_t);

// ------------------------------------------------------------ gemv
// y := alpha op(A) x + beta y.  'N': a workgroup owns 64 rows, its 4 waves
// split the columns and reduce through LDS (coalesced along rows).
// 'T': one wave per output element (column dot product, coalesced).
__global__ __launch_bounds__(256) void dgemv_n_kernel(int m, int n, double alpha,
                                                      const double* __restrict__ A, int lda,
                                                      const double* __restrict__ x, int incx,
                                                      double beta, double* __restrict__ y,
                                                      int incy, const int* __restrict__ abort) {
  __shared__ double part[4][64];
  if (abort && *abort) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  double acc = 0.0;
  if (i < m)
    for (int j = wave; j < n; j += 4) acc += A[(size_t)j * lda + i] * x[(size_t)j * incx];
  part[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && i < m) {
    const double v = alpha * (part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]);
    double* py = y + (size_t)i * incy;
    *py = (beta == 0.0) ? v : v + beta * *py;
  }
}

__global__ __launch_bounds__(256) void dgemv_t_kernel(int m, int n, double alpha,
                                                      const double* __restrict__ A, int lda,
                                                      const double* __restrict__ x, int incx,
                                                      double beta, double* __restrict__ y,
                                                      int incy, const int* __restrict__ abort) {
  if (abort && *abort) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = blockIdx.x * 4 + wave;
  if (j >= n) return;
  double acc = 0.0;
  for (int i = lane; i < m; i += 64) acc += A[(size_t)j * lda + i] * x[(size_t)i * incx];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) {
    const double v = alpha * acc;
    double* py = y + (size_t)j * incy;
    *py = (beta == 0.0) ? v : v + beta * *py;
  }
}

int dgemv_launch(bool trans, int m, int n, double alpha, const double* A, int lda,
                 const double* x, int incx, double beta, double* y, int incy, const int* abort,
                 hipStream_t s) {
  if (m <= 0 || n <= 0) return 0;
  if (!trans) {
    hipLaunchKernelGGL(dgemv_n_kernel, dim3((m + 63) / 64), dim3(256), 0, s, m, n, alpha, A, lda,
                       x, incx, beta, y, incy, abort);
  } else {
    hipLaunchKernelGGL(dgemv_t_kernel, dim3((n + 3) / 4), dim3(256), 0, s, m, n, alpha, A, lda, x,
                       incx, beta, y, incy, abort);
  }
  DM_LAUNCH_CHECK("dgemv_kernel");
  return 0;
}

// ------------------------------------------------------------ fill
__global__ void fill_i32_kernel(int* __restrict__ dst, int value, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = value;
}

int fill_i32_launch(int* dst, int value, size_t count, hipStream_t s) {
  if (count == 0) return 0;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, dst,
                     value, count);
  DM_LAUNCH_CHECK("fill_i32_kernel");
  return 0;
}

__global__ void copy_i32_kernel(int* __restrict__ dst, const int* __restrict__ src, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = src[i];
}

int copy_i32_launch(int* dst, const int* src, size_t count, hipStream_t s) {
  if (count == 0) return 0;
  hipLaunchKernelGGL(copy_i32_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, dst,
                     src, count);
  DM_LAUNCH_CHECK("copy_i32_kernel");
  return 0;
}

// ------------------------------------------------------------ generator
// Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2,
// 3", SC'11), counter = (i lo, i hi, 0, 0), key = (seed lo, seed hi).
// Two 53-bit uniforms in (0,1) from the four words, then Box-Muller.
__device__ __forceinline__ void philox_round(uint32_t (&c)[4], uint32_t k0, uint32_t k1) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k0;
  const uint32_t n2 = hi0 ^ c[3] ^ k1;
  c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

__global__ void randn_kernel(double* __restrict__ x, size_t count, uint64_t seed,
                             uint64_t offset) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint64_t idx = offset + i;
  uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0u, 0u};
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    philox_round(c, k0, k1);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  const double two53 = 9007199254740992.0;
  const double u = ((double)(c[0] >> 5) * 67108864.0 + (double)(c[1] >> 6) + 0.5) / two53;
  const double v = ((double)(c[2] >> 5) * 67108864.0 + (double)(c[3] >> 6) + 0.5) / two53;
  x[i] = sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}

int randn_launch(double* x, size_t count, uint64_t seed, uint64_t offset, hipStream_t s) {
  if (count == 0) return 0;
  hipLaunchKernelGGL(randn_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, x,
                     count, seed, offset);
  DM_LAUNCH_CHECK("randn_kernel");
  return 0;
}

}  // namespace dm

// Cholesky factorization of a diagonal tile (LAPACK potrf) on CDNA4.
//
// Replaces LAPACK.potrf! as called by potrf_checked! (src/array/cholesky.jl:
// 2-18, spawned at :39/:84).  info follows LAPACK: 0, or k > 0 when the
// leading minor of order k is not positive definite (pivot <= 0 or NaN).
//
// Structure: blocked right-looking factorization with 64-wide panels, two
// launches per panel:
//  1. potrf_row_kernel (one wave per 64 columns): the panel row
//     U12 := U11^-T A12 (U) / A21 := A21 L11^-T (L), one column (row) per
//     lane against the factored block held in LDS;
//  2. the trailing update A22 -= U12^T U12 (A21 A21^T for L) on MFMA, fused
//     with the factorization of the next 64 x 64 diagonal block
//     (potrf_update_kernel, gemm.hip).
// The first diagonal block is factored by potrf_diag_kernel (one wave, the
// block in registers: lane c owns column c of U, or row c of L = column c
// of L^T, broadcasting the current row through LDS -- no barriers).
// Every launch first checks the abort flag (and info), so after a failed
// pivot the remaining launches are no-ops.
#include <type_traits>
#include "dm_internal.h"
#include "potrf_diag.h"

namespace dm {


// The diagonal block of panel j0 (see potrf_diag.h), one wave.
template <typename T, bool UPPER, int PW>
__global__ __launch_bounds__(64) void potrf_diag_kernel(T* __restrict__ A, int n, int lda, int j0,
                                                        int* __restrict__ info,
                                                        int* __restrict__ abort,
                                                        int global_base) {
  __shared__ T rowb[2 * 64];
  __shared__ T pv;
  if (abort && *abort) return;
  if (info && *info) return;
  raise_wave_priority();  // always on the critical path of a factorization
  potrf_diag_block<T, UPPER, PW>(A, n, lda, j0, info, abort, global_base, rowb, &pv);
}

// Panel row beside a factored full PW x PW block: U12 := U11^-T A12 (U) or
// A21 := A21 L11^-T (L).  One wave per 64 columns of A12 (rows of A21), one
// column (row) per lane (potrf_row_stage / potrf_row_solve, potrf_diag.h).
template <typename T, bool UPPER, int PW>
__global__ __launch_bounds__(64) void potrf_row_kernel(T* __restrict__ A, int n, int lda, int j0,
                                                       const int* __restrict__ info,
                                                       const int* __restrict__ abort) {
  __shared__ T Us[PW * PW + PW];
  __shared__ T invd[PW];
  if (abort && *abort) return;
  if (info && *info) return;
  raise_wave_priority();
  potrf_row_stage<T, UPPER, PW>(A, lda, j0, Us, invd);
  potrf_row_solve<T, UPPER, PW>(A, n, lda, j0, blockIdx.x * 64, Us, invd);
}

namespace {
template <typename T, int PW>
int potrf_panels(bool upper, int n, T* A, int lda, int* info, int* abort, int global_offset,
                 hipStream_t s) {
  // Launches after a failed pivot are no-ops: they watch the plan-level
  // abort flag when there is one, else info.
  static_assert(PW == 64, "the fused update kernel factors 64-wide diagonal blocks");
  bool diag_done = false;  // the previous panel's fused update factored this block
  for (int j0 = 0; j0 < n; j0 += PW) {
    if (!diag_done) {
      if (upper)
        hipLaunchKernelGGL((potrf_diag_kernel<T, true, PW>), dim3(1), dim3(64), 0, s, A, n, lda,
                           j0, info, abort, global_offset);
      else
        hipLaunchKernelGGL((potrf_diag_kernel<T, false, PW>), dim3(1), dim3(64), 0, s, A, n, lda,
                           j0, info, abort, global_offset);
      DM_LAUNCH_CHECK("potrf_diag_kernel");
    }
    diag_done = false;
    const int kb = n - j0 < PW ? n - j0 : PW;
    const int rest = n - j0 - kb;
    if (rest <= 0) break;  // a block narrower than PW is the last one
    const unsigned nwg = (unsigned)((rest + 63) / 64);
    if (upper)
      hipLaunchKernelGGL((potrf_row_kernel<T, true, PW>), dim3(nwg), dim3(64), 0, s, A, n, lda, j0,
                         info, abort);
    else
      hipLaunchKernelGGL((potrf_row_kernel<T, false, PW>), dim3(nwg), dim3(64), 0, s, A, n, lda,
                         j0, info, abort);
    DM_LAUNCH_CHECK("potrf_row_kernel");
    // trailing update + next diagonal block in one launch
    const int rc = potrf_update_launch<T>(upper, n, A, lda, j0, info, abort, global_offset, s);
    if (rc) return rc;
    diag_done = true;
  }
  return 0;
}
}  // namespace

template <typename T>
int potrf_full(bool upper, int n, T* A, int lda, int* info, int* abort, int global_offset, T* ws,
               hipStream_t s) {
  (void)ws;
  return potrf_panels<T, 64>(upper, n, A, lda, info, abort, global_offset, s);
}

template int potrf_full<double>(bool, int, double*, int, int*, int*, int, double*, hipStream_t);
template int potrf_full<float>(bool, int, float*, int, int*, int*, int, float*, hipStream_t);

// ------------------------------------------------------------ gemv
// y := alpha op(A) x + beta y.  'N': a workgroup owns 64 rows, its 4 waves
// split the columns and reduce through LDS (coalesced along rows).
// 'T': one wave per output element (column dot product, coalesced).
// An empty inner dimension (n = 0 for 'N', m = 0 for 'T') leaves
// y := beta y, as BLAS does (beta == 0 writes zeros without reading y).
template <typename T>
__global__ __launch_bounds__(256) void gemv_n_kernel(int m, int n, T alpha,
                                                     const T* __restrict__ A, int lda,
                                                     const T* __restrict__ x, int incx,
                                                     T beta, T* __restrict__ y,
                                                     int incy, const int* __restrict__ abort) {
  __shared__ T part[4][64];
  if (abort && *abort) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  T acc = T(0);
  if (i < m)
    for (int j = wave; j < n; j += 4) acc += A[(size_t)j * lda + i] * x[(size_t)j * incx];
  part[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && i < m) {
    const T v = alpha * (part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]);
    T* py = y + (size_t)i * incy;
    *py = (beta == T(0)) ? v : v + beta * *py;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void gemv_t_kernel(int m, int n, T alpha,
                                                     const T* __restrict__ A, int lda,
                                                     const T* __restrict__ x, int incx,
                                                     T beta, T* __restrict__ y,
                                                     int incy, const int* __restrict__ abort) {
  if (abort && *abort) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = blockIdx.x * 4 + wave;
  if (j >= n) return;
  T acc = T(0);
  for (int i = lane; i < m; i += 64) acc += A[(size_t)j * lda + i] * x[(size_t)i * incx];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) {
    const T v = alpha * acc;
    T* py = y + (size_t)j * incy;
    *py = (beta == T(0)) ? v : v + beta * *py;
  }
}

template <typename T>
int gemv_launch(bool trans, int m, int n, T alpha, const T* A, int lda, const T* x, int incx,
                T beta, T* y, int incy, const int* abort, hipStream_t s) {
  const int ly = trans ? n : m;  // output length
  if (ly <= 0) return 0;
  // an empty inner dimension still scales y (both kernels then skip the
  // products: the 'N' kernel's column loop and the 'T' kernel's row loop
  // run zero times)
  if (!trans) {
    hipLaunchKernelGGL(gemv_n_kernel<T>, dim3((m + 63) / 64), dim3(256), 0, s, m, n, alpha, A,
                       lda, x, incx, beta, y, incy, abort);
  } else {
    hipLaunchKernelGGL(gemv_t_kernel<T>, dim3((n + 3) / 4), dim3(256), 0, s, m, n, alpha, A, lda,
                       x, incx, beta, y, incy, abort);
  }
  DM_LAUNCH_CHECK("gemv_kernel");
  return 0;
}

template int gemv_launch<double>(bool, int, int, double, const double*, int, const double*, int,
                                 double, double*, int, const int*, hipStream_t);
template int gemv_launch<float>(bool, int, int, float, const float*, int, const float*, int,
                                float, float*, int, const int*, hipStream_t);

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

// Batched uneven SPD solves: one persistent kernel over many systems.
//
// The "batching" form beside the plain drop-in: many SPD systems of uneven
// size (A_m n_m x n_m column-major with ld = n_m, B_m n_m x nrhs_m), each
// factorized A = U^T U (the default cholesky! uplo) and solved in place.
//
// One launch of persistent workgroups (512 threads = 8 waves, one per CU).  A
// workgroup pops the next system from a device work queue (an atomic
// counter over a largest-first order, so long factorizations start early)
// and walks that system's blocked right-looking Cholesky with panel width P
// (per system, chosen by the autotuner per size bucket):
//   1. the P x P diagonal block is factored in LDS (rank-1 steps) and its
//      inverse formed in LDS (one thread per column);
//   2. the panel row U12 = U11^-T A12 is one product with that inverse
//      (one thread per column, the P-vector in registers);
//   3. the trailing upper triangle A22 -= U12^T U12 runs on MFMA (FP64
//      16x16x4 / FP32 16x16x4), 32 x 32 tiles dealt round-robin to waves.
// Then U^T y = b and U x = y by P-row blocks (block update + inverse of the
// diagonal block).  Every workgroup leaves the loop once the counter passes
// the number of systems, so the grid always drains.  info[m] follows LAPACK
// (global order of the first non-positive pivot, 0 on success); a failed
// system is not solved.
#include <cstdlib>
#include <type_traits>
#include "dm_internal.h"

namespace dm {

struct BatchItem {
  int64_t a_off;  // element offset of A_m in the A slab
  int64_t b_off;  // element offset of B_m in the B slab
  int n, nrhs;
  int panel;      // P in {16, 32, 64}
  int pad_;
};

constexpr int BMAXP = 64;
constexpr int BMAXR = 8;  // right-hand sides per system (the batched configuration)


template <typename T> struct BMfma;
template <> struct BMfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * r + lk; }
};
template <> struct BMfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

__device__ __forceinline__ void btri_decode(int idx, int& tm, int& tn) {
  int c = (int)((sqrtf(8.0f * (float)idx + 1.0f) - 1.0f) * 0.5f);
  while ((c + 1) * (c + 2) / 2 <= idx) ++c;
  while (c * (c + 1) / 2 > idx) --c;
  tn = c;
  tm = idx - c * (c + 1) / 2;
}

// iv = inverse of the upper-triangular kb x kb block held in s (one thread
// per column; the lower part of iv is zeroed).
template <typename T>
__device__ __forceinline__ void upper_inverse(const T (*s)[BMAXP + 1], T (*iv)[BMAXP + 1], int kb) {
  const int c = threadIdx.x;
  if (c < kb) {
    for (int r = c; r >= 0; --r) {
      T acc = (r == c) ? T(1) : T(0);
      for (int p = r + 1; p <= c; ++p) acc -= s[r][p] * iv[p][c];
      iv[r][c] = acc / s[r][r];
    }
    for (int r = c + 1; r < kb; ++r) iv[r][c] = T(0);
  }
}

template <typename T, int BNT>
__global__ __launch_bounds__(BNT) void batched_posv_kernel(T* __restrict__ Aall,
                                                           T* __restrict__ Ball,
                                                           const BatchItem* __restrict__ items,
                                                           const int* __restrict__ order,
                                                           int count, int* __restrict__ next,
                                                           int* __restrict__ info,
                                                           int factor_only) {
  typedef typename BMfma<T>::v4 v4;
  __shared__ T s[BMAXP][BMAXP + 1];
  __shared__ T iv[BMAXP][BMAXP + 1];
  __shared__ T tb[BMAXP][BMAXR + 1];
  __shared__ int sidx;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  for (;;) {
    if (tid == 0) sidx = atomicAdd(next, 1);
    __syncthreads();
    const int idx = sidx;
    __syncthreads();
    if (idx >= count) break;
    const int mi = order[idx];
    const BatchItem it = items[mi];
    T* __restrict__ A = Aall + it.a_off;
    const int n = it.n;
    const int P = it.panel < 1 ? 1 : (it.panel > BMAXP ? BMAXP : it.panel);
    const size_t ld = (size_t)n;
    int fail = 0;
    // ---------------------------------------------------------- factor
    for (int k0 = 0; k0 < n; k0 += P) {
      const int kb = min(P, n - k0);
      for (int e = tid; e < kb * kb; e += BNT) {
        const int r = e % kb, c = e / kb;
        if (r <= c) s[r][c] = A[(k0 + r) + (size_t)(k0 + c) * ld];
      }
      __syncthreads();
      for (int j = 0; j < kb; ++j) {
        const T d = s[j][j];
        if (!(d > T(0))) {
          fail = k0 + j + 1;
          break;
        }
        const T sd = sqrt(d);
        if (tid > j && tid < kb) s[j][tid] = s[j][tid] / sd;
        __syncthreads();
        if (tid == j) s[j][j] = sd;
        const int w = kb - j - 1;
        for (int e = tid; e < w * w; e += BNT) {
          const int r = j + 1 + e % w, c = j + 1 + e / w;
          if (r <= c) s[r][c] -= s[j][r] * s[j][c];
        }
        __syncthreads();
      }
      for (int e = tid; e < kb * kb; e += BNT) {
        const int r = e % kb, c = e / kb;
        if (r <= c) A[(k0 + r) + (size_t)(k0 + c) * ld] = s[r][c];
      }
      if (fail) break;  // uniform
      __syncthreads();
      upper_inverse<T>(s, iv, kb);
      __syncthreads();
      // panel row: U12 = iv^T A12, 16 columns per wave step on MFMA:
      // D'[c][r] = sum_p A12(p, c) iv(p, r), lane holds c = row(lk, q), r = lr
      // (so a quarter-wave stores 16 consecutive rows of one column).  Each
      // wave reads its whole chunk before writing it back.
      const int j0 = k0 + kb, w = n - j0;
      for (int c0 = wave * 16; c0 < w; c0 += (BNT / 64) * 16) {
        v4 acc[4];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) acc[rb] = v4{0, 0, 0, 0};
        const int cc = c0 + lr;
        for (int p = 0; p < kb; p += 4) {
          const int pp = p + lk;
          const T a = (pp < kb && cc < w) ? A[(k0 + pp) + (size_t)(j0 + cc) * ld] : T(0);
#pragma unroll
          for (int rb = 0; rb < 4; ++rb) {
            const int r = rb * 16 + lr;
            const T b = (pp < kb && r < kb) ? iv[pp][r] : T(0);
            acc[rb] = BMfma<T>::mac(a, b, acc[rb]);
          }
        }
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int c = c0 + BMfma<T>::row(lk, q);
            const int r = rb * 16 + lr;
            if (r < kb && c < w) A[(k0 + r) + (size_t)(j0 + c) * ld] = acc[rb][q];
          }
      }
      __syncthreads();
      // trailing update of the upper triangle: A22 -= U12^T U12
      const int t32 = (w + 31) / 32;
      const int ntile = t32 * (t32 + 1) / 2;
      for (int t = wave; t < ntile; t += BNT / 64) {
        int ti, tj;
        btri_decode(t, ti, tj);
        const int i0 = j0 + ti * 32, jj0 = j0 + tj * 32;
        // The tile's current values are read first and U12's k-strip in
        // chunks of four k-steps, the next chunk loaded while the current
        // one feeds the MFMAs: sixteen loads in flight instead of four.
        // The k order of the accumulation is unchanged.
        T cold[2][2][4];
#pragma unroll
        for (int hi = 0; hi < 2; ++hi)
#pragma unroll
          for (int hj = 0; hj < 2; ++hj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int i = i0 + hi * 16 + lr;
              const int j = jj0 + hj * 16 + BMfma<T>::row(lk, r);
              cold[hi][hj][r] = (i <= j && j < n) ? A[i + (size_t)j * ld] : T(0);
            }
        v4 acc[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a) acc[a][0] = acc[a][1] = v4{0, 0, 0, 0};
        T xa[4][2][2], xb[4][2][2];  // [k-step][i/j][h]: two chunk buffers
        auto load = [&](T (&x)[4][2][2], int p0) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int pp = p0 + 4 * q + lk;
            const bool pok = pp < kb;
            const size_t pr = (size_t)(k0 + pp);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const int ci = i0 + h * 16 + lr, cj = jj0 + h * 16 + lr;
              x[q][0][h] = (pok && ci < n) ? A[pr + (size_t)ci * ld] : T(0);
              x[q][1][h] = (pok && cj < n) ? A[pr + (size_t)cj * ld] : T(0);
            }
          }
        };
        // D'[a][b] = sum_k X(k, j a) X(k, i b): lane holds j = .. + row(lk,r), i = .. + lr
        auto mma = [&](const T (&x)[4][2][2], int p0) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (p0 + 4 * q >= kb) break;  // uniform
#pragma unroll
            for (int hi = 0; hi < 2; ++hi)
#pragma unroll
              for (int hj = 0; hj < 2; ++hj)
                acc[hi][hj] = BMfma<T>::mac(x[q][1][hj], x[q][0][hi], acc[hi][hj]);
          }
        };
        load(xa, 0);
        for (int p0 = 0; p0 < kb; p0 += 32) {
          if (p0 + 16 < kb) load(xb, p0 + 16);
          mma(xa, p0);
          if (p0 + 32 < kb) load(xa, p0 + 32);
          if (p0 + 16 < kb) mma(xb, p0 + 16);
        }
#pragma unroll
        for (int hi = 0; hi < 2; ++hi)
#pragma unroll
          for (int hj = 0; hj < 2; ++hj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int i = i0 + hi * 16 + lr;
              const int j = jj0 + hj * 16 + BMfma<T>::row(lk, r);
              if (i <= j && j < n) A[i + (size_t)j * ld] = cold[hi][hj][r] - acc[hi][hj][r];
            }
      }
      __syncthreads();
    }
    if (tid == 0) info[mi] = fail;
    const int nrhs = it.nrhs < BMAXR ? it.nrhs : BMAXR;
    if (fail || factor_only || nrhs <= 0) {
      __syncthreads();
      continue;
    }
    T* __restrict__ B = Ball + it.b_off;
    // ---------------------------------------------------------- U^T y = b
    for (int k0 = 0; k0 < n; k0 += P) {
      const int kb = min(P, n - k0);
      for (int e = tid; e < kb * nrhs; e += BNT) {
        const int r = e % kb, c = e / kb;
        const T* ucol = A + (size_t)(k0 + r) * ld;
        const T* bc = B + (size_t)c * ld;
        T acc = bc[k0 + r];
        for (int p = 0; p < k0; ++p) acc -= ucol[p] * bc[p];
        tb[r][c] = acc;
      }
      for (int e = tid; e < kb * kb; e += BNT) {
        const int r = e % kb, c = e / kb;
        if (r <= c) s[r][c] = A[(k0 + r) + (size_t)(k0 + c) * ld];
      }
      __syncthreads();
      upper_inverse<T>(s, iv, kb);
      __syncthreads();
      for (int e = tid; e < kb * nrhs; e += BNT) {
        const int r = e % kb, c = e / kb;
        T acc = T(0);
        for (int p = 0; p <= r; ++p) acc += iv[p][r] * tb[p][c];
        B[(k0 + r) + (size_t)c * ld] = acc;
      }
      __syncthreads();
    }
    // ---------------------------------------------------------- U x = y
    const int nblk = (n + P - 1) / P;
    for (int bi = nblk - 1; bi >= 0; --bi) {
      const int k0 = bi * P;
      const int kb = min(P, n - k0);
      for (int e = tid; e < kb * nrhs; e += BNT) {
        const int r = e % kb, c = e / kb;
        const T* bc = B + (size_t)c * ld;
        T acc = bc[k0 + r];
        for (int q = k0 + kb; q < n; ++q) acc -= A[(k0 + r) + (size_t)q * ld] * bc[q];
        tb[r][c] = acc;
      }
      for (int e = tid; e < kb * kb; e += BNT) {
        const int r = e % kb, c = e / kb;
        if (r <= c) s[r][c] = A[(k0 + r) + (size_t)(k0 + c) * ld];
      }
      __syncthreads();
      upper_inverse<T>(s, iv, kb);
      __syncthreads();
      for (int e = tid; e < kb * nrhs; e += BNT) {
        const int r = e % kb, c = e / kb;
        T acc = T(0);
        for (int p = r; p < kb; ++p) acc += iv[r][p] * tb[p][c];
        B[(k0 + r) + (size_t)c * ld] = acc;
      }
      __syncthreads();
    }
  }
}

}  // namespace dm

using namespace dm;

// Factor (and, unless factor_only, solve) the `count` systems listed in
// order_host (indices into the num_items entries of items_host) in one
// launch.  items/order: host arrays (copied to the device with the queue
// counter); info_dev: num_items ints.  workgroups: persistent grid size
// (0 = one per CU).
extern "C" int dm_batched_posv(int dtype, int num_items, void* A_base, void* B_base,
                               const void* items_host, int count, const int* order_host,
                               int* info_dev, int factor_only, int workgroups, dm_stream_t s) {
  if (dtype != DM_F64 && dtype != DM_F32) return arg_fail(1, "dtype must be DM_F64 or DM_F32");
  if (num_items < 0) return arg_fail(2, "num_items < 0");
  if (count < 0) return arg_fail(6, "count < 0");
  if (count > 0 && (!items_host || !order_host || !info_dev))
    return arg_fail(5, "NULL items/order/info");
  int rc = ensure_init();
  if (rc) return rc;
  if (count == 0) return 0;
  hipStream_t st = stream_of(s);
  const BatchItem* items = (const BatchItem*)items_host;
  for (int i = 0; i < num_items; ++i) {
    if (items[i].n < 0 || items[i].nrhs < 0 || items[i].nrhs > BMAXR)
      return arg_fail(5, "item with n < 0 or nrhs outside [0, 8]");
    if (items[i].panel < 1 || items[i].panel > BMAXP) return arg_fail(5, "panel outside [1, 64]");
  }
  for (int i = 0; i < count; ++i)
    if (order_host[i] < 0 || order_host[i] >= num_items)
      return arg_fail(7, "order entry outside [0, num_items)");
  // device copy: [counter | pad][items][order]
  const size_t ib = sizeof(BatchItem) * (size_t)num_items, ob = sizeof(int) * (size_t)count;
  const size_t bytes = 256 + ib + ob;
  void* mem = nullptr;
  DM_HIP(hipMallocAsync(&mem, bytes, st));
  char* base = (char*)mem;
  DM_HIP(hipMemcpyAsync(base + 256, items, ib, hipMemcpyHostToDevice, st));
  DM_HIP(hipMemcpyAsync(base + 256 + ib, order_host, ob, hipMemcpyHostToDevice, st));
  rc = fill_i32_launch((int*)base, 0, 1, st);
  if (rc) return rc;
  int dev_id = 0, cus = 256;
  (void)hipGetDevice(&dev_id);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id);
  // workgroup size 512 threads (8 waves), one workgroup per CU; measured on
  // 4096 systems n in [32, 2048]: 170 ms against 198 ms for 2 x 256 per CU.
  int grid = workgroups > 0 ? workgroups : cus;
  if (grid > count) grid = count;
  if (dtype == DM_F64)
    hipLaunchKernelGGL((batched_posv_kernel<double, 512>), dim3(grid), dim3(512), 0, st,
                       (double*)A_base, (double*)B_base, (const BatchItem*)(base + 256),
                       (const int*)(base + 256 + ib), count, (int*)base, info_dev, factor_only);
  else
    hipLaunchKernelGGL((batched_posv_kernel<float, 512>), dim3(grid), dim3(512), 0, st,
                       (float*)A_base, (float*)B_base, (const BatchItem*)(base + 256),
                       (const int*)(base + 256 + ib), count, (int*)base, info_dev, factor_only);
  DM_LAUNCH_CHECK("batched_posv_kernel");
  DM_HIP(hipFreeAsync(mem, st));
  return 0;
}

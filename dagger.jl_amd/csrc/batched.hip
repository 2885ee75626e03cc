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
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>
#include "dm_internal.h"
#include "potrf_diag.h"

namespace dm {

struct BatchItem {
  int64_t a_off;  // element offset of A_m in the A slab
  int64_t b_off;  // element offset of B_m in the B slab
  int n, nrhs;
  int panel;      // P in {16, 32, 64}
  int lda;        // A_m's leading dimension (0: n, the packed layout)
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
    const size_t ld = it.lda > 0 ? (size_t)it.lda : (size_t)n;  // A's
    const size_t ldb = (size_t)n;                                 // B's
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
        const T* bc = B + (size_t)c * ldb;
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
        B[(k0 + r) + (size_t)c * ldb] = acc;
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
        const T* bc = B + (size_t)c * ldb;
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
        B[(k0 + r) + (size_t)c * ldb] = acc;
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
    if (items[i].lda != 0 && items[i].lda < items[i].n) return arg_fail(5, "item lda < n");
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

// ===========================================================================
// Large systems of a batch, tiled and level-synchronous.  One workgroup per
// system (above) leaves the largest systems -- which hold most of the flops
// of a log-uniform size mix -- to a few CUs at a time, at the rate of one
// CU's rank-1 and 32 x 32 MFMA steps.  Here every listed system advances
// through the same blocked right-looking Cholesky in lockstep, each step one
// launch over all systems still active at it (systems sorted largest first,
// so those are a prefix of the system table):
//   for each 64-column block jj:
//     1. the 64 x 64 diagonal block of every system is factored by one wave
//        (potrf_diag_block, the tile POTRF's diagonal step);
//     2. its panel row U[jj, jj+64:n] := U_jj^-T A[jj, jj+64:n] over the
//        system's full width, one wave per 64 columns (potrf_row_solve);
//     3. the rows left in the current outer tile of nb columns get the
//        block's update (a batched SYRK-shaped GEMM with K = 64);
//   after an outer tile's last block, the trailing matrix below it gets the
//   whole tile's update A22 -= U12^T U12 in one batched GEMM with K = nb
//   (the bulk of the flops, at the matrix cores' K-blocked rate).  Each
//   system may have its own outer tile (item_nb: picked per size bucket by
//   the host autotuner); the solves use the plan's nb for every system.
// Then U^T y = b and U x = y by 64-row blocks, right-looking, in place on a
// working copy of B: a batched in-register triangular solve of each
// system's diagonal block, then one batched GEMM subtracting the solved
// block from the rows still to solve.  The working copy goes back to B only
// for systems with info = 0, so a failed system's B is left as it was (as
// LAPACK posv does).  A non-positive pivot sets the system's info (its
// global column, 1-based); its later blocks are skipped and its remaining
// updates compute on its own data only.
namespace dm {
namespace {

// dst (rows x cols, leading dimension ldd) = src (leading dimension lds);
// skipped when gate >= 0 and info[gate] != 0.
struct BlockCopy {
  const void* src;
  void* dst;
  int64_t lds, ldd;
  int rows, cols;
  int gate, pad_;
};

template <typename T>
__global__ __launch_bounds__(256) void block_copy_kernel(const BlockCopy* __restrict__ jobs,
                                                         const int* __restrict__ info) {
  const BlockCopy j = jobs[blockIdx.x];
  if (j.gate >= 0 && info[j.gate] != 0) return;
  const auto src = gptr((const T*)j.src);
  const auto dst = gptr((T*)j.dst);
  const int total = j.rows * j.cols;
  for (int e = threadIdx.x; e < total; e += 256) {
    const int r = e % j.rows, c = e / j.rows;
    dst[r + c * j.ldd] = src[r + c * j.lds];
  }
}

// One system of a tiled plan.
struct TiledSys {
  int64_t a_off;  // A_m in the A slab
  int64_t z_off;  // its right-hand sides in the plan's working slab (ld = n)
  int n, ld, nrhs;
  int sys;        // index into the batch (and into info)
  int tnb;        // the system's outer tile of the factorization (multiple of 64)
  int pad_;
};

__global__ void zero_info_kernel(const TiledSys* __restrict__ d, int count,
                                 int* __restrict__ info) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) info[d[i].sys] = 0;
}

// Diagonal block (jj, jj) of every active system, one wave each; a system
// whose info is set (an earlier pivot failed) is skipped, so info keeps the
// first failure: jj + local column.  A full block with columns to its right
// also gets W = U_jj^-T (64 x 64) in slot blockIdx.x of Wall, for the panel
// row U[jj, jj+64:n] := W A[jj, jj+64:n] that follows as a batched GEMM.
template <typename T>
__global__ __launch_bounds__(64) void bpotrf_diag_kernel(T* __restrict__ A,
                                                         const TiledSys* __restrict__ d, int jj,
                                                         int* __restrict__ info,
                                                         T* __restrict__ Wall) {
  __shared__ T rowb[2 * 64];
  __shared__ T pv;
  __shared__ T Us[64 * 64 + 64];
  __shared__ T invd[64];
  const TiledSys s = d[blockIdx.x];
  if (info[s.sys]) return;
  T* As = A + s.a_off;
  potrf_diag_block<T, true, 64>(As, s.n, s.ld, jj, info + s.sys, nullptr, 0, rowb, &pv);
  if (s.n - jj <= 64) return;  // no panel row (uniform)
  __syncthreads();             // the factored block is visible to the whole workgroup
  if (info[s.sys]) return;     // uniform
  potrf_row_stage<T, true, 64>(As, s.ld, jj, Us, invd);
  potrf_row_inverse<T, 64>(Wall + (size_t)blockIdx.x * 64 * 64, 64, Us, invd);
}

// Z := U_k^-T Z (FWD) or U_k^-1 Z for the kb x kb diagonal tile at
// (c0, c0), kb = min(nbk, n - c0), of every active system, in place on its
// right-hand sides, one 64-row sub-block at a time: the sub-block's
// triangle is solved in LDS (wave w: column w, lane i: row i), then the
// rows of the tile still to solve are updated from it.
template <typename T, bool FWD>
__global__ __launch_bounds__(512) void btrsv_kernel(const T* __restrict__ A, T* __restrict__ Z,
                                                    const TiledSys* __restrict__ d, int c0,
                                                    int nbk) {
  __shared__ T Us[64][65];
  __shared__ T ys[BMAXR][64];
  const TiledSys s = d[blockIdx.x];
  const int kb = min(nbk, s.n - c0), nsub = (kb + 63) / 64;
  const T* As = A + s.a_off;
  const size_t ld = (size_t)s.ld;
  T* zb = Z + s.z_off + c0;
  const int w = threadIdx.x >> 6, i = threadIdx.x & 63;
  for (int qq = 0; qq < nsub; ++qq) {
    const int q = FWD ? qq : nsub - 1 - qq;
    const int s0 = q * 64, sb = min(64, kb - s0);
    for (int e = threadIdx.x; e < 64 * 64; e += 512) {
      const int r = e & 63, c = e >> 6;
      Us[r][c] = (r <= c && c < sb) ? As[(size_t)(c0 + s0 + c) * ld + c0 + s0 + r] : T(0);
    }
    __syncthreads();
    if (w < s.nrhs) {
      T* zc = zb + (size_t)w * s.n + s0;
      T z = i < sb ? zc[i] : T(0);
      if (FWD) {  // U^T y = z: row j of y from rows 0..j-1
        for (int j = 0; j < sb; ++j) {
          const T yj = __shfl(z, j, 64) / Us[j][j];
          if (i == j) z = yj;
          else if (i > j) z -= Us[j][i] * yj;
        }
      } else {    // U x = y: row j of x from rows j+1..sb-1
        for (int j = sb - 1; j >= 0; --j) {
          const T xj = __shfl(z, j, 64) / Us[j][j];
          if (i == j) z = xj;
          else if (i < j) z -= Us[i][j] * xj;
        }
      }
      if (i < sb) zc[i] = z;
      ys[w][i] = z;
    }
    __syncthreads();
    // rows of the tile still to solve (below the sub-block for FWD, above it
    // otherwise), 64 at a time, their part of U staged through LDS so the
    // reads run down columns
    const int r0 = FWD ? s0 + sb : 0, r1 = FWD ? kb : s0;
    for (int rc = r0; rc < r1; rc += 64) {
      const int nr = min(64, r1 - rc);
      for (int e = threadIdx.x; e < 64 * 64; e += 512) {
        const int a = e & 63, b = e >> 6;  // a runs down a column of U
        if (FWD)  // Us[k][c] = U(s0 + k, rc + c)
          Us[a][b] = (a < sb && b < nr) ? As[(size_t)(c0 + rc + b) * ld + c0 + s0 + a] : T(0);
        else      // Us[c][k] = U(rc + c, s0 + k)
          Us[a][b] = (a < nr && b < sb) ? As[(size_t)(c0 + s0 + b) * ld + c0 + rc + a] : T(0);
      }
      __syncthreads();
      for (int e = threadIdx.x; e < nr * s.nrhs; e += 512) {
        const int c = e % nr, ww = e / nr;
        T acc = T(0);
        if (FWD)
          for (int k = 0; k < sb; ++k) acc += Us[k][c] * ys[ww][k];
        else
          for (int k = 0; k < sb; ++k) acc += Us[c][k] * ys[ww][k];
        zb[(size_t)ww * s.n + rc + c] -= acc;
      }
      __syncthreads();
    }
  }
}

}  // namespace

struct BatchTiledBase {
  virtual ~BatchTiledBase() {}
  virtual int run(int* info_dev, int factor_only, hipStream_t s) = 0;
};

template <typename T>
struct BatchTiled : BatchTiledBase {
  struct Launch {
    bool ta = false, tb = false, vec = false;
    GemmCfgId cfg{1, 64, 64};
    long total = 0;
    std::vector<GemmJob<T>> jobs;  // host copy (the device table is uploaded once)
    const GemmJob<T>* dev = nullptr;
  };
  struct Copies {
    std::vector<BlockCopy> jobs;
    const BlockCopy* dev = nullptr;
  };
  struct Step {            // 64-column block jj = 64 b
    int count = 0;         // systems with n > jj (a prefix of the system table)
    Launch row;            // panel rows U[jj, jj+64:n] := W A[jj, jj+64:n], in place
    Launch inner, trail;   // rows left in the outer tile; after its last block, A22
  };
  struct Tile {            // outer tile c0 = nb k of the sweeps
    int count = 0;         // systems with right-hand sides and n > c0 (a prefix)
    Launch fwd, bwd;       // updates after the tile's triangular solve
  };
  T* A = nullptr;
  int nb = 256, nsys = 0, nrhs_sys = 0;
  void* data = nullptr;    // working slab of right-hand sides, then the W slots
  void* tables = nullptr;  // system tables, copy jobs, GEMM jobs
  const TiledSys* sys_dev = nullptr;
  const TiledSys* rhs_dev = nullptr;
  T* Z = nullptr;
  T* W = nullptr;          // one 64 x 64 block inverse per system
  std::vector<Step> steps;
  std::vector<Tile> tiles;
  Copies z_copy, x_copy;

  ~BatchTiled() override {
    if (data || tables) (void)hipDeviceSynchronize();
    if (data) (void)hipFree(data);
    if (tables) (void)hipFree(tables);
  }

  static int gemm(const Launch& L, hipStream_t s) {
    if (L.jobs.empty()) return 0;
    GemmJob<T> none{};
    return gemm_launch<T>(L.ta, L.tb, L.vec, L.cfg, L.dev, (int)L.jobs.size(), none, L.total,
                          nullptr, s);
  }
  static int copy(const Copies& C, const int* info, hipStream_t s) {
    if (C.jobs.empty()) return 0;
    const BlockCopy* jobs = C.dev;
    hipLaunchKernelGGL((block_copy_kernel<T>), dim3((unsigned)C.jobs.size()), dim3(256), 0, s,
                       jobs, info);
    DM_LAUNCH_CHECK("block_copy_kernel");
    return 0;
  }

  int run(int* info, int factor_only, hipStream_t s) override {
    if (nsys == 0) return 0;
    const TiledSys* sd = sys_dev;
    const TiledSys* rd = rhs_dev;
    T* a = A;
    T* z = Z;
    hipLaunchKernelGGL(zero_info_kernel, dim3((nsys + 255) / 256), dim3(256), 0, s, sd, nsys,
                       info);
    DM_LAUNCH_CHECK("zero_info_kernel");
    int rc = 0;
    for (size_t b = 0; b < steps.size(); ++b) {
      const Step& S = steps[b];
      const int jj = (int)b * 64, cnt = S.count;
      T* w = W;
      hipLaunchKernelGGL((bpotrf_diag_kernel<T>), dim3(cnt), dim3(64), 0, s, a, sd, jj, info, w);
      DM_LAUNCH_CHECK("bpotrf_diag_kernel");
      if ((rc = gemm(S.row, s))) return rc;
      if ((rc = gemm(S.inner, s))) return rc;
      if ((rc = gemm(S.trail, s))) return rc;
    }
    if (factor_only || nrhs_sys == 0) return 0;
    if ((rc = copy(z_copy, nullptr, s))) return rc;
    const int nbk = nb;
    for (size_t k = 0; k < tiles.size(); ++k) {
      const Tile& K = tiles[k];
      const int c0 = (int)k * nb, cnt = K.count;
      hipLaunchKernelGGL((btrsv_kernel<T, true>), dim3(cnt), dim3(512), 0, s, (const T*)a, z, rd,
                         c0, nbk);
      DM_LAUNCH_CHECK("btrsv_kernel");
      if ((rc = gemm(K.fwd, s))) return rc;
    }
    for (size_t k = tiles.size(); k-- > 0;) {
      const Tile& K = tiles[k];
      const int c0 = (int)k * nb, cnt = K.count;
      hipLaunchKernelGGL((btrsv_kernel<T, false>), dim3(cnt), dim3(512), 0, s, (const T*)a, z,
                         rd, c0, nbk);
      DM_LAUNCH_CHECK("btrsv_kernel");
      if ((rc = gemm(K.bwd, s))) return rc;
    }
    return copy(x_copy, info, s);  // gated by each system's info
  }

  int build(T* A_base, T* B_base, const BatchItem* items, const int* order, int count, int nb_,
            const int* item_nb) {
    A = A_base;
    nb = nb_;
    // systems largest first (stable), so the systems active at a block are a prefix
    std::vector<int> sys(order, order + count);
    std::stable_sort(sys.begin(), sys.end(),
                     [&](int a, int b) { return items[a].n > items[b].n; });
    nsys = (int)sys.size();
    if (nsys == 0) return 0;
    std::vector<TiledSys> td(nsys), rd;
    int64_t z_elems = 0;
    for (int i = 0; i < nsys; ++i) {
      const BatchItem& it = items[sys[i]];
      TiledSys t{};
      t.a_off = it.a_off;
      t.n = it.n;
      t.ld = it.lda > 0 ? it.lda : it.n;
      t.nrhs = it.nrhs;
      t.sys = sys[i];
      t.tnb = item_nb && item_nb[sys[i]] > 0 ? item_nb[sys[i]] : nb;
      if (it.nrhs > 0 && it.n > 0) {
        t.z_off = z_elems;
        z_elems += (int64_t)it.n * it.nrhs;
        rd.push_back(t);
      }
      td[i] = t;
    }
    nrhs_sys = (int)rd.size();
    const size_t z_bytes = (sizeof(T) * (size_t)z_elems + 255) / 256 * 256;
    DM_HIP(hipMalloc(&data, z_bytes + sizeof(T) * 64 * 64 * (size_t)nsys));
    Z = (T*)data;
    W = (T*)((char*)data + z_bytes);
    const int nblk = (td[0].n + 63) / 64;
    steps.assign(nblk, Step());
    auto a_at = [&](const TiledSys& t, int64_t r, int64_t c) {
      return A + t.a_off + r + c * t.ld;
    };
    for (int b = 0; b < nblk; ++b) {
      Step& S = steps[b];
      const int jj = 64 * b;
      while (S.count < nsys && td[S.count].n > jj) ++S.count;
      S.inner.ta = S.trail.ta = true;
      for (int i = 0; i < S.count; ++i) {
        const TiledSys& t = td[i];
        const int nrem = t.n - jj - 64;
        if (nrem <= 0) continue;
        const int tnb = t.tnb;
        const int64_t c0 = (int64_t)(jj / tnb) * tnb;  // the outer tile holding this block
        const bool last_in_tile = jj + 64 == c0 + tnb;
        {  // in place: one workgroup holds all 64 rows of its columns and
           // reads them (K = 64) before it stores (beta = 0: C is not read)
          GemmJob<T> g{};
          g.A = W + (size_t)i * 64 * 64; g.lda = 64;
          g.B = a_at(t, jj, jj + 64); g.ldb = t.ld;
          g.C = a_at(t, jj, jj + 64); g.ldc = t.ld;
          g.M = 64; g.N = nrem; g.K = 64;
          g.alpha = T(1); g.beta = T(0);
          S.row.jobs.push_back(g);
        }
        const int c_end = (int)std::min<int64_t>(c0 + tnb, t.n);
        const int rows = c_end - jj - 64;
        if (rows > 0) {  // rows of this outer tile below the block, K = 64:
          // their upper triangle (a square triangular job) ...
          GemmJob<T> g{};
          g.A = g.B = a_at(t, jj, jj + 64);
          g.lda = g.ldb = t.ld;
          g.C = a_at(t, jj + 64, jj + 64);
          g.ldc = t.ld;
          g.M = rows; g.N = rows; g.K = 64;
          g.alpha = T(-1); g.beta = T(1);
          g.tri = 1;
          S.inner.jobs.push_back(g);
          if (t.n > c_end) {  // ... and the columns right of the outer tile
            g.B = a_at(t, jj, c_end);
            g.C = a_at(t, jj + 64, c_end);
            g.N = t.n - c_end;
            g.tri = 0;
            S.inner.jobs.push_back(g);
          }
        }
        if (last_in_tile && t.n > c0 + tnb) {  // the tile's update of A22: K = tnb
          const int m = (int)(t.n - c0 - tnb);
          GemmJob<T> g{};
          g.A = g.B = a_at(t, c0, c0 + tnb);
          g.lda = g.ldb = t.ld;
          g.C = a_at(t, c0 + tnb, c0 + tnb);
          g.ldc = t.ld;
          g.M = m; g.N = m; g.K = tnb;
          g.alpha = T(-1); g.beta = T(1);
          g.tri = 1;
          S.trail.jobs.push_back(g);
        }
      }
    }
    tiles.assign((td[0].n + nb - 1) / nb, Tile());
    for (size_t k = 0; k < tiles.size(); ++k) {
      Tile& K = tiles[k];
      const int c0 = (int)k * nb;
      while (K.count < nrhs_sys && rd[K.count].n > c0) ++K.count;
      K.fwd.ta = true;
      for (int i = 0; i < K.count; ++i) {
        const TiledSys& t = rd[i];
        const int kb = std::min(nb, t.n - c0), nrem = t.n - c0 - kb;
        T* zs = Z + t.z_off;
        if (nrem > 0) {  // Z[c0+kb:] -= U[c0:c0+kb, c0+kb:]^T Z[c0:c0+kb]
          GemmJob<T> g{};
          g.A = a_at(t, c0, c0 + kb); g.lda = t.ld;
          g.B = zs + c0; g.ldb = t.n;
          g.C = zs + c0 + kb; g.ldc = t.n;
          g.M = nrem; g.N = t.nrhs; g.K = kb;
          g.alpha = T(-1); g.beta = T(1);
          K.fwd.jobs.push_back(g);
        }
        if (c0 > 0) {    // Z[0:c0] -= U[0:c0, c0:c0+kb] Z[c0:c0+kb]
          GemmJob<T> g{};
          g.A = a_at(t, 0, c0); g.lda = t.ld;
          g.B = zs + c0; g.ldb = t.n;
          g.C = zs; g.ldc = t.n;
          g.M = c0; g.N = t.nrhs; g.K = kb;
          g.alpha = T(-1); g.beta = T(1);
          K.bwd.jobs.push_back(g);
        }
      }
    }
    for (const TiledSys& t : rd) {
      BlockCopy c{};
      c.src = B_base + items[t.sys].b_off;
      c.dst = Z + t.z_off;
      c.lds = c.ldd = t.n;
      c.rows = t.n;
      c.cols = t.nrhs;
      c.gate = -1;
      z_copy.jobs.push_back(c);
      c.src = Z + t.z_off;
      c.dst = B_base + items[t.sys].b_off;
      c.gate = t.sys;
      x_copy.jobs.push_back(c);
    }
    // ---- device tables
    std::vector<Launch*> gl;
    for (Step& S : steps)
      for (Launch* L : {&S.inner, &S.trail}) gl.push_back(L);
    std::vector<Launch*> rows;  // the in-place panel rows: 64-row tiles only
    for (Step& S : steps) rows.push_back(&S.row);
    const size_t n_factor = gl.size();
    for (Tile& K : tiles)
      for (Launch* L : {&K.fwd, &K.bwd}) gl.push_back(L);
    const size_t n_pick = gl.size();
    for (Launch* L : rows) gl.push_back(L);
    size_t off = 0;
    auto take = [&](size_t bytes) {
      const size_t o = off;
      off += (bytes + 255) / 256 * 256;
      return o;
    };
    const size_t o_sys = take(sizeof(TiledSys) * td.size());
    const size_t o_rhs = take(sizeof(TiledSys) * std::max<size_t>(rd.size(), 1));
    const size_t o_zc = take(sizeof(BlockCopy) * z_copy.jobs.size());
    const size_t o_xc = take(sizeof(BlockCopy) * x_copy.jobs.size());
    std::vector<size_t> o_gl(gl.size(), 0);
    for (size_t q = 0; q < gl.size(); ++q) {
      Launch& L = *gl[q];
      if (L.jobs.empty()) continue;
      int mM = 0, mN = 0, mK = 0;
      L.vec = true;
      for (const auto& j : L.jobs) {
        mM = std::max(mM, j.M); mN = std::max(mN, j.N); mK = std::max(mK, j.K);
        L.vec = L.vec && gemm_job_vec_ok<T>(L.ta, L.tb, j);
      }
      // the factorization's updates (K = 64 and K = nb, mostly triangular)
      // stay on 128 x 128 tiles: measured 30 ms against 45 ms on 256 x 128
      // tiles for the 4096-system mix (the picker takes the wider tile only
      // for vector-load launches)
      L.cfg = q >= n_pick ? GemmCfgId{1, 64, 64}
                          : gemm_pick_cfg<T>(mM, mN, mK, (int)L.jobs.size(), L.vec && q >= n_factor);
      L.total = gemm_prepare_jobs<T>(L.jobs.data(), (int)L.jobs.size(), L.cfg);
      o_gl[q] = take(sizeof(GemmJob<T>) * L.jobs.size());
    }
    DM_HIP(hipMalloc(&tables, std::max<size_t>(off, 256)));
    char* tb = (char*)tables;
    std::vector<unsigned char> h(off, 0);
    auto put = [&](size_t o, const void* p, size_t bytes) {
      if (bytes) std::memcpy(h.data() + o, p, bytes);
    };
    put(o_sys, td.data(), sizeof(TiledSys) * td.size());
    put(o_rhs, rd.data(), sizeof(TiledSys) * rd.size());
    put(o_zc, z_copy.jobs.data(), sizeof(BlockCopy) * z_copy.jobs.size());
    put(o_xc, x_copy.jobs.data(), sizeof(BlockCopy) * x_copy.jobs.size());
    sys_dev = (const TiledSys*)(tb + o_sys);
    rhs_dev = (const TiledSys*)(tb + o_rhs);
    z_copy.dev = (const BlockCopy*)(tb + o_zc);
    x_copy.dev = (const BlockCopy*)(tb + o_xc);
    for (size_t q = 0; q < gl.size(); ++q) {
      Launch& L = *gl[q];
      if (L.jobs.empty()) continue;
      put(o_gl[q], L.jobs.data(), sizeof(GemmJob<T>) * L.jobs.size());
      L.dev = (const GemmJob<T>*)(tb + o_gl[q]);
    }
    DM_HIP(hipMemcpy(tables, h.data(), off, hipMemcpyHostToDevice));
    return 0;
  }
};

}  // namespace dm

struct dm_batch_plan_s {
  std::unique_ptr<dm::BatchTiledBase> p;
};

extern "C" int dm_batched_tiled_create(dm_batch_plan_t* plan, int dtype, int num_items,
                                       void* A_base, void* B_base, const void* items_host,
                                       int count, const int* order_host, int nb,
                                       const int* item_nb) {
  if (!plan) return arg_fail(1, "plan is NULL");
  *plan = nullptr;
  if (dtype != DM_F64 && dtype != DM_F32) return arg_fail(2, "dtype must be DM_F64 or DM_F32");
  if (num_items < 0) return arg_fail(3, "num_items < 0");
  if (count < 0) return arg_fail(7, "count < 0");
  if (count > 0 && (!items_host || !order_host)) return arg_fail(6, "NULL items/order");
  if (nb < 64 || nb > 1024 || nb % 64) return arg_fail(9, "nb must be a multiple of 64 in [64, 1024]");
  int rc = ensure_init();
  if (rc) return rc;
  const BatchItem* items = (const BatchItem*)items_host;
  for (int i = 0; i < num_items; ++i) {
    if (items[i].n < 0 || items[i].nrhs < 0 || items[i].nrhs > BMAXR)
      return arg_fail(6, "item with n < 0 or nrhs outside [0, 8]");
    if (items[i].lda != 0 && items[i].lda < items[i].n) return arg_fail(6, "item lda < n");
  }
  for (int i = 0; i < count; ++i)
    if (order_host[i] < 0 || order_host[i] >= num_items)
      return arg_fail(8, "order entry outside [0, num_items)");
  if (item_nb)
    for (int i = 0; i < num_items; ++i)
      if (item_nb[i] != 0 && (item_nb[i] < 64 || item_nb[i] > 1024 || item_nb[i] % 64))
        return arg_fail(10, "item_nb entry not 0 or a multiple of 64 in [64, 1024]");
  std::unique_ptr<dm_batch_plan_s> h(new dm_batch_plan_s());
  if (dtype == DM_F64) {
    auto* t = new BatchTiled<double>();
    h->p.reset(t);
    rc = t->build((double*)A_base, (double*)B_base, items, order_host, count, nb, item_nb);
  } else {
    auto* t = new BatchTiled<float>();
    h->p.reset(t);
    rc = t->build((float*)A_base, (float*)B_base, items, order_host, count, nb, item_nb);
  }
  if (rc) return rc;
  *plan = h.release();
  return 0;
}

extern "C" int dm_batched_tiled_run(dm_batch_plan_t plan, int* info_dev, int factor_only,
                                    dm_stream_t s) {
  if (!plan) return arg_fail(1, "plan is NULL");
  if (!info_dev) return arg_fail(2, "info is NULL");
  return plan->p->run(info_dev, factor_only, stream_of(s));
}

extern "C" int dm_batched_tiled_destroy(dm_batch_plan_t plan) {
  delete plan;
  return 0;
}

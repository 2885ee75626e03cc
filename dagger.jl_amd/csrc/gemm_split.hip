// FP32 GEMM / SYRK tile updates computed on the BF16 matrix cores.
//
// Used for the off-diagonal trailing updates of the mixed-precision
// Cholesky (FP32 tiles; the FP64 residual and refinement recover full
// accuracy).  gfx950 runs dense BF16 MFMA at 16x the FP32 MFMA rate, so an
// FP32 product evaluated as a short sum of BF16 products is still several
// times faster than v_mfma_f32_16x16x4f32.
//
// Splitting (Henry, Tang & Heinecke, "Leveraging the bfloat16 Artificial
// Intelligence Datatype For Higher-Precision Computations", ARITH 2019):
//   x = h + m + l,  h = bf16(x), m = bf16(x - h), l = bf16(x - h - m)
// Each difference is exact in FP32, so h + m + l carries 24 significant bits.
//   TERMS = 6 ("bf16x6"): a*b ~ hh + (hm + mh) + (hl + lh + mm)  (FP32-like)
//   TERMS = 3 ("bf16x3"): a*b ~ hh + (hm + mh)                   (~2^-16)
// BF16 x BF16 products are exact in the FP32 accumulator; the small terms
// are issued first.
//
// Kernel shape (wave64, v_mfma_f32_16x16x32_bf16):
//  * 16 waves (4 x 4) over a 256x128 (or 128x128) C tile, BK = 32 (one MFMA
//    k); four waves per SIMD hide LDS and global latency.
//  * Operands must be contiguous along k (C = alpha op(A) op(B) + beta C with
//    op(A) = A^T and op(B) = B, the shapes of the upper tiled Cholesky's
//    SYRK('U','T') and GEMM('T','N')).  FP32 rows go global -> registers,
//    are split into their BF16 terms once per workgroup, and are stored as
//    BF16 planes in LDS (double buffered, 144 KB for 256x128 and three
//    terms), 16-byte chunks XOR-swizzled by row so that the fragment reads
//    of a quarter-wave spread over the banks.  Splitting when the tile is
//    staged (not when each wave reads a fragment) does the conversion once
//    instead of once per wave that consumes the row.
//  * MFMA operand layout (measured on gfx950): lane l holds row l%16,
//    k = 8*(l/16) .. 8*(l/16)+7 of both operands; the 16x16 result D[i][j]
//    sits in lane j + 16*(i/4), item i%4.  The MFMA computes C^T (operand
//    roles swapped) so that a quarter-wave stores 16 consecutive rows of a
//    column-major C.
#include "dm_internal.h"
#include "gemm_tiles.h"

namespace dm {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <int TERMS, int BM_, int BN_>
struct SplitGemm {
  static constexpr int BM = BM_, BN = BN_, BK = 32, WM = 4, WN = 4;
  static constexpr int NT = 64 * WM * WN;
  static constexpr int WTM = BM / WM, WTN = BN / WN;
  static constexpr int FM = WTM / 16, FN = WTN / 16;
  static constexpr int P = TERMS == 6 ? 3 : 2;                // BF16 terms per operand
  static constexpr int A_PLANE = BM * BK, B_PLANE = BN * BK;  // BF16 elements per plane
  static constexpr int STAGE = P * (A_PLANE + B_PLANE);
  static constexpr int LA = BM * BK / (4 * NT);  // float4 loads per thread
  static constexpr int LB = BN * BK / (4 * NT);
  static_assert(LA * 4 * NT == BM * BK && LB * 4 * NT == BN * BK, "tile split");

  // 16-byte chunk c (k = 8c .. 8c+7) of row `row` inside a plane
  static __device__ __forceinline__ int chunk_off(int row, int c) {
    return row * BK + ((c ^ ((row >> 1) & 3)) << 3);
  }

  // Row r (of op(X) with k contiguous: X[k + r*ld]) chunk k .. k+3.
  static __device__ __forceinline__ f4 load4(const float* __restrict__ X, int ld, int r, int k,
                                             int R, int K) {
    if (r >= R) return f4{0, 0, 0, 0};
    const auto p = gptr(X + (size_t)r * ld + k);
    if (k + 3 < K) return *(const DM_GLOBAL_AS f4*)p;
    f4 v{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (k + q < K) v[q] = p[q];
    return v;
  }

  static __device__ __forceinline__ void load(const GemmJob<float>& job, int m0, int n0, int k0,
                                              f4 (&ra)[LA], f4 (&rb)[LB]) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int e = threadIdx.x + i * NT;
      ra[i] = load4(job.A, job.lda, m0 + e / (BK / 4), k0 + (e % (BK / 4)) * 4, job.M, job.K);
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int e = threadIdx.x + i * NT;
      rb[i] = load4(job.B, job.ldb, n0 + e / (BK / 4), k0 + (e % (BK / 4)) * 4, job.N, job.K);
    }
  }

  // Split 4 consecutive k (k = 4 kq ..) of row `row` into the P planes.
  static __device__ __forceinline__ void put(__bf16* __restrict__ planes, int plane, int row,
                                             int kq, f4 x) {
    bf4 h, m, l;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const __bf16 hh = (__bf16)x[e];
      float r = x[e] - (float)hh;
      const __bf16 mm = (__bf16)r;
      h[e] = hh;
      m[e] = mm;
      if (P == 3) {
        r -= (float)mm;
        l[e] = (__bf16)r;
      }
    }
    const int off = chunk_off(row, kq >> 1) + (kq & 1) * 4;
    *(bf4*)(planes + off) = h;
    *(bf4*)(planes + plane + off) = m;
    if (P == 3) *(bf4*)(planes + 2 * plane + off) = l;
  }

  static __device__ __forceinline__ void store(__bf16* __restrict__ st, const f4 (&ra)[LA],
                                               const f4 (&rb)[LB]) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int e = threadIdx.x + i * NT;
      put(st, A_PLANE, e / (BK / 4), e % (BK / 4), ra[i]);
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int e = threadIdx.x + i * NT;
      put(st + P * A_PLANE, B_PLANE, e / (BK / 4), e % (BK / 4), rb[i]);
    }
  }

  // 8 consecutive k of one row, all P terms.
  static __device__ __forceinline__ void frag(const __bf16* __restrict__ planes, int plane, int row,
                                              int lk, bf8 (&t)[P]) {
    const int off = chunk_off(row, lk);
#pragma unroll
    for (int p = 0; p < P; ++p) t[p] = *(const bf8*)(planes + p * plane + off);
  }

  static __device__ __forceinline__ void run(const GemmJob<float>& job, int tm, int tn,
                                             __bf16* smem) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm0 = (wave % WM) * WTM, wn0 = (wave / WM) * WTN;
    const int m0 = tm * BM, n0 = tn * BN;
    const int lr = lane & 15, lk = lane >> 4;

    f4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = f4{0, 0, 0, 0};

    f4 ra[LA], rb[LB];
    const int nk = (job.K + BK - 1) / BK;
    if (nk > 0) {
      load(job, m0, n0, 0, ra, rb);
      store(smem, ra, rb);
    }
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const __bf16* st = smem + (kt & 1) * STAGE;
      const bool more = kt + 1 < nk;
      if (more) load(job, m0, n0, (kt + 1) * BK, ra, rb);
      bf8 af[FM][P], bf[FN][P];
#pragma unroll
      for (int i = 0; i < FM; ++i) frag(st, A_PLANE, wm0 + i * 16 + lr, lk, af[i]);
#pragma unroll
      for (int j = 0; j < FN; ++j) frag(st + P * A_PLANE, B_PLANE, wn0 + j * 16 + lr, lk, bf[j]);
      // (A term, B term) products, smallest first
      constexpr int pa6[6] = {2, 0, 1, 1, 0, 0}, pb6[6] = {0, 2, 1, 0, 1, 0};
      constexpr int pa3[3] = {1, 0, 0}, pb3[3] = {0, 1, 0};
#pragma unroll
      for (int q = 0; q < TERMS; ++q) {
        const int pa = TERMS == 6 ? pa6[q] : pa3[q];
        const int pb = TERMS == 6 ? pb6[q] : pb3[q];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[j][pb], af[i][pa], acc[i][j],
                                                                0, 0, 0);
      }
      if (more) store(smem + ((kt + 1) & 1) * STAGE, ra, rb);
      __syncthreads();
    }

    // lane holds C[m][n] with m = .. + lr, n = .. + 4*lk + r
    const float alpha = job.alpha, beta = job.beta;
    const int tri = job.tri;
    // A tile wholly inside C and the triangle stores without per-element
    // branches; per 16-row strip the old values are loaded before any store.
    const bool whole = m0 + BM <= job.M && n0 + BN <= job.N &&
                       (tri == 0 || (tri == 1 ? m0 + BM - 1 <= n0 : m0 >= n0 + BN - 1));
    if (whole) {
      const size_t ldc = job.ldc;
      auto cbase = gptr(job.C + (size_t)(n0 + wn0 + 4 * lk) * ldc + m0 + wm0 + lr);
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        float old[FN][4];
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            old[j][r] = beta == 0.0f ? 0.0f : cbase[(size_t)(j * 16 + r) * ldc + i * 16];
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float v = alpha * acc[i][j][r];
            cbase[(size_t)(j * 16 + r) * ldc + i * 16] = beta == 0.0f ? v : v + beta * old[j][r];
          }
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int m = m0 + wm0 + i * 16 + lr;
#pragma unroll
      for (int j = 0; j < FN; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = n0 + wn0 + j * 16 + 4 * lk + r;
          bool ok = m < job.M && n < job.N;
          if (tri == 1) ok = ok && m <= n;
          else if (tri == 2) ok = ok && m >= n;
          if (ok) {
            auto c = gptr(job.C + (size_t)n * job.ldc + m);
            const float v = alpha * acc[i][j][r];
            *c = (beta == 0.0f) ? v : v + beta * *c;
          }
        }
      }
    }
  }
};

template <int TERMS, int BM>
__global__ __launch_bounds__(1024) void gemm_split_kernel(const GemmJob<float>* __restrict__ jobs,
                                                          int njobs, GemmJob<float> single,
                                                          int total,
                                                          const int* __restrict__ abort) {
  typedef SplitGemm<TERMS, BM, 128> G;
  __shared__ __attribute__((aligned(16))) __bf16 smem[2 * G::STAGE];
  if (abort && *abort) return;
  if (single.prio) raise_wave_priority();
  constexpr bool packed = BM == 128;  // SYRK jobs enumerate only triangle tiles
  for (TileRange r = tile_range(total); r.t < r.end; r.t += r.step) {
    GemmJob<float> job;
    int tm, tn;
    locate_tile(jobs, njobs, single, r.t, job, tm, tn, packed);
    if (!packed && job.tri) {  // rectangular tiles: skip those outside the triangle
      const int m0 = tm * BM, n0 = tn * 128;
      if (job.tri == 1 ? m0 > n0 + 128 - 1 : n0 > m0 + BM - 1) continue;
    }
    G::run(job, tm, tn, smem);
    __syncthreads();  // the next tile's prologue rewrites LDS
  }
}

}  // namespace

bool gemm_split_ok(bool ta, bool tb, const GemmJob<float>& j) {
  // op(A) = A^T and op(B) = B (both contiguous along k), 16-byte rows.
  if (!ta || tb) return false;
  if (((uintptr_t)j.A & 15) || ((uintptr_t)j.B & 15)) return false;
  if ((j.lda & 3) || (j.ldb & 3)) return false;
  return true;
}

int gemm_split_launch(int terms, int bm, const GemmJob<float>* jobs_dev, int njobs,
                      const GemmJob<float>& single, long total_wg, long grid, const int* abort,
                      hipStream_t s) {
  if (total_wg <= 0) return 0;
  if (total_wg > 0x7fffffffL) return arg_fail(0, "gemm: too many workgroups");
  if (grid <= 0 || grid > total_wg) grid = total_wg;
  if (bm != 128 && bm != 256) return arg_fail(2, "split GEMM tile must be 128 or 256 rows");
  if (terms != 3 && terms != 6) return arg_fail(1, "split GEMM terms must be 3 or 6");
  const dim3 g((unsigned)grid), b(1024);
  if (terms == 6) {
    if (bm == 256)
      hipLaunchKernelGGL((gemm_split_kernel<6, 256>), g, b, 0, s, jobs_dev, njobs, single,
                         (int)total_wg, abort);
    else
      hipLaunchKernelGGL((gemm_split_kernel<6, 128>), g, b, 0, s, jobs_dev, njobs, single,
                         (int)total_wg, abort);
  } else {
    if (bm == 256)
      hipLaunchKernelGGL((gemm_split_kernel<3, 256>), g, b, 0, s, jobs_dev, njobs, single,
                         (int)total_wg, abort);
    else
      hipLaunchKernelGGL((gemm_split_kernel<3, 128>), g, b, 0, s, jobs_dev, njobs, single,
                         (int)total_wg, abort);
  }
  DM_LAUNCH_CHECK("gemm_split_kernel");
  return 0;
}

}  // namespace dm

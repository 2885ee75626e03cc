// Triangular solves with many right-hand sides (BLAS trsm) on CDNA4.
//
// Replaces BLAS.trsm! tile calls at src/array/cholesky.jl:41,86 and
// src/array/trsm.jl:95-183.
//
// Every trsm is reduced to one "effective" left solve  Tm X = alpha V:
//   * Tm is op(A) (side L) or op(A)^T (side R), read from A directly or
//     transposed, lower or upper triangular;
//   * the right-hand-side vectors are the columns (side L) or rows (side R)
//     of B, solved in place.
// The kernel works on 32-row blocks of Tm in solve order.  A workgroup owns
// 16 vectors and keeps their solved part in LDS (so later blocks read it at
// LDS speed and no global write->read ordering is needed).  For each block
// the four waves split the k-range of  alpha*V_I - Tm[I,solved] X[solved]
// on FP64/FP32 MFMA, reduce through LDS, and one MFMA product with the
// pre-inverted 32x32 diagonal block gives X_I.  The diagonal blocks are
// inverted up front by a small kernel (one thread per column, exact
// substitution), so the serial part of each block is a short MFMA chain.
// Lengths above 1024 (FP64) / 2048 (FP32) are split on the host into
// solve / GEMM update / solve.
#include "dm_internal.h"

namespace dm {

template <typename T> struct MfmaT;
// Output layouts (lane l, accumulator item r; 16x16 block D[i][j], j = l%16):
//   f64 16x16x4: i = 4*r + l/16        f32 16x16x4: i = 4*(l/16) + r
// (the f64 layout was measured on gfx950 with a one-hot lane probe).
template <> struct MfmaT<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * r + lk; }
  static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct MfmaT<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
  static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

TrsmShape trsm_shape(char side, char uplo, char transa, char diag) {
  TrsmShape s;
  s.side_l = is_left(side);
  const bool up = is_upper(uplo);
  const bool tr = is_trans(transa);
  if (s.side_l) {
    s.ttrans = tr;
    s.lower = tr ? up : !up;
  } else {
    // X op(A) = B  <=>  op(A)^T X^T = B^T
    s.ttrans = !tr;
    s.lower = tr ? !up : up;
  }
  s.unit = is_unit(diag);
  return s;
}

template <typename T, bool TTRANS>
__device__ __forceinline__ T tm_at(const T* __restrict__ A, int lda, int i, int k) {
  return TTRANS ? A[(size_t)i * lda + k] : A[(size_t)k * lda + i];
}

// inv[b*1024 + c*32 + r] = (Tm_bb^-1)(r, c); rows/cols past L act as identity.
// One workgroup per diagonal block, one thread per column of the inverse.
template <typename T, bool LOWER, bool TTRANS, bool UNIT>
__global__ __launch_bounds__(32) void trinv_blocks_kernel(const T* __restrict__ A, int lda, int L,
                                                          T* __restrict__ inv,
                                                          const int* __restrict__ abort) {
  __shared__ T tb[TRSM_BS][TRSM_BS + 1];  // Tm_bb, row-major
  __shared__ T xs[TRSM_BS][TRSM_BS + 1];  // column c of the inverse in xs[.][c]
  if (abort && *abort) return;
  const int b = blockIdx.x;
  const int c = threadIdx.x;
  const int r0 = b * TRSM_BS;
  const int nb = min(TRSM_BS, L - r0);
  for (int r = 0; r < TRSM_BS; ++r) {
    T v = (r == c) ? T(1) : T(0);
    if (r < nb && c < nb) v = tm_at<T, TTRANS>(A, lda, r0 + r, r0 + c);
    tb[r][c] = v;
    xs[r][c] = T(0);
  }
  __syncthreads();
  if (LOWER) {
    for (int r = c; r < TRSM_BS; ++r) {
      T s = (r == c) ? T(1) : T(0);
      for (int p = c; p < r; ++p) s -= tb[r][p] * xs[p][c];
      xs[r][c] = (UNIT || r >= nb) ? s : s / tb[r][r];
    }
  } else {
    for (int r = c; r >= 0; --r) {
      T s = (r == c) ? T(1) : T(0);
      for (int p = r + 1; p <= c; ++p) s -= tb[r][p] * xs[p][c];
      xs[r][c] = (UNIT || r >= nb) ? s : s / tb[r][r];
    }
  }
  __syncthreads();
  T* out = inv + (size_t)b * TRSM_BS * TRSM_BS;
  for (int r = 0; r < TRSM_BS; ++r) out[(size_t)c * TRSM_BS + r] = xs[r][c];
}

template <typename T, bool LOWER, bool TTRANS, bool SIDE_L>
__global__ __launch_bounds__(256) void trsm_kernel(const TrsmJob<T>* __restrict__ jobs, int njobs,
                                                   TrsmJob<T> single,
                                                   const int* __restrict__ abort) {
  typedef typename MfmaT<T>::v4 v4;
  DM_DYNAMIC_LDS(dm_trsm_smem);
  if (abort && *abort) return;
  TrsmJob<T> job;
  int local;
  if (njobs == 0) {
    job = single;
    local = blockIdx.x;
  } else {
    const int wg = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].wg_begin <= wg) lo = mid; else hi = mid - 1;
    }
    job = jobs[lo];
    local = wg - job.wg_begin;
  }
  const int L = job.L, NV = job.NV;
  const int nblk = (L + TRSM_BS - 1) / TRSM_BS;
  const int Lp = nblk * TRSM_BS;
  T* X = reinterpret_cast<T*>(dm_trsm_smem);      // [Lp][16]
  T* red = X + (size_t)Lp * 16;                    // [4][32][16]
  T* bt = red + 4 * TRSM_BS * 16;                  // [32][16]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int j0 = local * 16;
  const T* __restrict__ A = job.A;
  T* __restrict__ B = job.B;
  const int lda = job.lda, ldb = job.ldb;
  const T alpha = job.alpha;

  for (int e = tid; e < Lp * 16; e += 256) X[e] = T(0);
  __syncthreads();

  for (int bi = 0; bi < nblk; ++bi) {
    const int b = LOWER ? bi : nblk - 1 - bi;
    const int r0 = b * TRSM_BS;
    // ---- phase 1: partial products with already-solved rows
    const int kbeg = LOWER ? 0 : r0 + TRSM_BS;
    const int kend = LOWER ? r0 : L;
    const int nsteps = (kend - kbeg + 3) >> 2;
    v4 acc[2][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) acc[h][0] = acc[h][1] = v4{0, 0, 0, 0};
    const int ra = r0 + lr, rb = r0 + 16 + lr;
    for (int st = wave, par = 0; st < nsteps; st += 4, par ^= 1) {
      const int k = kbeg + 4 * st + lk;
      const bool kok = k < kend;
      const T t0 = (kok && ra < L) ? tm_at<T, TTRANS>(A, lda, ra, k) : T(0);
      const T t1 = (kok && rb < L) ? tm_at<T, TTRANS>(A, lda, rb, k) : T(0);
      const T x = kok ? X[k * 16 + lr] : T(0);
      if (par == 0) {
        acc[0][0] = MfmaT<T>::mac(t0, x, acc[0][0]);
        acc[1][0] = MfmaT<T>::mac(t1, x, acc[1][0]);
      } else {
        acc[0][1] = MfmaT<T>::mac(t0, x, acc[0][1]);
        acc[1][1] = MfmaT<T>::mac(t1, x, acc[1][1]);
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        red[(wave * TRSM_BS + h * 16 + MfmaT<T>::row(lk, r)) * 16 + lr] = acc[h][0][r] + acc[h][1][r];
    __syncthreads();
    // ---- phase 2: bt = alpha*V_I - sum of partials
    for (int e = tid; e < TRSM_BS * 16; e += 256) {
      int i, j;
      if (SIDE_L) { i = e % TRSM_BS; j = e / TRSM_BS; }  // rows contiguous in B
      else { j = e % 16; i = e / 16; }                   // vectors contiguous in B
      const int gi = r0 + i, gj = j0 + j;
      T v = T(0);
      if (gi < L && gj < NV) {
        const T bv = SIDE_L ? B[(size_t)gj * ldb + gi] : B[(size_t)gi * ldb + gj];
        const int q = i * 16 + j;
        v = alpha * bv - (red[q] + red[TRSM_BS * 16 + q] + red[2 * TRSM_BS * 16 + q] +
                          red[3 * TRSM_BS * 16 + q]);
      }
      bt[i * 16 + j] = v;
    }
    __syncthreads();
    // ---- phase 3: X_I = inv(Tm_II) * bt  (waves 0,1 each own 16 rows)
    if (wave < 2) {
      const T* __restrict__ iv = job.inv + (size_t)b * TRSM_BS * TRSM_BS;
      v4 xa = v4{0, 0, 0, 0}, xb = v4{0, 0, 0, 0};
#pragma unroll
      for (int kk = 0; kk < TRSM_BS / 4; kk += 2) {
        const int c0 = 4 * kk + lk, c1 = 4 * (kk + 1) + lk;
        xa = MfmaT<T>::mac(iv[c0 * TRSM_BS + wave * 16 + lr], bt[c0 * 16 + lr], xa);
        xb = MfmaT<T>::mac(iv[c1 * TRSM_BS + wave * 16 + lr], bt[c1 * 16 + lr], xb);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = wave * 16 + MfmaT<T>::row(lk, r);
        const int gi = r0 + i, gj = j0 + lr;
        const T v = xa[r] + xb[r];
        X[gi * 16 + lr] = v;
        if (gi < L && gj < NV) {
          if (SIDE_L) B[(size_t)gj * ldb + gi] = v;
          else B[(size_t)gi * ldb + gj] = v;
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ host side
template <typename T>
int trinv_blocks_launch(const TrsmShape& sh, const T* A, int lda, int L, T* inv,
                        const int* abort, hipStream_t s) {
  const int nblk = (L + TRSM_BS - 1) / TRSM_BS;
  if (nblk <= 0) return 0;
#define DM_TRINV(LO, TT, UN)                                                          \
  hipLaunchKernelGGL((trinv_blocks_kernel<T, LO, TT, UN>), dim3(nblk), dim3(TRSM_BS), 0, s, A, \
                     lda, L, inv, abort)
  if (sh.lower) {
    if (sh.ttrans) { if (sh.unit) DM_TRINV(true, true, true); else DM_TRINV(true, true, false); }
    else { if (sh.unit) DM_TRINV(true, false, true); else DM_TRINV(true, false, false); }
  } else {
    if (sh.ttrans) { if (sh.unit) DM_TRINV(false, true, true); else DM_TRINV(false, true, false); }
    else { if (sh.unit) DM_TRINV(false, false, true); else DM_TRINV(false, false, false); }
  }
#undef DM_TRINV
  DM_LAUNCH_CHECK("trinv_blocks_kernel");
  return 0;
}

namespace {
template <typename T>
size_t trsm_smem_bytes(int L) {
  const int Lp = ((L + TRSM_BS - 1) / TRSM_BS) * TRSM_BS;
  return sizeof(T) * ((size_t)Lp * 16 + 4 * TRSM_BS * 16 + TRSM_BS * 16);
}
template <typename T, bool LO, bool TT, bool SL>
int trsm_go(const TrsmJob<T>* jobs, int njobs, const TrsmJob<T>& single, long total, size_t smem,
            const int* abort, hipStream_t s) {
  hipLaunchKernelGGL((trsm_kernel<T, LO, TT, SL>), dim3((unsigned)total), dim3(256), smem, s, jobs,
                     njobs, single, abort);
  DM_LAUNCH_CHECK("trsm_kernel");
  return 0;
}
template <typename T, bool LO, bool TT, bool SL>
int trsm_attr() {
  DM_HIP(hipFuncSetAttribute((const void*)trsm_kernel<T, LO, TT, SL>,
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)trsm_smem_bytes<T>(trsm_max_len<T>())));
  return 0;
}
template <typename T>
int trsm_attr_all() {
  int rc = 0;
  rc = rc ? rc : trsm_attr<T, true, true, true>();
  rc = rc ? rc : trsm_attr<T, true, true, false>();
  rc = rc ? rc : trsm_attr<T, true, false, true>();
  rc = rc ? rc : trsm_attr<T, true, false, false>();
  rc = rc ? rc : trsm_attr<T, false, true, true>();
  rc = rc ? rc : trsm_attr<T, false, true, false>();
  rc = rc ? rc : trsm_attr<T, false, false, true>();
  rc = rc ? rc : trsm_attr<T, false, false, false>();
  return rc;
}
}  // namespace

// Raises the dynamic LDS limit of every trsm kernel (called once at init,
// outside any stream capture).
int trsm_init_attributes() {
  int rc = trsm_attr_all<double>();
  return rc ? rc : trsm_attr_all<float>();
}

template <typename T>
int trsm_launch(const TrsmShape& sh, const TrsmJob<T>* jobs_dev, int njobs,
                const TrsmJob<T>& single, long total_wg, const int* abort, hipStream_t s) {
  if (total_wg <= 0) return 0;
  // LDS is sized for the longest job; batched callers pass the max L in single.L.
  const size_t smem = trsm_smem_bytes<T>(single.L);
  if (sh.lower) {
    if (sh.ttrans) {
      if (sh.side_l) return trsm_go<T, true, true, true>(jobs_dev, njobs, single, total_wg, smem, abort, s);
      return trsm_go<T, true, true, false>(jobs_dev, njobs, single, total_wg, smem, abort, s);
    }
    if (sh.side_l) return trsm_go<T, true, false, true>(jobs_dev, njobs, single, total_wg, smem, abort, s);
    return trsm_go<T, true, false, false>(jobs_dev, njobs, single, total_wg, smem, abort, s);
  }
  if (sh.ttrans) {
    if (sh.side_l) return trsm_go<T, false, true, true>(jobs_dev, njobs, single, total_wg, smem, abort, s);
    return trsm_go<T, false, true, false>(jobs_dev, njobs, single, total_wg, smem, abort, s);
  }
  if (sh.side_l) return trsm_go<T, false, false, true>(jobs_dev, njobs, single, total_wg, smem, abort, s);
  return trsm_go<T, false, false, false>(jobs_dev, njobs, single, total_wg, smem, abort, s);
}

// Pointer to Tm[r, c] sub-block origin and whether it reads as op 'T' on A.
template <typename T>
static inline const T* tm_sub(const TrsmShape& sh, const T* A, int lda, int r, int c) {
  return sh.ttrans ? A + (size_t)r * lda + c : A + (size_t)c * lda + r;
}

// Effective solve of L-long vectors with host recursion over trsm_max_len.
template <typename T>
static int trsm_eff(const TrsmShape& sh, int L, int NV, T alpha, const T* A, int lda, T* B,
                    int ldb, T* ws, const int* abort, hipStream_t s) {
  if (L <= 0 || NV <= 0) return 0;
  if (L <= trsm_max_len<T>()) {
    if (!ws) return arg_fail(0, "trsm: no workspace");
    T* inv = ws;
    int rc = trinv_blocks_launch<T>(sh, A, lda, L, inv, abort, s);
    if (rc) return rc;
    TrsmJob<T> j{};
    j.A = A; j.B = B; j.inv = inv; j.L = L; j.NV = NV; j.lda = lda; j.ldb = ldb; j.alpha = alpha;
    return trsm_launch<T>(sh, nullptr, 0, j, trsm_wgs(NV), abort, s);
  }
  // Split at a multiple of the block size: first solve the leading part in
  // solve order, update the rest, then solve the rest.
  int L1 = ((L / 2) + TRSM_BS - 1) / TRSM_BS * TRSM_BS;
  const int L2 = L - L1;
  // Vectors sub-block origin: side L -> row offset, side R -> column offset.
  auto vsub = [&](int r) -> T* { return sh.side_l ? B + r : B + (size_t)r * ldb; };
  // Lower: leading L1 rows first; upper: trailing L1 rows first.
  const int first_n = L1, second_n = L2;
  const int first_r = sh.lower ? 0 : L2;
  const int second_r = sh.lower ? L1 : 0;
  int rc = trsm_eff<T>(sh, first_n, NV, alpha, tm_sub(sh, A, lda, first_r, first_r), lda,
                       vsub(first_r), ldb, ws, abort, s);
  if (rc) return rc;
  // V2 := alpha*V2 - Tm[second, first] X1
  const T* t21 = tm_sub(sh, A, lda, second_r, first_r);
  if (sh.side_l) {
    // rows: B[second] (second_n x NV) -= Tm21 (second_n x first_n) * X1 (first_n x NV)
    rc = gemm_single<T>(sh.ttrans, false, second_n, NV, first_n, T(-1), t21, lda, vsub(first_r),
                        ldb, alpha, vsub(second_r), ldb, 0, abort, s);
  } else {
    // columns: B[:, second] (NV x second_n) -= X1^T (NV x first_n) * Tm21^T (first_n x second_n)
    rc = gemm_single<T>(false, !sh.ttrans, NV, second_n, first_n, T(-1), vsub(first_r), ldb, t21,
                        lda, alpha, vsub(second_r), ldb, 0, abort, s);
  }
  if (rc) return rc;
  return trsm_eff<T>(sh, second_n, NV, T(1), tm_sub(sh, A, lda, second_r, second_r), lda,
                     vsub(second_r), ldb, ws, abort, s);
}

template <typename T>
int trsm_full(char side, char uplo, char transa, char diag, int m, int n, T alpha, const T* A,
              int lda, T* B, int ldb, T* ws, const int* abort, hipStream_t s) {
  const TrsmShape sh = trsm_shape(side, uplo, transa, diag);
  const int L = sh.side_l ? m : n;
  const int NV = sh.side_l ? n : m;
  if (m <= 0 || n <= 0) return 0;
  if (alpha == T(0)) {
    // B := 0 (BLAS: A not referenced)
    return gemm_single<T>(false, false, m, n, 0, T(0), B, ldb, B, ldb, T(0), B, ldb, 0, abort, s);
  }
  return trsm_eff<T>(sh, L, NV, alpha, A, lda, B, ldb, ws, abort, s);
}

template int trinv_blocks_launch<double>(const TrsmShape&, const double*, int, int, double*,
                                         const int*, hipStream_t);
template int trinv_blocks_launch<float>(const TrsmShape&, const float*, int, int, float*,
                                        const int*, hipStream_t);
template int trsm_launch<double>(const TrsmShape&, const TrsmJob<double>*, int,
                                 const TrsmJob<double>&, long, const int*, hipStream_t);
template int trsm_launch<float>(const TrsmShape&, const TrsmJob<float>*, int,
                                const TrsmJob<float>&, long, const int*, hipStream_t);
template int trsm_full<double>(char, char, char, char, int, int, double, const double*, int,
                               double*, int, double*, const int*, hipStream_t);
template int trsm_full<float>(char, char, char, char, int, int, float, const float*, int, float*,
                              int, float*, const int*, hipStream_t);

}  // namespace dm

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
// Solves longer than 256 are split on the host at a multiple of 32 from the
// top of the triangle: solve the first part, update the second part with one
// (batched) GEMM on the MFMA GEMM kernel, solve the second part -- so most of
// the flops of a long solve run at GEMM speed.
// A leaf (L <= 256) is one kernel: a workgroup owns 64 vectors, each of its
// four waves 16 of them, with no cross-wave synchronisation.  A wave keeps its
// solved part in LDS and walks 32-row blocks of Tm in solve order: MFMA
// products of the block's rows with the solved part, then one MFMA product
// with the pre-inverted 32x32 diagonal block.  The diagonal blocks are
// inverted up front by a small kernel (one thread per column, exact
// substitution), so the serial part of each block is a short MFMA chain.
#include "dm_internal.h"
#include "trsm_plan.h"

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

template <typename T, bool TTRANS>
__device__ __forceinline__ T tm_at(const T* __restrict__ A, int lda, int i, int k) {
  const auto g = gptr(A);
  return TTRANS ? g[(size_t)i * lda + k] : g[(size_t)k * lda + i];
}

// inv[b*1024 + c*32 + r] = (Tm_bb^-1)(r, c); rows/cols past L act as identity.
// One workgroup per diagonal block, one thread per column of the inverse.
template <typename T, bool LOWER, bool TTRANS, bool UNIT>
__global__ __launch_bounds__(32) void trinv_blocks_kernel(const T* __restrict__ A, int lda, int L,
                                                          T* __restrict__ inv,
                                                          const int* __restrict__ abort,
                                                          int prio) {
  __shared__ T tb[TRSM_BS][TRSM_BS + 1];  // Tm_bb, row-major
  __shared__ T xs[TRSM_BS][TRSM_BS + 1];  // column c of the inverse in xs[.][c]
  if (abort && *abort) return;
  if (prio) raise_wave_priority();
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

// Orders this wave's LDS writes before its later LDS reads.
__device__ __forceinline__ void trsm_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, bool LOWER, bool TTRANS, bool SIDE_L>
__global__ __launch_bounds__(256) void trsm_kernel(const TrsmJob<T>* __restrict__ jobs, int njobs,
                                                   TrsmJob<T> single, int total,
                                                   const int* __restrict__ abort) {
  typedef typename MfmaT<T>::v4 v4;
  DM_DYNAMIC_LDS(dm_trsm_smem);
  if (abort && *abort) return;
  if (single.prio) raise_wave_priority();
  // A launch may have fewer workgroups than units of 64 vectors: each
  // workgroup then loops over its share (waves never synchronise, so a wave
  // simply moves on to its next unit).
  for (int wg = blockIdx.x; wg < total; wg += gridDim.x) {
  TrsmJob<T> job;
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
  // The LDS layout follows the launch's longest solve (single.L).
  const int Lmax = ((single.L + TRSM_BS - 1) / TRSM_BS) * TRSM_BS;
  const int L = job.L, NV = job.NV;
  const int nblk = (L + TRSM_BS - 1) / TRSM_BS;
  const int Lp = nblk * TRSM_BS;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  T* X = reinterpret_cast<T*>(dm_trsm_smem) + (size_t)wave * (Lmax + TRSM_BS) * 16;  // [Lp][16]
  T* bt = X + (size_t)Lmax * 16;                                                     // [32][16]
  const int j0 = local * TRSM_VPW + wave * 16;  // first vector of this wave
  if (j0 >= NV) continue;                       // waves never synchronise with each other
  const T* __restrict__ A = job.A;
  const auto B = gptr(job.B);
  const int lda = job.lda, ldb = job.ldb;
  const T alpha = job.alpha;
  const int gj = j0 + lr;  // this lane's vector in the MFMA output layout

  for (int e = lane; e < Lp * 16; e += 64) X[e] = T(0);
  trsm_wave_sync();

  for (int bi = 0; bi < nblk; ++bi) {
    const int b = LOWER ? bi : nblk - 1 - bi;
    const int r0 = b * TRSM_BS;
    // ---- acc = Tm[I, solved] X[solved]  (rows r0 + 0..15 and r0 + 16..31)
    const int kbeg = LOWER ? 0 : r0 + TRSM_BS;
    const int kend = LOWER ? r0 : L;
    const int nsteps = (kend - kbeg + 3) >> 2;
    v4 acc[2][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) acc[h][0] = acc[h][1] = v4{0, 0, 0, 0};
    const int ra = r0 + lr, rb = r0 + 16 + lr;
    const bool raok = ra < L, rbok = rb < L;
    int st = 0;
    // four k-steps per iteration: eight Tm loads in flight per lane (a leaf
    // workgroup runs one wave per SIMD, so latency is hidden by ILP only)
    for (; st + 4 <= nsteps; st += 4) {
      T ta[4], tb[4], xv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = kbeg + 4 * (st + q) + lk;
        const bool ok = k < kend;
        ta[q] = (ok && raok) ? tm_at<T, TTRANS>(A, lda, ra, k) : T(0);
        tb[q] = (ok && rbok) ? tm_at<T, TTRANS>(A, lda, rb, k) : T(0);
        xv[q] = ok ? X[k * 16 + lr] : T(0);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        acc[0][q & 1] = MfmaT<T>::mac(ta[q], xv[q], acc[0][q & 1]);
        acc[1][q & 1] = MfmaT<T>::mac(tb[q], xv[q], acc[1][q & 1]);
      }
    }
    for (; st + 2 <= nsteps; st += 2) {
      const int k0 = kbeg + 4 * st + lk, k1 = k0 + 4;
      const bool ok0 = k0 < kend, ok1 = k1 < kend;
      const T t00 = (ok0 && raok) ? tm_at<T, TTRANS>(A, lda, ra, k0) : T(0);
      const T t10 = (ok0 && rbok) ? tm_at<T, TTRANS>(A, lda, rb, k0) : T(0);
      const T t01 = (ok1 && raok) ? tm_at<T, TTRANS>(A, lda, ra, k1) : T(0);
      const T t11 = (ok1 && rbok) ? tm_at<T, TTRANS>(A, lda, rb, k1) : T(0);
      const T x0 = ok0 ? X[k0 * 16 + lr] : T(0);
      const T x1 = ok1 ? X[k1 * 16 + lr] : T(0);
      acc[0][0] = MfmaT<T>::mac(t00, x0, acc[0][0]);
      acc[1][0] = MfmaT<T>::mac(t10, x0, acc[1][0]);
      acc[0][1] = MfmaT<T>::mac(t01, x1, acc[0][1]);
      acc[1][1] = MfmaT<T>::mac(t11, x1, acc[1][1]);
    }
    if (st < nsteps) {
      const int k0 = kbeg + 4 * st + lk;
      const bool ok0 = k0 < kend;
      const T t00 = (ok0 && raok) ? tm_at<T, TTRANS>(A, lda, ra, k0) : T(0);
      const T t10 = (ok0 && rbok) ? tm_at<T, TTRANS>(A, lda, rb, k0) : T(0);
      const T x0 = ok0 ? X[k0 * 16 + lr] : T(0);
      acc[0][0] = MfmaT<T>::mac(t00, x0, acc[0][0]);
      acc[1][0] = MfmaT<T>::mac(t10, x0, acc[1][0]);
    }
    // ---- bt = alpha V_I - acc  (lane: rows h*16 + row(lk, r), vector lr)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = h * 16 + MfmaT<T>::row(lk, r);
        const int gi = r0 + i;
        T v = T(0);
        if (gi < L && gj < NV) {
          const T bv = SIDE_L ? B[(size_t)gj * ldb + gi] : B[(size_t)gi * ldb + gj];
          v = alpha * bv - (acc[h][0][r] + acc[h][1][r]);
        }
        bt[i * 16 + lr] = v;
      }
    trsm_wave_sync();
    // ---- X_I = inv(Tm_II) bt
    const auto iv = gptr(job.inv + (size_t)b * TRSM_BS * TRSM_BS);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      v4 xa = v4{0, 0, 0, 0}, xb = v4{0, 0, 0, 0};
#pragma unroll
      for (int kk = 0; kk < TRSM_BS / 4; kk += 2) {
        const int c0 = 4 * kk + lk, c1 = 4 * (kk + 1) + lk;
        xa = MfmaT<T>::mac(iv[c0 * TRSM_BS + h * 16 + lr], bt[c0 * 16 + lr], xa);
        xb = MfmaT<T>::mac(iv[c1 * TRSM_BS + h * 16 + lr], bt[c1 * 16 + lr], xb);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = h * 16 + MfmaT<T>::row(lk, r);
        const int gi = r0 + i;
        const T v = xa[r] + xb[r];
        X[gi * 16 + lr] = v;
        if (gi < L && gj < NV) {
          if (SIDE_L) B[(size_t)gj * ldb + gi] = v;
          else B[(size_t)gi * ldb + gj] = v;
        }
      }
    }
    trsm_wave_sync();
  }
  }
}

// ------------------------------------------------------------ host side
template <typename T>
int trinv_blocks_launch(const TrsmShape& sh, const T* A, int lda, int L, T* inv,
                        const int* abort, hipStream_t s, int prio) {
  const int nblk = (L + TRSM_BS - 1) / TRSM_BS;
  if (nblk <= 0) return 0;
#define DM_TRINV(LO, TT, UN)                                                          \
  hipLaunchKernelGGL((trinv_blocks_kernel<T, LO, TT, UN>), dim3(nblk), dim3(TRSM_BS), 0, s, A, \
                     lda, L, inv, abort, prio)
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
  return sizeof(T) * 4 * (size_t)(Lp + TRSM_BS) * 16;  // per wave: X [Lp][16], bt [32][16]
}
template <typename T, bool LO, bool TT, bool SL>
int trsm_go(const TrsmJob<T>* jobs, int njobs, const TrsmJob<T>& single, long total, long grid,
            size_t smem, const int* abort, hipStream_t s) {
  hipLaunchKernelGGL((trsm_kernel<T, LO, TT, SL>), dim3((unsigned)grid), dim3(256), smem, s, jobs,
                     njobs, single, (int)total, abort);
  DM_LAUNCH_CHECK("trsm_kernel");
  return 0;
}
template <typename T, bool LO, bool TT, bool SL>
int trsm_attr() {
  DM_HIP(hipFuncSetAttribute((const void*)trsm_kernel<T, LO, TT, SL>,
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)trsm_smem_bytes<T>(TRSM_LEAF)));
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
                const TrsmJob<T>& single, long total_wg, const int* abort, hipStream_t s,
                long max_wg) {
  if (total_wg <= 0) return 0;
  if (total_wg > 0x7fffffffL) return arg_fail(0, "trsm: too many workgroups");
  const long grid = (max_wg > 0 && max_wg < total_wg) ? max_wg : total_wg;
  // LDS is sized for the longest job; batched callers pass the max L in single.L.
  const size_t smem = trsm_smem_bytes<T>(single.L);
#define DM_TRSM_GO(LO, TT, SL) \
  trsm_go<T, LO, TT, SL>(jobs_dev, njobs, single, total_wg, grid, smem, abort, s)
  if (sh.lower) {
    if (sh.ttrans) return sh.side_l ? DM_TRSM_GO(true, true, true) : DM_TRSM_GO(true, true, false);
    return sh.side_l ? DM_TRSM_GO(true, false, true) : DM_TRSM_GO(true, false, false);
  }
  if (sh.ttrans) return sh.side_l ? DM_TRSM_GO(false, true, true) : DM_TRSM_GO(false, true, false);
  return sh.side_l ? DM_TRSM_GO(false, false, true) : DM_TRSM_GO(false, false, false);
#undef DM_TRSM_GO
}

// Unbatched solve with the diagonal blocks of each leaf inverted into ws.
template <typename T>
static int trsm_eff(const TrsmShape& sh, int L, int NV, T alpha, const T* A, int lda, T* B,
                    int ldb, T* ws, const int* abort, hipStream_t s) {
  if (L <= 0 || NV <= 0) return 0;
  if (L <= trsm_leaf_len()) {
    if (!ws) return arg_fail(0, "trsm: no workspace");
    T* inv = ws;
    int rc = trinv_blocks_launch<T>(sh, A, lda, L, inv, abort, s);
    if (rc) return rc;
    TrsmJob<T> j{};
    j.A = A; j.B = B; j.inv = inv; j.L = L; j.NV = NV; j.lda = lda; j.ldb = ldb; j.alpha = alpha;
    return trsm_launch<T>(sh, nullptr, 0, j, trsm_wgs(NV), abort, s);
  }
  const TrsmSplit sp = trsm_split(sh, L);
  int rc = trsm_eff<T>(sh, sp.first_n, NV, alpha, tm_sub(sh, A, lda, sp.first_r, sp.first_r), lda,
                       vsub(sh, B, ldb, sp.first_r), ldb, ws, abort, s);
  if (rc) return rc;
  bool ta, tb;
  const GemmJob<T> g = trsm_update_job<T>(sh, sp, NV, alpha, A, lda, B, ldb, ta, tb);
  rc = gemm_single<T>(ta, tb, g.M, g.N, g.K, g.alpha, g.A, g.lda, g.B, g.ldb, g.beta, g.C, g.ldc, 0,
                      abort, s);
  if (rc) return rc;
  return trsm_eff<T>(sh, sp.second_n, NV, T(1), tm_sub(sh, A, lda, sp.second_r, sp.second_r), lda,
                     vsub(sh, B, ldb, sp.second_r), ldb, ws, abort, s);
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
                                         const int*, hipStream_t, int);
template int trinv_blocks_launch<float>(const TrsmShape&, const float*, int, int, float*,
                                        const int*, hipStream_t, int);
template int trsm_launch<double>(const TrsmShape&, const TrsmJob<double>*, int,
                                 const TrsmJob<double>&, long, const int*, hipStream_t, long);
template int trsm_launch<float>(const TrsmShape&, const TrsmJob<float>*, int,
                                const TrsmJob<float>&, long, const int*, hipStream_t, long);
template int trsm_full<double>(char, char, char, char, int, int, double, const double*, int,
                               double*, int, double*, const int*, hipStream_t);
template int trsm_full<float>(char, char, char, char, int, int, float, const float*, int, float*,
                              int, float*, const int*, hipStream_t);

}  // namespace dm

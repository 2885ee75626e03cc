// The diagonal-block and panel-row steps of the tile POTRF (one wave each),
// shared by potrf_diag_kernel / potrf_row_kernel (potrf.hip), the fused
// trailing-update kernel that factors the next diagonal block (gemm.hip) and
// the batched level-synchronous factorization (batched.hip).
#pragma once
#include <type_traits>
#include "dm_internal.h"

namespace dm {

// Both panel kernels keep their code small enough for the instruction
// cache: a single wave runs them, so every instruction fetched from L2
// costs the whole kernel.  Instead of unrolling the 64 steps (a triangle
// of static register indices, tens of KB of code), each lane keeps a
// window of its column that slides by one row per step (the update of
// row i+1 lands in register i, so the row being eliminated is always in
// register 0), and the step loop stays rolled.  The window shrinks in
// phases of 16 steps, so the update length follows the triangle in
// steps of 16 rows.

// sd = sqrt(d) and inv = 1/sqrt(d) for a pivot d > 0.  In the normal range
// the reciprocal square root comes from the hardware estimate refined by
// Newton steps (y += y/2 (1 - d y^2)), and sd = d y is corrected once by
// the residual d - sd^2: about a quarter of the dependent latency of a
// correctly rounded sqrt followed by a division, which sits on the serial
// chain of every column step.  Both results are within an ulp or two.
__device__ __forceinline__ void sqrt_inv(double d, double& sd, double& inv) {
  if (d > 1e-290 && d < 1e290) {
    double y = __builtin_amdgcn_rsq(d);
    double e = __builtin_fma(-d * y, y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
    e = __builtin_fma(-d * y, y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
    double s = d * y;
    const double r = __builtin_fma(-s, s, d);
    sd = __builtin_fma(0.5 * y, r, s);
    inv = y;
  } else {  // tiny, huge, zero, negative or NaN: the exact forms
    sd = sqrt(d);
    inv = 1.0 / sd;
  }
}
__device__ __forceinline__ void sqrt_inv(float d, float& sd, float& inv) {
  if (d > 1e-30f && d < 1e30f) {
    float y = __builtin_amdgcn_rsqf(d);
    const float e = __builtin_fmaf(-d * y, y, 1.0f);
    y = __builtin_fmaf(0.5f * y, e, y);
    const float s = d * y;
    const float r = __builtin_fmaf(-s, s, d);
    sd = __builtin_fmaf(0.5f * y, r, s);
    inv = y;
  } else {
    sd = sqrtf(d);
    inv = 1.0f / sd;
  }
}

// Factor the kb x kb diagonal block at (j0, j0), kb = min(PW, n - j0), in one
// wave.  Element (r, c), r <= c, of the block's upper view is A[j0+r, j0+c]
// (U) or A[j0+c, j0+r] (L, i.e. L^T); lane c holds rows j.. of column c of
// that view at step j.  On a failed pivot at local column j: info =
// j0 + j + 1 and *abort = global_base + j0 + j + 1.
template <typename T, bool UPPER, int PW>
struct PotrfDiag {
  T c[PW];     // c[i] = element (j + i, t) of the upper view at step j
  T d, sd, inv;  // pivot of the current step, its square root and reciprocal
  T* rowb;     // LDS: row j of the factor (PW entries) then PW zeros
  T* pv;       // LDS: next pivot
  T* A;
  int lda, j0, kb, t;

  __device__ __forceinline__ void put(int r, T v) {
    if (UPPER) A[(size_t)(j0 + t) * lda + j0 + r] = v;
    else A[(size_t)(j0 + r) * lda + j0 + t] = v;
  }

  // Steps j0s .. j0s+15 (those below kb) with a window of L live rows.
  // Returns the failed local column + 1, or 0.
  template <int L>
  __device__ __forceinline__ int phase(int js) {
    const int je = min(js + 16, kb);
#pragma unroll 1
    for (int j = js; j < je; ++j) {
      if (!(d > T(0))) return j + 1;
      T u = T(0), fin = c[0];
      if (t == j) fin = sd;
      else if (t > j) {
        u = c[0] * inv;
        fin = u;
      }
      if (t >= j && t < kb) put(j, fin);
      if (t == j + 1) *pv = c[1] - u * u;
      rowb[t] = u;  // zero for lanes <= j: their columns are left unchanged
      __builtin_amdgcn_wave_barrier();
      d = *pv;
      sqrt_inv(d, sd, inv);
      const T* rb = rowb + j + 1;  // reads past PW land in the zero half
#pragma unroll
      for (int i = 0; i < L - 1; ++i) c[i] = c[i + 1] - rb[i] * u;
      c[L - 1] = T(0);
      __builtin_amdgcn_wave_barrier();  // rowb/pv are rewritten by the next step
    }
    return 0;
  }
};

// Factor the kb x kb block at (j0, j0) of A, kb = min(PW, n - j0), with the
// first wave of the calling workgroup (threadIdx.x < 64 run it; others
// return at once).  rowb: 128 LDS elements, pv: one.  On a failed pivot at
// local column j: info = j0 + j + 1 and *abort = global_base + j0 + j + 1.
template <typename T, bool UPPER, int PW>
__device__ __forceinline__ void potrf_diag_block(T* __restrict__ A, int n, int lda, int j0,
                                                 int* __restrict__ info, int* __restrict__ abort,
                                                 int global_base, T* rowb, T* pv_) {
  static_assert(PW == 32 || PW == 64, "panel width");
  if (threadIdx.x >= 64) return;
  PotrfDiag<T, UPPER, PW> f;
  f.t = threadIdx.x;
  f.kb = min(PW, n - j0);
  f.A = A;
  f.lda = lda;
  f.j0 = j0;
  f.rowb = rowb;
  f.pv = pv_;
  const int t = f.t;
#pragma unroll
  for (int r = 0; r < PW; ++r) {
    const bool in = t < f.kb && r <= t;
    f.c[r] = in ? (UPPER ? A[(size_t)(j0 + t) * lda + j0 + r] : A[(size_t)(j0 + r) * lda + j0 + t])
                : T(0);
  }
  rowb[64 + t] = T(0);
  if (t == 0) *pv_ = f.c[0];
  __builtin_amdgcn_wave_barrier();
  f.d = *pv_;
  sqrt_inv(f.d, f.sd, f.inv);
  __builtin_amdgcn_wave_barrier();  // every lane has read pv before step 0 rewrites it
  int fail = f.template phase<PW>(0);
  if (!fail && PW > 16) fail = f.template phase<PW - 16>(16);
  if (PW == 64) {
    if (!fail) fail = f.template phase<(PW > 32 ? PW - 32 : 16)>(32);
    if (!fail) fail = f.template phase<(PW > 48 ? PW - 48 : 16)>(48);
  }
  if (t == 0 && fail) {
    if (info) *info = j0 + fail;
    if (abort) *abort = global_base + j0 + fail;
  }
}

// Panel row beside a factored full PW x PW block at (j0, j0): U12 :=
// U11^-T A12 (U) or A21 := A21 L11^-T (L), one column of A12 (row of A21)
// per lane.  potrf_row_stage (every thread of a one-wave workgroup) puts
// the block in LDS as its upper view with row stride PW -- so reading row p
// past column PW-1 lands in the (zero) strictly lower part of row p+1,
// followed by PW zeros -- and the reciprocals of its diagonal in invd.
// Us: PW * PW + PW elements, invd: PW.
template <typename T, bool UPPER, int PW>
__device__ __forceinline__ void potrf_row_stage(const T* __restrict__ A, int lda, int j0,
                                                T* Us, T* invd) {
  static_assert(PW == 32 || PW == 64, "panel width");
  const int t = threadIdx.x;
  if (t < PW) {
    // lane = column c; U: element (r, c) = A[j0+r, j0+c]; L: A[j0+c, j0+r]
    for (int r = 0; r < PW; ++r) {
      T v = T(0);
      if (r <= t) v = UPPER ? A[(size_t)(j0 + t) * lda + j0 + r] : A[(size_t)(j0 + r) * lda + j0 + t];
      Us[r * PW + t] = v;
    }
    Us[PW * PW + t] = T(0);
  }
  __syncthreads();
  if (t < PW) invd[t] = T(1) / Us[t * PW + t];
  __syncthreads();
}

// U11^T x = b for one column per lane, forward over the rows of the staged
// block; x[i] holds row p + i at step p, and row p's result goes to
// base[p * step].
template <typename T, int PW>
__device__ __forceinline__ void potrf_row_phases(T (&x)[PW], T* base, size_t step, const T* Us,
                                                 const T* invd) {
  auto phase = [&](auto len, int ps) {
    constexpr int L = decltype(len)::value;
#pragma unroll 1
    for (int p = ps; p < ps + 16; ++p) {
      const T xp = x[0] * invd[p];
      base[p * step] = xp;
      const T* u = Us + p * PW + p + 1;
#pragma unroll
      for (int i = 0; i < L - 1; ++i) x[i] = x[i + 1] - u[i] * xp;
      x[L - 1] = T(0);
    }
  };
  phase(std::integral_constant<int, PW>{}, 0);
  phase(std::integral_constant<int, PW - 16>{}, 16);
  if constexpr (PW == 64) {
    phase(std::integral_constant<int, 32>{}, 32);
    phase(std::integral_constant<int, 16>{}, 48);
  }
}

// Columns c0 + lane of A12 (rows of A21), counted from j0 + PW, after
// potrf_row_stage.  Lanes past the panel return (no barrier follows).
template <typename T, bool UPPER, int PW>
__device__ __forceinline__ void potrf_row_solve(T* __restrict__ A, int n, int lda, int j0, int c0,
                                                const T* Us, const T* invd) {
  const int t = threadIdx.x;
  const int c_begin = j0 + PW, rest = n - c_begin;
  if (c0 + t >= rest) return;
  const int cc = c_begin + c0 + t;
  const size_t step = UPPER ? 1 : (size_t)lda;
  T* const base = UPPER ? A + (size_t)cc * lda + j0 : A + (size_t)j0 * lda + cc;
  T x[PW];
  const T* px = base;
#pragma unroll
  for (int r = 0; r < PW; ++r) {
    x[r] = *px;
    px += step;
    if ((r & 15) == 15) __builtin_amdgcn_sched_barrier(0);  // bound the loads in flight
  }
  potrf_row_phases<T, PW>(x, base, step, Us, invd);
}

// W = U11^-T (PW x PW, column-major, leading dimension ldw) of the block
// staged by potrf_row_stage (upper view): lane c solves U11^T w = e_c.
template <typename T, int PW>
__device__ __forceinline__ void potrf_row_inverse(T* __restrict__ W, int ldw, const T* Us,
                                                  const T* invd) {
  const int t = threadIdx.x;
  if (t >= PW) return;
  T x[PW];
#pragma unroll
  for (int r = 0; r < PW; ++r) x[r] = (r == t) ? T(1) : T(0);
  potrf_row_phases<T, PW>(x, W + (size_t)t * ldw, 1, Us, invd);
}

}  // namespace dm

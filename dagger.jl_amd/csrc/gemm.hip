// GEMM / SYRK tile updates on CDNA4 matrix cores.
//
// Replaces the BLAS.gemm!/BLAS.syrk! tile calls of the tiled Cholesky
// (src/array/cholesky.jl:47,50,92,95) and of trsm! (src/array/trsm.jl:99-183).
//
// Design (gfx950, wave64):
//  * FP64 uses v_mfma_f64_16x16x4_f64, FP32 uses v_mfma_f32_16x16x4_f32.
//    Each wave owns a (BM/WM) x (BN/WN) block of C as 16x16 MFMA tiles.
//  * The MFMA computes C^T tiles (operand roles swapped) so a lane holds 16
//    consecutive rows of C across a quarter-wave: column-major stores of C
//    are 128-byte contiguous per quarter-wave.
//  * Operand tiles are staged global -> registers -> LDS with two LDS
//    buffers (one barrier per BK step).  The LDS layout follows the global
//    contiguous axis: k-contiguous operands are stored [mn][BK+pad] and read
//    as 16 rows x 4 k per MFMA; mn-contiguous operands are stored
//    [k][MN+16] so a quarter-wave reads 128 contiguous bytes.
//  * A launch is a batch of independent jobs (one per Dagger tile task);
//    workgroups find their job by binary search over wg_begin, and each XCD
//    (blockIdx is dealt round-robin over the 8 XCDs) owns a contiguous range
//    of tiles, so consecutive tiles of one job share an L2.  A launch may use
//    fewer workgroups than tiles (gemm_launch's max_wg; each workgroup then
//    loops over its share).
//  * SYRK is the same kernel with a triangle mask: only workgroup tiles that
//    touch the requested triangle are launched, and diagonal tiles store
//    masked.
#include "dm_internal.h"
#include "gemm_tiles.h"
#include "potrf_diag.h"
#include <type_traits>

namespace dm {

template <typename T> struct Mfma;
// Output layouts (lane l, accumulator item r; 16x16 block D[i][j], j = l%16):
//   f64 16x16x4: i = 4*r + l/16        f32 16x16x4: i = 4*(l/16) + r
// (the f64 layout was measured on gfx950 with a one-hot lane probe).
template <> struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * r + lk; }
  static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
  static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <typename T, int V> struct VecT;
template <> struct VecT<double, 1> { typedef double t; };
template <> struct VecT<double, 2> { typedef double t __attribute__((ext_vector_type(2))); };
template <> struct VecT<float, 1> { typedef float t; };
template <> struct VecT<float, 2> { typedef float t __attribute__((ext_vector_type(2))); };

// PF: tiles of operands in registers ahead of the one being multiplied (1:
// the next tile is loaded at the start of a k-step and goes into LDS
// halfway through it; >= 2: see kloop_pf).
template <typename T, int BM, int BN, int BK, int WM, int WN, bool TA, bool TB, int VEC,
          int PF = 1>
struct GemmKernel {
  static constexpr int NT = 64 * WM * WN;
  static constexpr int WTM = BM / WM, WTN = BN / WN;
  static constexpr int FM = WTM / 16, FN = WTN / 16;
  // Pad of k-contiguous rows.  A quarter-wave's 8-byte fragment reads hit
  // rows lr = 0..15 at the same k: a row stride of an odd number of doubles
  // (2 x odd dwords) puts them on 16 distinct bank pairs of the 32 banks;
  // an FP32 stride of 18 dwords does the same for 4-byte reads.
  static constexpr int KP = sizeof(T) == 8 ? 1 : 2;
  static constexpr int SK = BK + KP;   // stride of [mn][k] layout
  static constexpr int SMA = BM + 16;  // stride of [k][m] layout
  static constexpr int SNB = BN + 16;  // stride of [k][n] layout
  static constexpr int A_ELEMS = TA ? BM * SK : BK * SMA;
  static constexpr int B_ELEMS = TB ? BK * SNB : BN * SK;
  static constexpr int LA = (BM * BK) / (VEC * NT);
  static constexpr int LB = (BN * BK) / (VEC * NT);
  static_assert(LA * VEC * NT == BM * BK, "A tile not divisible");
  static_assert(LB * VEC * NT == BN * BK, "B tile not divisible");
  static_assert(FM * 16 * WM == BM && FN * 16 * WN == BN, "wave tiling");
  static_assert(BK % 4 == 0, "BK");
  typedef typename Mfma<T>::v4 v4;
  typedef typename VecT<T, VEC>::t vt;

  // VEC elements at p (always a valid address); bit i of *okm records
  // whether they lie inside the operand (they are zeroed when stored to LDS,
  // so no instruction reads the loaded registers before that store)
  static __device__ __forceinline__ void clamped_load(const T* p, bool ok, int i, T (&r)[VEC],
                                                      unsigned* okm) {
    vt v = *gptr((const vt*)p);
#pragma unroll
    for (int q = 0; q < VEC; ++q) r[q] = reinterpret_cast<const T*>(&v)[q];
    *okm |= (ok ? 1u : 0u) << i;
  }

  // CLAMP: a lane outside the operand loads the operand's first element(s)
  // instead and flags them in *okm (zeroed by store_a/store_b), so the loads
  // are unconditional: no branch around each load, and the compiler can wait
  // for one tile's loads while a later tile's are in flight.  M, K >= 1
  // whenever a tile is loaded, and 2-wide loads imply an even extent along
  // the contiguous axis.
  template <bool CLAMP = false>
  static __device__ __forceinline__ void load_a(const T* __restrict__ A, int lda, int m0, int k0,
                                                int M, int K, T (&ra)[LA][VEC],
                                                unsigned* okm = nullptr) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int e = threadIdx.x + i * NT;
      int gm, gk;
      if (TA) {  // A[k + m*lda], contiguous along k
        gk = k0 + (e % (BK / VEC)) * VEC;
        gm = m0 + e / (BK / VEC);
        const bool ok = gm < M && gk < K;
        const T* p = A + (size_t)gm * lda + gk;
        if (CLAMP) {
          clamped_load(ok ? p : A, ok, i, ra[i], okm);
        } else if (VEC == 2) {
          vt v = ok ? *gptr((const vt*)p) : vt{};
#pragma unroll
          for (int q = 0; q < VEC; ++q) ra[i][q] = reinterpret_cast<const T*>(&v)[q];
        } else {
          ra[i][0] = ok ? *gptr(p) : T(0);
        }
      } else {   // A[m + k*lda], contiguous along m
        gm = m0 + (e % (BM / VEC)) * VEC;
        gk = k0 + e / (BM / VEC);
        const bool ok = gm < M && gk < K;
        const T* p = A + (size_t)gk * lda + gm;
        if (CLAMP) {
          clamped_load(ok ? p : A, ok, i, ra[i], okm);
        } else if (VEC == 2) {
          vt v = ok ? *gptr((const vt*)p) : vt{};
#pragma unroll
          for (int q = 0; q < VEC; ++q) ra[i][q] = reinterpret_cast<const T*>(&v)[q];
        } else {
          ra[i][0] = ok ? *gptr(p) : T(0);
        }
      }
    }
  }
  // okm: in-range flags of clamped loads (out-of-range elements store 0)
  static __device__ __forceinline__ void store_a(T* __restrict__ As, const T (&ra)[LA][VEC],
                                                 unsigned okm = ~0u) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int e = threadIdx.x + i * NT;
      T v[VEC];
#pragma unroll
      for (int q = 0; q < VEC; ++q) v[q] = ((okm >> i) & 1u) ? ra[i][q] : T(0);
      if (TA) {
        const int k = (e % (BK / VEC)) * VEC, m = e / (BK / VEC);
        if (VEC == 2 && SK % 2 == 0) *(vt*)&As[m * SK + k] = *(const vt*)&v[0];
        else
#pragma unroll
          for (int q = 0; q < VEC; ++q) As[m * SK + k + q] = v[q];
      } else {
        const int m = (e % (BM / VEC)) * VEC, k = e / (BM / VEC);
        if (VEC == 2) *(vt*)&As[k * SMA + m] = *(const vt*)&v[0];
        else As[k * SMA + m] = v[0];
      }
    }
  }
  template <bool CLAMP = false>
  static __device__ __forceinline__ void load_b(const T* __restrict__ B, int ldb, int n0, int k0,
                                                int N, int K, T (&rb)[LB][VEC],
                                                unsigned* okm = nullptr) {
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int e = threadIdx.x + i * NT;
      int gn, gk;
      if (!TB) {  // B[k + n*ldb], contiguous along k
        gk = k0 + (e % (BK / VEC)) * VEC;
        gn = n0 + e / (BK / VEC);
        const bool ok = gn < N && gk < K;
        const T* p = B + (size_t)gn * ldb + gk;
        if (CLAMP) {
          clamped_load(ok ? p : B, ok, i, rb[i], okm);
        } else if (VEC == 2) {
          vt v = ok ? *gptr((const vt*)p) : vt{};
#pragma unroll
          for (int q = 0; q < VEC; ++q) rb[i][q] = reinterpret_cast<const T*>(&v)[q];
        } else {
          rb[i][0] = ok ? *gptr(p) : T(0);
        }
      } else {    // B[n + k*ldb], contiguous along n
        gn = n0 + (e % (BN / VEC)) * VEC;
        gk = k0 + e / (BN / VEC);
        const bool ok = gn < N && gk < K;
        const T* p = B + (size_t)gk * ldb + gn;
        if (CLAMP) {
          clamped_load(ok ? p : B, ok, i, rb[i], okm);
        } else if (VEC == 2) {
          vt v = ok ? *gptr((const vt*)p) : vt{};
#pragma unroll
          for (int q = 0; q < VEC; ++q) rb[i][q] = reinterpret_cast<const T*>(&v)[q];
        } else {
          rb[i][0] = ok ? *gptr(p) : T(0);
        }
      }
    }
  }
  static __device__ __forceinline__ void store_b(T* __restrict__ Bs, const T (&rb)[LB][VEC],
                                                 unsigned okm = ~0u) {
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int e = threadIdx.x + i * NT;
      T v[VEC];
#pragma unroll
      for (int q = 0; q < VEC; ++q) v[q] = ((okm >> i) & 1u) ? rb[i][q] : T(0);
      if (!TB) {
        const int k = (e % (BK / VEC)) * VEC, n = e / (BK / VEC);
        if (VEC == 2 && SK % 2 == 0) *(vt*)&Bs[n * SK + k] = *(const vt*)&v[0];
        else
#pragma unroll
          for (int q = 0; q < VEC; ++q) Bs[n * SK + k + q] = v[q];
      } else {
        const int n = (e % (BN / VEC)) * VEC, k = e / (BN / VEC);
        if (VEC == 2) *(vt*)&Bs[k * SNB + n] = *(const vt*)&v[0];
        else Bs[k * SNB + n] = v[0];
      }
    }
  }
  // op(A)[m][k] from LDS
  static __device__ __forceinline__ T frag_a(const T* __restrict__ As, int m, int k) {
    return TA ? As[m * SK + k] : As[k * SMA + m];
  }
  // op(B)[k][n] from LDS
  static __device__ __forceinline__ T frag_b(const T* __restrict__ Bs, int k, int n) {
    return TB ? Bs[k * SNB + n] : Bs[n * SK + k];
  }

  // the MFMA operands of sub-step kk (k = 4kk + lk) for this wave's 16-row
  // blocks of op(A) and 16-column blocks of op(B)
  static __device__ __forceinline__ void read_frags(const T* __restrict__ As,
                                                    const T* __restrict__ Bs, int kk, int wm0,
                                                    int wn0, int lr, int lk, T (&af)[FM],
                                                    T (&bf)[FN]) {
#pragma unroll
    for (int i = 0; i < FM; ++i) af[i] = frag_a(As, wm0 + i * 16 + lr, kk * 4 + lk);
#pragma unroll
    for (int j = 0; j < FN; ++j) bf[j] = frag_b(Bs, kk * 4 + lk, wn0 + j * 16 + lr);
  }

  // K-loop with PF (>= 2) tiles of operands in registers: the loads of tile
  // kt+PF are issued at the start of k-step kt and tile kt+1 goes into LDS
  // at the end of it, so a load has about PF k-steps of MFMAs to arrive.
  // For launches of one or two workgroups per CU whose k-steps wait on
  // memory (the 64-column bulk updates of a blocked substitution).
  // Register sets: tile t in set t % PF; the set of tile kt is free once
  // tile kt is in LDS (end of k-step kt-1).  The k-loop is unrolled PF
  // times so that every set index is a compile-time constant.
  // LDS: buffer kt&1 is read in k-step kt up to its barrier; buffer
  // (kt+1)&1 is written in k-step kt before that barrier, after the barrier
  // of k-step kt-1 that ended every read of it.
  static __device__ __forceinline__ void kloop_pf(const GemmJob<T>& job, int m0, int n0, int wm0,
                                                  int wn0, int lr, int lk, T* smem,
                                                  v4 (&acc)[FM][FN]) {
    static_assert(PF >= 2, "kloop_pf");
    static_assert(LA <= 32 && LB <= 32, "in-range flags");
    T* As0 = smem;
    T* Bs0 = smem + A_ELEMS;
    T* As1 = smem + A_ELEMS + B_ELEMS;
    T* Bs1 = As1 + A_ELEMS;
    const int M = job.M, N = job.N, K = job.K;
    const int nk = (K + BK - 1) / BK;
    T ra[PF][LA][VEC], rb[PF][LB][VEC];
    unsigned ma[PF], mb[PF];
    if (nk > 0) {
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        ma[u] = 0;
        mb[u] = 0;
        load_a<true>(job.A, job.lda, m0, u * BK, M, K, ra[u], &ma[u]);
        load_b<true>(job.B, job.ldb, n0, u * BK, N, K, rb[u], &mb[u]);
      }
      store_a(As0, ra[0], ma[0]);
      store_b(Bs0, rb[0], mb[0]);
    }
    __syncthreads();
    constexpr int KS = BK / 4;
    T af[FM], bf[FN];
    if (nk > 0) read_frags(As0, Bs0, 0, wm0, wn0, lr, lk, af, bf);
    // k-step kt: tile kt+PF is loaded into set (la, lb), tile kt+1 stored
    // from set (sa, sb); past the last tile every lane loads the operand's
    // first element(s), and a k-step past the last one only loads
    auto step = [&](int kt, T (&la)[LA][VEC], T (&lb)[LB][VEC], unsigned& mla, unsigned& mlb,
                    const T (&sa)[LA][VEC], const T (&sb)[LB][VEC], unsigned msa,
                    unsigned msb) {
      const bool odd = kt & 1;
      const T* As = odd ? As1 : As0;
      const T* Bs = odd ? Bs1 : Bs0;
      T* An = odd ? As0 : As1;
      T* Bn = odd ? Bs0 : Bs1;
      mla = 0;
      mlb = 0;
      load_a<true>(job.A, job.lda, m0, (kt + PF) * BK, M, K, la, &mla);
      load_b<true>(job.B, job.ldb, n0, (kt + PF) * BK, N, K, lb, &mlb);
      if (kt >= nk) return;
      const bool more = kt + 1 < nk;
#pragma unroll
      for (int kk = 0; kk < KS; ++kk) {
        T an[FM] = {}, bn[FN] = {};
        if (kk + 1 < KS) read_frags(As, Bs, kk + 1, wm0, wn0, lr, lk, an, bn);
        if (kk == KS - 1) {
          if (more) {
            store_a(An, sa, msa);
            store_b(Bn, sb, msb);
          }
          __syncthreads();
          if (more) read_frags(An, Bn, 0, wm0, wn0, lr, lk, an, bn);
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::mac(bf[j], af[i], acc[i][j]);
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = an[i];
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[j] = bn[j];
      }
    };
    for (int kt = 0; kt < nk; kt += PF) {
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        step(kt + u, ra[u], rb[u], ma[u], mb[u], ra[(u + 1) % PF], rb[(u + 1) % PF],
             ma[(u + 1) % PF], mb[(u + 1) % PF]);
      }
    }
  }

  static __device__ __forceinline__ void run(const GemmJob<T>& job, int tm, int tn, T* smem) {
    T* As0 = smem;
    T* Bs0 = smem + A_ELEMS;
    T* As1 = smem + A_ELEMS + B_ELEMS;
    T* Bs1 = As1 + A_ELEMS;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm0 = (wave % WM) * WTM;
    const int wn0 = (wave / WM) * WTN;
    const int m0 = tm * BM, n0 = tn * BN;
    const int M = job.M, N = job.N, K = job.K;

    v4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = v4{0, 0, 0, 0};

    const int lr = lane & 15, lk = lane >> 4;
    if constexpr (PF >= 2) {
      kloop_pf(job, m0, n0, wm0, wn0, lr, lk, smem, acc);
    } else {
    T ra[LA][VEC], rb[LB][VEC];
    const int nk = (K + BK - 1) / BK;
    if (nk > 0) {
      load_a(job.A, job.lda, m0, 0, M, K, ra);
      load_b(job.B, job.ldb, n0, 0, N, K, rb);
      store_a(As0, ra);
      store_b(Bs0, rb);
    }
    __syncthreads();
    // Software pipeline over the BK/4 sub-steps of a k-step: the fragments
    // of sub-step kk+1 are read from LDS before the MFMAs of sub-step kk; the
    // next k-step's tile goes from registers into the spare LDS buffer
    // halfway through; the one barrier comes before the last sub-step's
    // MFMAs, so the next k-step's first fragments are read while those run
    // and the matrix pipe does not wait on LDS latency at every k-step.
    // LDS hazards: buffer kt&1 is read in k-step kt up to its barrier; buffer
    // (kt+1)&1 is written in k-step kt after the barrier of k-step kt-1
    // (which ended every read of it) and read after the barrier of k-step kt.
    constexpr int KS = BK / 4;
    T af[FM], bf[FN];
    if (nk > 0) read_frags(As0, Bs0, 0, wm0, wn0, lr, lk, af, bf);
    for (int kt = 0; kt < nk; ++kt) {
      const bool odd = kt & 1;
      const T* As = odd ? As1 : As0;
      const T* Bs = odd ? Bs1 : Bs0;
      T* An = odd ? As0 : As1;
      T* Bn = odd ? Bs0 : Bs1;
      const bool more = kt + 1 < nk;
      if (more) {
        load_a(job.A, job.lda, m0, (kt + 1) * BK, M, K, ra);
        load_b(job.B, job.ldb, n0, (kt + 1) * BK, N, K, rb);
      }
#pragma unroll
      for (int kk = 0; kk < KS; ++kk) {
        T an[FM] = {}, bn[FN] = {};
        if (kk + 1 < KS) read_frags(As, Bs, kk + 1, wm0, wn0, lr, lk, an, bn);
        if (kk == KS / 2 && more) {
          store_a(An, ra);
          store_b(Bn, rb);
        }
        if (kk == KS - 1) {
          __syncthreads();
          if (more) read_frags(An, Bn, 0, wm0, wn0, lr, lk, an, bn);
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::mac(bf[j], af[i], acc[i][j]);
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = an[i];
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[j] = bn[j];
      }
    }
    }  // PF == 1

    // Epilogue: lane holds C[m][n] with m = .. + lr, n = .. + row(lk, r).
    // Per 16-row strip, all old values of C are loaded before any store, so
    // the strip's loads are in flight together (one memory latency per
    // strip, not one per element).
    const T alpha = job.alpha, beta = job.beta;
    const int tri = job.tri;
    // A tile wholly inside C and inside the triangle (nearly every tile of a
    // big update) stores without per-element branches, which would make
    // every store wait for the ones before it.
    const bool whole = m0 + BM <= M && n0 + BN <= N &&
                       (tri == 0 || (tri == 1 ? m0 + BM - 1 <= n0 : m0 >= n0 + BN - 1));
    if (whole) {
      const size_t ldc = job.ldc;
      auto cbase = gptr(job.C + (size_t)(n0 + wn0) * ldc + m0 + wm0 + lr);
      if (beta == T(0)) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              cbase[(size_t)(j * 16 + Mfma<T>::row(lk, r)) * ldc + i * 16] = alpha * acc[i][j][r];
      } else {
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          T old[FN][4];
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              old[j][r] = cbase[(size_t)(j * 16 + Mfma<T>::row(lk, r)) * ldc + i * 16];
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              cbase[(size_t)(j * 16 + Mfma<T>::row(lk, r)) * ldc + i * 16] =
                  alpha * acc[i][j][r] + beta * old[j][r];
        }
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int m = m0 + wm0 + i * 16 + lr;
      T old[FN][4];
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = n0 + wn0 + j * 16 + Mfma<T>::row(lk, r);
          const bool ok = in_c(m, n, M, N, tri);
          old[j][r] = (ok && beta != T(0)) ? *gptr(job.C + (size_t)n * job.ldc + m) : T(0);
        }
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = n0 + wn0 + j * 16 + Mfma<T>::row(lk, r);
          if (in_c(m, n, M, N, tri)) {
            const T v = alpha * acc[i][j][r];
            *gptr(job.C + (size_t)n * job.ldc + m) = (beta == T(0)) ? v : v + beta * old[j][r];
          }
        }
    }
  }
  // (m, n) inside C and inside the requested triangle (tri 1 upper, 2 lower)
  static __device__ __forceinline__ bool in_c(int m, int n, int M, int N, int tri) {
    bool ok = m < M && n < N;
    if (tri == 1) ok = ok && m <= n;
    else if (tri == 2) ok = ok && m >= n;
    return ok;
  }
};

template <typename T, int BM, int BN, int BK, int WM, int WN, bool TA, bool TB, int VEC,
          int PF = 1>
__global__ __launch_bounds__(64 * WM * WN) void gemm_kernel(const GemmJob<T>* __restrict__ jobs,
                                                            int njobs, GemmJob<T> single,
                                                            int total,
                                                            const int* __restrict__ abort) {
  typedef GemmKernel<T, BM, BN, BK, WM, WN, TA, TB, VEC, PF> G;
  __shared__ T smem[2 * (G::A_ELEMS + G::B_ELEMS)];
  if (abort && *abort) return;
  if (single.prio) raise_wave_priority();
  constexpr bool packed = BM == BN;  // SYRK jobs enumerate only triangle tiles
  for (TileRange r = tile_range(total); r.t < r.end; r.t += r.step) {
    GemmJob<T> job;
    int tm, tn;
    locate_tile(jobs, njobs, single, r.t, job, tm, tn, packed);
    if (!packed && job.tri) {  // rectangular tiles: skip those outside the triangle
      const int m0 = tm * BM, n0 = tn * BN;
      if (job.tri == 1 ? m0 > n0 + BN - 1 : n0 > m0 + BM - 1) continue;
    }
    G::run(job, tm, tn, smem);
    __syncthreads();  // the next tile's prologue rewrites LDS
  }
}


// Skinny products (N <= 64: solves with a few right-hand sides).  One
// workgroup owns 32 rows x all N columns of C; its four waves split K
// (wave w takes the 16-k blocks w, w+4, ...) and their partial sums are
// added in wave order through LDS, so the result is deterministic.  Inside
// a 16-k block, lane group l/16 feeds k = 4(l/16) + s to MFMA step s: both
// operands use the same k slots, and each lane walks 4 consecutive k,
// which is contiguous memory for A^T and for a column-major X.  A K-long
// chain per workgroup shrinks by 4 against the 64 x 64 tile, which matters
// for the latency-bound launches of a blocked substitution.
template <typename T, bool TA, bool TB, int NB>
__global__ __launch_bounds__(256) void gemm_skinny_kernel(const GemmJob<T>* __restrict__ jobs,
                                                          int njobs, GemmJob<T> single, int total,
                                                          const int* __restrict__ abort) {
  typedef typename Mfma<T>::v4 v4;
  __shared__ T red[3][2 * NB * 4][64];
  if (abort && *abort) return;
  if (single.prio) raise_wave_priority();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  for (TileRange rg = tile_range(total); rg.t < rg.end; rg.t += rg.step) {
    GemmJob<T> job;
    int tm, tn;
    locate_tile(jobs, njobs, single, rg.t, job, tm, tn, false);
    const int M = job.M, N = job.N, K = job.K;
    const int m0 = tm * 32, n0 = tn * 16 * NB;
    v4 acc[2][NB];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[rb][b] = v4{0, 0, 0, 0};
    // A wave takes the 16-k blocks w, w+4, w+8, ... in that order; the loads
    // of D consecutive blocks are issued together before their MFMAs, so a
    // K-long chain waits for memory once per D blocks (the accumulation
    // order is unchanged).
    constexpr int D = NB == 1 ? 4 : (NB == 2 ? 2 : 1);
    const int nkb = (K + 15) / 16;
    for (int kb0 = w; kb0 < nkb; kb0 += 4 * D) {
      T a[D][2][4], x[D][NB][4];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const int kbase = (kb0 + 4 * d) * 16 + 4 * lk;  // past K when the block is past nkb
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
          const int i = m0 + 16 * rb + lr;
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int k = kbase + s;
            const bool ok = i < M && k < K;
            a[d][rb][s] = ok ? (TA ? gptr(job.A)[(size_t)i * job.lda + k]
                                   : gptr(job.A)[(size_t)k * job.lda + i])
                             : T(0);
          }
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const int j = n0 + 16 * b + lr;
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int k = kbase + s;
            const bool ok = j < N && k < K;
            x[d][b][s] = ok ? (TB ? gptr(job.B)[(size_t)k * job.ldb + j]
                                  : gptr(job.B)[(size_t)j * job.ldb + k])
                            : T(0);
          }
        }
      }
#pragma unroll
      for (int d = 0; d < D; ++d)
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
          for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int b = 0; b < NB; ++b)
              acc[rb][b] = Mfma<T>::mac(a[d][rb][s], x[d][b][s], acc[rb][b]);
    }
    // partial sums of waves 1..3 through LDS, added to wave 0's in wave order
    if (w > 0) {
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) red[w - 1][(rb * NB + b) * 4 + r][lane] = acc[rb][b][r];
    }
    __syncthreads();
    if (w == 0) {
      const T alpha = job.alpha, beta = job.beta;
      const int tri = job.tri;
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int q = (rb * NB + b) * 4 + r;
            const T v = ((acc[rb][b][r] + red[0][q][lane]) + red[1][q][lane]) + red[2][q][lane];
            // result lane l, item r: row Mfma<T>::row(l/16, r), column l%16 of the 16x16 block
            const int m = m0 + 16 * rb + Mfma<T>::row(lk, r);
            const int n = n0 + 16 * b + lr;
            bool ok = m < M && n < N;
            if (tri == 1) ok = ok && m <= n;
            else if (tri == 2) ok = ok && m >= n;
            if (ok) {
              auto c = gptr(job.C + (size_t)n * job.ldc + m);
              *c = (beta == T(0)) ? alpha * v : alpha * v + beta * *c;
            }
          }
    }
    __syncthreads();  // LDS is rewritten by the next tile
  }
}

// Trailing update of one POTRF panel fused with the factorization of the
// next diagonal block.  The SYRK's packed triangle tiles are 64 x 64, so
// tile 0 is exactly the next 64 x 64 diagonal block: workgroup 0 updates it
// and then its first wave factors it (potrf_diag.h) while the other
// workgroups update the rest of the triangle -- one launch per panel
// fewer on the critical path, and the block factorization overlaps the
// update instead of following it.
template <typename T, bool UPPER, int VEC>
__global__ __launch_bounds__(256) void potrf_update_kernel(GemmJob<T> job, T* __restrict__ A, int n,
                                                           int lda, int j1,
                                                           int* __restrict__ info,
                                                           int* __restrict__ abort,
                                                           int global_base) {
  typedef GemmKernel<T, 64, 64, 16, 2, 2, UPPER, !UPPER, VEC> G;
  __shared__ T smem[2 * (G::A_ELEMS + G::B_ELEMS)];
  static_assert(2 * (G::A_ELEMS + G::B_ELEMS) >= 129, "LDS reuse for the block step");
  if (abort && *abort) return;
  if (info && *info) return;
  raise_wave_priority();  // always on the critical path of a factorization
  const int t = blockIdx.x;
  GemmJob<T> jb;
  int tm, tn;
  locate_tile<T>(nullptr, 0, job, t, jb, tm, tn, true);
  G::run(jb, tm, tn, smem);
  if (t != 0) return;
  __syncthreads();  // the block's update is complete and visible to wave 0; LDS is free
  potrf_diag_block<T, UPPER, 64>(A, n, lda, j1, info, abort, global_base, smem, smem + 128);
}

// ------------------------------------------------------------ host side
// Configurations (FP64; measured on the k-contiguous Cholesky update shape):
//   id 4: 256x128 tile, 16 waves of 64x32, BK 16 (110 KB LDS) -- big launches
//   id 0: 128x128 tile, 16 waves of 32x32, BK 32 (139 KB LDS)
//   id 1:  64x64  tile,  4 waves of 32x32, BK 16 -- small launches and
//          products of at most 64 columns
//   id 6:  64x64  tile,  4 waves of 32x32, BK 32, two tiles of operands in
//          registers ahead -- products of 17..64 columns with K >= 512 in
//          launches of >= 64 such tiles
// FP32 keeps 128x128 with 4 waves of 64x64 (id 0) and 64x64 (id 1).
// 16-wave workgroups put four waves on each SIMD, which hides LDS and
// global latency that a single 64x64-per-wave workgroup leaves exposed.
namespace {
constexpr int GEMM_CFG_256x128 = 4;
constexpr int GEMM_CFG_SKINNY = 5;  // 32 rows x bn (16..64) columns, K split over 4 waves
constexpr int GEMM_CFG_64x64_K32 = 6;  // 64 x 64 tiles, 32-deep k-blocks, 2-deep prefetch
}

template <typename T>
GemmCfgId gemm_pick_cfg(int M, int N, int K, int njobs, bool vec) {
  const long nj = njobs > 0 ? njobs : 1;
  // A launch of at most 64 columns too small to fill the device (fewer than
  // 64 tiles of 64 x 64, e.g. the critical-path products of the ldiv! sweeps)
  // runs the K-split skinny kernel on 32 x 16 tiles: four times the
  // workgroups, each with a quarter of the columns.  An element's arithmetic
  // does not depend on the column split.
  if (N <= 64 && (long)((M + 63) / 64) * ((N + 63) / 64) * nj < 64)
    return GemmCfgId{GEMM_CFG_SKINNY, 32, 16};
  // 17..48 columns with a long K in a large launch: the 64 x 64 tile with
  // two-deep prefetch, although it computes up to half its columns for
  // nothing (the 32-column bulk updates of the refinement solves of an FP32
  // factor at N = 32768: the mixed solve 126.7 -> 122.7 ms)
  if (N > 16 && N <= 48 && K >= 512) return GemmCfgId{GEMM_CFG_64x64_K32, 64, 64};
  // few columns: the K-split skinny kernel (at 64 columns the 64 x 64 tile
  // is faster in large launches: ldiv! with 64 right-hand sides at N = 32768
  // took 9.2 ms with the skinny kernel for its bulk updates, 7.0 ms with
  // the tile)
  if (N <= 48) return GemmCfgId{GEMM_CFG_SKINNY, 32, 16 * ((N + 15) / 16)};
  // at most 64 columns: the 64 x 64 tile (a 128-wide tile would compute
  // half its columns for nothing).  With a long K a workgroup's k-loop is a
  // chain of dependent memory latencies, so it moves 32-deep k-blocks: the
  // ldiv! sweeps' bulk updates (K = 1024, 64 columns) took 127 us instead of
  // about 140 us per launch (64-deep blocks were slower again).
  if (N <= 64) return GemmCfgId{K >= 512 ? GEMM_CFG_64x64_K32 : 1, 64, 64};
  if (sizeof(T) == 8 && vec) {
    // Largest tile whose launch still puts about one workgroup on every CU
    // (scalar-load launches keep the 128 x 128 tile: the 256 x 128 one does
    // not fit its registers without spilling inside the k-loop).
    const long t256 = (long)((M + 255) / 256) * ((N + 127) / 128) * nj;
    if (t256 >= 256) return GemmCfgId{GEMM_CFG_256x128, 256, 128};
  }
  const long t128 = (long)((M + 127) / 128) * ((N + 127) / 128) * nj;
  if (t128 >= 192) return GemmCfgId{0, 128, 128};
  return GemmCfgId{1, 64, 64};
}

template <typename T>
long gemm_prepare_jobs(GemmJob<T>* jobs, int njobs, GemmCfgId cfg) {
  long total = 0;
  for (int i = 0; i < njobs; ++i) {
    GemmJob<T>& j = jobs[i];
    j.tiles_m = (j.M + cfg.bm - 1) / cfg.bm;
    j.tiles_n = (j.N + cfg.bn - 1) / cfg.bn;
    j.wg_begin = (int)total;
    long cnt;
    if (j.tri == 0 || cfg.bm != cfg.bn || cfg.id == GEMM_CFG_SKINNY) cnt = (long)j.tiles_m * j.tiles_n;
    else cnt = (long)j.tiles_m * (j.tiles_m + 1) / 2;  // square tiles of a square (M == N) C
    if (j.M <= 0 || j.N <= 0) cnt = 0;
    total += cnt;
  }
  return total;
}

namespace {
template <typename T, int BM, int BN, int BK, int WM, int WN, bool TA, bool TB, int VEC,
          int PF = 1>
int launch_one(const GemmJob<T>* jobs, int njobs, const GemmJob<T>& single, long total, long grid,
               const int* abort, hipStream_t s) {
  hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, TA, TB, VEC, PF>), dim3((unsigned)grid),
                     dim3(64 * WM * WN), 0, s, jobs, njobs, single, (int)total, abort);
  DM_LAUNCH_CHECK("gemm_kernel");
  return 0;
}
template <typename T, bool TA, bool TB, int VEC>
int launch_cfg(int id, const GemmJob<T>* jobs, int njobs, const GemmJob<T>& single, long total,
               long grid, const int* abort, hipStream_t s) {
  if constexpr (sizeof(T) == 8) {
    if (id == GEMM_CFG_256x128)
      return launch_one<T, 256, 128, 16, 4, 4, TA, TB, VEC>(jobs, njobs, single, total, grid, abort,
                                                            s);
    if (id == 0)
      return launch_one<T, 128, 128, 32, 4, 4, TA, TB, VEC>(jobs, njobs, single, total, grid, abort,
                                                            s);
  } else {
    if (id == 0)
      return launch_one<T, 128, 128, 16, 2, 2, TA, TB, VEC>(jobs, njobs, single, total, grid, abort,
                                                            s);
  }
  if (id == 1)
    return launch_one<T, 64, 64, 16, 2, 2, TA, TB, VEC>(jobs, njobs, single, total, grid, abort, s);
  if (id == GEMM_CFG_64x64_K32)
    return launch_one<T, 64, 64, 32, 2, 2, TA, TB, VEC, 2>(jobs, njobs, single, total, grid, abort,
                                                          s);
  return arg_fail(0, "gemm: unknown configuration");
}
template <typename T, bool TA, bool TB>
int launch_skinny(int bn, const GemmJob<T>* jobs, int njobs, const GemmJob<T>& single, long total,
                  long grid, const int* abort, hipStream_t s) {
  switch (bn) {
    case 16:
      hipLaunchKernelGGL((gemm_skinny_kernel<T, TA, TB, 1>), dim3((unsigned)grid), dim3(256), 0, s,
                         jobs, njobs, single, (int)total, abort);
      break;
    case 32:
      hipLaunchKernelGGL((gemm_skinny_kernel<T, TA, TB, 2>), dim3((unsigned)grid), dim3(256), 0, s,
                         jobs, njobs, single, (int)total, abort);
      break;
    case 48:
      hipLaunchKernelGGL((gemm_skinny_kernel<T, TA, TB, 3>), dim3((unsigned)grid), dim3(256), 0, s,
                         jobs, njobs, single, (int)total, abort);
      break;
    case 64:
      hipLaunchKernelGGL((gemm_skinny_kernel<T, TA, TB, 4>), dim3((unsigned)grid), dim3(256), 0, s,
                         jobs, njobs, single, (int)total, abort);
      break;
    default:
      return arg_fail(0, "gemm: skinny width must be 16, 32, 48 or 64");
  }
  DM_LAUNCH_CHECK("gemm_skinny_kernel");
  return 0;
}
template <typename T, int VEC>
int launch_tt(bool ta, bool tb, int id, const GemmJob<T>* jobs, int njobs, const GemmJob<T>& single,
              long total, long grid, const int* abort, hipStream_t s) {
  if (ta && tb) return launch_cfg<T, true, true, VEC>(id, jobs, njobs, single, total, grid, abort, s);
  if (ta) return launch_cfg<T, true, false, VEC>(id, jobs, njobs, single, total, grid, abort, s);
  if (tb) return launch_cfg<T, false, true, VEC>(id, jobs, njobs, single, total, grid, abort, s);
  return launch_cfg<T, false, false, VEC>(id, jobs, njobs, single, total, grid, abort, s);
}
// Vector (2-wide) loads need: even leading dimensions, 2-element aligned
// pointers and even extents along each operand's contiguous axis.
template <typename T>
bool job_vec_ok(bool ta, bool tb, const GemmJob<T>& j) {
  const uintptr_t al = 2 * sizeof(T) - 1;
  if (((uintptr_t)j.A & al) || ((uintptr_t)j.B & al)) return false;
  if ((j.lda & 1) || (j.ldb & 1)) return false;
  if (ta ? (j.K & 1) : (j.M & 1)) return false;
  if (tb ? (j.N & 1) : (j.K & 1)) return false;
  return true;
}
}  // namespace

template <typename T>
int gemm_launch(bool ta, bool tb, bool vec, GemmCfgId cfg, const GemmJob<T>* jobs_dev, int njobs,
                const GemmJob<T>& single, long total_wg, const int* abort, hipStream_t s,
                long max_wg) {
  if (total_wg <= 0) return 0;
  if (total_wg > 0x7fffffffL) return arg_fail(0, "gemm: too many workgroups");
  const long grid = (max_wg > 0 && max_wg < total_wg) ? max_wg : total_wg;
  if (cfg.id == GEMM_CFG_BF16X3 || cfg.id == GEMM_CFG_BF16X6) {
    if constexpr (std::is_same<T, float>::value)
      return gemm_split_launch(cfg.id == GEMM_CFG_BF16X6 ? 6 : 3, cfg.bm, jobs_dev, njobs,
                               single, total_wg, grid, abort, s);
    return arg_fail(0, "gemm: BF16 split configurations are FP32 only");
  }
  if (cfg.id == GEMM_CFG_SKINNY) {
    if (ta && tb) return launch_skinny<T, true, true>(cfg.bn, jobs_dev, njobs, single, total_wg, grid, abort, s);
    if (ta) return launch_skinny<T, true, false>(cfg.bn, jobs_dev, njobs, single, total_wg, grid, abort, s);
    if (tb) return launch_skinny<T, false, true>(cfg.bn, jobs_dev, njobs, single, total_wg, grid, abort, s);
    return launch_skinny<T, false, false>(cfg.bn, jobs_dev, njobs, single, total_wg, grid, abort, s);
  }
  if (vec) return launch_tt<T, 2>(ta, tb, cfg.id, jobs_dev, njobs, single, total_wg, grid, abort, s);
  return launch_tt<T, 1>(ta, tb, cfg.id, jobs_dev, njobs, single, total_wg, grid, abort, s);
}

template <typename T>
bool gemm_job_vec_ok(bool ta, bool tb, const GemmJob<T>& j) { return job_vec_ok(ta, tb, j); }

int gemm_wgs_per_cu(GemmCfgId cfg) {
  // the 128x128 and 256x128 workgroups and the 16-wave BF16-split ones
  // fill a CU's LDS (or, FP32 128x128, its registers); 64x64 fits two
  return (cfg.id == 0 || cfg.id == GEMM_CFG_256x128 || cfg.id == GEMM_CFG_BF16X3 ||
          cfg.id == GEMM_CFG_BF16X6) ? 1 : 2;
}

template <typename T>
int gemm_single(bool ta, bool tb, int M, int N, int K, T alpha, const T* A, int lda, const T* B,
                int ldb, T beta, T* C, int ldc, int tri, const int* abort, hipStream_t s) {
  if (M <= 0 || N <= 0) return 0;
  GemmJob<T> j{};
  j.A = A; j.B = B; j.C = C;
  j.M = M; j.N = N; j.K = K;
  j.lda = lda; j.ldb = ldb; j.ldc = ldc;
  j.alpha = alpha; j.beta = beta; j.tri = tri;
  if (K <= 0 || alpha == T(0)) { j.alpha = T(0); j.K = 0; }
  const bool vec = job_vec_ok(ta, tb, j);
  GemmCfgId cfg = gemm_pick_cfg<T>(M, N, K, 1, vec);
  long total = gemm_prepare_jobs<T>(&j, 1, cfg);
  return gemm_launch<T>(ta, tb, vec, cfg, nullptr, 0, j, total, abort, s, 0);
}

template GemmCfgId gemm_pick_cfg<double>(int, int, int, int, bool);
template GemmCfgId gemm_pick_cfg<float>(int, int, int, int, bool);
template long gemm_prepare_jobs<double>(GemmJob<double>*, int, GemmCfgId);
template long gemm_prepare_jobs<float>(GemmJob<float>*, int, GemmCfgId);
template int gemm_launch<double>(bool, bool, bool, GemmCfgId, const GemmJob<double>*, int,
                                 const GemmJob<double>&, long, const int*, hipStream_t, long);
template int gemm_launch<float>(bool, bool, bool, GemmCfgId, const GemmJob<float>*, int,
                                const GemmJob<float>&, long, const int*, hipStream_t, long);
template bool gemm_job_vec_ok<double>(bool, bool, const GemmJob<double>&);
template bool gemm_job_vec_ok<float>(bool, bool, const GemmJob<float>&);
template int gemm_single<double>(bool, bool, int, int, int, double, const double*, int,
                                 const double*, int, double, double*, int, int, const int*,
                                 hipStream_t);
template int gemm_single<float>(bool, bool, int, int, int, float, const float*, int, const float*,
                                int, float, float*, int, int, const int*, hipStream_t);


template <typename T>
int potrf_update_launch(bool upper, int n, T* A, int lda, int j0, int* info, int* abort,
                        int global_base, hipStream_t s) {
  const int kb = 64, rest = n - j0 - kb;
  if (rest <= 0) return 0;
  GemmJob<T> j{};
  if (upper) {  // A22 -= U12^T U12
    j.A = j.B = A + (size_t)(j0 + kb) * lda + j0;
    j.tri = 1;
  } else {      // A22 -= A21 A21^T
    j.A = j.B = A + (size_t)j0 * lda + j0 + kb;
    j.tri = 2;
  }
  j.C = A + (size_t)(j0 + kb) * lda + j0 + kb;
  j.M = j.N = rest;
  j.K = kb;
  j.lda = j.ldb = j.ldc = lda;
  j.alpha = T(-1);
  j.beta = T(1);
  const long total = gemm_prepare_jobs<T>(&j, 1, GemmCfgId{1, 64, 64});
  const bool vec = job_vec_ok(upper, !upper, j);
  const int j1 = j0 + kb;
  if (upper) {
    if (vec)
      hipLaunchKernelGGL((potrf_update_kernel<T, true, 2>), dim3((unsigned)total), dim3(256), 0, s, j,
                         A, n, lda, j1, info, abort, global_base);
    else
      hipLaunchKernelGGL((potrf_update_kernel<T, true, 1>), dim3((unsigned)total), dim3(256), 0, s, j,
                         A, n, lda, j1, info, abort, global_base);
  } else {
    if (vec)
      hipLaunchKernelGGL((potrf_update_kernel<T, false, 2>), dim3((unsigned)total), dim3(256), 0, s,
                         j, A, n, lda, j1, info, abort, global_base);
    else
      hipLaunchKernelGGL((potrf_update_kernel<T, false, 1>), dim3((unsigned)total), dim3(256), 0, s,
                         j, A, n, lda, j1, info, abort, global_base);
  }
  DM_LAUNCH_CHECK("potrf_update_kernel");
  return 0;
}
template int potrf_update_launch<double>(bool, int, double*, int, int, int*, int*, int, hipStream_t);
template int potrf_update_launch<float>(bool, int, float*, int, int, int*, int*, int, hipStream_t);

}  // namespace dm

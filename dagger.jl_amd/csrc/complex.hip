// Complex (ComplexF64 / ComplexF32) tile operations of the tiled Cholesky
// and its solves: LAPACK.potrf!, BLAS.trsm!, BLAS.herk! and BLAS.gemm! with
// conjugate transposes, as _chol! calls them for T <: Complex
// (src/array/cholesky.jl:27-28,44-45,48-50,89-95) and trsm! passes 'C'
// through to its GEMM tasks (src/array/trsm.jl:103-111,157-163).
//
// Elements are (re, im) pairs of R (Julia's layout).  Design:
//  * GEMM/HERK: 64 x 64 tiles of C, 4 waves of 32 x 32, on the real matrix
//    cores (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32): each complex
//    product takes four real ones, re += ar*br - ai*bi, im += ar*bi + ai*br.
//    op(A) and op(B) are staged into LDS as separate re/im planes with the
//    conjugation of 'C' applied while staging.  HERK is the same kernel
//    restricted to one triangle of C with the imaginary part of C's diagonal
//    set to zero (BLAS herk semantics).
//  * POTRF: the recursion n -> (n1, n2) of LAPACK's potrf2 (factor A11;
//    A12 := A11^-H A12 / A21 := A21 A11^-H; A22 -= herk; factor A22) down to
//    64 x 64 leaves factored in LDS by one workgroup.
//  * TRSM: the same recursion on the triangle (solve, GEMM update, solve)
//    down to 64 x 64 leaves, where one thread solves one right-hand-side
//    vector by substitution against the triangle in LDS.
// Every kernel returns at once when the abort word is set (a failed pivot),
// so the rest of a failed factorization becomes a no-op.
#include "dm_internal.h"
#include "complex_ops.h"

namespace dm {

namespace {

template <typename R> struct Mf;
// output layout (lane l, item r; 16x16 block D[i][j], j = l%16):
//   f64 16x16x4: i = 4*r + l/16        f32 16x16x4: i = 4*(l/16) + r
template <> struct Mf<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * r + lk; }
  static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct Mf<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
  static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

constexpr int ZT = 64;        // C tile
constexpr int ZBK = 8;        // k per step
constexpr int ZS = ZT + 4;    // LDS row stride of the [k][mn] planes
constexpr int ZLEAF = 64;     // POTRF / TRSM leaf size

// element (r, c) of op(X) for a column-major X with leading dimension ld:
// op 0 'N' X[r, c], 1 'T' X[c, r], 2 'C' conj(X[c, r]), 3 conj(X[r, c])
template <typename R>
__device__ __forceinline__ void op_elem(const R* __restrict__ X, int64_t ld, int op, int r, int c,
                                        R& re, R& im) {
  const int64_t idx = (op == 0 || op == 3) ? (int64_t)c * ld + r : (int64_t)r * ld + c;
  typedef R v2 __attribute__((ext_vector_type(2)));
  const v2 v = *gptr((const v2*)(X + 2 * idx));
  re = v.x;
  im = (op >= 2) ? -v.y : v.y;
}

// C = alpha op(A) op(B) + beta C on one 64 x 64 tile per workgroup.
template <typename R>
__global__ __launch_bounds__(256) void zgemm_kernel(ZGemmArgs<R> g, const int* __restrict__ abort) {
  typedef typename Mf<R>::v4 v4;
  __shared__ R As[2][2][ZBK * ZS];  // [buffer][re/im][k * ZS + m]
  __shared__ R Bs[2][2][ZBK * ZS];  // [buffer][re/im][k * ZS + n]
  if (abort && *abort) return;
  const int tiles_m = (g.M + ZT - 1) / ZT;
  const int m0 = (blockIdx.x % tiles_m) * ZT, n0 = (blockIdx.x / tiles_m) * ZT;
  if (g.tri == 1 && m0 > n0 + ZT - 1) return;  // tile below the upper triangle
  if (g.tri == 2 && n0 > m0 + ZT - 1) return;  // tile above the lower triangle
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int wm0 = (wave & 1) * 32, wn0 = (wave >> 1) * 32;
  v4 cre[2][2], cim[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) cre[i][j] = cim[i][j] = v4{0, 0, 0, 0};

  // staging: 64 x 8 elements of each operand per step, two per thread.  For
  // an op that reads X[c, r] (k contiguous in memory) lanes walk k first.
  auto stage = [&](int buf, int k0) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = tid + 256 * q;
      const bool kfa = g.opa == 1 || g.opa == 2;
      const int am = kfa ? e / ZBK : e % ZT, ak = kfa ? e % ZBK : e / ZT;
      R re = 0, im = 0;
      if (m0 + am < g.M && k0 + ak < g.K) op_elem(g.A, g.lda, g.opa, m0 + am, k0 + ak, re, im);
      As[buf][0][ak * ZS + am] = re;
      As[buf][1][ak * ZS + am] = im;
      // op(B)[k, n]: 'N' reads B[k, n] (k contiguous)
      const bool kfb = g.opb == 0 || g.opb == 3;
      const int bn = kfb ? e / ZBK : e % ZT, bk = kfb ? e % ZBK : e / ZT;
      re = 0; im = 0;
      if (n0 + bn < g.N && k0 + bk < g.K) op_elem(g.B, g.ldb, g.opb, k0 + bk, n0 + bn, re, im);
      Bs[buf][0][bk * ZS + bn] = re;
      Bs[buf][1][bk * ZS + bn] = im;
    }
  };
  const int nk = (g.K + ZBK - 1) / ZBK;
  if (nk > 0) stage(0, 0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) stage(cur ^ 1, (kt + 1) * ZBK);  // the other buffer: read in step kt-1, ended by its barrier
#pragma unroll
    for (int kk = 0; kk < ZBK / 4; ++kk) {
      const int k = kk * 4 + lk;
      R ar[2], ai[2], br[2], bi[2], nbi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = As[cur][0][k * ZS + wm0 + i * 16 + lr];
        ai[i] = As[cur][1][k * ZS + wm0 + i * 16 + lr];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        br[j] = Bs[cur][0][k * ZS + wn0 + j * 16 + lr];
        bi[j] = Bs[cur][1][k * ZS + wn0 + j * 16 + lr];
        nbi[j] = -bi[j];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          cre[i][j] = Mf<R>::mac(br[j], ar[i], cre[i][j]);
          cre[i][j] = Mf<R>::mac(nbi[j], ai[i], cre[i][j]);
          cim[i][j] = Mf<R>::mac(bi[j], ar[i], cim[i][j]);
          cim[i][j] = Mf<R>::mac(br[j], ai[i], cim[i][j]);
        }
    }
    __syncthreads();
  }
  // lane holds C[m][n], m = m0 + wm0 + 16i + lr, n = n0 + wn0 + 16j + row(lk, r)
  typedef R v2 __attribute__((ext_vector_type(2)));
  const bool bz = g.beta_re == R(0) && g.beta_im == R(0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm0 + i * 16 + lr, n = n0 + wn0 + j * 16 + Mf<R>::row(lk, r);
        bool ok = m < g.M && n < g.N;
        if (g.tri == 1) ok = ok && m <= n;
        else if (g.tri == 2) ok = ok && m >= n;
        if (!ok) continue;
        const R pr = cre[i][j][r], pi = cim[i][j][r];
        R vr = g.alpha_re * pr - g.alpha_im * pi;
        R vi = g.alpha_re * pi + g.alpha_im * pr;
        auto c = gptr((v2*)(g.C + 2 * ((int64_t)n * g.ldc + m)));
        if (!bz) {
          const v2 o = *c;
          vr += g.beta_re * o.x - g.beta_im * o.y;
          vi += g.beta_re * o.y + g.beta_im * o.x;
        }
        if (g.herk && m == n) vi = R(0);
        *c = v2{vr, vi};
      }
}

// Cholesky of an n x n (n <= 64) Hermitian block in place, one workgroup:
// the 'U' factor in the upper triangle (A = U^H U) or the 'L' factor in the
// lower one (A = L L^H); the other strict triangle is left alone.  The
// diagonal's imaginary part is ignored and set to zero (LAPACK potrf2).  A
// non-positive or NaN pivot at column j sets *info = info_off + j + 1 (when
// *info is 0) and the abort word to global_base + info_off + j + 1 (the
// global column, as the real kernels report it), and leaves the block
// partly factored.
template <typename R, bool UPPER>
__global__ __launch_bounds__(256) void zpotrf_leaf_kernel(R* __restrict__ A, int n, int lda,
                                                          int* __restrict__ info, int info_off,
                                                          int* __restrict__ abort,
                                                          int global_base) {
  __shared__ R t[2][ZLEAF][ZLEAF + 1];  // [re/im][row][col]
  __shared__ int failed;
  if (abort && *abort) return;
  const int tid = threadIdx.x;
  typedef R v2 __attribute__((ext_vector_type(2)));
  for (int e = tid; e < n * n; e += 256) {
    const int r = e % n, c = e / n;
    const v2 v = *gptr((const v2*)(A + 2 * ((int64_t)c * lda + r)));
    t[0][r][c] = v.x;
    t[1][r][c] = v.y;
  }
  if (tid == 0) failed = 0;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {
      const R d = t[0][j][j];
      if (!(d > R(0))) {
        failed = 1;
        if (info && *info == 0) *info = info_off + j + 1;
        if (abort) *abort = global_base + info_off + j + 1;
      } else {
        t[0][j][j] = sqrt(d);
        t[1][j][j] = R(0);
      }
    }
    __syncthreads();
    if (failed) break;
    const R d = t[0][j][j];
    // scale the rest of row j (U) / column j (L)
    for (int k = j + 1 + tid; k < n; k += 256) {
      if (UPPER) { t[0][j][k] /= d; t[1][j][k] /= d; }
      else { t[0][k][j] /= d; t[1][k][j] /= d; }
    }
    __syncthreads();
    // trailing update of the stored triangle
    const int w = n - j - 1;
    for (int e = tid; e < w * w; e += 256) {
      const int a = j + 1 + e % w, b = j + 1 + e / w;  // a row, b column
      if (UPPER) {
        if (a > b) continue;
        // A[a][b] -= conj(U[j][a]) U[j][b]
        const R ur = t[0][j][a], ui = t[1][j][a], vr = t[0][j][b], vi = t[1][j][b];
        t[0][a][b] -= ur * vr + ui * vi;
        t[1][a][b] -= ur * vi - ui * vr;
      } else {
        if (a < b) continue;
        // A[a][b] -= L[a][j] conj(L[b][j])
        const R lr = t[0][a][j], li = t[1][a][j], mr = t[0][b][j], mi = t[1][b][j];
        t[0][a][b] -= lr * mr + li * mi;
        t[1][a][b] -= li * mr - lr * mi;
      }
    }
    __syncthreads();
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += 256) {
    const int r = e % n, c = e / n;
    if (UPPER ? r > c : r < c) continue;
    *gptr((v2*)(A + 2 * ((int64_t)c * lda + r))) = v2{t[0][r][c], t[1][r][c]};
  }
}

// Solves op(T) x = alpha b for the vectors of B, T an n x n (n <= 64)
// triangle: one thread per vector, the triangle in LDS.  Vector v's element
// i is B[v * vstride + i * estride].  op: 0 T, 1 T^T, 2 T^H, 3 conj(T).
template <typename R>
__global__ __launch_bounds__(64) void ztrsm_leaf_kernel(const R* __restrict__ T, int n, int ldt,
                                                        bool lower, int op, bool unit, R alpha_re,
                                                        R alpha_im, R* __restrict__ B,
                                                        int64_t estride, int64_t vstride,
                                                        int nvec, const int* __restrict__ abort) {
  __shared__ R t[2][ZLEAF][ZLEAF + 1];  // op(T)[i][k]
  __shared__ R x[2][ZLEAF][64];         // this workgroup's vectors
  if (abort && *abort) return;
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += 64) {
    const int i = e % n, k = e / n;
    R re, im;
    op_elem(T, ldt, op, i, k, re, im);
    t[0][i][k] = re;
    t[1][i][k] = im;
  }
  __syncthreads();
  const int v = blockIdx.x * 64 + tid;
  if (v >= nvec) return;
  typedef R v2 __attribute__((ext_vector_type(2)));
  R* bv = B + 2 * (int64_t)v * vstride;
  for (int i = 0; i < n; ++i) {
    const v2 b = *gptr((const v2*)(bv + 2 * (int64_t)i * estride));
    x[0][i][tid] = alpha_re * b.x - alpha_im * b.y;
    x[1][i][tid] = alpha_re * b.y + alpha_im * b.x;
  }
  // op(T) is lower for (lower, N / conj) and (upper, T / H)
  const bool lo = (op == 0 || op == 3) ? lower : !lower;
  for (int s = 0; s < n; ++s) {
    const int i = lo ? s : n - 1 - s;
    R sr = x[0][i][tid], si = x[1][i][tid];
    const int kb = lo ? 0 : i + 1, ke = lo ? i : n;
    for (int k = kb; k < ke; ++k) {
      const R tr = t[0][i][k], ti = t[1][i][k], xr = x[0][k][tid], xi = x[1][k][tid];
      sr -= tr * xr - ti * xi;
      si -= tr * xi + ti * xr;
    }
    if (!unit) {  // s / d by the textbook complex quotient
      const R dr = t[0][i][i], di = t[1][i][i];
      const R den = dr * dr + di * di;
      const R qr = (sr * dr + si * di) / den, qi = (si * dr - sr * di) / den;
      sr = qr;
      si = qi;
    }
    x[0][i][tid] = sr;
    x[1][i][tid] = si;
  }
  for (int i = 0; i < n; ++i)
    *gptr((v2*)(bv + 2 * (int64_t)i * estride)) = v2{x[0][i][tid], x[1][i][tid]};
}

int op_code(char t) { return is_notrans(t) ? 0 : (t == 'T' || t == 't') ? 1 : 2; }

int split_point(int n) { return ((n / 2 + 31) / 32) * 32; }

}  // namespace

template <typename R>
int zgemm_launch(const ZGemmArgs<R>& g, const int* abort, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return 0;
  const long tiles = (long)((g.M + ZT - 1) / ZT) * ((g.N + ZT - 1) / ZT);
  if (tiles > 0x7fffffffL) return arg_fail(0, "zgemm: too many tiles");
  hipLaunchKernelGGL((zgemm_kernel<R>), dim3((unsigned)tiles), dim3(256), 0, s, g, abort);
  DM_LAUNCH_CHECK("zgemm_kernel");
  return 0;
}

template <typename R>
int zgemm_call(char ta, char tb, int m, int n, int k, R ar, R ai, const R* A, int lda, const R* B,
               int ldb, R br, R bi, R* C, int ldc, int tri, int herk, const int* abort,
               hipStream_t s) {
  ZGemmArgs<R> g{};
  g.A = A; g.B = B; g.C = C;
  g.M = m; g.N = n; g.K = (ar == R(0) && ai == R(0)) ? 0 : k;
  g.lda = lda; g.ldb = ldb; g.ldc = ldc;
  g.alpha_re = ar; g.alpha_im = ai; g.beta_re = br; g.beta_im = bi;
  g.opa = op_code(ta); g.opb = op_code(tb);
  g.tri = tri; g.herk = herk;
  return zgemm_launch<R>(g, abort, s);
}

// trsm(side, uplo, trans, diag, alpha, T, B) by recursion on the triangle
template <typename R>
int ztrsm_rec(bool left, bool lower, int op, bool unit, int L, R ar, R ai, const R* T, int ldt,
              R* B, int ldb, int nv, const int* abort, hipStream_t s) {
  if (L <= 0 || nv <= 0) return 0;
  if (L <= ZLEAF) {
    // left: vectors are B's columns; right: x op(T) = b as op(T)^T x^T = b^T
    // on B's rows (op N<->T, H<->conj)
    const int lop = left ? op : (op == 0 ? 1 : op == 1 ? 0 : op == 2 ? 3 : 2);
    const int64_t es = left ? 1 : ldb, vs = left ? ldb : 1;
    hipLaunchKernelGGL((ztrsm_leaf_kernel<R>), dim3((nv + 63) / 64), dim3(64), 0, s, T, L, ldt,
                       lower, lop, unit, ar, ai, B, es, vs, nv, abort);
    DM_LAUNCH_CHECK("ztrsm_leaf_kernel");
    return 0;
  }
  const int P = split_point(L), Q = L - P;
  const R* T11 = T;
  const R* T22 = T + 2 * ((int64_t)P * ldt + P);
  const R* T21 = T + 2 * (int64_t)P;               // rows P.., cols 0..P-1
  const R* T12 = T + 2 * ((int64_t)P * ldt);       // rows 0..P-1, cols P..
  const bool lo = (op == 0) ? lower : !lower;      // op(T) lower
  const char opc = op == 0 ? 'N' : op == 1 ? 'T' : 'C';
  // op(T)21 = op(T21) for N, op(T12) for T/H; op(T)12 the mirror
  const R* O21 = op == 0 ? T21 : T12;
  const R* O12 = op == 0 ? T12 : T21;
  int rc;
  if (left) {
    R* B1 = B;
    R* B2 = B + 2 * (int64_t)P;  // rows P..
    if (lo) {
      if ((rc = ztrsm_rec<R>(true, lower, op, unit, P, ar, ai, T11, ldt, B1, ldb, nv, abort, s))) return rc;
      if ((rc = zgemm_call<R>(opc, 'N', Q, nv, P, R(-1), R(0), O21, ldt, B1, ldb, ar, ai, B2, ldb, 0, 0, abort, s))) return rc;
      return ztrsm_rec<R>(true, lower, op, unit, Q, R(1), R(0), T22, ldt, B2, ldb, nv, abort, s);
    }
    if ((rc = ztrsm_rec<R>(true, lower, op, unit, Q, ar, ai, T22, ldt, B2, ldb, nv, abort, s))) return rc;
    if ((rc = zgemm_call<R>(opc, 'N', P, nv, Q, R(-1), R(0), O12, ldt, B2, ldb, ar, ai, B1, ldb, 0, 0, abort, s))) return rc;
    return ztrsm_rec<R>(true, lower, op, unit, P, R(1), R(0), T11, ldt, B1, ldb, nv, abort, s);
  }
  // right: X op(T) = alpha B, B = [B1 B2] by columns, nv = rows of B
  R* B1 = B;
  R* B2 = B + 2 * ((int64_t)P * ldb);
  if (!lo) {  // op(T) upper: X1 first, then X2 op(T)22 = alpha B2 - X1 op(T)12
    if ((rc = ztrsm_rec<R>(false, lower, op, unit, P, ar, ai, T11, ldt, B1, ldb, nv, abort, s))) return rc;
    if ((rc = zgemm_call<R>('N', opc, nv, Q, P, R(-1), R(0), B1, ldb, O12, ldt, ar, ai, B2, ldb, 0, 0, abort, s))) return rc;
    return ztrsm_rec<R>(false, lower, op, unit, Q, R(1), R(0), T22, ldt, B2, ldb, nv, abort, s);
  }
  if ((rc = ztrsm_rec<R>(false, lower, op, unit, Q, ar, ai, T22, ldt, B2, ldb, nv, abort, s))) return rc;
  if ((rc = zgemm_call<R>('N', opc, nv, P, Q, R(-1), R(0), B2, ldb, O21, ldt, ar, ai, B1, ldb, 0, 0, abort, s))) return rc;
  return ztrsm_rec<R>(false, lower, op, unit, P, R(1), R(0), T11, ldt, B1, ldb, nv, abort, s);
}

template <typename R>
int ztrsm_call(char side, char uplo, char trans, char diag, int m, int n, R ar, R ai, const R* A,
               int lda, R* B, int ldb, const int* abort, hipStream_t s) {
  const bool left = is_left(side);
  return ztrsm_rec<R>(left, is_lower(uplo), op_code(trans), is_unit(diag), left ? m : n, ar, ai,
                      A, lda, B, ldb, left ? n : m, abort, s);
}

// potrf(uplo, A) by potrf2's recursion; info gets the global 1-based column
// of the first failed pivot (offset info_off), abort is set with it.
template <typename R>
int zpotrf_rec(bool upper, int n, R* A, int lda, int* info, int info_off, int* abort,
               int global_base, hipStream_t s) {
  if (n <= 0) return 0;
  if (n <= ZLEAF) {
    if (upper)
      hipLaunchKernelGGL((zpotrf_leaf_kernel<R, true>), dim3(1), dim3(256), 0, s, A, n, lda, info,
                         info_off, abort, global_base);
    else
      hipLaunchKernelGGL((zpotrf_leaf_kernel<R, false>), dim3(1), dim3(256), 0, s, A, n, lda, info,
                         info_off, abort, global_base);
    DM_LAUNCH_CHECK("zpotrf_leaf_kernel");
    return 0;
  }
  const int P = split_point(n), Q = n - P;
  R* A11 = A;
  R* A22 = A + 2 * ((int64_t)P * lda + P);
  int rc = zpotrf_rec<R>(upper, P, A11, lda, info, info_off, abort, global_base, s);
  if (rc) return rc;
  if (upper) {
    R* A12 = A + 2 * ((int64_t)P * lda);
    if ((rc = ztrsm_call<R>('L', 'U', 'C', 'N', P, Q, R(1), R(0), A11, lda, A12, lda, abort, s))) return rc;
    if ((rc = zgemm_call<R>('C', 'N', Q, Q, P, R(-1), R(0), A12, lda, A12, lda, R(1), R(0), A22, lda, 1, 1, abort, s))) return rc;
  } else {
    R* A21 = A + 2 * (int64_t)P;
    if ((rc = ztrsm_call<R>('R', 'L', 'C', 'N', Q, P, R(1), R(0), A11, lda, A21, lda, abort, s))) return rc;
    if ((rc = zgemm_call<R>('N', 'C', Q, Q, P, R(-1), R(0), A21, lda, A21, lda, R(1), R(0), A22, lda, 2, 1, abort, s))) return rc;
  }
  return zpotrf_rec<R>(upper, Q, A22, lda, info, info_off + P, abort, global_base, s);
}

template int zgemm_call<double>(char, char, int, int, int, double, double, const double*, int,
                                const double*, int, double, double, double*, int, int, int,
                                const int*, hipStream_t);
template int zgemm_call<float>(char, char, int, int, int, float, float, const float*, int,
                               const float*, int, float, float, float*, int, int, int, const int*,
                               hipStream_t);
template int ztrsm_call<double>(char, char, char, char, int, int, double, double, const double*,
                                int, double*, int, const int*, hipStream_t);
template int ztrsm_call<float>(char, char, char, char, int, int, float, float, const float*, int,
                               float*, int, const int*, hipStream_t);
template int zpotrf_rec<double>(bool, int, double*, int, int*, int, int*, int, hipStream_t);
template int zpotrf_rec<float>(bool, int, float*, int, int*, int, int*, int, hipStream_t);

}  // namespace dm

// ------------------------------------------------------------ C ABI
using namespace dm;

namespace {

template <typename R>
int zpotrf_api(char uplo, int n, void* A, int lda, int* info_dev, dm_stream_t s) {
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(1, "uplo must be 'U' or 'L'");
  if (n < 0) return arg_fail(2, "n < 0");
  if (lda < (n > 1 ? n : 1)) return arg_fail(4, "lda < max(1,n)");
  if (!info_dev) return arg_fail(5, "info is NULL");
  int rc = ensure_init();
  if (rc) return rc;
  hipStream_t st = stream_of(s);
  // the abort word of this call: a failed pivot makes the rest of the
  // recursion a no-op (per-stream scratch; calls on one stream are ordered)
  int* abort = (int*)stream_workspace(st, 256);
  if (!abort) return arg_fail(0, "potrf: workspace allocation failed");
  if ((rc = fill_i32_launch(info_dev, 0, 1, st))) return rc;
  if ((rc = fill_i32_launch(abort, 0, 1, st))) return rc;
  return zpotrf_rec<R>(is_upper(uplo), n, (R*)A, lda, info_dev, 0, abort, 0, st);
}

template <typename R>
int ztrsm_api(char side, char uplo, char transa, char diag, int m, int n, const R* alpha,
              const void* A, int lda, void* B, int ldb, dm_stream_t s) {
  if (!is_left(side) && !is_right(side)) return arg_fail(1, "side must be 'L' or 'R'");
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(2, "uplo must be 'U' or 'L'");
  if (!is_trans(transa) && !is_notrans(transa)) return arg_fail(3, "transa must be 'N','T','C'");
  if (!is_unit(diag) && !is_nonunit(diag)) return arg_fail(4, "diag must be 'N' or 'U'");
  if (m < 0) return arg_fail(5, "m < 0");
  if (n < 0) return arg_fail(6, "n < 0");
  if (!alpha) return arg_fail(7, "alpha is NULL");
  const int ka = is_left(side) ? m : n;
  if (lda < (ka > 1 ? ka : 1)) return arg_fail(9, "lda too small");
  if (ldb < (m > 1 ? m : 1)) return arg_fail(11, "ldb < max(1,m)");
  int rc = ensure_init();
  if (rc) return rc;
  return ztrsm_call<R>(side, uplo, transa, diag, m, n, alpha[0], alpha[1], (const R*)A, lda,
                       (R*)B, ldb, nullptr, stream_of(s));
}

template <typename R>
int zherk_api(char uplo, char trans, int n, int k, R alpha, const void* A, int lda, R beta,
              void* C, int ldc, dm_stream_t s) {
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(1, "uplo must be 'U' or 'L'");
  if (!is_notrans(trans) && trans != 'C' && trans != 'c') return arg_fail(2, "trans must be 'N' or 'C'");
  if (n < 0) return arg_fail(3, "n < 0");
  if (k < 0) return arg_fail(4, "k < 0");
  const int rows = is_notrans(trans) ? n : k;
  if (lda < (rows > 1 ? rows : 1)) return arg_fail(7, "lda too small");
  if (ldc < (n > 1 ? n : 1)) return arg_fail(10, "ldc < max(1,n)");
  int rc = ensure_init();
  if (rc) return rc;
  if (n == 0 || ((alpha == R(0) || k == 0) && beta == R(1))) return 0;  // BLAS quick return
  const bool tr = !is_notrans(trans);
  // 'C': C = alpha A^H A + beta C;  'N': C = alpha A A^H + beta C
  return zgemm_call<R>(tr ? 'C' : 'N', tr ? 'N' : 'C', n, n, k, alpha, R(0), (const R*)A, lda,
                       (const R*)A, lda, beta, R(0), (R*)C, ldc, is_upper(uplo) ? 1 : 2, 1,
                       nullptr, stream_of(s));
}

template <typename R>
int zgemm_api(char transa, char transb, int m, int n, int k, const R* alpha, const void* A,
              int lda, const void* B, int ldb, const R* beta, void* C, int ldc, dm_stream_t s) {
  if (!is_trans(transa) && !is_notrans(transa)) return arg_fail(1, "transa must be 'N','T','C'");
  if (!is_trans(transb) && !is_notrans(transb)) return arg_fail(2, "transb must be 'N','T','C'");
  if (m < 0) return arg_fail(3, "m < 0");
  if (n < 0) return arg_fail(4, "n < 0");
  if (k < 0) return arg_fail(5, "k < 0");
  if (!alpha) return arg_fail(6, "alpha is NULL");
  if (!beta) return arg_fail(11, "beta is NULL");
  const int ra = is_notrans(transa) ? m : k, rb = is_notrans(transb) ? k : n;
  if (lda < (ra > 1 ? ra : 1)) return arg_fail(8, "lda too small");
  if (ldb < (rb > 1 ? rb : 1)) return arg_fail(10, "ldb too small");
  if (ldc < (m > 1 ? m : 1)) return arg_fail(13, "ldc < max(1,m)");
  int rc = ensure_init();
  if (rc) return rc;
  return zgemm_call<R>(transa, transb, m, n, k, alpha[0], alpha[1], (const R*)A, lda,
                       (const R*)B, ldb, beta[0], beta[1], (R*)C, ldc, 0, 0, nullptr,
                       stream_of(s));
}

}  // namespace

extern "C" {
int dm_zpotrf(char uplo, int n, void* A, int lda, int* info_dev, dm_stream_t s) {
  return zpotrf_api<double>(uplo, n, A, lda, info_dev, s);
}
int dm_cpotrf(char uplo, int n, void* A, int lda, int* info_dev, dm_stream_t s) {
  return zpotrf_api<float>(uplo, n, A, lda, info_dev, s);
}
int dm_ztrsm(char side, char uplo, char transa, char diag, int m, int n, const double* alpha,
             const void* A, int lda, void* B, int ldb, dm_stream_t s) {
  return ztrsm_api<double>(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, s);
}
int dm_ctrsm(char side, char uplo, char transa, char diag, int m, int n, const float* alpha,
             const void* A, int lda, void* B, int ldb, dm_stream_t s) {
  return ztrsm_api<float>(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, s);
}
int dm_zherk(char uplo, char trans, int n, int k, double alpha, const void* A, int lda,
             double beta, void* C, int ldc, dm_stream_t s) {
  return zherk_api<double>(uplo, trans, n, k, alpha, A, lda, beta, C, ldc, s);
}
int dm_cherk(char uplo, char trans, int n, int k, float alpha, const void* A, int lda,
             float beta, void* C, int ldc, dm_stream_t s) {
  return zherk_api<float>(uplo, trans, n, k, alpha, A, lda, beta, C, ldc, s);
}
int dm_zgemm(char transa, char transb, int m, int n, int k, const double* alpha, const void* A,
             int lda, const void* B, int ldb, const double* beta, void* C, int ldc,
             dm_stream_t s) {
  return zgemm_api<double>(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}
int dm_cgemm(char transa, char transb, int m, int n, int k, const float* alpha, const void* A,
             int lda, const void* B, int ldb, const float* beta, void* C, int ldc,
             dm_stream_t s) {
  return zgemm_api<float>(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}
}  // extern "C"

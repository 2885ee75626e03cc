// Mixed precision: low-precision tiled Cholesky + FP64 iterative refinement.
//
// The "precision" form beside the plain FP64 drop-in: the factorization runs
// the same tile DAG (src/array/cholesky.jl:38-53) on an FP32 copy of A
// (TRSM/SYRK/GEMM on FP32 MFMA), while the residual r = b - A x, its norms and
// the solution update stay in FP64 on the original A -- the structure of
// LAPACK's dsposv, on resident tiles:
//   x0 = A32^-1 b;  repeat  r = b - A x (FP64);  stop if
//   max_j ||r_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) <= tol;
//   d = A32^-1 r (FP32 solve);  x += d (FP64).
// Kernels here: precision conversion of tile slabs, column infinity norms,
// the infinity norm of A, and x += d.  Deterministic: one owner per output,
// fixed reduction order.
#include <cmath>
#include <memory>
#include <vector>
#include "dag.h"

namespace dm {

__global__ void convert_d2s_kernel(float* __restrict__ dst, const double* __restrict__ src,
                                   size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = (float)src[i];
}

__global__ void convert_s2d_kernel(double* __restrict__ dst, const float* __restrict__ src,
                                   size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = (double)src[i];
}

// dst (n x n, leading dimension ldd) = src (leading dimension lds), one
// column per blockIdx.y, converted to D.
template <typename D, typename S>
__global__ __launch_bounds__(256) void convert_tile_kernel(D* __restrict__ dst, int ldd,
                                                           const S* __restrict__ src, int lds,
                                                           int n) {
  const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (i < n) dst[(size_t)j * ldd + i] = (D)src[(size_t)j * lds + i];
}

// The precision form's FP64 diagonal POTRF of an FP32 tile: the tile goes
// to the FP64 workspace W (n x n), is factored there in FP64 (potrf_full:
// same info / abort semantics), and the factor is rounded back to FP32.  The
// triangle POTRF leaves alone makes an exact round trip.
int potrf_f64_of_f32(bool upper, int n, float* A, int lda, double* W, int* info, int* abort,
                     int global_base, hipStream_t s) {
  if (n <= 0) return 0;
  const dim3 grid((unsigned)((n + 255) / 256), (unsigned)n);
  hipLaunchKernelGGL((convert_tile_kernel<double, float>), grid, dim3(256), 0, s, W, n,
                     (const float*)A, lda, n);
  DM_LAUNCH_CHECK("convert_tile_kernel");
  int rc = potrf_full<double>(upper, n, W, n, info, abort, global_base, nullptr, s);
  if (rc) return rc;
  hipLaunchKernelGGL((convert_tile_kernel<float, double>), grid, dim3(256), 0, s, A, lda,
                     (const double*)W, n, n);
  DM_LAUNCH_CHECK("convert_tile_kernel");
  return 0;
}

// A[i,i] *= factor for i < n on a square tiled matrix with square blocks
// (tile (t,t) of a Blocks(nb, nb) m x m slab starts at t*nb*m + t*nb*cols(t)
// and has ld = rows(t) = cols(t)).
template <typename T>
__global__ void scale_diag_kernel(T* __restrict__ A, int64_t m, int nb, double factor) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t t = i / nb, l = i - t * nb;
  const int64_t ext = (m - t * nb) < nb ? (m - t * nb) : nb;  // rows = cols of tile t
  T* d = A + t * nb * m + t * nb * ext + l * ext + l;
  *d = (T)((double)*d * factor);
}

// x += (double)d over a packed slab.
__global__ void axpy_s2d_kernel(double* __restrict__ x, const float* __restrict__ d,
                                size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) x[i] += (double)d[i];
}

// Partial row sums of |A| for a tile-packed m x n matrix: workgroup
// (blockIdx.x, blockIdx.y) sums the 256 rows of row block blockIdx.x over the
// columns of column tile blockIdx.y (consecutive threads read consecutive
// rows of a tile: coalesced).  part[tj * m + i] = sum_{j in tile tj} |A_ij|.
__global__ __launch_bounds__(256) void row_abs_part_kernel(const double* __restrict__ A,
                                                           int64_t m, int64_t n, int mb, int nb,
                                                           double* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tj = blockIdx.y;
  if (i >= m) return;
  const int64_t ti = i / mb, ii = i - ti * mb;
  const int64_t rows = (ti * mb + mb <= m) ? mb : m - ti * mb;
  const int64_t cols = (tj * nb + nb <= n) ? nb : n - tj * nb;
  const double* t = A + tj * nb * m + ti * mb * cols;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int64_t jj = 0;
  for (; jj + 4 <= cols; jj += 4) {
    s0 += fabs(t[(jj + 0) * rows + ii]);
    s1 += fabs(t[(jj + 1) * rows + ii]);
    s2 += fabs(t[(jj + 2) * rows + ii]);
    s3 += fabs(t[(jj + 3) * rows + ii]);
  }
  for (; jj < cols; ++jj) s0 += fabs(t[jj * rows + ii]);
  part[tj * m + i] = (s0 + s1) + (s2 + s3);
}

// rowsum[i] = sum over column tiles of part (fixed order).
__global__ void row_abs_sum_kernel(const double* __restrict__ part, int64_t m, int64_t nt,
                                   double* __restrict__ rowsum) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int64_t tj = 0; tj < nt; ++tj) s += part[tj * m + i];
  rowsum[i] = s;
}

// Max over a vector (single workgroup, fixed order): out = max_i v[i].
__global__ __launch_bounds__(256) void max_reduce_kernel(const double* __restrict__ v,
                                                         int64_t count,
                                                         double* __restrict__ out) {
  __shared__ double part[256];
  double mx = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += 256) mx = fmax(mx, v[i]);
  part[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0];
}

// Column infinity norms of a tile-packed m x n matrix: one workgroup per
// column, max |X_ij| over i.
__global__ __launch_bounds__(256) void col_abs_max_kernel(const double* __restrict__ X,
                                                          int64_t m, int64_t n, int mb, int nb,
                                                          double* __restrict__ out) {
  __shared__ double part[256];
  const int64_t j = blockIdx.x;
  const int64_t tj = j / nb, jj = j - tj * nb;
  const int64_t cols = (tj * nb + nb <= n) ? nb : n - tj * nb;
  double mx = 0.0;
  for (int64_t i = threadIdx.x; i < m; i += 256) {
    const int64_t ti = i / mb, ii = i - ti * mb;
    const int64_t rows = (ti * mb + mb <= m) ? mb : m - ti * mb;
    mx = fmax(mx, fabs(X[tj * nb * m + ti * mb * cols + jj * rows + ii]));
  }
  part[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = part[0];
}

// Column dot products of two tile-packed m x n matrices of the same shape
// and blocks: one workgroup per column, out[j] = sum_i X_ij Y_ij in a fixed
// order (strided partial sums, then a tree), so the result is deterministic.
__global__ __launch_bounds__(256) void col_dot_kernel(const double* __restrict__ X,
                                                      const double* __restrict__ Y, int64_t m,
                                                      int64_t n, int mb, int nb,
                                                      double* __restrict__ out) {
  __shared__ double part[256];
  const int64_t j = blockIdx.x;
  const int64_t tj = j / nb, jj = j - tj * nb;
  const int64_t cols = (tj * nb + nb <= n) ? nb : n - tj * nb;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < m; i += 256) {
    const int64_t ti = i / mb, ii = i - ti * mb;
    const int64_t rows = (ti * mb + mb <= m) ? mb : m - ti * mb;
    const int64_t e = tj * nb * m + ti * mb * cols + jj * rows + ii;
    acc = fma(X[e], Y[e], acc);
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = part[0];
}

// Y[:, j] = (scale * a[j]) X[:, j] + beta Y[:, j] on tile-packed matrices of
// the same shape and blocks (a == nullptr: a[j] = 1; beta == 0: Y is not
// read).  Grid: (row blocks of 256, columns); X may be Y.
__global__ __launch_bounds__(256) void col_axpby_kernel(const double* __restrict__ a,
                                                        double scale, const double* X,
                                                        double beta, double* Y, int64_t m,
                                                        int64_t n, int mb, int nb) {
  const int64_t j = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const int64_t tj = j / nb, jj = j - tj * nb;
  const int64_t cols = (tj * nb + nb <= n) ? nb : n - tj * nb;
  const int64_t ti = i / mb, ii = i - ti * mb;
  const int64_t rows = (ti * mb + mb <= m) ? mb : m - ti * mb;
  const int64_t e = tj * nb * m + ti * mb * cols + jj * rows + ii;
  const double c = scale * (a ? a[j] : 1.0);
  Y[e] = beta == 0.0 ? c * X[e] : c * X[e] + beta * Y[e];
}

static unsigned nblocks(size_t count) { return (unsigned)((count + 255) / 256); }

int convert_slab(void* dst, int dst_dtype, const void* src, int src_dtype, size_t count,
                 hipStream_t s) {
  if (!count) return 0;
  if (dst_dtype == DM_F32 && src_dtype == DM_F64)
    hipLaunchKernelGGL(convert_d2s_kernel, dim3(nblocks(count)), dim3(256), 0, s, (float*)dst,
                       (const double*)src, count);
  else if (dst_dtype == DM_F64 && src_dtype == DM_F32)
    hipLaunchKernelGGL(convert_s2d_kernel, dim3(nblocks(count)), dim3(256), 0, s, (double*)dst,
                       (const float*)src, count);
  else
    return arg_fail(0, "convert: unsupported dtype pair");
  DM_LAUNCH_CHECK("convert_kernel");
  return 0;
}

// ------------------------------------------------------------ tiled GEMM DAG
// C = alpha op(A) op(B) + beta C on tile grids (A: m x k tiles, B: k x n
// tiles, all with the same block size), one GEMM task per (i, j, l) with beta
// applied at l = 0 -- the tile loop of Dagger's matmul (src/array/mul.jl).
void build_gemm_tiled(Dag& d, char ta, char tb, double alpha, const dm_tiled& A,
                      const dm_tiled& B, double beta, const dm_tiled& Cm);


// ---------------------------------------------------------------- residual
// C = beta C + alpha A X for a tiled FP64 A (m x k, square blocks mb with
// mb % 128 == 0) and X, C of one block column (n <= 64): the FP64 residual
// r = b - A x of iterative refinement in one launch.  The tile DAG form
// accumulates C_i += A_ik X_k one k after another (a chain of mt launches);
// here each workgroup owns 64 rows of C over the whole k range, so A is
// streamed from HBM once (about 1 ms for N = 32768) at 2n flops per element.
//  * 4 waves, each 16 rows x n columns on v_mfma_f64_16x16x4 (A operand:
//    lane l holds A[row l%16][k l/16]; X operand: X[k l/16][col l%16];
//    result lane l, item r = D[4r + l/16][l%16]);
//  * X is staged through LDS in chunks of 128 k (double-buffered, row
//    stride n_pad + 1), A is loaded straight into registers one chunk ahead.
// Each output has one owner and a fixed summation order (deterministic).
constexpr int SK_KC = 128;

template <int NB16>
struct Skinny {
  static constexpr int NP = 16 * NB16, SX = NP + 1;
  static constexpr int XPT = SK_KC * NP / 256;  // X elements per thread per chunk
};

template <int NB16>
__global__ __launch_bounds__(256) void tiled_skinny_gemm_kernel(const double* __restrict__ A,
                                                                const double* __restrict__ X,
                                                                double* __restrict__ Cm,
                                                                int64_t m, int64_t k, int n,
                                                                int mb, double alpha,
                                                                double beta) {
  typedef Skinny<NB16> S;
  typedef double v4 __attribute__((ext_vector_type(4)));
  __shared__ double xs[2][SK_KC * S::SX];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int64_t ti = r0 / mb;
  const int lrow0 = (int)(r0 - ti * mb);
  const int rows_ti = (int)(m - ti * mb < mb ? m - ti * mb : mb);
  const int arow = lrow0 + 16 * w + lr;  // this lane's A row (local to tile row ti)
  const bool rowok = arow < rows_ti;
  const int64_t nch = (k + SK_KC - 1) / SK_KC;

  // chunk c: global k [c*KC, c*KC + KC) inside tile column tk = c*KC / mb
  auto a_load = [&](int64_t c, double (&a)[SK_KC / 4]) {
    const int64_t kg0 = c * SK_KC, tk = kg0 / mb;
    const int kl0 = (int)(kg0 - tk * mb);
    const int cols = (int)(k - tk * mb < mb ? k - tk * mb : mb);
    const double* Ap = A + tk * (int64_t)mb * m + ti * (int64_t)mb * cols;
#pragma unroll
    for (int q = 0; q < SK_KC / 4; ++q) {
      const int kl = kl0 + 4 * q + lk;
      a[q] = (rowok && kl < cols) ? Ap[(int64_t)kl * rows_ti + arow] : 0.0;
    }
  };
  auto x_load = [&](int64_t c, double (&xr)[S::XPT]) {
    const int64_t kg0 = c * SK_KC, tk = kg0 / mb;
    const int kl0 = (int)(kg0 - tk * mb);
    const int ldx = (int)(k - tk * mb < mb ? k - tk * mb : mb);
    const double* Xp = X + tk * (int64_t)mb * n;
#pragma unroll
    for (int q = 0; q < S::XPT; ++q) {
      const int e = tid + 256 * q;
      const int kl = e % SK_KC, j = e / SK_KC;
      xr[q] = (kl0 + kl < ldx && j < n) ? Xp[(int64_t)j * ldx + kl0 + kl] : 0.0;
    }
  };
  auto x_store = [&](double* dst, const double (&xr)[S::XPT]) {
#pragma unroll
    for (int q = 0; q < S::XPT; ++q) {
      const int e = tid + 256 * q;
      dst[(e % SK_KC) * S::SX + e / SK_KC] = xr[q];
    }
  };

  v4 acc[NB16];
#pragma unroll
  for (int b = 0; b < NB16; ++b) acc[b] = v4{0, 0, 0, 0};
  double an[SK_KC / 4], ac[SK_KC / 4], xr[S::XPT];
  a_load(0, an);
  x_load(0, xr);
  x_store(xs[0], xr);
  __syncthreads();
  for (int64_t c = 0; c < nch; ++c) {
#pragma unroll
    for (int q = 0; q < SK_KC / 4; ++q) ac[q] = an[q];
    const bool more = c + 1 < nch;
    if (more) {
      a_load(c + 1, an);
      x_load(c + 1, xr);
    }
    const double* xb = xs[c & 1];
#pragma unroll
    for (int q = 0; q < SK_KC / 4; ++q) {
#pragma unroll
      for (int b = 0; b < NB16; ++b) {
        const double xf = xb[(4 * q + lk) * S::SX + 16 * b + lr];
        acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[q], xf, acc[b], 0, 0, 0);
      }
    }
    if (more) x_store(xs[(c + 1) & 1], xr);
    __syncthreads();
  }
  double* Cp = Cm + ti * (int64_t)mb * n;  // C tile (ti, 0), ld = rows_ti
#pragma unroll
  for (int b = 0; b < NB16; ++b) {
    const int col = 16 * b + lr;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = lrow0 + 16 * w + 4 * r + lk;
      if (row < rows_ti && col < n) {
        double* p = Cp + (int64_t)col * rows_ti + row;
        const double v = alpha * acc[b][r];
        *p = beta == 0.0 ? v : v + beta * *p;
      }
    }
  }
}

bool tiled_skinny_gemm_ok(char ta, char tb, const dm_tiled& A, const dm_tiled& X,
                          const dm_tiled& Cm) {
  return ta == 'N' && tb == 'N' && A.dtype == DM_F64 && X.dtype == DM_F64 && Cm.dtype == DM_F64 &&
         A.mb == A.nb && A.mb % SK_KC == 0 && X.mb == A.mb && X.nb == A.mb && Cm.mb == A.mb &&
         X.n == Cm.n && Cm.n >= 1 && Cm.n <= 64 && X.n <= X.nb && Cm.n <= Cm.nb &&
         A.m == Cm.m && A.n == X.m;
}

int tiled_skinny_gemm_launch(const dm_tiled& A, const dm_tiled& X, const dm_tiled& Cm,
                             double alpha, double beta, hipStream_t s) {
  const int n = (int)Cm.n;
  const unsigned grid = (unsigned)((A.m + 63) / 64);
  if (grid == 0) return 0;
  const double* a = (const double*)A.base;
  const double* x = (const double*)X.base;
  double* c = (double*)Cm.base;
  switch ((n + 15) / 16) {
    case 1:
      hipLaunchKernelGGL(tiled_skinny_gemm_kernel<1>, dim3(grid), dim3(256), 0, s, a, x, c, A.m,
                         A.n, n, A.mb, alpha, beta);
      break;
    case 2:
      hipLaunchKernelGGL(tiled_skinny_gemm_kernel<2>, dim3(grid), dim3(256), 0, s, a, x, c, A.m,
                         A.n, n, A.mb, alpha, beta);
      break;
    case 3:
      hipLaunchKernelGGL(tiled_skinny_gemm_kernel<3>, dim3(grid), dim3(256), 0, s, a, x, c, A.m,
                         A.n, n, A.mb, alpha, beta);
      break;
    default:
      hipLaunchKernelGGL(tiled_skinny_gemm_kernel<4>, dim3(grid), dim3(256), 0, s, a, x, c, A.m,
                         A.n, n, A.mb, alpha, beta);
      break;
  }
  DM_LAUNCH_CHECK("tiled_skinny_gemm_kernel");
  return 0;
}

}  // namespace dm

using namespace dm;

extern "C" {

int dm_tiled_convert(const dm_tiled* dst, const dm_tiled* src, dm_stream_t s) {
  if (!dst || !src) return arg_fail(1, "NULL tiled matrix");
  if (dst->m != src->m || dst->n != src->n || dst->mb != src->mb || dst->nb != src->nb)
    return arg_fail(2, "shapes/blocks differ");
  int rc = ensure_init();
  if (rc) return rc;
  return convert_slab(dst->base, dst->dtype, src->base, src->dtype, (size_t)(dst->m * dst->n),
                      stream_of(s));
}

int dm_tiled_norm_inf(const dm_tiled* A, double* out_dev, double* work_dev, dm_stream_t s) {
  if (!A || A->dtype != DM_F64) return arg_fail(1, "A must be an FP64 tiled matrix");
  if (!out_dev || !work_dev) return arg_fail(2, "NULL output/work");
  int rc = ensure_init();
  if (rc) return rc;
  hipStream_t st = stream_of(s);
  if (A->m <= 0 || A->n <= 0) return fill_i32_launch((int*)out_dev, 0, 2, st);
  const int64_t nt = (A->n + A->nb - 1) / A->nb;
  double* part = work_dev + A->m;  // work: m row sums, then nt * m partial sums
  hipLaunchKernelGGL(row_abs_part_kernel, dim3(nblocks((size_t)A->m), (unsigned)nt), dim3(256), 0,
                     st, (const double*)A->base, A->m, A->n, A->mb, A->nb, part);
  DM_LAUNCH_CHECK("row_abs_part_kernel");
  hipLaunchKernelGGL(row_abs_sum_kernel, dim3(nblocks((size_t)A->m)), dim3(256), 0, st,
                     (const double*)part, A->m, nt, work_dev);
  DM_LAUNCH_CHECK("row_abs_sum_kernel");
  hipLaunchKernelGGL(max_reduce_kernel, dim3(1), dim3(256), 0, st, (const double*)work_dev, A->m,
                     out_dev);
  DM_LAUNCH_CHECK("max_reduce_kernel");
  return 0;
}

int dm_tiled_col_norm_inf(const dm_tiled* X, double* out_dev, dm_stream_t s) {
  if (!X || X->dtype != DM_F64) return arg_fail(1, "X must be an FP64 tiled matrix");
  if (!out_dev) return arg_fail(2, "NULL output");
  int rc = ensure_init();
  if (rc) return rc;
  if (X->n <= 0) return 0;
  hipLaunchKernelGGL(col_abs_max_kernel, dim3((unsigned)X->n), dim3(256), 0, stream_of(s),
                     (const double*)X->base, X->m, X->n, X->mb, X->nb, out_dev);
  DM_LAUNCH_CHECK("col_abs_max_kernel");
  return 0;
}

static bool same_layout(const dm_tiled* X, const dm_tiled* Y) {
  return X->m == Y->m && X->n == Y->n && X->mb == Y->mb && X->nb == Y->nb;
}

int dm_tiled_col_dot(const dm_tiled* X, const dm_tiled* Y, double* out_dev, dm_stream_t s) {
  if (!X || !Y || X->dtype != DM_F64 || Y->dtype != DM_F64)
    return arg_fail(1, "X and Y must be FP64 tiled matrices");
  if (!same_layout(X, Y)) return arg_fail(2, "X and Y differ in shape or blocks");
  if (!out_dev) return arg_fail(3, "NULL output");
  int rc = ensure_init();
  if (rc) return rc;
  if (X->n <= 0) return 0;
  if (X->m <= 0) return fill_i32_launch((int*)out_dev, 0, (int)(2 * X->n), stream_of(s));
  hipLaunchKernelGGL(col_dot_kernel, dim3((unsigned)X->n), dim3(256), 0, stream_of(s),
                     (const double*)X->base, (const double*)Y->base, X->m, X->n, X->mb, X->nb,
                     out_dev);
  DM_LAUNCH_CHECK("col_dot_kernel");
  return 0;
}

int dm_tiled_col_axpby(const double* a_dev, double scale, const dm_tiled* X, double beta,
                       const dm_tiled* Y, dm_stream_t s) {
  if (!X || !Y || X->dtype != DM_F64 || Y->dtype != DM_F64)
    return arg_fail(3, "X and Y must be FP64 tiled matrices");
  if (!same_layout(X, Y)) return arg_fail(5, "X and Y differ in shape or blocks");
  if (X->n > 65535) return arg_fail(3, "at most 65535 columns");
  int rc = ensure_init();
  if (rc) return rc;
  if (X->m <= 0 || X->n <= 0) return 0;
  hipLaunchKernelGGL(col_axpby_kernel, dim3(nblocks((size_t)X->m), (unsigned)X->n), dim3(256), 0,
                     stream_of(s), a_dev, scale, (const double*)X->base, beta,
                     (double*)Y->base, X->m, X->n, X->mb, X->nb);
  DM_LAUNCH_CHECK("col_axpby_kernel");
  return 0;
}

int dm_tiled_scale_diag(const dm_tiled* A, double factor, dm_stream_t s) {
  if (!A || !A->base) return arg_fail(1, "NULL tiled matrix");
  if (A->m != A->n || A->mb != A->nb) return arg_fail(1, "A must be square with square blocks");
  if (A->dtype != DM_F64 && A->dtype != DM_F32) return arg_fail(1, "unknown dtype");
  if (!(factor > 0.0)) return arg_fail(2, "factor must be positive");
  int rc = ensure_init();
  if (rc) return rc;
  if (A->m == 0) return 0;
  const dim3 grid(nblocks((size_t)A->m));
  if (A->dtype == DM_F64)
    hipLaunchKernelGGL(scale_diag_kernel<double>, grid, dim3(256), 0, stream_of(s),
                       (double*)A->base, A->m, A->nb, factor);
  else
    hipLaunchKernelGGL(scale_diag_kernel<float>, grid, dim3(256), 0, stream_of(s),
                       (float*)A->base, A->m, A->nb, factor);
  DM_LAUNCH_CHECK("scale_diag_kernel");
  return 0;
}

int dm_tiled_axpy_s2d(const dm_tiled* X, const dm_tiled* D, dm_stream_t s) {
  if (!X || !D || X->dtype != DM_F64 || D->dtype != DM_F32)
    return arg_fail(1, "X must be FP64 and D FP32");
  if (X->m != D->m || X->n != D->n || X->mb != D->mb || X->nb != D->nb)
    return arg_fail(2, "shapes/blocks differ");
  int rc = ensure_init();
  if (rc) return rc;
  const size_t count = (size_t)(X->m * X->n);
  if (!count) return 0;
  hipLaunchKernelGGL(axpy_s2d_kernel, dim3(nblocks(count)), dim3(256), 0, stream_of(s),
                     (double*)X->base, (const float*)D->base, count);
  DM_LAUNCH_CHECK("axpy_s2d_kernel");
  return 0;
}

}  // extern "C"

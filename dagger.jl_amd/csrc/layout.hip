// Layout kernels for resident DArrays: a full column-major matrix <-> the
// tile-packed layout of dm_tiled.  Device side of copyto!/darray_copyto!
// (src/array/copy.jl:39-151) and of the re-blocking copy_buffered
// (src/array/copy.jl:11-34).  One thread per element; consecutive threads
// walk down a column so both sides of the copy are coalesced within a tile.
#include "dm_internal.h"

namespace dm {

template <typename T, bool PACK>
__global__ void tile_copy_kernel(T* __restrict__ tiled, T* __restrict__ full, int64_t ld,
                                 int64_t m, int64_t n, int mb, int nb, T diag_add) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * n) return;
  const int64_t i = idx % m, j = idx / m;
  const int64_t ti = i / mb, tj = j / nb;
  const int64_t ii = i - ti * mb, jj = j - tj * nb;
  const int64_t cols = (tj * nb + nb <= n) ? nb : n - tj * nb;
  const int64_t rows = (ti * mb + mb <= m) ? mb : m - ti * mb;
  const int64_t off = tj * nb * m + ti * mb * cols + jj * rows + ii;
  if (PACK) tiled[off] = full[j * ld + i] + (i == j ? diag_add : T(0));
  else full[j * ld + i] = tiled[off];
}

template <typename T>
static int tile_copy(bool pack, const dm_tiled* t, void* full, int64_t ld, double diag_add,
                     hipStream_t s) {
  const int64_t total = t->m * t->n;
  if (total <= 0) return 0;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  if (pack)
    hipLaunchKernelGGL((tile_copy_kernel<T, true>), dim3(blocks), dim3(256), 0, s, (T*)t->base,
                       (T*)full, ld, t->m, t->n, t->mb, t->nb, (T)diag_add);
  else
    hipLaunchKernelGGL((tile_copy_kernel<T, false>), dim3(blocks), dim3(256), 0, s, (T*)t->base,
                       (T*)full, ld, t->m, t->n, t->mb, t->nb, (T)0);
  DM_LAUNCH_CHECK("tile_copy_kernel");
  return 0;
}

}  // namespace dm

using namespace dm;

// The real view of a complex tiled matrix: each (re, im) pair is two rows, so
// an m x n complex Blocks(mb, nb) matrix is, bit for bit, the 2m x n real
// Blocks(2mb, nb) one (each tile's ld doubles with its rows).
static dm_tiled real_view(const dm_tiled& t) {
  dm_tiled r = t;
  if (is_cplx(t.dtype)) {
    r.m *= 2;
    r.mb *= 2;
    r.dtype = t.dtype == DM_C64 ? DM_F64 : DM_F32;
  }
  return r;
}

extern "C" int dm_tiled_pack(const dm_tiled* dst, const void* src, int64_t ld, double diag_add,
                             dm_stream_t s) {
  if (!dst) return arg_fail(1, "dst is NULL");
  if (dst->mb <= 0 || dst->nb <= 0) return arg_fail(1, "block size must be positive");
  if (ld < (dst->m > 1 ? dst->m : 1)) return arg_fail(3, "ld < max(1,m)");
  if (is_cplx(dst->dtype)) {  // copies of the real view (no complex diagonal shift)
    if (diag_add != 0.0) return arg_fail(4, "diag_add is for real element types");
    const dm_tiled r = real_view(*dst);
    return dm_tiled_pack(&r, src, 2 * ld, 0.0, s);
  }
  int rc = ensure_init();
  if (rc) return rc;
  if (dst->dtype == DM_F64) return tile_copy<double>(true, dst, (void*)src, ld, diag_add, stream_of(s));
  if (dst->dtype == DM_F32) return tile_copy<float>(true, dst, (void*)src, ld, diag_add, stream_of(s));
  return arg_fail(1, "unknown dtype");
}

extern "C" int dm_tiled_unpack(void* dst, int64_t ld, const dm_tiled* src, dm_stream_t s) {
  if (!src) return arg_fail(3, "src is NULL");
  if (src->mb <= 0 || src->nb <= 0) return arg_fail(3, "block size must be positive");
  if (ld < (src->m > 1 ? src->m : 1)) return arg_fail(2, "ld < max(1,m)");
  if (is_cplx(src->dtype)) {
    const dm_tiled r = real_view(*src);
    return dm_tiled_unpack(dst, 2 * ld, &r, s);
  }
  int rc = ensure_init();
  if (rc) return rc;
  if (src->dtype == DM_F64) return tile_copy<double>(false, src, dst, ld, 0.0, stream_of(s));
  if (src->dtype == DM_F32) return tile_copy<float>(false, src, dst, ld, 0.0, stream_of(s));
  return arg_fail(3, "unknown dtype");
}

namespace dm {
__global__ void add_diag_kernel(double* __restrict__ A, int64_t ld, int64_t n,
                                const double* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[i * ld + i] += d[i];
}
}  // namespace dm

namespace dm {
// Tile k of a one-tile-column matrix (m x nb, Blocks(nb, nb)) := [I_r 0],
// r = rows of tile k (its ld).
template <typename T>
__global__ void identity_tiles_kernel(T* __restrict__ D, int64_t m, int nb) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * nb) return;
  const int64_t k = e / ((int64_t)nb * nb), w = e - k * nb * nb;
  const int64_t r = (k * nb + nb <= m) ? nb : m - k * nb;  // rows (= ld) of tile k
  const int64_t j = w / r, i = w - j * r;
  if (j < nb) D[e] = (i == j) ? T(1) : T(0);
}
}  // namespace dm

extern "C" int dm_tiled_identity(const dm_tiled* D, dm_stream_t s) {
  if (!D) return arg_fail(1, "D is NULL");
  if (D->n != D->nb || D->mb != D->nb) return arg_fail(1, "D must be m x nb with Blocks(nb, nb)");
  int rc = ensure_init();
  if (rc) return rc;
  const int64_t count = D->m * D->nb;
  if (count <= 0) return 0;
  const unsigned grid = (unsigned)((count + 255) / 256);
  if (D->dtype == DM_F64)
    hipLaunchKernelGGL((identity_tiles_kernel<double>), dim3(grid), dim3(256), 0, stream_of(s),
                       (double*)D->base, D->m, D->nb);
  else if (D->dtype == DM_F32)
    hipLaunchKernelGGL((identity_tiles_kernel<float>), dim3(grid), dim3(256), 0, stream_of(s),
                       (float*)D->base, D->m, D->nb);
  else
    return arg_fail(1, "unsupported dtype");
  DM_LAUNCH_CHECK("identity_tiles_kernel");
  return 0;
}

// A[i,i] += d[i] on a full column-major FP64 matrix (generator helper).
extern "C" int dm_add_diag_d(double* A, int64_t ld, int64_t n, const double* d_dev,
                             dm_stream_t s) {
  if (!A && n > 0) return arg_fail(1, "A is NULL");
  if (ld < (n > 1 ? n : 1)) return arg_fail(2, "ld < max(1,n)");
  int rc = ensure_init();
  if (rc) return rc;
  if (n <= 0) return 0;
  hipLaunchKernelGGL(add_diag_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     stream_of(s), A, ld, n, d_dev);
  DM_LAUNCH_CHECK("add_diag_kernel");
  return 0;
}

// ---------------------------------------------------------------------------
// ishermitian(::DArray) (src/array/linalg.jl:90-124): diagonal tiles must be
// exactly Hermitian (Julia's ishermitian on a dense block); an off-diagonal
// pair (i > j) passes when ||A_ij - A_ji^H||_F <= sqrt(eps) *
// max(||A_ij||_F, ||A_ji||_F) (is_cross_hermitian, src/array/linalg.jl:38-48).
// One workgroup per tile pair; fixed-order reductions.
namespace dm {
// One workgroup (8 waves) per tile pair (i >= j), walking the pair in 64 x 64
// sub-blocks: the sub-block of A_ij is read along its columns, the matching
// sub-block of A_ji along its columns too and transposed through LDS, so
// both reads are contiguous (A is streamed from HBM once).  CPLX: elements
// are (re, im) pairs of T and the partner is conjugated (A_ji^H), so a
// diagonal tile must also have an exactly zero imaginary diagonal.
// The sub-blocks are software-pipelined: both tiles' loads of the next
// sub-block are issued before the current one is compared, so two
// sub-blocks' reads are in flight per workgroup.
constexpr int HERM_NT = 1024;
template <typename T, bool CPLX>
__global__ __launch_bounds__(HERM_NT) void herm_pair_kernel(const T* __restrict__ A, int64_t m,
                                                           int mb, int mt,
                                                           int* __restrict__ ok_pairs) {
  constexpr int S = CPLX ? 2 : 1;  // T values per element
  constexpr int NW = HERM_NT / 64;  // waves: rows hi, hi + NW, ... of a sub-block
  __shared__ T ts[S][64 * 65];
  __shared__ double s0[HERM_NT], s1[HERM_NT], s2[HERM_NT];
  // decode pair index -> (i, j) with i >= j, column by column of the lower triangle
  int p = blockIdx.x, j = 0;
  while (p >= mt - j) { p -= mt - j; ++j; }
  const int i = j + p;
  const int ri = (int)((i * (int64_t)mb + mb <= m) ? mb : m - i * (int64_t)mb);
  const int rj = (int)((j * (int64_t)mb + mb <= m) ? mb : m - j * (int64_t)mb);
  // tile (r, c) origin for a square m x m Blocks(mb, mb) matrix
  auto tile = [&](int r, int c, int64_t cols) {
    return A + S * (c * (int64_t)mb * m + r * (int64_t)mb * cols);
  };
  const T* Aij = tile(i, j, rj);  // ri x rj, ld = ri
  const T* Aji = tile(j, i, ri);  // rj x ri, ld = rj
  const int tid = threadIdx.x, lo = tid & 63, hi = tid >> 6;
  double d2 = 0, n1 = 0, n2 = 0;
  constexpr int Q = 64 / NW;
  const int ncb = (rj + 63) / 64, nsub = ((ri + 63) / 64) * ncb;
  // registers of one sub-block: A_ji elements (transposed into LDS) and the
  // matching A_ij elements (compared against LDS)
  T xji[Q][S], xij[Q][S];
  auto load = [&](int sb) {
    const int r0 = (sb / ncb) * 64, c0 = (sb % ncb) * 64;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      {  // A_ji(c0 + lo, r0 + rr): column r0 + rr of A_ji, rows along lanes
        const int rr = hi + NW * q, cc = lo;
        const bool in = r0 + rr < ri && c0 + cc < rj;
#pragma unroll
        for (int h = 0; h < S; ++h)
          xji[q][h] = in ? Aji[S * ((int64_t)(r0 + rr) * rj + c0 + cc) + h] : T(0);
      }
      {  // A_ij(r0 + lo, c0 + cc): column c0 + cc of A_ij, rows along lanes
        const int rr = lo, cc = hi + NW * q;
        const bool in = r0 + rr < ri && c0 + cc < rj;
#pragma unroll
        for (int h = 0; h < S; ++h)
          xij[q][h] = in ? Aij[S * ((int64_t)(c0 + cc) * ri + r0 + rr) + h] : T(0);
      }
    }
  };
  if (nsub > 0) load(0);
  for (int sb = 0; sb < nsub; ++sb) {
    const int r0 = (sb / ncb) * 64, c0 = (sb % ncb) * 64;
    // ts[rr * 65 + cc] = A_ji(c0 + cc, r0 + rr) = (A_ji^T)(r0 + rr, c0 + cc)
    __syncthreads();  // the previous sub-block's LDS reads are done
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int h = 0; h < S; ++h) ts[h][(hi + NW * q) * 65 + lo] = xji[q][h];
    T cur[Q][S];
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int h = 0; h < S; ++h) cur[q][h] = xij[q][h];
    if (sb + 1 < nsub) load(sb + 1);  // in flight during the comparison below
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int rr = lo, cc = hi + NW * q;
      if (r0 + rr < ri && c0 + cc < rj) {
        const double a = (double)cur[q][0];
        const double b = (double)ts[0][rr * 65 + cc];
        double ai = 0, bi = 0;
        if (CPLX) {
          ai = (double)cur[q][S - 1];
          bi = -(double)ts[S - 1][rr * 65 + cc];  // conj
        }
        if (i == j) {
          d2 += (a != b || ai != bi) ? 1.0 : 0.0;
        } else {
          d2 += (a - b) * (a - b) + (ai - bi) * (ai - bi);
          n1 += a * a + ai * ai;
          n2 += b * b + bi * bi;
        }
      }
    }
  }
  s0[tid] = d2; s1[tid] = n1; s2[tid] = n2;
  __syncthreads();
  for (int w = HERM_NT / 2; w > 0; w >>= 1) {
    if (tid < w) {
      s0[tid] += s0[tid + w];
      s1[tid] += s1[tid + w];
      s2[tid] += s2[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    int ok;
    if (i == j) ok = s0[0] == 0.0;
    else {
      const double rtol = sizeof(T) == 8 ? 1.4901161193847656e-08 : 3.4526698300124393e-04;
      ok = sqrt(s0[0]) <= rtol * fmax(sqrt(s1[0]), sqrt(s2[0]));
    }
    ok_pairs[blockIdx.x] = ok;
  }
}

__global__ void all_ok_kernel(const int* __restrict__ ok, int count, int* __restrict__ out) {
  __shared__ int s[256];
  int v = 1;
  for (int i = threadIdx.x; i < count; i += 256) v &= ok[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] &= s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}
}  // namespace dm

// *result_dev = 1 if A (square, Blocks(nb, nb)) is Hermitian by the rule
// above, else 0.  work_dev: mt*(mt+1)/2 ints.
extern "C" int dm_ishermitian_tiled(const dm_tiled* A, int* result_dev, int* work_dev,
                                    dm_stream_t s) {
  if (!A) return arg_fail(1, "A is NULL");
  if (A->m != A->n || A->mb != A->nb) return arg_fail(1, "A must be square with square tiles");
  if (!result_dev || !work_dev) return arg_fail(2, "NULL result/work");
  int rc = ensure_init();
  if (rc) return rc;
  hipStream_t st = stream_of(s);
  const int mt = (int)((A->m + A->mb - 1) / A->mb);
  const int pairs = mt * (mt + 1) / 2;
  if (pairs == 0) return fill_i32_launch(result_dev, 1, 1, st);
  switch (A->dtype) {
    case DM_F64:
      hipLaunchKernelGGL((herm_pair_kernel<double, false>), dim3(pairs), dim3(HERM_NT), 0, st,
                         (const double*)A->base, A->m, A->mb, mt, work_dev);
      break;
    case DM_F32:
      hipLaunchKernelGGL((herm_pair_kernel<float, false>), dim3(pairs), dim3(HERM_NT), 0, st,
                         (const float*)A->base, A->m, A->mb, mt, work_dev);
      break;
    case DM_C64:
      hipLaunchKernelGGL((herm_pair_kernel<double, true>), dim3(pairs), dim3(HERM_NT), 0, st,
                         (const double*)A->base, A->m, A->mb, mt, work_dev);
      break;
    case DM_C32:
      hipLaunchKernelGGL((herm_pair_kernel<float, true>), dim3(pairs), dim3(HERM_NT), 0, st,
                         (const float*)A->base, A->m, A->mb, mt, work_dev);
      break;
    default:
      return arg_fail(1, "unknown dtype");
  }
  DM_LAUNCH_CHECK("herm_pair_kernel");
  hipLaunchKernelGGL(all_ok_kernel, dim3(1), dim3(256), 0, st, work_dev, pairs, result_dev);
  DM_LAUNCH_CHECK("all_ok_kernel");
  return 0;
}

// ------------------------------------------------ norm2 and chkfinite
// One pass over the stored tiles: per-workgroup partial sums of squares
// (accumulated in FP64, fixed order) of the elements in the requested part
// (0 whole, 1 upper triangle with the diagonal, 2 lower triangle with the
// diagonal), and a flag set when any element of the part is not finite.
// The reference sums norm_sqr per chunk and then over chunks
// (src/array/linalg.jl:1-34); chkfinite is all(isfinite) per chunk
// (src/array/linalg.jl:127-148).
namespace dm {

constexpr int kSumsqThreads = 256;
constexpr int kSumsqMaxPerTile = 8;

// sixteen bytes of one tile, read by one vector load
template <typename T>
struct alignas(16) Pack16 {
  T v[16 / sizeof(T)];
};

template <typename T>
__global__ __launch_bounds__(kSumsqThreads) void tiled_sumsq_kernel(
    const T* __restrict__ base, int64_t m, int64_t n, int mb, int nb, int part, int bpt,
    double* __restrict__ partial, int* __restrict__ nonfinite) {
  __shared__ double red[kSumsqThreads];
  const int64_t mt = (m + mb - 1) / mb;
  const int64_t tile = blockIdx.x / bpt;
  const int sub = (int)(blockIdx.x - tile * bpt);
  const int64_t ti = tile % mt, tj = tile / mt;
  const int64_t r0 = ti * mb, c0 = tj * nb;
  const int h = (int)min((int64_t)mb, m - r0), w = (int)min((int64_t)nb, n - c0);
  const T* __restrict__ t = base + tj * (int64_t)nb * m + ti * (int64_t)mb * w;
  // tiles wholly outside the part contribute nothing
  const bool skip = (part == 1 && r0 > c0 + w - 1) || (part == 2 && c0 > r0 + h - 1);
  const bool whole = part == 0 || (part == 1 && r0 + h - 1 <= c0) ||
                     (part == 2 && c0 + w - 1 <= r0);
  double acc = 0.0;
  bool bad = false;
  if (!skip && whole) {
    // 16-byte loads (two FP64 or four FP32 elements) while the tile start is
    // 16-byte aligned, two in flight per thread, then the elements the
    // packs leave over one at a time; two partial sums in a fixed order
    constexpr int kPer = (int)(16 / sizeof(T));
    const int count = h * w, stride = bpt * kSumsqThreads;
    const int first = sub * kSumsqThreads + threadIdx.x;
    double a0 = 0.0, a1 = 0.0;
    int tail = 0;
    if ((reinterpret_cast<uintptr_t>(t) & 15) == 0) {
      const Pack16<T>* __restrict__ tv = reinterpret_cast<const Pack16<T>*>(t);
      const int packs = count / kPer;
      int e = first;
      for (; e + stride < packs; e += 2 * stride) {
        const Pack16<T> p0 = tv[e], p1 = tv[e + stride];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
          const double x0 = (double)p0.v[q], x1 = (double)p1.v[q];
          bad = bad || !isfinite(x0) || !isfinite(x1);
          a0 += x0 * x0;
          a1 += x1 * x1;
        }
      }
      for (; e < packs; e += stride) {
        const Pack16<T> p0 = tv[e];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
          const double x0 = (double)p0.v[q];
          bad = bad || !isfinite(x0);
          a0 += x0 * x0;
        }
      }
      tail = packs * kPer;
    }
    for (int e = tail + first; e < count; e += stride) {
      const double x = (double)t[e];
      bad = bad || !isfinite(x);
      a0 += x * x;
    }
    acc = a0 + a1;
  } else if (!skip) {
    const int count = h * w;
    for (int e = sub * kSumsqThreads + threadIdx.x; e < count; e += bpt * kSumsqThreads) {
      bool in = whole;
      if (!in) {
        const int cl = e / h, rl = e - cl * h;
        in = part == 1 ? (r0 + rl <= c0 + cl) : (r0 + rl >= c0 + cl);
      }
      if (in) {
        const double x = (double)t[e];
        bad = bad || !isfinite(x);
        acc += x * x;
      }
    }
  }
  if (bad) *nonfinite = 1;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kSumsqThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kSumsqThreads) void sum_sqrt_kernel(const double* __restrict__ partial,
                                                                 int64_t count,
                                                                 double* __restrict__ out) {
  __shared__ double red[kSumsqThreads];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += kSumsqThreads) acc += partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kSumsqThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sqrt(red[0]);
}

}  // namespace dm

// *out_dev = sqrt(sum of squares over the part), *nonfinite_dev = 1 if an
// element of the part is Inf or NaN, else 0.  work_dev: 8 * mt * nt doubles.
extern "C" int dm_tiled_norm2(const dm_tiled* A, int part, double* out_dev, int* nonfinite_dev,
                              double* work_dev, dm_stream_t s) {
  if (!A) return arg_fail(1, "A is NULL");
  if (A->m < 0 || A->n < 0 || A->mb < 1 || A->nb < 1) return arg_fail(1, "bad tiled matrix");
  if (part < 0 || part > 2) return arg_fail(2, "part must be 0 (whole), 1 (upper) or 2 (lower)");
  if (is_cplx(A->dtype)) {  // sum of |z|^2 = sum over the real view's entries
    if (part != 0) return arg_fail(2, "complex element types: part must be 0 (whole)");
    const dm_tiled r = real_view(*A);
    return dm_tiled_norm2(&r, 0, out_dev, nonfinite_dev, work_dev, s);
  }
  if (A->dtype != DM_F64 && A->dtype != DM_F32) return arg_fail(1, "unknown dtype");
  if (!out_dev || !nonfinite_dev || !work_dev) return arg_fail(3, "NULL output/flag/work");
  int rc = ensure_init();
  if (rc) return rc;
  hipStream_t st = stream_of(s);
  rc = fill_i32_launch(nonfinite_dev, 0, 1, st);
  if (rc) return rc;
  const int64_t mt = (A->m + A->mb - 1) / A->mb, nt = (A->n + A->nb - 1) / A->nb;
  const int64_t tiles = mt * nt;
  if (tiles == 0) return fill_i32_launch((int*)out_dev, 0, 2, st);
  if ((int64_t)A->mb * A->nb > (int64_t)1 << 30) return arg_fail(1, "tiles above 2^30 elements");
  if (tiles * kSumsqMaxPerTile > ((int64_t)1 << 31) - 1) return arg_fail(1, "too many tiles");
  // about 2048 workgroups in all, at most 8 per tile
  int bpt = (int)((2048 + tiles - 1) / tiles);
  bpt = bpt < 1 ? 1 : (bpt > kSumsqMaxPerTile ? kSumsqMaxPerTile : bpt);
  const dim3 grid((unsigned)(tiles * bpt));
  if (A->dtype == DM_F64)
    hipLaunchKernelGGL((tiled_sumsq_kernel<double>), grid, dim3(kSumsqThreads), 0, st,
                       (const double*)A->base, A->m, A->n, A->mb, A->nb, part, bpt, work_dev,
                       nonfinite_dev);
  else
    hipLaunchKernelGGL((tiled_sumsq_kernel<float>), grid, dim3(kSumsqThreads), 0, st,
                       (const float*)A->base, A->m, A->n, A->mb, A->nb, part, bpt, work_dev,
                       nonfinite_dev);
  DM_LAUNCH_CHECK("tiled_sumsq_kernel");
  hipLaunchKernelGGL(sum_sqrt_kernel, dim3(1), dim3(kSumsqThreads), 0, st, (const double*)work_dev,
                     tiles * bpt, out_dev);
  DM_LAUNCH_CHECK("sum_sqrt_kernel");
  return 0;
}

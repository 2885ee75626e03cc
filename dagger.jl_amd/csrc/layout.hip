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

extern "C" int dm_tiled_pack(const dm_tiled* dst, const void* src, int64_t ld, double diag_add,
                             dm_stream_t s) {
  if (!dst) return arg_fail(1, "dst is NULL");
  if (dst->mb <= 0 || dst->nb <= 0) return arg_fail(1, "block size must be positive");
  if (ld < (dst->m > 1 ? dst->m : 1)) return arg_fail(3, "ld < max(1,m)");
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
  int rc = ensure_init();
  if (rc) return rc;
  if (src->dtype == DM_F64) return tile_copy<double>(false, src, dst, ld, 0.0, stream_of(s));
  if (src->dtype == DM_F32) return tile_copy<float>(false, src, dst, ld, 0.0, stream_of(s));
  return arg_fail(3, "unknown dtype");
}

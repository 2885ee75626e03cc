// Inverted diagonal tiles of a tiled triangular factor: the first stage of
// the streaming ldiv! (dm_potrs_tiled).
//
// ldiv!(C::Cholesky, B) (src/array/linalg.jl:385-401) is two triangular
// sweeps over the factor.  Solved tile by tile as Dagger's trsm! does
// (src/array/trsm.jl:90-114), every step waits for a triangular solve of one
// diagonal tile with the right-hand sides, a serial chain of short,
// latency-bound launches.  Here the inverse of every diagonal tile is formed
// once per call (batched over all tiles, on the MFMA GEMM), after which both
// sweeps are GEMM-only: Y_k = inv(T_kk) R_k and the updates R_j -= T_jk Y_k
// of the following block rows -- wide batched launches that stream the
// factor from HBM.
//
// Inversion of one triangular tile (LAPACK trtri's recursion, bottom-up):
//   leaves: the 64 x 64 diagonal blocks, inverted by substitution in LDS
//           (one workgroup per block, one thread per column);
//   levels s = 64, 128, ...: adjacent s-blocks (X11, X22 already inverted)
//           get their coupling block from two batched GEMMs,
//             upper: X12 = -X11 (U12 X22)     lower: X21 = -X22 (L21 X11).
// The inverse tiles are stored dense (the other triangle zeroed), one
// rows(k) x rows(k) column-major array per tile.
#include <algorithm>
#include <cstring>
#include <functional>
#include <vector>
#include "potrs.h"

namespace dm {

// Inverse of one n x n (n <= 64) triangular block: thread c computes column
// c of the inverse by substitution against the block staged in LDS.
template <typename T, bool UPPER>
__global__ __launch_bounds__(64) void trtri_leaf_kernel(const TrtriLeaf* __restrict__ leaves,
                                                        const int* __restrict__ abort) {
  __shared__ T t[64][65];
  __shared__ T x[64][65];
  if (abort && *abort) return;
  const TrtriLeaf L = leaves[blockIdx.x];
  const auto src = gptr((const T*)L.src);
  const auto dst = gptr((T*)L.dst);
  const int n = L.n, r = threadIdx.x;
  for (int cc = 0; cc < 64; ++cc) {  // thread r reads row r: coalesced per column
    T v = (r == cc) ? T(1) : T(0);
    if (r < n && cc < n) v = src[(size_t)cc * L.lds + r];
    t[r][cc] = v;
    x[r][cc] = T(0);
  }
  __syncthreads();
  const int c = threadIdx.x;
  if (UPPER) {
    for (int i = c; i >= 0; --i) {
      T s = (i == c) ? T(1) : T(0);
      for (int p = i + 1; p <= c; ++p) s -= t[i][p] * x[p][c];
      x[i][c] = s / t[i][i];
    }
  } else {
    for (int i = c; i < 64; ++i) {
      T s = (i == c) ? T(1) : T(0);
      for (int p = c; p < i; ++p) s -= t[i][p] * x[p][c];
      x[i][c] = s / t[i][i];
    }
  }
  __syncthreads();
  for (int cc = 0; cc < n; ++cc)
    if (r < n) dst[(size_t)cc * L.ldd + r] = x[r][cc];
}

namespace {
template <typename T>
size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
}  // namespace

template <typename T>
TrtriPlan<T>::~TrtriPlan() {
  if (mem) (void)hipFree(mem);
}

template <typename T>
int TrtriPlan<T>::build(bool up, int64_t m_, int nb_, const std::function<const T*(int)>& tile) {
  const int count = (int)((m_ + nb_ - 1) / nb_);
  auto rws = [m_, nb_](int k) { return (int)std::min<int64_t>(nb_, m_ - (int64_t)k * nb_); };
  int rc = build_list(up, count, nb_, tile, rws, rws);
  m = m_;
  return rc;
}

template <typename T>
int TrtriPlan<T>::build_list(bool up, int count, int nb_, const std::function<const T*(int)>& tile,
                             const std::function<int(int)>& rows_of,
                             const std::function<int(int)>& ld_of) {
  upper = up;
  nb = nb_;
  mt = count;
  m = 0;
  rowsv.assign(mt, 0);
  for (int k = 0; k < mt; ++k) {
    rowsv[k] = rows_of(k);
    if (rowsv[k] < 0 || rowsv[k] > nb || ld_of(k) < rowsv[k])
      return arg_fail(0, "trtri: tile rows outside [0, nb] or ld < rows");
    m += rowsv[k];
  }
  inv_bytes = align_up<T>(sizeof(T) * (size_t)mt * nb * nb, 256);
  std::vector<TrtriLeaf> lv;
  std::vector<size_t> leaf_off;  // element offset of each leaf in the inverse tiles
  for (int k = 0; k < mt; ++k) {
    const int r = rows(k), lt = ld_of(k);
    for (int l0 = 0; l0 < r; l0 += 64) {
      TrtriLeaf L{};
      L.src = tile(k) + (size_t)l0 * lt + l0;
      L.dst = nullptr;  // patched below (offset into the inverse tiles)
      L.lds = lt;
      L.ldd = r;
      L.n = std::min(64, r - l0);
      leaf_off.push_back((size_t)k * nb * nb + (size_t)l0 * r + l0);
      lv.push_back(L);
    }
  }
  nleaves = (int)lv.size();
  // levels: pairs (p0, p0 + s) of already-inverted blocks
  struct Pair { int k, p0, s, s2; };
  std::vector<std::vector<Pair>> levels;
  for (int s = 64; s < nb; s *= 2) {
    std::vector<Pair> ps;
    for (int k = 0; k < mt; ++k) {
      const int r = rows(k);
      for (int p0 = 0; p0 + s < r; p0 += 2 * s) ps.push_back(Pair{k, p0, s, std::min(s, r - p0 - s)});
    }
    if (!ps.empty()) levels.push_back(ps);
  }
  size_t scratch_elems = 0;
  for (auto& ps : levels) {
    size_t e = 0;
    for (auto& p : ps) e += (size_t)p.s * p.s2;
    scratch_elems = std::max(scratch_elems, e);
  }
  const size_t scratch_bytes = align_up<T>(sizeof(T) * scratch_elems, 256);
  // job tables: every level's two batches, built against offsets first
  std::vector<std::vector<GemmJob<T>>> tabs;
  for (auto& ps : levels) {
    std::vector<GemmJob<T>> j1, j2;
    size_t off = 0;
    for (auto& p : ps) {
      const int r = rows(p.k);
      const size_t base = (size_t)p.k * nb * nb;
      const size_t d11 = base + (size_t)p.p0 * r + p.p0;
      const size_t d22 = base + (size_t)(p.p0 + p.s) * r + p.p0 + p.s;
      GemmJob<T> a{}, b{};
      const int lt = ld_of(p.k);
      if (upper) {
        // S = U12 X22 (s x s2); X12 = -X11 S
        a.A = tile(p.k) + (size_t)(p.p0 + p.s) * lt + p.p0; a.lda = lt;
        a.B = (const T*)(uintptr_t)(d22 + 1); a.ldb = r;   // inverse-tile offsets + 1
        a.M = p.s; a.N = p.s2; a.K = p.s2;
        b.A = (const T*)(uintptr_t)(d11 + 1); b.lda = r;
        b.M = p.s; b.N = p.s2; b.K = p.s;
        b.C = (T*)(uintptr_t)(base + (size_t)(p.p0 + p.s) * r + p.p0 + 1); b.ldc = r;
      } else {
        // S = L21 X11 (s2 x s); X21 = -X22 S
        a.A = tile(p.k) + (size_t)p.p0 * lt + p.p0 + p.s; a.lda = lt;
        a.B = (const T*)(uintptr_t)(d11 + 1); a.ldb = r;
        a.M = p.s2; a.N = p.s; a.K = p.s;
        b.A = (const T*)(uintptr_t)(d22 + 1); b.lda = r;
        b.M = p.s2; b.N = p.s; b.K = p.s2;
        b.C = (T*)(uintptr_t)(base + (size_t)p.p0 * r + p.p0 + p.s + 1); b.ldc = r;
      }
      // S in the level scratch (marked by a high tag bit until patched)
      a.C = (T*)(uintptr_t)(((size_t)1 << 62) | off); a.ldc = a.M;
      b.B = (const T*)(uintptr_t)(((size_t)1 << 62) | off); b.ldb = a.M;
      a.alpha = T(1); a.beta = T(0);
      b.alpha = T(-1); b.beta = T(0);
      off += (size_t)a.M * a.N;
      j1.push_back(a);
      j2.push_back(b);
    }
    tabs.push_back(j1);
    tabs.push_back(j2);
  }
  size_t tab_bytes = align_up<T>(sizeof(TrtriLeaf) * lv.size(), 256);
  std::vector<size_t> tab_off;
  for (auto& t : tabs) {
    tab_off.push_back(tab_bytes);
    tab_bytes += align_up<T>(sizeof(GemmJob<T>) * t.size(), 256);
  }
  DM_HIP(hipMalloc(&mem, inv_bytes + scratch_bytes + tab_bytes));
  // The inverse tiles' other triangle stays zero: the leaves write their
  // whole square (zeros included) and the level products only blocks inside
  // the triangle, so it is cleared once here instead of on every run.
  DM_HIP(hipMemset(mem, 0, inv_bytes));
  char* base = (char*)mem;
  T* invb = (T*)base;
  T* scr = (T*)(base + inv_bytes);
  char* tb = base + inv_bytes + scratch_bytes;
  for (size_t i = 0; i < lv.size(); ++i) lv[i].dst = invb + leaf_off[i];
  auto fix_in = [&](const T* p) -> const T* {
    const uintptr_t v = (uintptr_t)p;
    if (v & ((uintptr_t)1 << 62)) return scr + (v & (((uintptr_t)1 << 62) - 1));
    return invb + (v - 1);
  };
  std::vector<unsigned char> blob(tab_bytes, 0);
  std::memcpy(blob.data(), lv.data(), sizeof(TrtriLeaf) * lv.size());
  launches.clear();
  for (size_t t = 0; t < tabs.size(); ++t) {
    auto& jobs = tabs[t];
    const bool first = (t % 2) == 0;
    for (auto& j : jobs) {
      if (first) {
        j.B = fix_in(j.B);
        j.C = (T*)fix_in(j.C);
      } else {
        j.A = fix_in(j.A);
        j.B = fix_in(j.B);
        j.C = (T*)fix_in(j.C);
      }
    }
    Launch L;
    int maxM = 0, maxN = 0, maxK = 0;
    for (auto& j : jobs) {
      maxM = std::max(maxM, j.M); maxN = std::max(maxN, j.N); maxK = std::max(maxK, j.K);
    }
    L.njobs = (int)jobs.size();
    L.vec = true;
    for (auto& j : jobs) L.vec = L.vec && gemm_job_vec_ok<T>(false, false, j);
    L.cfg = gemm_pick_cfg<T>(maxM, maxN, maxK, L.njobs, L.vec);
    L.total = gemm_prepare_jobs<T>(jobs.data(), L.njobs, L.cfg);
    std::memcpy(blob.data() + tab_off[t], jobs.data(), sizeof(GemmJob<T>) * jobs.size());
    L.jobs = (const GemmJob<T>*)(tb + tab_off[t]);
    launches.push_back(L);
  }
  leaves_dev = (const TrtriLeaf*)tb;
  DM_HIP(hipMemcpy(tb, blob.data(), blob.size(), hipMemcpyHostToDevice));
  return 0;
}

template <typename T>
int TrtriPlan<T>::run(hipStream_t s, const int* abort) const {
  int rc = 0;
  if (nleaves) {
    if (upper)
      hipLaunchKernelGGL((trtri_leaf_kernel<T, true>), dim3(nleaves), dim3(64), 0, s, leaves_dev,
                         abort);
    else
      hipLaunchKernelGGL((trtri_leaf_kernel<T, false>), dim3(nleaves), dim3(64), 0, s,
                         leaves_dev, abort);
    DM_LAUNCH_CHECK("trtri_leaf_kernel");
  }
  GemmJob<T> dummy{};
  for (const Launch& L : launches) {
    rc = gemm_launch<T>(false, false, L.vec, L.cfg, L.jobs, L.njobs, dummy, L.total, abort, s);
    if (rc) return rc;
  }
  return 0;
}

template struct TrtriPlan<double>;
template struct TrtriPlan<float>;

}  // namespace dm

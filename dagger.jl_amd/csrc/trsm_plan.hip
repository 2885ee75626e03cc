// Host planning of batched recursive triangular solves (no device code):
// the effective-solve shape of a BLAS trsm call and the leaf/GEMM launch
// sequence of a group of solves.  Used by the datadeps executor.
#include <map>
#include "trsm_plan.h"

namespace dm {

int trsm_leaf_len() {
  // Leaves of 128 rows need half the LDS of 256-row leaves, so two leaf
  // workgroups share a CU (measured at N = 32768: ldiv! 33.5 ms against
  // 36.6 ms, cholesky! unchanged).
  return 128;
}

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

// Batched planning: jobs of equal length L.
template <typename T>
static void plan_rec(const TrsmShape& sh, int L, const std::vector<TrsmJob<T>>& jobs,
                     std::vector<TrsmStep<T>>& steps, int leaf) {
  if (jobs.empty() || L <= 0) return;
  if (L <= leaf) {
    TrsmStep<T> st;
    st.leaf = true;
    st.maxL = L;
    long wg = 0;
    for (TrsmJob<T> j : jobs) {
      j.L = L;
      j.wg_begin = (int)wg;
      wg += trsm_wgs(j.NV);
      st.tjobs.push_back(j);
    }
    st.total_wg = wg;
    steps.push_back(std::move(st));
    return;
  }
  const TrsmSplit sp = trsm_split(sh, L);
  std::vector<TrsmJob<T>> j1, j2;
  TrsmStep<T> up;
  up.leaf = false;
  for (const TrsmJob<T>& j : jobs) {
    TrsmJob<T> a = j;
    a.A = tm_sub(sh, j.A, j.lda, sp.first_r, sp.first_r);
    a.B = vsub(sh, j.B, j.ldb, sp.first_r);
    a.inv = j.inv + (size_t)(sp.first_r / TRSM_BS) * TRSM_BS * TRSM_BS;
    j1.push_back(a);
    up.gjobs.push_back(trsm_update_job<T>(sh, sp, j.NV, j.alpha, j.A, j.lda, j.B, j.ldb, up.ta,
                                          up.tb));
    TrsmJob<T> c = j;
    c.A = tm_sub(sh, j.A, j.lda, sp.second_r, sp.second_r);
    c.B = vsub(sh, j.B, j.ldb, sp.second_r);
    c.inv = j.inv + (size_t)(sp.second_r / TRSM_BS) * TRSM_BS * TRSM_BS;
    c.alpha = T(1);
    j2.push_back(c);
  }
  plan_rec<T>(sh, sp.first_n, j1, steps, leaf);
  steps.push_back(std::move(up));
  plan_rec<T>(sh, sp.second_n, j2, steps, leaf);
}

template <typename T>
void trsm_plan_steps(const TrsmShape& sh, const std::vector<TrsmJob<T>>& jobs,
                     std::vector<TrsmStep<T>>& steps, int leaf) {
  if (leaf < TRSM_BS || leaf > TRSM_LEAF || leaf % TRSM_BS) leaf = trsm_leaf_len();
  // one recursion per distinct length (a group normally has one)
  std::map<int, std::vector<TrsmJob<T>>> byL;
  for (const TrsmJob<T>& j : jobs)
    if (j.L > 0 && j.NV > 0) byL[j.L].push_back(j);
  for (auto& kv : byL) plan_rec<T>(sh, kv.first, kv.second, steps, leaf);
}

template void trsm_plan_steps<double>(const TrsmShape&, const std::vector<TrsmJob<double>>&,
                                      std::vector<TrsmStep<double>>&, int);
template void trsm_plan_steps<float>(const TrsmShape&, const std::vector<TrsmJob<float>>&,
                                     std::vector<TrsmStep<float>>&, int);

}  // namespace dm

// Host-side geometry of the recursive triangular solve, shared by the
// unbatched path (trsm.hip) and the executor's batched planner
// (trsm_plan.hip).
#pragma once
#include "dm_internal.h"

namespace dm {

// Origin of Tm[r, c] in A.
template <typename T>
inline const T* tm_sub(const TrsmShape& sh, const T* A, int lda, int r, int c) {
  return sh.ttrans ? A + (size_t)r * lda + c : A + (size_t)c * lda + r;
}
// Origin of vector element r (side L: row r of B; side R: column r).
template <typename T>
inline T* vsub(const TrsmShape& sh, T* B, int ldb, int r) {
  return sh.side_l ? B + r : B + (size_t)r * ldb;
}

// Split of a solve of length L > trsm_leaf_len() at P (a multiple of TRSM_BS from
// the top of the triangle): rows [0, P) and [P, L).  Lower triangles solve
// [0, P) first, upper ones [P, L) first.
struct TrsmSplit {
  int first_r, first_n, second_r, second_n;
};
inline TrsmSplit trsm_split(const TrsmShape& sh, int L) {
  const int P = ((L / 2) + TRSM_BS - 1) / TRSM_BS * TRSM_BS;
  if (sh.lower) return TrsmSplit{0, P, P, L - P};
  return TrsmSplit{P, L - P, 0, P};
}

// The GEMM update V2 := alpha*V2 - Tm21 X1 of one job (side L: rows of B;
// side R: columns) and its operand transposes.
template <typename T>
inline GemmJob<T> trsm_update_job(const TrsmShape& sh, const TrsmSplit& sp, int NV, T alpha,
                                  const T* A, int lda, T* B, int ldb, bool& ta, bool& tb) {
  const T* t21 = tm_sub(sh, A, lda, sp.second_r, sp.first_r);
  GemmJob<T> g{};
  if (sh.side_l) {
    // B[second] (second_n x NV) -= Tm21 (second_n x first_n) * X1 (first_n x NV)
    ta = sh.ttrans;
    tb = false;
    g.A = t21; g.lda = lda;
    g.B = vsub(sh, B, ldb, sp.first_r); g.ldb = ldb;
    g.M = sp.second_n; g.N = NV;
  } else {
    // B[:, second] (NV x second_n) -= X1^T (NV x first_n) * Tm21^T (first_n x second_n)
    ta = false;
    tb = !sh.ttrans;
    g.A = vsub(sh, B, ldb, sp.first_r); g.lda = ldb;
    g.B = t21; g.ldb = lda;
    g.M = NV; g.N = sp.second_n;
  }
  g.K = sp.first_n;
  g.C = vsub(sh, B, ldb, sp.second_r);
  g.ldc = ldb;
  g.alpha = T(-1);
  g.beta = alpha;
  g.tri = 0;
  return g;
}

}  // namespace dm

// Complex (ComplexF64 / ComplexF32) tile operations (complex.hip).  Pointers
// are to interleaved (re, im) pairs of R; sizes and leading dimensions count
// complex elements.
#pragma once
#include <hip/hip_runtime.h>

namespace dm {

template <typename R>
struct ZGemmArgs {
  const R* A;
  const R* B;
  R* C;
  int M, N, K;
  int lda, ldb, ldc;
  R alpha_re, alpha_im, beta_re, beta_im;
  int opa, opb;  // 0 'N', 1 'T', 2 'C' (conjugate transpose), 3 conjugate only
  int tri;       // 0 all of C, 1 upper triangle, 2 lower triangle
  int herk;      // 1: the imaginary part of C's diagonal is set to zero
};

// C = alpha op(A) op(B) + beta C (BLAS gemm!; tri/herk as above give herk!)
template <typename R>
int zgemm_call(char ta, char tb, int m, int n, int k, R ar, R ai, const R* A, int lda,
               const R* B, int ldb, R br, R bi, R* C, int ldc, int tri, int herk,
               const int* abort, hipStream_t s);

// BLAS trsm!(side, uplo, trans, diag, alpha, A, B) on an m x n B
template <typename R>
int ztrsm_call(char side, char uplo, char trans, char diag, int m, int n, R ar, R ai,
               const R* A, int lda, R* B, int ldb, const int* abort, hipStream_t s);

// LAPACK potrf!(uplo, A) of an n x n tile: the first failed pivot's 1-based
// column within the tile (info_off + local) goes to *info (if 0) and
// global_base + that column to *abort
template <typename R>
int zpotrf_rec(bool upper, int n, R* A, int lda, int* info, int info_off, int* abort,
               int global_base, hipStream_t s);

}  // namespace dm

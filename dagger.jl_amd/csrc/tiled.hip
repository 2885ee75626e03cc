// Whole-matrix forms: the exact task DAGs of Dagger's tiled Cholesky and
// tiled triangular solves, built on resident tiles and run through the
// datadeps executor with a cached plan (and hipGraph) per matrix.
//
//   _chol!(A, UpperTriangular)   src/array/cholesky.jl:19-63 (loop :38-53)
//   _chol!(A, LowerTriangular)   src/array/cholesky.jl:64-108 (loop :83-98)
//   trsm!(side, uplo, trans, diag, alpha, A, B)   src/array/trsm.jl:59-192
//   ldiv!(C::Cholesky, B)        src/array/linalg.jl:385-401
//
// Tiles follow Dagger's chunk layout (each chunk a dense column-major array
// with ld = its row count, src/array/darray.jl:568-581) packed in one HBM
// allocation per matrix, so every task argument is a resident device
// pointer and no per-region copies exist (the role of generate_slot!/
// move_rewrap in src/datadeps/aliasing.jl:920-1060 disappears).
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <sstream>
#include <unordered_map>
#include <vector>
#include "dag.h"
#include "potrs.h"

namespace dm {

int copy_i32_launch(int* dst, const int* src, size_t count, hipStream_t s);

namespace {

struct TileView {
  char* base;
  int64_t m, n;
  int mb, nb;
  int dtype;
  size_t es;
  int mt, nt;
  explicit TileView(const dm_tiled& t)
      : base((char*)t.base), m(t.m), n(t.n), mb(t.mb), nb(t.nb), dtype(t.dtype),
        es(t.dtype == DM_F32 ? 4 : t.dtype == DM_C64 ? 16 : 8) {
    mt = (int)((m + mb - 1) / mb);
    nt = (int)((n + nb - 1) / nb);
  }
  int rows(int i) const { return (int)std::min<int64_t>(mb, m - (int64_t)i * mb); }
  int cols(int j) const { return (int)std::min<int64_t>(nb, n - (int64_t)j * nb); }
  void* ptr(int i, int j) const {
    return base + es * ((size_t)j * nb * m + (size_t)i * mb * cols(j));
  }
  int ld(int i) const { return rows(i); }
};

int check_tiled(const dm_tiled* A, int argno) {
  if (!A) return arg_fail(argno, "tiled matrix is NULL");
  if (!A->base && A->m > 0 && A->n > 0) return arg_fail(argno, "tiled base is NULL");
  if (A->m < 0 || A->n < 0) return arg_fail(argno, "negative size");
  if (A->mb <= 0 || A->nb <= 0) return arg_fail(argno, "block size must be positive");
  if (A->dtype != DM_F64 && A->dtype != DM_F32 && A->dtype != DM_C64 && A->dtype != DM_C32)
    return arg_fail(argno, "unknown dtype");
  return 0;
}

// BLAS transpose character of a tiled operation: 'C' stays the conjugate
// transpose on complex tiles and is 'T' on real ones.
char trans_char(char t, int dtype) {
  if (is_notrans(t)) return 'N';
  return (is_cplx(dtype) && (t == 'C' || t == 'c')) ? 'C' : 'T';
}

dm_task make(int op, int dtype) {
  dm_task t;
  std::memset(&t, 0, sizeof t);
  t.op = op;
  t.dtype = dtype;
  t.side = t.uplo = t.transa = t.transb = t.diag = 'N';
  return t;
}

void add_potrf(Dag& d, const TileView& A, int k, char uplo, int* info) {
  dm_task t = make(DM_OP_POTRF, A.dtype);
  t.uplo = uplo;
  t.n = A.rows(k);
  t.k = k * A.mb;  // global row offset (for the global info)
  t.c = A.ptr(k, k);
  t.ldc = A.ld(k);
  t.info = info;
  t.access[2] = DM_INOUT;
  d.add(t);
}

void add_trsm(Dag& d, int dtype, char side, char uplo, char trans, char diag, double alpha,
              const void* a, int lda, void* b, int m, int n, int ldb) {
  dm_task t = make(DM_OP_TRSM, dtype);
  t.side = side; t.uplo = uplo; t.transa = trans; t.diag = diag;
  t.m = m; t.n = n;
  t.alpha = alpha;
  t.a = a; t.lda = lda;
  t.c = b; t.ldc = ldb;
  t.access[0] = DM_IN;
  t.access[2] = DM_INOUT;
  d.add(t);
}

void add_syrk(Dag& d, int dtype, char uplo, char trans, int n, int k, double alpha, const void* a,
              int lda, double beta, void* c, int ldc) {
  dm_task t = make(DM_OP_SYRK, dtype);
  t.uplo = uplo; t.transa = trans;
  t.n = n; t.k = k;
  t.alpha = alpha; t.beta = beta;
  t.a = a; t.lda = lda;
  t.c = c; t.ldc = ldc;
  t.access[0] = DM_IN;
  t.access[2] = DM_INOUT;
  d.add(t);
}

void add_gemm(Dag& d, int dtype, char ta, char tb, int m, int n, int k, double alpha,
              const void* a, int lda, const void* b, int ldb, double beta, void* c, int ldc) {
  dm_task t = make(DM_OP_GEMM, dtype);
  t.transa = ta; t.transb = tb;
  t.m = m; t.n = n; t.k = k;
  t.alpha = alpha; t.beta = beta;
  t.a = a; t.lda = lda;
  t.b = b; t.ldb = ldb;
  t.c = c; t.ldc = ldc;
  t.access[0] = DM_IN;
  t.access[1] = DM_IN;
  t.access[2] = DM_INOUT;
  d.add(t);
}

// src/array/cholesky.jl:38-53 (U) and :83-98 (L).
void build_cholesky(Dag& d, const TileView& A, char uplo, int* info) {
  const int mt = A.mt, nt = A.nt;
  const int dt = A.dtype;
  const char tc = is_cplx(dt) ? 'C' : 'T';  // trans = iscomplex ? 'C' : 'T'; SYRK -> herk
  if (is_upper(uplo)) {
    for (int k = 0; k < mt; ++k) {
      add_potrf(d, A, k, 'U', info);
      for (int n = k + 1; n < nt; ++n)
        add_trsm(d, dt, 'L', 'U', tc, 'N', 1.0, A.ptr(k, k), A.ld(k), A.ptr(k, n), A.rows(k),
                 A.cols(n), A.ld(k));
      for (int m = k + 1; m < mt; ++m) {
        add_syrk(d, dt, 'U', tc, A.cols(m), A.rows(k), -1.0, A.ptr(k, m), A.ld(k), 1.0,
                 A.ptr(m, m), A.ld(m));
        for (int n = m + 1; n < nt; ++n)
          add_gemm(d, dt, tc, 'N', A.rows(m), A.cols(n), A.rows(k), -1.0, A.ptr(k, m), A.ld(k),
                   A.ptr(k, n), A.ld(k), 1.0, A.ptr(m, n), A.ld(m));
      }
    }
  } else {
    for (int k = 0; k < mt; ++k) {
      add_potrf(d, A, k, 'L', info);
      for (int m = k + 1; m < mt; ++m)
        add_trsm(d, dt, 'R', 'L', tc, 'N', 1.0, A.ptr(k, k), A.ld(k), A.ptr(m, k), A.rows(m),
                 A.cols(k), A.ld(m));
      for (int n = k + 1; n < nt; ++n) {
        add_syrk(d, dt, 'L', 'N', A.rows(n), A.cols(k), -1.0, A.ptr(n, k), A.ld(n), 1.0,
                 A.ptr(n, n), A.ld(n));
        for (int m = n + 1; m < mt; ++m)
          add_gemm(d, dt, 'N', tc, A.rows(m), A.rows(n), A.cols(k), -1.0, A.ptr(m, k), A.ld(m),
                   A.ptr(n, k), A.ld(n), 1.0, A.ptr(m, n), A.ld(m));
      }
    }
  }
}

void add_getrf(Dag& d, const TileView& A, int k, bool pivot, int* ipiv, int* info, int64_t rows) {
  dm_task t = make(DM_OP_GETRF, A.dtype);
  t.transa = pivot ? 'P' : 'N';
  t.m = (int)rows;
  t.n = A.cols(k);
  t.k = A.mb;
  t.c = A.ptr(k, k);
  t.b = ipiv + (size_t)k * A.nb;  // chunk k of the Blocks(nb) pivot vector
  t.ldb = k * A.mb;
  t.info = info;
  t.access[1] = DM_OUT;
  t.access[2] = DM_INOUT;
  d.add(t);
}

void add_laswp(Dag& d, const TileView& A, int k, int j, const int* ipiv) {
  dm_task t = make(DM_OP_LASWP, A.dtype);
  t.m = (int)(A.m - (int64_t)k * A.mb);
  t.n = A.cols(j);
  t.k = A.mb;
  t.c = A.ptr(k, j);
  t.b = ipiv + (size_t)k * A.nb;
  t.lda = (int)std::min<int64_t>(A.m - (int64_t)k * A.mb, A.cols(k));
  t.ldb = k * A.mb;
  t.access[1] = DM_IN;
  t.access[2] = DM_INOUT;
  d.add(t);
}

// lu!(A, RowMaximum()) (src/array/lu.jl:289-371): per block column k the
// panel with partial pivoting (tiles k..mt-1), its row swaps in every other
// block column, then trsm!('L','L','N','U') across the row and the GEMM
// trailing update.  lu!(A, NoPivot()) (src/array/lu.jl:10-40): the diagonal
// tile's LU without pivoting, trsm!('R','U','N','N') down the column,
// trsm!('L','L','N','U') across the row, GEMM.
void build_lu(Dag& d, const TileView& A, bool pivot, int* ipiv, int* info) {
  const int mt = A.mt, nt = A.nt, dt = A.dtype;
  for (int k = 0; k < std::min(mt, nt); ++k) {
    if (pivot) {
      add_getrf(d, A, k, true, ipiv, info, A.m - (int64_t)k * A.mb);
      for (int j = 0; j < nt; ++j)
        if (j != k) add_laswp(d, A, k, j, ipiv);
    } else {
      add_getrf(d, A, k, false, ipiv, info, A.rows(k));
      for (int m = k + 1; m < mt; ++m)
        add_trsm(d, dt, 'R', 'U', 'N', 'N', 1.0, A.ptr(k, k), A.ld(k), A.ptr(m, k), A.rows(m),
                 A.cols(k), A.ld(m));
    }
    for (int j = k + 1; j < nt; ++j) {
      add_trsm(d, dt, 'L', 'L', 'N', 'U', 1.0, A.ptr(k, k), A.ld(k), A.ptr(k, j), A.rows(k),
               A.cols(j), A.ld(k));
      for (int i = k + 1; i < mt; ++i)
        add_gemm(d, dt, 'N', 'N', A.rows(i), A.cols(j), A.cols(k), -1.0, A.ptr(i, k), A.ld(i),
                 A.ptr(k, j), A.ld(k), 1.0, A.ptr(i, j), A.ld(i));
    }
  }
}

// src/array/trsm.jl:88-190, 0-based (Dagger's index (Bmt-k)+1 is Bmt-k here
// with k counted from 1).
void build_trsm(Dag& d, char side, char uplo, char trans, char diag, double alpha,
                const TileView& A, const TileView& B) {
  const int dt = B.dtype;
  const int Bmt = B.mt, Bnt = B.nt;
  const char tr = trans_char(trans, dt);
  auto trsm = [&](double al, int ai, int bi, int bj) {
    add_trsm(d, dt, side, uplo, tr, diag, al, A.ptr(ai, ai), A.ld(ai), B.ptr(bi, bj), B.rows(bi),
             B.cols(bj), B.ld(bi));
  };
  // C[ci,cj] = -op(X) op(Y) + beta C, with X, Y tiles of A or B
  auto gemm = [&](char ta, char tb, const TileView& X, int xi, int xj, const TileView& Y, int yi,
                  int yj, double beta, int ci, int cj) {
    const int M = B.rows(ci), N = B.cols(cj);
    const int K = is_trans(ta) ? X.rows(xi) : X.cols(xj);
    add_gemm(d, dt, ta, tb, M, N, K, -1.0, X.ptr(xi, xj), X.ld(xi), Y.ptr(yi, yj), Y.ld(yi), beta,
             B.ptr(ci, cj), B.ld(ci));
  };
  if (is_left(side)) {
    if (is_upper(uplo)) {
      if (tr == 'N') {
        for (int k = 1; k <= Bmt; ++k) {
          const double la = k == 1 ? alpha : 1.0;
          const int kk = Bmt - k;
          for (int n = 0; n < Bnt; ++n) trsm(la, kk, kk, n);
          for (int m = k + 1; m <= Bmt; ++m)
            for (int n = 0; n < Bnt; ++n) gemm('N', 'N', A, Bmt - m, kk, B, kk, n, la, Bmt - m, n);
        }
      } else {
        for (int k = 0; k < Bmt; ++k) {
          const double la = k == 0 ? alpha : 1.0;
          for (int n = 0; n < Bnt; ++n) trsm(la, k, k, n);
          for (int m = k + 1; m < Bmt; ++m)
            for (int n = 0; n < Bnt; ++n) gemm(tr, 'N', A, k, m, B, k, n, la, m, n);
        }
      }
    } else {
      if (tr == 'N') {
        for (int k = 0; k < Bmt; ++k) {
          const double la = k == 0 ? alpha : 1.0;
          for (int n = 0; n < Bnt; ++n) trsm(la, k, k, n);
          for (int m = k + 1; m < Bmt; ++m)
            for (int n = 0; n < Bnt; ++n) gemm('N', 'N', A, m, k, B, k, n, la, m, n);
        }
      } else {
        for (int k = 1; k <= Bmt; ++k) {
          const double la = k == 1 ? alpha : 1.0;
          const int kk = Bmt - k;
          for (int n = 0; n < Bnt; ++n) trsm(la, kk, kk, n);
          for (int m = k + 1; m <= Bmt; ++m)
            for (int n = 0; n < Bnt; ++n) gemm(tr, 'N', A, kk, Bmt - m, B, kk, n, la, Bmt - m, n);
        }
      }
    }
  } else {
    if (is_upper(uplo)) {
      if (tr == 'N') {
        for (int k = 0; k < Bnt; ++k) {
          const double la = k == 0 ? alpha : 1.0;
          for (int m = 0; m < Bmt; ++m) trsm(la, k, m, k);
          for (int m = 0; m < Bmt; ++m)
            for (int n = k + 1; n < Bnt; ++n) gemm('N', 'N', B, m, k, A, k, n, la, m, n);
        }
      } else {
        for (int k = 1; k <= Bnt; ++k) {
          const int kk = Bnt - k;
          for (int m = 0; m < Bmt; ++m) {
            trsm(alpha, kk, m, kk);
            for (int n = k + 1; n <= Bnt; ++n)
              gemm('N', tr, B, m, kk, A, Bnt - n, kk, 1.0, m, Bnt - n);
          }
        }
      }
    } else {
      if (tr == 'N') {
        for (int k = 1; k <= Bnt; ++k) {
          const double la = k == 1 ? alpha : 1.0;
          const int kk = Bnt - k;
          for (int m = 0; m < Bmt; ++m) {
            trsm(la, kk, m, kk);
            for (int n = k + 1; n <= Bnt; ++n)
              gemm('N', 'N', B, m, kk, A, kk, Bnt - n, la, m, Bnt - n);
          }
        }
      } else {
        for (int k = 0; k < Bnt; ++k) {
          for (int m = 0; m < Bmt; ++m) {
            trsm(alpha, k, m, k);
            for (int n = k + 1; n < Bnt; ++n) gemm('N', tr, B, m, k, A, n, k, 1.0, m, n);
          }
        }
      }
    }
  }
}

}  // namespace

// C = alpha op(A) op(B) + beta C on tile grids with one block size: one GEMM
// task per (i, j, l), beta applied at l = 0 (the tile loop of Dagger's
// matmul, src/array/mul.jl).
void build_gemm_tiled(Dag& d, char ta, char tb, double alpha, const dm_tiled& A,
                      const dm_tiled& B, double beta, const dm_tiled& Cm) {
  TileView va(A), vb(B), vc(Cm);
  const bool tra = is_trans(ta), trb = is_trans(tb);
  const int K = tra ? va.mt : va.nt;
  for (int j = 0; j < vc.nt; ++j)
    for (int i = 0; i < vc.mt; ++i)
      for (int l = 0; l < K; ++l) {
        const int ai = tra ? l : i, aj = tra ? i : l;
        const int bi = trb ? j : l, bj = trb ? l : j;
        const int kk = tra ? va.rows(l) : va.cols(l);
        add_gemm(d, vc.dtype, trans_char(ta, vc.dtype), trans_char(tb, vc.dtype), vc.rows(i),
                 vc.cols(j), kk, alpha,
                 va.ptr(ai, aj), va.ld(ai), vb.ptr(bi, bj), vb.ld(bi), l == 0 ? beta : 1.0,
                 vc.ptr(i, j), vc.ld(i));
      }
}

namespace {

// ldiv!(C::Cholesky, B) (src/array/linalg.jl:385-401) as two GEMM-only
// substitution sweeps over the factor A with the inverses D_k of its
// diagonal tiles (dense rows(k) x rows(k) arrays at dinv(k)), for every
// block column c of B; Y (B's shape) holds the intermediate solution.
//   'U' (A = U):  Y_k = D_k^T R_k,  R_j -= U_kj^T Y_k (j > k);
//                 X_k = D_k Y_k,    Y_i -= U_ik X_k   (i < k, k descending)
//   'L' (A = L):  Y_k = D_k R_k,    R_j -= L_jk Y_k   (j > k);
//                 X_k = D_k^T Y_k,  Y_i -= L_ki^T X_k (i < k, k descending)
// R is B itself (updated in place by the first sweep) and X lands in B.
// The update order over k for each block row is the reference's: tile
// products are subtracted in increasing (first sweep) and decreasing
// (second sweep) k, as trsm! applies them (src/array/trsm.jl:90-114).
void build_potrs_inv(Dag& d, char uplo, const TileView& A, const TileView& B, const TileView& Y,
                     const std::function<void*(int)>& dinv) {
  const int dt = B.dtype, mt = A.mt;
  const bool up = is_upper(uplo);
  for (int c = 0; c < B.nt; ++c) {
    const int nc = B.cols(c);
    for (int k = 0; k < mt; ++k) {
      const int rk = A.rows(k);
      add_gemm(d, dt, up ? 'T' : 'N', 'N', rk, nc, rk, 1.0, dinv(k), rk, B.ptr(k, c), B.ld(k),
               0.0, Y.ptr(k, c), Y.ld(k));
      for (int j = k + 1; j < mt; ++j) {
        if (up)
          add_gemm(d, dt, 'T', 'N', A.rows(j), nc, rk, -1.0, A.ptr(k, j), A.ld(k), Y.ptr(k, c),
                   Y.ld(k), 1.0, B.ptr(j, c), B.ld(j));
        else
          add_gemm(d, dt, 'N', 'N', A.rows(j), nc, rk, -1.0, A.ptr(j, k), A.ld(j), Y.ptr(k, c),
                   Y.ld(k), 1.0, B.ptr(j, c), B.ld(j));
      }
    }
    for (int k = mt - 1; k >= 0; --k) {
      const int rk = A.rows(k);
      add_gemm(d, dt, up ? 'N' : 'T', 'N', rk, nc, rk, 1.0, dinv(k), rk, Y.ptr(k, c), Y.ld(k),
               0.0, B.ptr(k, c), B.ld(k));
      for (int i = 0; i < k; ++i) {
        if (up)
          add_gemm(d, dt, 'N', 'N', A.rows(i), nc, rk, -1.0, A.ptr(i, k), A.ld(i), B.ptr(k, c),
                   B.ld(k), 1.0, Y.ptr(i, c), Y.ld(i));
        else
          add_gemm(d, dt, 'T', 'N', A.rows(i), nc, rk, -1.0, A.ptr(k, i), A.ld(k), B.ptr(k, c),
                   B.ld(k), 1.0, Y.ptr(i, c), Y.ld(i));
      }
    }
  }
}

// Device memory owned by a cached plan (freed when the plan is dropped; the
// plan's destructor first waits for the device).
struct DevMem {
  void* p = nullptr;
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
};

// The streaming ldiv! plan: inverse diagonal tiles formed on the origin
// stream before the GEMM-only sweeps fork onto the executor's streams.
template <typename T>
int setup_potrs_inv(Dag& dag, char up, const TileView& ta, const TileView& tb) {
  auto plan = std::make_shared<TrtriPlan<T>>();
  int rc = plan->build(is_upper(up), ta.m, ta.mb, [&](int k) { return (const T*)ta.ptr(k, k); });
  if (rc) return rc;
  auto ybuf = std::make_shared<DevMem>();
  DM_HIP(hipMalloc(&ybuf->p, std::max<size_t>(sizeof(T) * (size_t)tb.m * tb.n, 256)));
  dm_tiled yt{ybuf->p, tb.m, tb.n, tb.mb, tb.nb, tb.dtype};
  TileView ty(yt);
  build_potrs_inv(dag, up, ta, tb, ty, [&](int k) { return (void*)plan->inv(k); });
  dag.pre_hook = [plan, ybuf](Dag* self, hipStream_t s) { return plan->run(s, self->abort_dev); };
  return 0;
}

// ------------------------------------------------------------ plan cache
struct CacheEntry {
  std::unique_ptr<Dag> dag;
  std::vector<const void*> bases;
};
std::mutex g_cache_mu;
std::unordered_map<std::string, CacheEntry> g_cache;

std::string key_of(const char* kind, std::initializer_list<const void*> ptrs,
                   std::initializer_list<int64_t> nums, const dm_plan_opts& o) {
  std::ostringstream os;
  os << kind;
  for (const void* p : ptrs) os << ':' << p;
  for (int64_t v : nums) os << ':' << v;
  os << ':' << o.num_streams << ',' << o.batch << ',' << o.lookahead << ',' << o.use_graph << ','
     << o.profile << ',' << o.f32_update << ',' << o.f64_diag;
  return os.str();
}

dm_plan_opts opts_or_default(const dm_plan_opts* o) {
  if (o) return *o;
  return dm_plan_opts{2, 1, 1, 1, 0, 0, 0};
}

// Looks up or builds (build(d) fills tasks) and plans a DAG.
template <typename F>
int get_plan(const std::string& key, std::vector<const void*> bases, const dm_plan_opts& o,
             F build, Dag** out) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto it = g_cache.find(key);
  if (it != g_cache.end()) {
    *out = it->second.dag.get();
    return 0;
  }
  std::unique_ptr<Dag> d(new Dag());
  int rc = build(*d);
  if (rc) return rc;
  rc = d->plan(o);
  if (rc) return rc;
  *out = d.get();
  CacheEntry e;
  e.dag = std::move(d);
  e.bases = std::move(bases);
  g_cache.emplace(key, std::move(e));
  return 0;
}

}  // namespace
}  // namespace dm

using namespace dm;

extern "C" {

int dm_cholesky_tiled(char uplo, const dm_tiled* A, int* info_dev, int* global_info_dev,
                      const dm_plan_opts* opts, dm_stream_t s) {
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(1, "uplo must be 'U' or 'L'");
  int rc = check_tiled(A, 2);
  if (rc) return rc;
  if (A->m != A->n) return arg_fail(2, "matrix must be square");
  if (A->mb != A->nb) return arg_fail(2, "tiles must be square (Blocks(nb, nb))");
  if (!info_dev) return arg_fail(3, "info is NULL");
  rc = ensure_init();
  if (rc) return rc;
  hipStream_t st = stream_of(s);
  // info = [0] (src/array/cholesky.jl:32)
  rc = fill_i32_launch(info_dev, 0, 1, st);
  if (rc) return rc;
  if (A->m == 0) {
    if (global_info_dev) return fill_i32_launch(global_info_dev, 0, 1, st);
    return 0;
  }
  const dm_plan_opts o = opts_or_default(opts);
  TileView tv(*A);
  const std::string key = key_of("chol", {A->base, info_dev, global_info_dev},
                                 {A->m, A->mb, A->dtype, is_upper(uplo)}, o);
  Dag* d = nullptr;
  rc = get_plan(key, {A->base, info_dev, global_info_dev}, o,
                [&](Dag& dag) {
                  build_cholesky(dag, tv, is_upper(uplo) ? 'U' : 'L', info_dev);
                  if (global_info_dev)
                    dag.post_hook = [global_info_dev](Dag* self, hipStream_t hs) {
                      return copy_i32_launch(global_info_dev, self->abort_dev, 1, hs);
                    };
                  return 0;
                },
                &d);
  if (rc) return rc;
  return d->run(st);
}

int dm_trsm_tiled(char side, char uplo, char trans, char diag, double alpha, const dm_tiled* A,
                  const dm_tiled* B, const dm_plan_opts* opts, dm_stream_t s) {
  if (!is_left(side) && !is_right(side)) return arg_fail(1, "side must be 'L' or 'R'");
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(2, "uplo must be 'U' or 'L'");
  if (!is_trans(trans) && !is_notrans(trans)) return arg_fail(3, "trans must be 'N','T','C'");
  if (!is_unit(diag) && !is_nonunit(diag)) return arg_fail(4, "diag must be 'N' or 'U'");
  int rc = check_tiled(A, 6);
  if (rc) return rc;
  rc = check_tiled(B, 7);
  if (rc) return rc;
  if (A->m != A->n || A->mb != A->nb) return arg_fail(6, "A must be square with square tiles");
  if (A->dtype != B->dtype) return arg_fail(7, "A and B dtypes differ");
  if (B->mb != A->mb || B->nb != A->nb) return arg_fail(7, "B must use A's block size");
  if (is_left(side) ? B->m != A->m : B->n != A->n) return arg_fail(7, "B does not conform");
  rc = ensure_init();
  if (rc) return rc;
  if (B->m == 0 || B->n == 0) return 0;
  const dm_plan_opts o = opts_or_default(opts);
  TileView ta(*A), tb(*B);
  const char up = is_upper(uplo) ? 'U' : 'L';
  const char sd = is_left(side) ? 'L' : 'R';
  const char tr = trans_char(trans, A->dtype);
  const char dg = is_unit(diag) ? 'U' : 'N';
  char tag[16];
  snprintf(tag, sizeof tag, "trsm%c%c%c%c", sd, up, tr, dg);
  int64_t abits;
  std::memcpy(&abits, &alpha, sizeof abits);
  const std::string key =
      key_of(tag, {A->base, B->base}, {A->m, A->mb, B->m, B->n, A->dtype, abits}, o);
  Dag* d = nullptr;
  rc = get_plan(key, {A->base, B->base}, o,
                [&](Dag& dag) {
                  build_trsm(dag, sd, up, tr, dg, alpha, ta, tb);
                  return 0;
                },
                &d);
  if (rc) return rc;
  return d->run(stream_of(s));
}

int dm_potrs_tiled(char uplo, const dm_tiled* A, const dm_tiled* B, const dm_plan_opts* opts,
                   dm_stream_t s) {
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(1, "uplo must be 'U' or 'L'");
  int rc = check_tiled(A, 2);
  if (rc) return rc;
  rc = check_tiled(B, 3);
  if (rc) return rc;
  if (A->m != A->n || A->mb != A->nb) return arg_fail(2, "A must be square with square tiles");
  if (A->dtype != B->dtype) return arg_fail(3, "A and B dtypes differ");
  if (B->mb != A->mb || B->nb != A->nb) return arg_fail(3, "B must use A's block size");
  if (B->m != A->m) return arg_fail(3, "B does not conform");
  rc = ensure_init();
  if (rc) return rc;
  if (B->n == 0 || B->m == 0) return 0;
  const dm_plan_opts o = opts_or_default(opts);
  TileView ta(*A), tb(*B);
  const char up = is_upper(uplo) ? 'U' : 'L';
  const std::string key =
      key_of(up == 'U' ? "potrsU" : "potrsL", {A->base, B->base}, {A->m, A->mb, B->n, A->dtype}, o);
  Dag* d = nullptr;
  rc = get_plan(key, {A->base, B->base}, o,
                [&](Dag& dag) {
                  // inverse diagonal tiles, then the two sweeps of GEMM tasks;
                  // complex factors: the two trsm! regions (linalg.jl:385-401)
                  if (is_cplx(A->dtype)) {
                    build_trsm(dag, 'L', up, up == 'U' ? 'C' : 'N', 'N', 1.0, ta, tb);
                    build_trsm(dag, 'L', up, up == 'U' ? 'N' : 'C', 'N', 1.0, ta, tb);
                    return 0;
                  }
                  return A->dtype == DM_F64 ? setup_potrs_inv<double>(dag, up, ta, tb)
                                            : setup_potrs_inv<float>(dag, up, ta, tb);
                },
                &d);
  if (rc) return rc;
  return d->run(stream_of(s));
}

int dm_gemm_tiled(char transa, char transb, double alpha, const dm_tiled* A, const dm_tiled* B,
                  double beta, const dm_tiled* Cm, const dm_plan_opts* opts, dm_stream_t s) {
  if (!is_trans(transa) && !is_notrans(transa)) return arg_fail(1, "transa must be 'N','T','C'");
  if (!is_trans(transb) && !is_notrans(transb)) return arg_fail(2, "transb must be 'N','T','C'");
  int rc = check_tiled(A, 4);
  if (rc) return rc;
  rc = check_tiled(B, 5);
  if (rc) return rc;
  rc = check_tiled(Cm, 7);
  if (rc) return rc;
  if (A->dtype != B->dtype || A->dtype != Cm->dtype) return arg_fail(7, "dtypes differ");
  const int64_t am = is_trans(transa) ? A->n : A->m, ak = is_trans(transa) ? A->m : A->n;
  const int64_t bk = is_trans(transb) ? B->n : B->m, bn = is_trans(transb) ? B->m : B->n;
  if (am != Cm->m || bn != Cm->n || ak != bk) return arg_fail(7, "shapes do not conform");
  // One square block size throughout; C may only differ in its column block
  // when it has a single block column that B's blocking also gives (n <= nb
  // of A), since the GEMM tasks take C's column count from B's tiles.
  if (A->mb != A->nb || B->mb != A->mb || B->nb != A->mb || Cm->mb != A->mb ||
      (Cm->nb != A->mb && Cm->n > A->mb))
    return arg_fail(7, "all operands must use one square block size");
  rc = ensure_init();
  if (rc) return rc;
  if (Cm->m == 0 || Cm->n == 0) return 0;
  const dm_plan_opts o = opts_or_default(opts);
  const char ta = trans_char(transa, A->dtype), tb = trans_char(transb, A->dtype);
  // A times a block column of at most 64 columns (the residual of
  // refinement): one streaming launch instead of the k-ordered tile DAG
  if (!is_cplx(A->dtype) && tiled_skinny_gemm_ok(ta, tb, *A, *B, *Cm))
    return tiled_skinny_gemm_launch(*A, *B, *Cm, alpha, beta, stream_of(s));
  int64_t ab, bb;
  std::memcpy(&ab, &alpha, sizeof ab);
  std::memcpy(&bb, &beta, sizeof bb);
  char tag[16];
  snprintf(tag, sizeof tag, "gemm%c%c", ta, tb);
  const std::string key = key_of(tag, {A->base, B->base, Cm->base},
                                 {A->m, A->n, B->m, B->n, Cm->m, Cm->n, A->mb, A->dtype, ab, bb},
                                 o);
  Dag* d = nullptr;
  rc = get_plan(key, {A->base, B->base, Cm->base}, o,
                [&](Dag& dag) {
                  build_gemm_tiled(dag, ta, tb, alpha, *A, *B, beta, *Cm);
                  return 0;
                },
                &d);
  if (rc) return rc;
  return d->run(stream_of(s));
}

int dm_getrf_tiled(int pivot, const dm_tiled* A, int* ipiv_dev, int* info_dev,
                   const dm_plan_opts* opts, dm_stream_t s) {
  if (pivot != 0 && pivot != 1) return arg_fail(1, "pivot must be 0 (NoPivot) or 1 (RowMaximum)");
  int rc = check_tiled(A, 2);
  if (rc) return rc;
  if (is_cplx(A->dtype)) return arg_fail(2, "LU of complex element types is not built");
  if (A->mb != A->nb) return arg_fail(2, "tiles must be square (Blocks(nb, nb))");
  if (!ipiv_dev && A->m > 0 && A->n > 0) return arg_fail(3, "ipiv is NULL");
  if (!info_dev) return arg_fail(4, "info is NULL");
  rc = ensure_init();
  if (rc) return rc;
  hipStream_t st = stream_of(s);
  rc = fill_i32_launch(info_dev, 0, 1, st);
  if (rc) return rc;
  if (A->m == 0 || A->n == 0) return 0;
  const dm_plan_opts o = opts_or_default(opts);
  TileView tv(*A);
  const std::string key = key_of(pivot ? "getrfP" : "getrfN", {A->base, ipiv_dev, info_dev},
                                 {A->m, A->n, A->mb, A->dtype}, o);
  Dag* d = nullptr;
  rc = get_plan(key, {A->base, ipiv_dev, info_dev}, o,
                [&](Dag& dag) {
                  build_lu(dag, tv, pivot == 1, ipiv_dev, info_dev);
                  return 0;
                },
                &d);
  if (rc) return rc;
  return d->run(st);
}

int dm_tiled_laswp(const dm_tiled* B, const int* ipiv_dev, int npiv, dm_stream_t s) {
  int rc = check_tiled(B, 1);
  if (rc) return rc;
  if (is_cplx(B->dtype)) return arg_fail(1, "row swaps of complex element types are not built");
  if (npiv < 0 || npiv > B->m) return arg_fail(3, "npiv must be in [0, rows of B]");
  if (!ipiv_dev && npiv > 0) return arg_fail(2, "ipiv is NULL");
  rc = ensure_init();
  if (rc) return rc;
  if (B->n == 0 || npiv == 0) return 0;
  TileView tb(*B);
  hipStream_t st = stream_of(s);
  for (int j = 0; j < tb.nt; ++j) {
    rc = B->dtype == DM_F64
             ? laswp_tiled<double>((double*)tb.ptr(0, j), B->m, tb.cols(j), B->mb, ipiv_dev, npiv,
                                   0, nullptr, st)
             : laswp_tiled<float>((float*)tb.ptr(0, j), B->m, tb.cols(j), B->mb, ipiv_dev, npiv,
                                  0, nullptr, st);
    if (rc) return rc;
  }
  return 0;
}

int dm_dag_build_cholesky(dm_dag_t* dag, char uplo, const dm_tiled* A, int* info_dev) {
  if (!dag) return arg_fail(1, "dag is NULL");
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(2, "uplo must be 'U' or 'L'");
  int rc = check_tiled(A, 3);
  if (rc) return rc;
  if (A->m != A->n || A->mb != A->nb) return arg_fail(3, "A must be square with square tiles");
  std::unique_ptr<Dag> d(new Dag());
  TileView tv(*A);
  build_cholesky(*d, tv, is_upper(uplo) ? 'U' : 'L', info_dev);
  *dag = new dm_dag_s{std::move(d)};
  return 0;
}

int dm_dag_build_potrs(dm_dag_t* dag, char uplo, const dm_tiled* A, const dm_tiled* B) {
  if (!dag) return arg_fail(1, "dag is NULL");
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(2, "uplo must be 'U' or 'L'");
  int rc = check_tiled(A, 3);
  if (rc) return rc;
  rc = check_tiled(B, 4);
  if (rc) return rc;
  if (A->m != A->n || A->mb != A->nb) return arg_fail(3, "A must be square with square tiles");
  if (B->mb != A->mb || B->nb != A->nb || B->m != A->m || B->dtype != A->dtype)
    return arg_fail(4, "B does not conform to A");
  std::unique_ptr<Dag> d(new Dag());
  TileView ta(*A), tb(*B);
  const char tc = is_cplx(A->dtype) ? 'C' : 'T';
  if (is_upper(uplo)) {
    build_trsm(*d, 'L', 'U', tc, 'N', 1.0, ta, tb);
    build_trsm(*d, 'L', 'U', 'N', 'N', 1.0, ta, tb);
  } else {
    build_trsm(*d, 'L', 'L', 'N', 'N', 1.0, ta, tb);
    build_trsm(*d, 'L', 'L', tc, 'N', 1.0, ta, tb);
  }
  *dag = new dm_dag_s{std::move(d)};
  return 0;
}

int dm_plan_cache_clear(void) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  g_cache.clear();
  return 0;
}

int dm_forget_matrix(const void* base) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  for (auto it = g_cache.begin(); it != g_cache.end();) {
    bool hit = false;
    for (const void* b : it->second.bases) hit = hit || (b == base);
    if (hit) it = g_cache.erase(it);
    else ++it;
  }
  return 0;
}

}  // extern "C"

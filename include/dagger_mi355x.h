/*
 * dagger_mi355x.h — C ABI of the MI355X (gfx950) backend for Dagger.jl's
 * tiled dense linear-algebra path: cholesky!/ldiv! on a Blocks-partitioned
 * DArray, driven by spawn_datadeps.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an int status:
 *     0      success
 *    <0      argument error, LAPACK xerbla style: -i names the i-th argument
 *    >0      HIP runtime error code (hipError_t value) + DM_HIP_ERROR_BASE
 * dm_last_error_string() returns a message for the last failure on the
 * calling host thread.
 *
 * Matrices are column-major with explicit leading dimensions, exactly as the
 * BLAS/LAPACK calls they replace.  Device pointers come from dm_malloc (or any
 * hipMalloc'd memory of the current device).  Every compute entry point is
 * asynchronous on the given stream (NULL = the library's default stream).
 * Calls on different streams from different host threads are safe; calls on
 * one stream are ordered by HIP.  D->H copies synchronize their stream.
 *
 * Reference interfaces replaced (paths are inside JuliaParallel/Dagger.jl):
 *   tile ops   <- the BLAS/LAPACK function objects that ext/ROCExt.jl:358-377
 *                 remaps for a device processor, called from
 *                 src/array/cholesky.jl:39-50,84-95 and src/array/trsm.jl:17-183
 *   executor   <- Dagger.execute!(proc, f, args...) (src/processor.jl:40,
 *                 ext/ROCExt.jl:338-356) and the datadeps region runner
 *                 spawn_datadeps / distribute_tasks! (src/datadeps/queue.jl:64-471)
 *   memory     <- move/move!/unsafe_free!/allocate_array_func for a device
 *                 memory space (ext/ROCExt.jl:74-77,169-330, src/memory-spaces.jl:660-671)
 *   streams    <- gpu_synchronize / with_context! (src/gpu.jl:88-120)
 */
#ifndef DAGGER_MI355X_H
#define DAGGER_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DM_HIP_ERROR_BASE 1000

typedef struct dm_stream_s* dm_stream_t;
typedef struct dm_event_s* dm_event_t;
typedef struct dm_graph_s* dm_graph_t;
typedef struct dm_dag_s* dm_dag_t;

/* element types.  Complex elements are (re, im) pairs, Julia's ComplexF64 /
 * ComplexF32 layout: a complex m x n column-major array is bit-identical to
 * a real 2m x n one, and a Blocks(mb, nb) tiling to Blocks(2mb, nb). */
enum { DM_F64 = 0, DM_F32 = 1, DM_C64 = 2, DM_C32 = 3 };

/* ---------------------------------------------------------------- runtime */

/* Select and initialise a device.  Replaces AMDGPU.device!/with_context!
 * inside ROCExt's execute! (ext/ROCExt.jl:338-356) and gpu_synchronize
 * (src/gpu.jl:88-117).  Idempotent. */
int dm_init(int device);
int dm_device_count(int* count);
const char* dm_last_error_string(void);
/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int dm_version(void);

/* Device memory.  Replaces ROCArray allocation and unsafe_free!
 * (ext/ROCExt.jl:74-77,169-179; src/memory-spaces.jl:660-671). */
int dm_malloc(void** ptr, size_t bytes);
int dm_free(void* ptr);
/* Async copies on a stream.  h2d/d2d return after enqueueing; d2h
 * synchronizes the stream before returning (move! contract,
 * ext/ROCExt.jl:181-192). */
int dm_memcpy_h2d(void* dst, const void* src, size_t bytes, dm_stream_t s);
int dm_memcpy_d2h(void* dst, const void* src, size_t bytes, dm_stream_t s);
int dm_memcpy_d2d(void* dst, const void* src, size_t bytes, dm_stream_t s);
/* Strided 2-D copies (column-major panels): width_bytes per column, height
 * columns.  Used to scatter a host matrix into resident tiles and back
 * (move(CPUProc, DevProc, ::Array), ext/ROCExt.jl:209-272). */
int dm_memcpy2d_h2d(void* dst, size_t dpitch, const void* src, size_t spitch,
                    size_t width_bytes, size_t height, dm_stream_t s);
int dm_memcpy2d_d2h(void* dst, size_t dpitch, const void* src, size_t spitch,
                    size_t width_bytes, size_t height, dm_stream_t s);
int dm_memcpy2d_d2d(void* dst, size_t dpitch, const void* src, size_t spitch,
                    size_t width_bytes, size_t height, dm_stream_t s);
/* Fill an int32 device buffer with a value (a kernel, so it is safe inside
 * graph capture). */
int dm_fill_i32(int* dst, int value, size_t count, dm_stream_t s);

/* Streams, events, graphs (ROCExt keeps one HIP stream per device,
 * ext/ROCExt.jl:93-97,532-553; this backend uses several). */
int dm_stream_create(dm_stream_t* s);
/* Also releases the executor streams that plans issued from s shared
 * (eager plans fork onto a pool of streams kept per caller stream). */
int dm_stream_destroy(dm_stream_t s);
int dm_stream_sync(dm_stream_t s);
int dm_device_sync(void);
int dm_event_create(dm_event_t* e);
int dm_event_destroy(dm_event_t e);
int dm_event_record(dm_event_t e, dm_stream_t s);
int dm_stream_wait_event(dm_stream_t s, dm_event_t e);
int dm_event_sync(dm_event_t e);
int dm_event_elapsed_ms(float* ms, dm_event_t start, dm_event_t stop);
/* Stream capture into a hipGraph; launch replays it on a stream. */
int dm_graph_begin(dm_stream_t s);
int dm_graph_end(dm_stream_t s, dm_graph_t* g);
int dm_graph_launch(dm_graph_t g, dm_stream_t s);
int dm_graph_destroy(dm_graph_t g);

/* ------------------------------------------------------------- tile ops
 * BLAS/LAPACK argument order and meaning, column-major, device pointers.
 * These are what `Dagger.move(::CPUProc, ::DevProc, ::typeof(BLAS.gemm!))`
 * and friends return for this backend (ext/ROCExt.jl:358-377).  Characters
 * are 'N','T','C' (C == T for real types), 'U','L', 'L','R', 'N','U'. */

/* LAPACK potrf: A = U^T U ('U') or L L^T ('L').  *info_dev (device int)
 * receives the LAPACK info: 0, or k>0 if the leading minor of order k is not
 * positive definite.  potrf_checked! (src/array/cholesky.jl:2-18) reads it. */
int dm_dpotrf(char uplo, int n, double* A, int lda, int* info_dev, dm_stream_t s);
int dm_spotrf(char uplo, int n, float* A, int lda, int* info_dev, dm_stream_t s);

/* BLAS trsm: B := alpha * op(A)^-1 B (side 'L') or alpha * B op(A)^-1 ('R').
 * Called at src/array/cholesky.jl:41,86 and src/array/trsm.jl:95-183. */
int dm_dtrsm(char side, char uplo, char transa, char diag, int m, int n,
             double alpha, const double* A, int lda, double* B, int ldb,
             dm_stream_t s);
int dm_strsm(char side, char uplo, char transa, char diag, int m, int n,
             float alpha, const float* A, int lda, float* B, int ldb,
             dm_stream_t s);
/* dm_dtrsm that allocates nothing (no hipMalloc, hipFree or allocating
 * library call on any call).  No counterpart in the reference: it is for
 * callers that cannot allocate -- work under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation.  Same arguments, results
 * bitwise equal to dm_dtrsm.  Its scratch (the inverted 32x32 diagonal
 * blocks of one leaf solve, 64 KiB) is a static per-GPU device array, held
 * until the process ends; its size does not depend on m or n (longer solves
 * recurse through the same buffer), so no size is rejected for it.  Every
 * call on one GPU shares that array: calls that overlap in time on one GPU
 * (on different streams, or from different threads) race and must not be
 * issued; calls ordered on one stream are safe.  Each call fully writes the
 * part of the scratch it reads, so nothing carries over between calls. */
int dm_dtrsm_noalloc(char side, char uplo, char transa, char diag, int m, int n,
                     double alpha, const double* A, int lda, double* B, int ldb,
                     dm_stream_t s);

/* BLAS syrk: C := alpha op(A) op(A)^T + beta C on the uplo triangle only.
 * trans 'T': op(A) = A^T (A is k x n); 'N': op(A) = A (A is n x k).
 * Called at src/array/cholesky.jl:47,92. */
int dm_dsyrk(char uplo, char trans, int n, int k, double alpha, const double* A,
             int lda, double beta, double* C, int ldc, dm_stream_t s);
int dm_ssyrk(char uplo, char trans, int n, int k, float alpha, const float* A,
             int lda, float beta, float* C, int ldc, dm_stream_t s);

/* BLAS gemm: C := alpha op(A) op(B) + beta C.  beta == 0 does not read C.
 * Called at src/array/cholesky.jl:50,95 and src/array/trsm.jl:99-183. */
int dm_dgemm(char transa, char transb, int m, int n, int k, double alpha,
             const double* A, int lda, const double* B, int ldb, double beta,
             double* C, int ldc, dm_stream_t s);
int dm_sgemm(char transa, char transb, int m, int n, int k, float alpha,
             const float* A, int lda, const float* B, int ldb, float beta,
             float* C, int ldc, dm_stream_t s);
/* Complex tile ops (ComplexF64: z, ComplexF32: c).  A, B, C point to
 * interleaved (re, im) pairs; sizes and leading dimensions count complex
 * elements.  The BLAS/LAPACK calls _chol! and trsm! make for T <: Complex
 * (src/array/cholesky.jl:40-50,85-95; src/array/trsm.jl:90-183):
 *   potrf: A = U^H U ('U') or L L^H ('L'), LAPACK info as dm_dpotrf; the
 *          imaginary part of A's diagonal is ignored and zeroed;
 *   trsm : transa 'N', 'T' or 'C' (conjugate transpose); alpha points to a
 *          host complex scalar {re, im};
 *   herk : C := alpha op(A) op(A)^H + beta C on the uplo triangle, trans
 *          'N' (op(A) = A) or 'C' (op(A) = A^H), real alpha and beta, the
 *          imaginary part of C's diagonal set to zero (BLAS herk!);
 *   gemm : C := alpha op(A) op(B) + beta C, op 'N', 'T' or 'C'; alpha and
 *          beta point to host complex scalars. */
int dm_zpotrf(char uplo, int n, void* A, int lda, int* info_dev, dm_stream_t s);
int dm_cpotrf(char uplo, int n, void* A, int lda, int* info_dev, dm_stream_t s);
int dm_ztrsm(char side, char uplo, char transa, char diag, int m, int n,
             const double* alpha, const void* A, int lda, void* B, int ldb, dm_stream_t s);
int dm_ctrsm(char side, char uplo, char transa, char diag, int m, int n,
             const float* alpha, const void* A, int lda, void* B, int ldb, dm_stream_t s);
int dm_zherk(char uplo, char trans, int n, int k, double alpha, const void* A, int lda,
             double beta, void* C, int ldc, dm_stream_t s);
int dm_cherk(char uplo, char trans, int n, int k, float alpha, const void* A, int lda,
             float beta, void* C, int ldc, dm_stream_t s);
int dm_zgemm(char transa, char transb, int m, int n, int k, const double* alpha,
             const void* A, int lda, const void* B, int ldb, const double* beta, void* C,
             int ldc, dm_stream_t s);
int dm_cgemm(char transa, char transb, int m, int n, int k, const float* alpha,
             const void* A, int lda, const void* B, int ldb, const float* beta, void* C,
             int ldc, dm_stream_t s);

/* FP32 GEMM on the BF16 matrix cores: every FP32 operand is split into
 * `terms` == 6: three BF16 parts (products hh,hm,mh,hl,lh,mm) or
 * `terms` == 3: two parts (hh,hm,mh), accumulated in FP32.  No reference
 * counterpart: the low-precision off-diagonal update of the mixed-precision
 * factorization (SYRK/GEMM of src/array/cholesky.jl:47,50 on FP32 tiles).
 * Requires transa = 'T', transb = 'N', lda and ldb multiples of 4 and
 * 16-byte aligned A and B. */
int dm_sgemm_bf16x(int terms, char transa, char transb, int m, int n, int k,
                   float alpha, const float* A, int lda, const float* B, int ldb,
                   float beta, float* C, int ldc, dm_stream_t stream);

/* BLAS trsv / gemv (vector right-hand sides), src/array/trsm.jl:17-51. */
int dm_dtrsv(char uplo, char trans, char diag, int n, const double* A, int lda,
             double* x, int incx, dm_stream_t s);
int dm_dgemv(char trans, int m, int n, double alpha, const double* A, int lda,
             const double* x, int incx, double beta, double* y, int incy,
             dm_stream_t s);

/* ------------------------------------------------------- datadeps executor
 * A static task graph in the datadeps model (src/datadeps/queue.jl:44-54):
 * tasks are added in submission order with In/Out/InOut access modes on
 * their array arguments; overlapping accesses are ordered exactly as
 * sequential execution would order them (In after earlier Out/InOut;
 * Out/InOut after every earlier overlapping access), In/In never
 * synchronize.  dm_dag_plan groups independent ready tasks of one kernel
 * shape into batched launches, maps the groups onto HIP streams with event
 * dependencies, and can capture the whole schedule into a hipGraph that
 * dm_dag_run replays.  Replaces the Julia-side planner + Sch execution of a
 * device datadeps region (src/datadeps/queue.jl:120-471,
 * src/sch/Sch.jl:1788-2067) and the per-region tile copies
 * (src/datadeps/aliasing.jl:920-1060): tiles stay resident. */

enum { DM_OP_POTRF = 1, DM_OP_TRSM = 2, DM_OP_SYRK = 3, DM_OP_GEMM = 4,
       DM_OP_TRSV = 5, DM_OP_GEMV = 6, DM_OP_GETRF = 7, DM_OP_LASWP = 8 };
enum { DM_IN = 1, DM_OUT = 2, DM_INOUT = 3 };

/* One tile task.  Operand slots by op:
 *   POTRF: c = A (InOut), info = per-task device info int (Out)
 *   TRSM : a = A, c = B          GEMM: a = A, b = B, c = C
 *   SYRK : a = A, c = C          GEMV: a = A, b = x, c = y
 *   TRSV : a = A, c = x
 *   GETRF: the LU panel of a tiled block column (src/array/lu.jl:160-204):
 *          c = its first tile (the panel: m rows in tiles of k rows, n
 *          columns, tiles packed below each other), b = int pivots (Out,
 *          global 1-based rows, ldb = global row of the panel's first
 *          row), info = int (InOut: the first zero pivot, global, 1-based),
 *          transa 'P' with partial pivoting, 'N' without
 *   LASWP: a block column's row swaps (swaprows_trail!, src/array/lu.jl:
 *          239-262): c = its tile at the panel's row (m rows in tiles of k,
 *          n columns), b = the panel's pivots (In), lda = their count,
 *          ldb = global row of c's first row
 * m,n,k follow the BLAS call (TRSM: m,n of B; SYRK: n,k; POTRF: n in n).
 * access[i] is DM_IN/DM_OUT/DM_INOUT for slots a,b,c (0 = unused). */
typedef struct dm_task {
    int op;
    int dtype;
    char side, uplo, transa, transb, diag, pad_[3];
    int m, n, k;
    double alpha, beta;
    const void* a; int lda;
    const void* b; int ldb;
    void* c; int ldc;
    int* info;
    int access[3];
} dm_task;

typedef struct dm_plan_opts {
    int num_streams;     /* streams to spread groups over (>=1) */
    int batch;           /* 1: batch independent same-shape tasks per launch */
    int lookahead;       /* slack (in levels) below which tasks go on the
                            critical stream; 0 disables the split */
    int use_graph;       /* 1: capture the schedule into a hipGraph (plans
                            with at most 2 streams and at most 2e11 flops;
                            other plans issue eagerly, and
                            dm_dag_stats.graph reports 0) */
    int profile;         /* 1 (eager runs only): time every launch group
                            with HIP events on its own stream */
    int f32_update;      /* FP32 GEMM/SYRK arithmetic: 0 FP32 matrix cores;
                            6 or 3: BF16 matrix cores with 3- or 2-term
                            operand splitting (6 or 3 products, FP32
                            accumulation) where operands are k-contiguous
                            (the upper variant), FP32 cores elsewhere */
    int f64_diag;        /* 1: FP32 plans factor their diagonal tiles in
                            FP64 (the tile converted to an FP64 workspace of
                            the plan, FP64 POTRF, the factor rounded back to
                            FP32) -- the precision form's FP64 diagonal
                            POTRF; FP64 plans ignore it */
} dm_plan_opts;

/* Per-launch-group timing of the last profiled run (dm_plan_opts.profile):
 * the hipEvent time of the group's launches on its stream and the
 * algorithmic flops of its tasks (GEMM 2mnk, SYRK n^2 k, TRSM m n ka,
 * POTRF n^3/3).  A batched TRSM group reports one record per launch: its
 * leaf solves (op TRSM, L^2 flops per vector) and the GEMM updates between
 * them (op GEMM, 2mnk, with their tile configuration).  The event-based counterpart of Dagger's timespan logging
 * (src/utils/logging.jl:26-143) for device work. */
typedef struct dm_group_profile {
    int op;        /* DM_OP_* of the group's tasks (SYRK groups report GEMM) */
    int ntasks;
    int stream;
    int level;
    double flops;
    float ms;
    int variant;   /* GEMM groups: the tile configuration launched, encoded
                      as id*1000000 + BM*1000 + BN (e.g. 4256128: FP64
                      256x128; 3128128: FP32 via BF16x6 128x128); -1 for
                      other groups */
} dm_group_profile;

typedef struct dm_dag_stats {
    int num_tasks, num_edges, num_groups, num_levels, num_launches;
    int num_streams, graph;
} dm_dag_stats;

int dm_dag_create(dm_dag_t* dag);
int dm_dag_destroy(dm_dag_t dag);
int dm_dag_add_task(dm_dag_t dag, const dm_task* t);
/* Dependency edges after planning, as (from, to) task-index pairs;
 * *count is the capacity in pairs on input, the number of edges on output. */
int dm_dag_edges(dm_dag_t dag, int* pairs, int* count);
int dm_dag_plan(dm_dag_t dag, const dm_plan_opts* opts);
/* Runs the planned DAG ordered after prior work on `s`; completes on `s`.
 * abort_dev (may be NULL): device int; every launch becomes a no-op once it
 * is non-zero (set by a failing POTRF), mirroring dependent tasks failing
 * after PosDefException in the reference. */
int dm_dag_run(dm_dag_t dag, dm_stream_t s);
int dm_dag_get_stats(dm_dag_t dag, dm_dag_stats* st);
/* Records of the last profiled run; *count: capacity in, records out. */
int dm_dag_profile(dm_dag_t dag, dm_group_profile* out, int* count);

/* ------------------------------------------------- whole-matrix forms
 * A tiled matrix resident in HBM: Blocks(mb, nb) partition of an m x n
 * matrix (src/array/alloc.jl:23-33), tile (i,j) is a dense column-major
 * array with ld = its row count (Dagger's per-chunk Array), tiles packed in
 * column-major tile order in one allocation starting at base:
 *   offset(i,j) = j*nb*m + i*mb*cols(j)   (elements). */
typedef struct dm_tiled {
    void* base;
    int64_t m, n;
    int mb, nb;
    int dtype;
} dm_tiled;

/* Tiled Cholesky, the exact task DAG of LinearAlgebra._chol!
 * (src/array/cholesky.jl:19-108) for uplo 'U' or 'L', run through the
 * datadeps executor with a cached plan/graph per (matrix, options).
 * info_dev: device int receiving the reference's info: the LAPACK info of
 * the first failing diagonal tile (tile-local, as potrf_checked! fills it),
 * 0 on success.  global_info_dev (may be NULL): the global order of the
 * failing leading minor. */
int dm_cholesky_tiled(char uplo, const dm_tiled* A, int* info_dev,
                      int* global_info_dev, const dm_plan_opts* opts,
                      dm_stream_t s);

/* ishermitian(::DArray) (src/array/linalg.jl:90-124), the pre-check the
 * stdlib cholesky! runs: *result_dev = 1 when every diagonal tile is exactly
 * symmetric and every off-diagonal pair satisfies ||A_ij - A_ji^T||_F <=
 * sqrt(eps) max(||A_ij||_F, ||A_ji||_F) (src/array/linalg.jl:38-48), else 0.
 * work_dev: mt*(mt+1)/2 ints. */
int dm_ishermitian_tiled(const dm_tiled* A, int* result_dev, int* work_dev, dm_stream_t s);

/* LinearAlgebra.norm2 of a DArray (src/array/linalg.jl:1-5), of its
 * UpperTriangular view (:7-19) or LowerTriangular view (:21-33), selected by
 * part = 0 / 1 / 2; and the finiteness test of LAPACK.chkfinite(::DArray)
 * (src/array/linalg.jl:127-148) over the same elements, in one HBM pass.
 * *out_dev = sqrt of the sum of squares (accumulated in FP64 in a fixed
 * order), *nonfinite_dev = 1 when an element is Inf or NaN, else 0.
 * FP64 or FP32.  work_dev: 8 * mt * nt doubles. */
int dm_tiled_norm2(const dm_tiled* A, int part, double* out_dev, int* nonfinite_dev,
                   double* work_dev, dm_stream_t s);

/* ldiv!(C::Cholesky, B) for a matrix right-hand side (src/array/linalg.jl:
 * 385-401): two tiled trsm! passes (src/array/trsm.jl:59-192) over the
 * resident factor.  B is tiled with Blocks(nb, nb) rows matching A. */
int dm_potrs_tiled(char uplo, const dm_tiled* A, const dm_tiled* B,
                   const dm_plan_opts* opts, dm_stream_t s);

/* Tiled LU: lu!(A, RowMaximum()) (pivot = 1, src/array/lu.jl:289-371) or
 * lu!(A, NoPivot()) (pivot = 0, src/array/lu.jl:10-40) of an m x n tiled
 * matrix with square blocks, the reference's task DAG (panel factorization,
 * row swaps, trsm!, gemm!) run through the executor with a cached plan.
 * ipiv_dev (min(m, n) ints): the reference's ipiv, global 1-based rows
 * (chunk k of Blocks(nb) holds panel k's pivots; NoPivot: 1..min(m, n)).
 * info_dev: 0, or the global 1-based column of the first exactly zero
 * pivot (LAPACK getrf's info; the factorization still completes). */
int dm_getrf_tiled(int pivot, const dm_tiled* A, int* ipiv_dev, int* info_dev,
                   const dm_plan_opts* opts, dm_stream_t s);
/* Rows of a tiled B permuted by the swaps ipiv[0..npiv) (global 1-based
 * rows) applied in order: _apply_ipiv_rows! of ldiv!(::LU, B)
 * (src/array/linalg.jl:266-310). */
int dm_tiled_laswp(const dm_tiled* B, const int* ipiv_dev, int npiv, dm_stream_t s);

/* Record (unplanned) the task DAG of the tiled Cholesky / ldiv! on resident
 * tiles into a new dag handle -- the same tasks dm_cholesky_tiled and
 * dm_potrs_tiled run -- for callers that plan, profile or replay it
 * themselves. */
int dm_dag_build_cholesky(dm_dag_t* dag, char uplo, const dm_tiled* A, int* info_dev);
int dm_dag_build_potrs(dm_dag_t* dag, char uplo, const dm_tiled* A, const dm_tiled* B);

/* Tiled trsm!, one datadeps region (src/array/trsm.jl:59-192). */
int dm_trsm_tiled(char side, char uplo, char trans, char diag, double alpha,
                  const dm_tiled* A, const dm_tiled* B,
                  const dm_plan_opts* opts, dm_stream_t s);

/* Tiled C = alpha op(A) op(B) + beta C over resident tiles of one square
 * block size (one GEMM task per tile triple, beta at the first k-tile): the
 * tile loop of Dagger's matmul (src/array/mul.jl), here used for the FP64
 * residual of iterative refinement.  Cached plan per operands. */
int dm_gemm_tiled(char transa, char transb, double alpha, const dm_tiled* A,
                  const dm_tiled* B, double beta, const dm_tiled* C,
                  const dm_plan_opts* opts, dm_stream_t s);

/* ---------------------------------------------------- mixed precision
 * Building blocks of the low-precision-factor + FP64-refinement solve
 * (dsposv structure; no reference counterpart -- the precision form beside
 * the plain drop-in). */
/* dst = (dst dtype) src, elementwise over the packed slab (same shape and
 * blocks, FP64 <-> FP32). */
int dm_tiled_convert(const dm_tiled* dst, const dm_tiled* src, dm_stream_t s);
/* *out_dev = ||A||_inf (max row sum, FP64); work_dev: m * (1 + ceil(n/nb))
 * doubles. */
int dm_tiled_norm_inf(const dm_tiled* A, double* out_dev, double* work_dev, dm_stream_t s);
/* out_dev[j] = max_i |X_ij| for each column j (FP64). */
int dm_tiled_col_norm_inf(const dm_tiled* X, double* out_dev, dm_stream_t s);
/* out_dev[j] = sum_i X_ij Y_ij for each column j (FP64, X and Y of the same
 * shape and blocks; deterministic order).  Column-blocked GMRES-IR: the
 * Gram-Schmidt coefficients and norms of n right-hand sides at once. */
int dm_tiled_col_dot(const dm_tiled* X, const dm_tiled* Y, double* out_dev, dm_stream_t s);
/* Y[:, j] = (scale * a_dev[j]) X[:, j] + beta Y[:, j] for each column j
 * (FP64, same shape and blocks; a_dev NULL: a_dev[j] = 1; beta = 0: Y is
 * not read; X may be Y; at most 65535 columns).  Column-blocked GMRES-IR:
 * orthogonalization, normalization and the solution update. */
int dm_tiled_col_axpby(const double* a_dev, double scale, const dm_tiled* X, double beta,
                       const dm_tiled* Y, dm_stream_t s);
/* A[i,i] *= factor on a square tiled matrix with square blocks (FP64 or
 * FP32): the scaled diagonal shift A + c u diag(A) that keeps a
 * low-precision copy of an ill-conditioned SPD matrix positive definite
 * (Higham & Pranesh, SISC 2021). */
int dm_tiled_scale_diag(const dm_tiled* A, double factor, dm_stream_t s);
/* X += D with X FP64 and D FP32 (same shape and blocks). */
int dm_tiled_axpy_s2d(const dm_tiled* X, const dm_tiled* D, dm_stream_t s);

/* Full column-major matrix (leading dimension ld) <-> tile-packed layout.
 * pack also adds diag_add to the diagonal (0 for a plain copy).  Device
 * side of copyto!(DA, A)/collect and of the re-blocking copy in
 * maybe_copy_buffered (src/array/copy.jl:3-151). */
int dm_tiled_pack(const dm_tiled* dst, const void* src, int64_t ld, double diag_add,
                  dm_stream_t s);
int dm_tiled_unpack(void* dst, int64_t ld, const dm_tiled* src, dm_stream_t s);
/* A[i,i] += d[i] on a full column-major FP64 matrix (input generators). */
int dm_add_diag_d(double* A, int64_t ld, int64_t n, const double* d_dev, dm_stream_t s);
/* Every tile of a one-tile-column matrix D (m x nb, Blocks(nb, nb)) set to
 * [I_r 0] with r its row count (the start of an inverse of diagonal tiles;
 * no reference counterpart). */
int dm_tiled_identity(const dm_tiled* D, dm_stream_t s);

/* Drop every cached plan/graph (call before freeing matrices whose plans
 * were cached, or rely on dm_forget_matrix for one base pointer). */
int dm_plan_cache_clear(void);
int dm_forget_matrix(const void* base);

/* ------------------------------------------------------------- batched
 * Many uneven SPD systems in one persistent launch (no reference
 * counterpart: the batching form beside the plain drop-in).  System m:
 * A_m = A_base + a_off (n x n, column-major, ld = n), B_m = B_base + b_off
 * (n x nrhs, ld = n, nrhs <= 8), panel width P in [1, 64] (autotuned per size
 * bucket by the host).  Factored A = U^T U in place (upper triangle) and,
 * unless factor_only, B overwritten with the solution.  items_host: the
 * num_items systems; order_host: the `count` systems to process, as indices
 * into items_host (each in [0, num_items), processing order: largest first
 * is best).  info_dev[m] (num_items ints): LAPACK info of processed system m.
 * workgroups: persistent grid (0 = one 512-thread workgroup per CU). */
typedef struct dm_batch_item {
    int64_t a_off, b_off;
    int n, nrhs, panel;
    int lda;       /* A_m's leading dimension, >= n (0: n, the packed layout) */
} dm_batch_item;
int dm_batched_posv(int dtype, int num_items, void* A_base, void* B_base,
                    const void* items_host, int count, const int* order_host, int* info_dev,
                    int factor_only, int workgroups, dm_stream_t s);

/* The same factorization and solve for the larger systems of a batch
 * (no reference counterpart), tiled and level-synchronous: every listed
 * system advances through one blocked right-looking Cholesky in lockstep,
 * each step one launch over all systems still active at it -- per 64-column
 * block the diagonal block (one wave per system), its panel row over the
 * system's full width and the update of the rows left in the current
 * outer tile of nb columns; after each outer tile the trailing update with
 * K = nb as one batched MFMA GEMM -- then both triangular sweeps by 64-row
 * blocks (batched in-register solves of the diagonal blocks and batched
 * GEMM updates) on a working copy of the right-hand sides.  Items and order
 * as for dm_batched_posv (each item's `panel` is ignored); nb a multiple of
 * 64 in [64, 1024]: the solves' tile, and the factorization's outer tile
 * of every item whose item_nb entry is 0 (item_nb: num_items entries, or
 * NULL; a nonzero entry, a multiple of 64 in [64, 1024], is that system's
 * outer tile -- the autotuner picks it per size bucket).  The plan holds its job tables and an n x nrhs working
 * slab per system; run it any number of times (on any stream, one run at a
 * time), each run factoring A in place and, unless factor_only, overwriting
 * B_m with the solution of every system whose info is 0 (a failed system's
 * B_m is left as it was).  info_dev: as for dm_batched_posv. */
typedef struct dm_batch_plan_s* dm_batch_plan_t;
int dm_batched_tiled_create(dm_batch_plan_t* plan, int dtype, int num_items, void* A_base,
                            void* B_base, const void* items_host, int count,
                            const int* order_host, int nb, const int* item_nb);
int dm_batched_tiled_run(dm_batch_plan_t plan, int* info_dev, int factor_only, dm_stream_t s);
int dm_batched_tiled_destroy(dm_batch_plan_t plan);

/* ---------------------------------------------------------- generators
 * Seeded counter-based normal variates: Philox4x32-10 keyed by `seed`,
 * counter = element index, Box-Muller on the first two 32-bit outputs.
 * Element i of the result depends only on (seed, offset + i). */
int dm_fill_randn_d(double* x, size_t count, uint64_t seed, uint64_t offset,
                    dm_stream_t s);

#ifdef __cplusplus
}
#endif
#endif /* DAGGER_MI355X_H */

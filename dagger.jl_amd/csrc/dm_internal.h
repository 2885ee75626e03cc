// Internal declarations shared by the HIP translation units of the
// MI355X backend.  Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>
#include "../../include/dagger_mi355x.h"

// The dynamic LDS buffer of a kernel (sized by the launch).
#ifndef DM_DYNAMIC_LDS
#define DM_DYNAMIC_LDS(name) extern __shared__ unsigned char name[]
#endif

// Hides a register value from the optimizer (no instruction is emitted).
#ifndef DM_OPAQUE
#define DM_OPAQUE(x) asm volatile("" : "+v"(x))
#endif

struct dm_stream_s { hipStream_t h; };
struct dm_event_s { hipEvent_t h; };
struct dm_graph_s { hipGraph_t g; hipGraphExec_t exec; };

namespace dm {

// ------------------------------------------------------------ error state
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int arg_fail(int argno, const char* what);

#define DM_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) return ::dm::hip_fail(e_, #call);                 \
  } while (0)

// Launch-error check after a kernel launch.
#define DM_LAUNCH_CHECK(name)                                               \
  do {                                                                      \
    hipError_t e_ = hipGetLastError();                                      \
    if (e_ != hipSuccess) return ::dm::hip_fail(e_, name);                  \
  } while (0)

int ensure_init();
hipStream_t stream_of(dm_stream_t s);

// Runs fn on a helper thread with a 1 GiB stack and returns its result
// (graph capture/instantiate/launch: the HIP runtime recurses over graph
// dependency chains).  Calls are serialized.
int run_big_stack(std::function<int()> fn);

// Per-stream scratch buffer (grown on demand, never shrunk while the
// library is loaded).  Calls ordered on one stream may reuse it.
void* stream_workspace(hipStream_t s, size_t bytes);

// Executor streams of eagerly issued plans, shared by every plan issued
// from one caller stream: the first `count` streams of that caller's pool
// (created on first use; at most 8).  Plans therefore add no streams of
// their own, and a process's streams stay few beside the hardware queues
// they are multiplexed onto.  Released with the caller stream
// (dm_stream_destroy); the default stream's pool lives until unload.
int exec_streams(hipStream_t caller, int count, hipStream_t* out);
void release_exec_streams(hipStream_t caller);

inline bool is_cplx(int dtype) { return dtype == DM_C64 || dtype == DM_C32; }
inline bool is_trans(char t) { return t == 'T' || t == 't' || t == 'C' || t == 'c'; }
inline bool is_notrans(char t) { return t == 'N' || t == 'n'; }
inline bool is_upper(char u) { return u == 'U' || u == 'u'; }
inline bool is_lower(char u) { return u == 'L' || u == 'l'; }
inline bool is_left(char s) { return s == 'L' || s == 'l'; }
inline bool is_right(char s) { return s == 'R' || s == 'r'; }
inline bool is_unit(char d) { return d == 'U' || d == 'u'; }
inline bool is_nonunit(char d) { return d == 'N' || d == 'n'; }

// ------------------------------------------------------------ GEMM family
// One C tile update of a (possibly batched) GEMM/SYRK launch:
//   C[0:M,0:N] = alpha * op(A) op(B) + beta * C   restricted to the
//   triangle `tri` (0 full, 1 upper m<=n, 2 lower m>=n).
template <typename T>
struct GemmJob {
  const T* A;
  const T* B;
  T* C;
  int M, N, K;
  int lda, ldb, ldc;
  T alpha, beta;
  int tri;
  int tiles_m, tiles_n;  // workgroup tiles along M and N (filled by host)
  int wg_begin;          // first workgroup of this job in a batched launch
  int prio;              // launch argument (`single`) only: 1 for critical-path
                         // launches, whose waves raise their issue priority
};

// Critical-path kernels (the diagonal POTRF, the panel TRSM and the next
// panel's update of a look-ahead schedule) share CUs with the bulk trailing
// update: raising their waves' instruction-issue priority (s_setprio, per
// wave; the HIP streams keep the default priority) lets their short
// dependent steps go ahead of the bulk's MFMA waves on the same SIMD.
__device__ __forceinline__ void raise_wave_priority() { __builtin_amdgcn_s_setprio(3); }

// Global-memory views of pointers that reach a kernel inside job structs.
// The compiler cannot prove such a pointer global and emits flat accesses,
// which count against lgkmcnt as well as vmcnt: the LDS wait in front of a
// k-step's MFMAs would then also wait for the next k-step's prefetch and
// expose its whole memory latency.  Through these views the accesses are
// global_load/global_store, tracked by vmcnt only.
#ifndef DM_GLOBAL_AS
#define DM_GLOBAL_AS __attribute__((address_space(1)))
#endif
template <typename P>
__device__ __forceinline__ const DM_GLOBAL_AS P* gptr(const P* p) {
  return (const DM_GLOBAL_AS P*)p;
}
template <typename P>
__device__ __forceinline__ DM_GLOBAL_AS P* gptr(P* p) {
  return (DM_GLOBAL_AS P*)p;
}

// Kernel configuration chosen by the host for a launch.
struct GemmCfgId {
  int id;     // index into the configuration table
  int bm, bn; // workgroup tile
};
template <typename T>
GemmCfgId gemm_pick_cfg(int M, int N, int K, int njobs, bool vec = true);

// Fills tiles_m/tiles_n/wg_begin for jobs under `cfg`; returns total WGs.
template <typename T>
long gemm_prepare_jobs(GemmJob<T>* jobs, int njobs, GemmCfgId cfg);

// Launches a batched GEMM over device-resident jobs (jobs_dev) or, when
// njobs == 0, over the single job passed by value.
// vec: every job satisfies gemm_job_vec_ok (2-wide operand loads).
template <typename T>
bool gemm_job_vec_ok(bool ta, bool tb, const GemmJob<T>& j);
// max_wg > 0 caps the grid (workgroups then loop over their tiles).
template <typename T>
int gemm_launch(bool ta, bool tb, bool vec, GemmCfgId cfg, const GemmJob<T>* jobs_dev,
                int njobs, const GemmJob<T>& single, long total_wg,
                const int* abort, hipStream_t s, long max_wg = 0);
// Workgroups of a GEMM configuration that fit on one CU at once.
int gemm_wgs_per_cu(GemmCfgId cfg);

// Convenience: one GEMM call (BLAS semantics incl. tri for SYRK).
template <typename T>
int gemm_single(bool ta, bool tb, int M, int N, int K, T alpha, const T* A,
                int lda, const T* B, int ldb, T beta, T* C, int ldc, int tri,
                const int* abort, hipStream_t s);

// FP32 updates on the BF16 matrix cores with operand splitting
// (gemm_split.hip).  GemmCfgId ids for FP32 launches; workgroup tiles of
// 256x128 for launches of at least about one tile per CU, else 128x128.
constexpr int GEMM_CFG_BF16X3 = 2, GEMM_CFG_BF16X6 = 3;
inline GemmCfgId gemm_split_cfg(int terms, int M, int N, int njobs) {
  const int id = terms == 6 ? GEMM_CFG_BF16X6 : GEMM_CFG_BF16X3;
  const long nj = njobs > 0 ? njobs : 1;
  const long t256 = (long)((M + 255) / 256) * ((N + 127) / 128) * nj;
  return t256 >= 256 ? GemmCfgId{id, 256, 128} : GemmCfgId{id, 128, 128};
}
// op(A) = A^T, op(B) = B (k-contiguous operands), 16-byte aligned rows.
bool gemm_split_ok(bool ta, bool tb, const GemmJob<float>& j);
int gemm_split_launch(int terms, int bm, const GemmJob<float>* jobs_dev, int njobs,
                      const GemmJob<float>& single, long total_wg, long grid,
                      const int* abort, hipStream_t s);

// ------------------------------------------------------------ TRSM family
// Effective left solve  Tm X = alpha V  on L-long vectors, NV of them.
//   Tm(i,k) = ttrans ? A[k + i*lda] : A[i + k*lda], lower or upper.
//   side L: V(i,j) = B[i + j*ldb];  side R: V(i,j) = B[j + i*ldb].
template <typename T>
struct TrsmJob {
  const T* A;
  T* B;
  const T* inv;  // inverted 32x32 diagonal blocks of Tm (device)
  int L, NV;
  int lda, ldb;
  T alpha;
  int wg_begin;
  int prio;  // launch argument (`single`) only: 1 for critical-path launches
};

struct TrsmShape {
  bool lower;   // Tm lower triangular
  bool ttrans;  // Tm read transposed from A
  bool side_l;  // vectors are columns of B
  bool unit;
};
TrsmShape trsm_shape(char side, char uplo, char transa, char diag);

constexpr int TRSM_BS = 32;     // diagonal block inverted up front
constexpr int TRSM_LEAF = 256;  // longest solve held in LDS (longer ones recurse)
int trsm_leaf_len();            // recursion leaf length in use (128)
constexpr int TRSM_VPW = 64;    // vectors per workgroup of the leaf kernel (16 per wave)
// Workspace (elements) for any trsm_full / potrf_full call: the inverted
// diagonal blocks of one leaf solve.
template <typename T>
constexpr size_t trsm_ws_elems() {
  return (size_t)(TRSM_LEAF / TRSM_BS) * TRSM_BS * TRSM_BS;
}

// Inverts the 32x32 diagonal blocks of Tm into inv (ceil(L/32)*1024 elems).
template <typename T>
int trinv_blocks_launch(const TrsmShape& sh, const T* A, int lda, int L, T* inv,
                        const int* abort, hipStream_t s, int prio = 0);
// Batched / single leaf solve (L <= TRSM_LEAF; jobs' inv point at the
// inverted blocks of their own triangle).  single.L: longest L of the launch.
template <typename T>
int trsm_launch(const TrsmShape& sh, const TrsmJob<T>* jobs_dev, int njobs,
                const TrsmJob<T>& single, long total_wg, const int* abort,
                hipStream_t s, long max_wg = 0);
int trsm_init_attributes();
inline int trsm_wgs(int NV) { return (NV + TRSM_VPW - 1) / TRSM_VPW; }

// Recursive batched solve: the launches (leaf solves and GEMM updates, in
// issue order) that solve every job of `jobs` (any lengths).  Split points
// are multiples of TRSM_BS from the top of each triangle, so a job's inverse
// blocks (computed once over its full triangle) serve every leaf.
template <typename T>
struct TrsmStep {
  bool leaf = true;
  std::vector<TrsmJob<T>> tjobs;  // leaf solves (wg_begin filled)
  std::vector<GemmJob<T>> gjobs;  // update C = alpha_j V2 - Tm21 X1
  bool ta = false, tb = false;    // update operand transposes
  int maxL = 0;
  long total_wg = 0;
};
// leaf: recursion leaf length (0: trsm_leaf_len())
template <typename T>
void trsm_plan_steps(const TrsmShape& sh, const std::vector<TrsmJob<T>>& jobs,
                     std::vector<TrsmStep<T>>& steps, int leaf = 0);

// Full BLAS trsm (any size; host recursion above TRSM_LEAF, inverse blocks
// formed per leaf in ws).
template <typename T>
int trsm_full(char side, char uplo, char transa, char diag, int m, int n,
              T alpha, const T* A, int lda, T* B, int ldb, T* ws,
              const int* abort, hipStream_t s);

// ------------------------------------------------------------ POTRF
// info: LAPACK info of this call (local order); abort (may be null): set to
// global_offset + local order on failure and checked by every launch.
// Trailing SYRK of the 64-wide POTRF panel at j0 fused with the
// factorization of the next diagonal block (gemm.hip).
template <typename T>
int potrf_update_launch(bool upper, int n, T* A, int lda, int j0, int* info, int* abort,
                        int global_base, hipStream_t s);
// C = beta C + alpha A X on tiled FP64 operands where X and C are one block
// column of at most 64 columns (mixed.hip): one launch, A read once.
bool tiled_skinny_gemm_ok(char ta, char tb, const dm_tiled& A, const dm_tiled& X,
                          const dm_tiled& Cm);
int tiled_skinny_gemm_launch(const dm_tiled& A, const dm_tiled& X, const dm_tiled& Cm,
                             double alpha, double beta, hipStream_t s);
template <typename T>
int potrf_full(bool upper, int n, T* A, int lda, int* info, int* abort,
               int global_offset, T* ws, hipStream_t s);
// FP64 POTRF of an FP32 tile through the FP64 workspace W (n x n elements):
// convert, potrf_full<double>, round the factor back (mixed.hip).
int potrf_f64_of_f32(bool upper, int n, float* A, int lda, double* W, int* info, int* abort,
                     int global_base, hipStream_t s);

// ------------------------------------------------------------ LU (lu.hip)
// Workspace bytes of one panel factorization (contiguous panel, TRSM
// inverse blocks, pivots).
template <typename T>
size_t getrf_panel_ws_bytes(int64_t Mp, int W);
// LU (partial pivoting when `pivot`) of a tiled panel of Mp rows in tiles of
// mb rows and W columns starting at `tiles`; ipiv: global 1-based pivot rows
// (row0 = the panel's first global row), *info: first zero pivot (global).
template <typename T>
int getrf_panel(T* tiles, int64_t Mp, int W, int mb, int64_t row0, int* ipiv, int* info,
                int* abort, bool pivot, void* ws, hipStream_t s);
// The swaps ipiv[0..npiv) (global 1-based rows) applied in order to the
// rows of a tiled block-column region whose first row is global row0.
template <typename T>
int laswp_tiled(T* base, int64_t Mr, int W, int mb, const int* ipiv, int npiv, int64_t row0,
                const int* abort, hipStream_t s);

// ------------------------------------------------------------ vectors
// y := alpha op(A) x + beta y (BLAS gemv; an empty inner dimension still
// scales y by beta).
template <typename T>
int gemv_launch(bool trans, int m, int n, T alpha, const T* A, int lda, const T* x, int incx,
                T beta, T* y, int incy, const int* abort, hipStream_t s);

int fill_i32_launch(int* dst, int value, size_t count, hipStream_t s);

}  // namespace dm

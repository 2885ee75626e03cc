// Stubs: HIP runtime (memory = malloc) and device launches (no-ops that
// validate pointers are in the plan's allocation where cheap).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <hip/hip_runtime.h>
#include "dm_internal.h"
#include "complex_ops.h"
#include "potrs.h"
static int g_streams = 0, g_events = 0;
// Optional trace of the stream operations (environment DM_HARNESS_TRACE =
// output path): one line per launch ("L stream"), event record ("R event
// stream"), stream wait ("W stream event") and capture begin/end, so a test
// can rebuild the graph a stream capture of the same issue sequence forms.
static FILE* trace_file() {
  static FILE* f = [] {
    const char* p = getenv("DM_HARNESS_TRACE");
    return p ? fopen(p, "w") : (FILE*)nullptr;
  }();
  return f;
}
static void trace_launch(hipStream_t s) {
  if (FILE* f = trace_file()) fprintf(f, "L %p\n", (void*)s);
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (hipStream_t)(uintptr_t)(0x1000 + ++g_streams); return 0; }
hipError_t hipGetDevice(int* d) { *d = 0; return 0; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) { *v = 256; return 0; }
hipError_t hipStreamDestroy(hipStream_t s) { if (!s) { fprintf(stderr, "destroy null stream\n"); abort(); } return 0; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = (hipEvent_t)(uintptr_t)(0x100000 + ++g_events); return 0; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t) { return 0; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  if (!e) abort();
  if (FILE* f = trace_file()) fprintf(f, "R %p %p\n", (void*)e, (void*)s);
  return 0;
}
hipError_t hipEventSynchronize(hipEvent_t) { return 0; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 1.f; return 0; }
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
  if (!e) abort();
  if (FILE* f = trace_file()) fprintf(f, "W %p %p\n", (void*)s, (void*)e);
  return 0;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (FILE* f = trace_file()) fprintf(f, "B %p\n", (void*)s);
  return 0;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  if (FILE* f = trace_file()) { fprintf(f, "E %p\n", (void*)s); fflush(f); }
  *g = (hipGraph_t)0x777;
  return 0;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t, hipGraphNode_t*, char*, size_t) { *e = (hipGraphExec_t)0x778; return 0; }
hipError_t hipGraphGetNodes(hipGraph_t, hipGraphNode_t*, size_t* n) { *n = 0; return 0; }
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return 0; }
hipError_t hipGraphDestroy(hipGraph_t) { return 0; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return 0; }
hipError_t hipMalloc(void** p, size_t n) { *p = malloc(n ? n : 1); return 0; }
hipError_t hipFree(void* p) { free(p); return 0; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { memcpy(d, s, n); return 0; }
hipError_t hipMemset(void* d, int v, size_t n) { memset(d, v, n); return 0; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) { memcpy(d, s, n); return 0; }
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { memset(d, v, n); return 0; }
hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
hipError_t hipDeviceSynchronize() { return 0; }
hipError_t hipGetLastError() { return 0; }
const char* hipGetErrorString(hipError_t) { return "stub"; }
namespace dm {
static std::string g_err;
void set_error(const std::string& m) { g_err = m; }
int hip_fail(hipError_t e, const char* w) { fprintf(stderr, "hip_fail %s\n", w); return 1000 + e; }
int arg_fail(int a, const char* w) { fprintf(stderr, "arg_fail %d %s\n", a, w); return a > 0 ? -a : -99; }
int ensure_init() { return 0; }
hipStream_t stream_of(dm_stream_t s) { return s ? s->h : (hipStream_t)0x42; }
void* stream_workspace(hipStream_t, size_t) { return nullptr; }
int exec_streams(hipStream_t, int count, hipStream_t* out) {
  for (int i = 0; i < count; ++i) out[i] = (hipStream_t)(uintptr_t)(0x100 + i);
  return 0;
}
void release_exec_streams(hipStream_t) {}
int trsm_init_attributes() { return 0; }
int run_big_stack(std::function<int()> fn) { return fn(); }
template <typename T> GemmCfgId gemm_pick_cfg(int, int, int, int, bool) { return GemmCfgId{0, 128, 128}; }
template <typename T> long gemm_prepare_jobs(GemmJob<T>* jobs, int nj, GemmCfgId c) {
  long t = 0; for (int i = 0; i < nj; ++i) { jobs[i].tiles_m = (jobs[i].M + 127) / 128; jobs[i].tiles_n = (jobs[i].N + 127) / 128; jobs[i].wg_begin = (int)t; t += (long)jobs[i].tiles_m * jobs[i].tiles_n; } return t; }
template <typename T> bool gemm_job_vec_ok(bool, bool, const GemmJob<T>&) { return true; }
static long g_launch = 0;
template <typename T> int gemm_launch(bool, bool, bool, GemmCfgId, const GemmJob<T>*, int, const GemmJob<T>&, long, const int*, hipStream_t s, long max_wg) { if (max_wg < 0) abort(); trace_launch(s); ++g_launch; return 0; }
int gemm_wgs_per_cu(GemmCfgId c) { return c.id == 0 ? 1 : 2; }
template <typename T> int gemm_single(bool, bool, int, int, int, T, const T*, int, const T*, int, T, T*, int, int, const int*, hipStream_t s) { trace_launch(s); ++g_launch; return 0; }
template <typename T> int trinv_blocks_launch(const TrsmShape&, const T*, int, int, T* inv, const int*, hipStream_t s, int) { if (!inv) abort(); trace_launch(s); ++g_launch; return 0; }
template <typename T> int trsm_launch(const TrsmShape&, const TrsmJob<T>*, int, const TrsmJob<T>&, long, const int*, hipStream_t s, long max_wg) { if (max_wg < 0) abort(); trace_launch(s); ++g_launch; return 0; }
template <typename T> int trsm_full(char, char, char, char, int, int, T, const T*, int, T*, int, T* ws, const int*, hipStream_t s) { if (!ws) abort(); trace_launch(s); ++g_launch; return 0; }
template <typename T> int potrf_full(bool, int, T*, int, int*, int*, int, T*, hipStream_t s) { trace_launch(s); ++g_launch; return 0; }
int potrf_f64_of_f32(bool, int, float*, int, double* W, int*, int*, int, hipStream_t s) { if (!W) abort(); trace_launch(s); ++g_launch; return 0; }
bool gemm_split_ok(bool ta, bool tb, const GemmJob<float>& j) {
  return ta && !tb && !((uintptr_t)j.A & 15) && !((uintptr_t)j.B & 15) && !(j.lda & 3) &&
         !(j.ldb & 3);
}
// the streaming block-column product is a device kernel: the planner tests
// take the tile DAG path
bool tiled_skinny_gemm_ok(char, char, const dm_tiled&, const dm_tiled&, const dm_tiled&) {
  return false;
}
int tiled_skinny_gemm_launch(const dm_tiled&, const dm_tiled&, const dm_tiled&, double, double,
                             hipStream_t) {
  ++g_launch;
  return 0;
}
int gemm_split_launch(int terms, int, const GemmJob<float>*, int, const GemmJob<float>&, long,
                      long, const int*, hipStream_t) {
  if (terms != 3 && terms != 6) abort();
  ++g_launch;
  return 0;
}
template <typename T> int gemv_launch(bool, int, int, T, const T*, int, const T*, int, T, T*, int, const int*, hipStream_t) { ++g_launch; return 0; }
int fill_i32_launch(int*, int, size_t, hipStream_t s) { trace_launch(s); ++g_launch; return 0; }
int copy_i32_launch(int*, const int*, size_t, hipStream_t s) { trace_launch(s); ++g_launch; return 0; }
// The diagonal-tile inversion: the plan's inverse tiles are real host
// memory (their addresses become GEMM task operands); run() is one launch.
template <typename T> TrtriPlan<T>::~TrtriPlan() { free(mem); }
template <typename T>
int TrtriPlan<T>::build(bool up, int64_t m_, int nb_, const std::function<const T*(int)>& tile) {
  upper = up; m = m_; nb = nb_;
  mt = (int)((m + nb - 1) / nb);
  for (int k = 0; k < mt; ++k) if (!tile(k)) abort();
  inv_bytes = sizeof(T) * (size_t)mt * nb * nb;
  mem = malloc(inv_bytes ? inv_bytes : 1);
  return 0;
}
template <typename T> int TrtriPlan<T>::run(hipStream_t s, const int*) const {
  if (!mem) abort();
  trace_launch(s);
  ++g_launch;
  return 0;
}
template struct TrtriPlan<double>;
template struct TrtriPlan<float>;
template <typename T> size_t getrf_panel_ws_bytes(int64_t Mp, int W) {
  return sizeof(T) * (size_t)Mp * W + 65536;
}
template <typename T>
int getrf_panel(T* tiles, int64_t Mp, int W, int mb, int64_t, int* ipiv, int* info, int*, bool,
                void* ws, hipStream_t s) {
  if (!tiles || !ipiv || !info || !ws || Mp <= 0 || W <= 0 || mb <= 0) abort();
  trace_launch(s);
  ++g_launch;
  return 0;
}
template <typename T>
int laswp_tiled(T* base, int64_t Mr, int W, int mb, const int* ipiv, int npiv, int64_t,
                const int*, hipStream_t s) {
  if (!base || !ipiv || Mr <= 0 || W <= 0 || mb <= 0 || npiv < 0) abort();
  trace_launch(s);
  ++g_launch;
  return 0;
}
// complex tile operations (complex.hip): one traced launch each
template <typename R> int zgemm_call(char, char, int, int, int, R, R, const R*, int, const R*, int, R, R, R*, int, int, int, const int*, hipStream_t s) { trace_launch(s); ++g_launch; return 0; }
template <typename R> int ztrsm_call(char, char, char, char, int, int, R, R, const R*, int, R*, int, const int*, hipStream_t s) { trace_launch(s); ++g_launch; return 0; }
template <typename R> int zpotrf_rec(bool, int, R*, int, int*, int, int*, int, hipStream_t s) { trace_launch(s); ++g_launch; return 0; }
#define INST(T) \
  template int zgemm_call<T>(char, char, int, int, int, T, T, const T*, int, const T*, int, T, T, T*, int, int, int, const int*, hipStream_t); \
  template int ztrsm_call<T>(char, char, char, char, int, int, T, T, const T*, int, T*, int, const int*, hipStream_t); \
  template int zpotrf_rec<T>(bool, int, T*, int, int*, int, int*, int, hipStream_t); \
  template GemmCfgId gemm_pick_cfg<T>(int, int, int, int, bool); \
  template long gemm_prepare_jobs<T>(GemmJob<T>*, int, GemmCfgId); \
  template bool gemm_job_vec_ok<T>(bool, bool, const GemmJob<T>&); \
  template int gemm_launch<T>(bool, bool, bool, GemmCfgId, const GemmJob<T>*, int, const GemmJob<T>&, long, const int*, hipStream_t, long); \
  template int gemm_single<T>(bool, bool, int, int, int, T, const T*, int, const T*, int, T, T*, int, int, const int*, hipStream_t); \
  template int trinv_blocks_launch<T>(const TrsmShape&, const T*, int, int, T*, const int*, hipStream_t, int); \
  template int trsm_launch<T>(const TrsmShape&, const TrsmJob<T>*, int, const TrsmJob<T>&, long, const int*, hipStream_t, long); \
  template int trsm_full<T>(char, char, char, char, int, int, T, const T*, int, T*, int, T*, const int*, hipStream_t); \
  template int potrf_full<T>(bool, int, T*, int, int*, int*, int, T*, hipStream_t); \
  template int gemv_launch<T>(bool, int, int, T, const T*, int, const T*, int, T, T*, int, const int*, hipStream_t); \
  template size_t getrf_panel_ws_bytes<T>(int64_t, int); \
  template int getrf_panel<T>(T*, int64_t, int, int, int64_t, int*, int*, int*, bool, void*, hipStream_t); \
  template int laswp_tiled<T>(T*, int64_t, int, int, const int*, int, int64_t, const int*, hipStream_t);
INST(double)
INST(float)
}  // namespace dm
extern "C" const char* dm_last_error_string(void) { return dm::g_err.c_str(); }

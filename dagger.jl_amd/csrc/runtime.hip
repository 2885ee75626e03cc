// C ABI: runtime (errors, device, memory, streams, events, graphs) and the
// BLAS/LAPACK-shaped tile entry points.
#include <pthread.h>
#include <condition_variable>
#include <functional>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <unordered_map>
#include "dm_internal.h"

// Device scratch of the allocation-free TRSM entry point (dm_dtrsm_noalloc):
// the inverted 32x32 diagonal blocks of one leaf solve, which every leaf
// writes in full before reading it.  One copy per GPU, held until the
// process ends.  Host code never passes this name to a kernel or a copy: it
// is a host-side stand-in, and the device address of each GPU's copy comes
// from hipGetSymbolAddress at initialisation (scratch_dev below).
constexpr size_t kScratchElems = 8192;  // 64 KiB of FP64
static __device__ __align__(256) double scratch[kScratchElems];

namespace dm {
int randn_launch(double* x, size_t count, uint64_t seed, uint64_t offset, hipStream_t s);
static_assert(trsm_ws_elems<double>() <= kScratchElems,
              "the static TRSM scratch must hold one leaf's inverted diagonal blocks");

static thread_local std::string g_last_error;
static std::mutex g_init_mu;
static bool g_inited = false;
static int g_device = -1;
static hipStream_t g_default_stream = nullptr;
// Device address of `scratch` on each device the library initialised
// (indexed by device id; written once under g_init_mu).
constexpr int kMaxDevices = 64;
static void* g_scratch_dev[kMaxDevices] = {};

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* what) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
  set_error(buf);
  return DM_HIP_ERROR_BASE + (int)e;
}

int arg_fail(int argno, const char* what) {
  char buf[512];
  snprintf(buf, sizeof buf, "invalid argument %d: %s", argno, what);
  set_error(buf);
  return argno > 0 ? -argno : -99;
}

int ensure_init() {
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_inited) {
    DM_HIP(hipSetDevice(g_device));
    // Launch errors are read with hipGetLastError after each launch; drop
    // anything left over from an earlier, unrelated call.
    (void)hipGetLastError();
    return 0;
  }
  int n = 0;
  DM_HIP(hipGetDeviceCount(&n));
  if (n <= 0) {
    set_error("no HIP device visible");
    return DM_HIP_ERROR_BASE + (int)hipErrorNoDevice;
  }
  if (g_device < 0) g_device = 0;
  DM_HIP(hipSetDevice(g_device));
  DM_HIP(hipStreamCreateWithFlags(&g_default_stream, hipStreamNonBlocking));
  int rc = trsm_init_attributes();
  if (rc) return rc;
  // this device's copy of the static scratch (the device is current)
  if (g_device >= kMaxDevices) {
    set_error("device id beyond the library's per-device table");
    return DM_HIP_ERROR_BASE + (int)hipErrorInvalidDevice;
  }
  void* sp = nullptr;
  DM_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  g_scratch_dev[g_device] = sp;
  g_inited = true;
  return 0;
}

// Device address of the static scratch on the library's device (after
// ensure_init), or null.
static double* scratch_dev() {
  std::lock_guard<std::mutex> lk(g_init_mu);
  return (g_device >= 0 && g_device < kMaxDevices) ? (double*)g_scratch_dev[g_device] : nullptr;
}

hipStream_t stream_of(dm_stream_t s) { return s ? s->h : g_default_stream; }

// ----------------------------------------------------------- big-stack runner
// The HIP runtime walks a captured graph recursively (recursion depth grows
// with the longest dependency chain of the graph), which overflows a default
// 8 MiB thread stack for a large tiled factorization.  Graph capture,
// instantiation and launch therefore run on one helper thread with a 1 GiB
// stack (reserved, committed lazily); callers block until their call returns.
namespace {
struct BigStackRunner {
  std::mutex mu;
  std::condition_variable cv;
  std::function<int()>* job = nullptr;
  int result = 0;
  std::string err;
  bool done = false, started = false;
  pthread_t th;

  static void* main_fn(void* self) {
    auto* r = (BigStackRunner*)self;
    if (g_device >= 0) (void)hipSetDevice(g_device);
    std::unique_lock<std::mutex> lk(r->mu);
    for (;;) {
      r->cv.wait(lk, [&] { return r->job != nullptr; });
      std::function<int()>* j = r->job;
      lk.unlock();
      const int res = (*j)();
      lk.lock();
      r->err = res ? g_last_error : std::string();
      r->result = res;
      r->job = nullptr;
      r->done = true;
      r->cv.notify_all();
    }
    return nullptr;
  }

  int run(std::function<int()>& fn) {
    std::unique_lock<std::mutex> lk(mu);
    if (!started) {
      pthread_attr_t attr;
      pthread_attr_init(&attr);
      pthread_attr_setstacksize(&attr, (size_t)1 << 30);
      if (pthread_create(&th, &attr, &BigStackRunner::main_fn, this) != 0) {
        pthread_attr_destroy(&attr);
        lk.unlock();
        return fn();  // fall back to the caller's stack
      }
      pthread_attr_destroy(&attr);
      pthread_detach(th);
      started = true;
    }
    cv.wait(lk, [&] { return job == nullptr; });
    job = &fn;
    done = false;
    cv.notify_all();
    cv.wait(lk, [&] { return done; });
    done = false;
    const int res = result;
    if (res) set_error(err);  // the message was recorded on the helper thread
    cv.notify_all();
    return res;
  }
};
// Heap-allocated and never destroyed: the detached helper blocks on the
// condition variable for the life of the process, and destroying a
// condition variable that a thread still waits on can hang process exit.
BigStackRunner* g_big_stack = new BigStackRunner;
std::mutex* g_big_stack_serial = new std::mutex;
}  // namespace

int run_big_stack(std::function<int()> fn) {
  // one job at a time through the helper thread
  std::lock_guard<std::mutex> lk(*g_big_stack_serial);
  return g_big_stack->run(fn);
}

// --------------------------------------------------------------- workspace
namespace {
struct WsEntry { void* p; size_t bytes; };
std::mutex g_ws_mu;
std::unordered_map<hipStream_t, WsEntry> g_ws;
}  // namespace

void* stream_workspace(hipStream_t s, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  WsEntry& e = g_ws[s];
  if (e.bytes >= bytes) return e.p;
  // Growing: the old buffer may still be in use by queued work on s, so
  // wait for s before releasing it.
  if (e.p) {
    if (hipStreamSynchronize(s) != hipSuccess) return nullptr;
    (void)hipFree(e.p);
    e.p = nullptr;
    e.bytes = 0;
  }
  size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
  if (hipMalloc(&e.p, want) != hipSuccess) {
    e.p = nullptr;
    return nullptr;
  }
  e.bytes = want;
  return e.p;
}

// ----------------------------------------------------- executor streams
namespace {
std::mutex g_exec_mu;
std::unordered_map<hipStream_t, std::vector<hipStream_t>> g_exec;
}  // namespace

int exec_streams(hipStream_t caller, int count, hipStream_t* out) {
  std::lock_guard<std::mutex> lk(g_exec_mu);
  std::vector<hipStream_t>& v = g_exec[caller];
  while ((int)v.size() < count) {
    hipStream_t h = nullptr;
    DM_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    v.push_back(h);
  }
  for (int i = 0; i < count; ++i) out[i] = v[i];
  return 0;
}

void release_exec_streams(hipStream_t caller) {
  std::vector<hipStream_t> v;
  {
    std::lock_guard<std::mutex> lk(g_exec_mu);
    auto it = g_exec.find(caller);
    if (it == g_exec.end()) return;
    v.swap(it->second);
    g_exec.erase(it);
  }
  for (hipStream_t h : v) {
    (void)hipStreamSynchronize(h);
    (void)hipStreamDestroy(h);
  }
}

}  // namespace dm

using namespace dm;

#define DM_INIT()                 \
  do {                            \
    int rc_ = ::dm::ensure_init(); \
    if (rc_) return rc_;          \
  } while (0)

extern "C" {

int dm_version(void) { return 100; }

const char* dm_last_error_string(void) { return g_last_error.c_str(); }

int dm_device_count(int* count) {
  if (!count) return arg_fail(1, "count is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return 0;
}

int dm_init(int device) {
  {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_inited && device != g_device) {
      set_error("dm_init: library already bound to another device");
      return arg_fail(1, "device differs from the initialised one");
    }
    if (!g_inited) g_device = device;
  }
  return ensure_init();
}

int dm_malloc(void** ptr, size_t bytes) {
  if (!ptr) return arg_fail(1, "ptr is NULL");
  DM_INIT();
  *ptr = nullptr;
  if (bytes == 0) return 0;
  DM_HIP(hipMalloc(ptr, bytes));
  return 0;
}

int dm_free(void* ptr) {
  if (!ptr) return 0;
  DM_INIT();
  DM_HIP(hipFree(ptr));
  return 0;
}

int dm_memcpy_h2d(void* dst, const void* src, size_t bytes, dm_stream_t s) {
  DM_INIT();
  if (!bytes) return 0;
  DM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_of(s)));
  return 0;
}

int dm_memcpy_d2h(void* dst, const void* src, size_t bytes, dm_stream_t s) {
  DM_INIT();
  if (!bytes) return 0;
  DM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_of(s)));
  DM_HIP(hipStreamSynchronize(stream_of(s)));
  return 0;
}

int dm_memcpy_d2d(void* dst, const void* src, size_t bytes, dm_stream_t s) {
  DM_INIT();
  if (!bytes) return 0;
  DM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_of(s)));
  return 0;
}

int dm_memcpy2d_h2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width_bytes,
                    size_t height, dm_stream_t s) {
  DM_INIT();
  if (!width_bytes || !height) return 0;
  DM_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width_bytes, height, hipMemcpyHostToDevice,
                          stream_of(s)));
  return 0;
}

int dm_memcpy2d_d2h(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width_bytes,
                    size_t height, dm_stream_t s) {
  DM_INIT();
  if (!width_bytes || !height) return 0;
  DM_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width_bytes, height, hipMemcpyDeviceToHost,
                          stream_of(s)));
  DM_HIP(hipStreamSynchronize(stream_of(s)));
  return 0;
}

int dm_memcpy2d_d2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width_bytes,
                    size_t height, dm_stream_t s) {
  DM_INIT();
  if (!width_bytes || !height) return 0;
  DM_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width_bytes, height, hipMemcpyDeviceToDevice,
                          stream_of(s)));
  return 0;
}

int dm_fill_i32(int* dst, int value, size_t count, dm_stream_t s) {
  DM_INIT();
  return fill_i32_launch(dst, value, count, stream_of(s));
}

int dm_stream_create(dm_stream_t* s) {
  if (!s) return arg_fail(1, "stream is NULL");
  DM_INIT();
  auto* st = new dm_stream_s;
  hipError_t e = hipStreamCreateWithFlags(&st->h, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete st;
    return hip_fail(e, "hipStreamCreateWithFlags");
  }
  *s = st;
  return 0;
}

int dm_stream_destroy(dm_stream_t s) {
  if (!s) return 0;
  DM_INIT();
  release_exec_streams(s->h);
  hipError_t e = hipStreamDestroy(s->h);
  delete s;
  if (e != hipSuccess) return hip_fail(e, "hipStreamDestroy");
  return 0;
}

int dm_stream_sync(dm_stream_t s) {
  DM_INIT();
  DM_HIP(hipStreamSynchronize(stream_of(s)));
  return 0;
}

int dm_device_sync(void) {
  DM_INIT();
  DM_HIP(hipDeviceSynchronize());
  return 0;
}

int dm_event_create(dm_event_t* e) {
  if (!e) return arg_fail(1, "event is NULL");
  DM_INIT();
  auto* ev = new dm_event_s;
  hipError_t err = hipEventCreate(&ev->h);
  if (err != hipSuccess) {
    delete ev;
    return hip_fail(err, "hipEventCreate");
  }
  *e = ev;
  return 0;
}

int dm_event_destroy(dm_event_t e) {
  if (!e) return 0;
  DM_INIT();
  hipError_t err = hipEventDestroy(e->h);
  delete e;
  if (err != hipSuccess) return hip_fail(err, "hipEventDestroy");
  return 0;
}

int dm_event_record(dm_event_t e, dm_stream_t s) {
  if (!e) return arg_fail(1, "event is NULL");
  DM_INIT();
  DM_HIP(hipEventRecord(e->h, stream_of(s)));
  return 0;
}

int dm_stream_wait_event(dm_stream_t s, dm_event_t e) {
  if (!e) return arg_fail(2, "event is NULL");
  DM_INIT();
  DM_HIP(hipStreamWaitEvent(stream_of(s), e->h, 0));
  return 0;
}

int dm_event_sync(dm_event_t e) {
  if (!e) return arg_fail(1, "event is NULL");
  DM_INIT();
  DM_HIP(hipEventSynchronize(e->h));
  return 0;
}

int dm_event_elapsed_ms(float* ms, dm_event_t start, dm_event_t stop) {
  if (!ms) return arg_fail(1, "ms is NULL");
  if (!start) return arg_fail(2, "start is NULL");
  if (!stop) return arg_fail(3, "stop is NULL");
  DM_INIT();
  DM_HIP(hipEventElapsedTime(ms, start->h, stop->h));
  return 0;
}

int dm_graph_begin(dm_stream_t s) {
  DM_INIT();
  DM_HIP(hipStreamBeginCapture(stream_of(s), hipStreamCaptureModeThreadLocal));
  return 0;
}

int dm_graph_end(dm_stream_t s, dm_graph_t* g) {
  if (!g) return arg_fail(2, "graph is NULL");
  DM_INIT();
  hipGraph_t graph = nullptr;
  DM_HIP(hipStreamEndCapture(stream_of(s), &graph));
  hipGraphExec_t exec = nullptr;
  int rc = run_big_stack([&]() -> int {
    DM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    return 0;
  });
  if (rc) {
    (void)hipGraphDestroy(graph);
    return rc;
  }
  auto* gg = new dm_graph_s;
  gg->g = graph;
  gg->exec = exec;
  *g = gg;
  return 0;
}

int dm_graph_launch(dm_graph_t g, dm_stream_t s) {
  if (!g) return arg_fail(1, "graph is NULL");
  DM_INIT();
  hipGraphExec_t ex = g->exec;
  hipStream_t st = stream_of(s);
  return run_big_stack([ex, st]() -> int {
    DM_HIP(hipGraphLaunch(ex, st));
    return 0;
  });
}

int dm_graph_destroy(dm_graph_t g) {
  if (!g) return 0;
  DM_INIT();
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->g);
  delete g;
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------ tile ops
template <typename T>
static int potrf_api(char uplo, int n, T* A, int lda, int* info_dev, dm_stream_t s) {
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(1, "uplo must be 'U' or 'L'");
  if (n < 0) return arg_fail(2, "n < 0");
  if (lda < (n > 1 ? n : 1)) return arg_fail(4, "lda < max(1,n)");
  if (!info_dev) return arg_fail(5, "info is NULL");
  DM_INIT();
  hipStream_t st = stream_of(s);
  int rc = fill_i32_launch(info_dev, 0, 1, st);
  if (rc) return rc;
  if (n == 0) return 0;
  return potrf_full<T>(is_upper(uplo), n, A, lda, info_dev, nullptr, 0, nullptr, st);
}

extern "C" int dm_dpotrf(char uplo, int n, double* A, int lda, int* info_dev, dm_stream_t s) {
  return potrf_api<double>(uplo, n, A, lda, info_dev, s);
}
extern "C" int dm_spotrf(char uplo, int n, float* A, int lda, int* info_dev, dm_stream_t s) {
  return potrf_api<float>(uplo, n, A, lda, info_dev, s);
}

template <typename T>
static int trsm_api(char side, char uplo, char transa, char diag, int m, int n, T alpha,
                    const T* A, int lda, T* B, int ldb, dm_stream_t s) {
  if (!is_left(side) && !is_right(side)) return arg_fail(1, "side must be 'L' or 'R'");
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(2, "uplo must be 'U' or 'L'");
  if (!is_trans(transa) && !is_notrans(transa)) return arg_fail(3, "transa must be 'N','T','C'");
  if (!is_unit(diag) && !is_nonunit(diag)) return arg_fail(4, "diag must be 'N' or 'U'");
  if (m < 0) return arg_fail(5, "m < 0");
  if (n < 0) return arg_fail(6, "n < 0");
  const int ka = is_left(side) ? m : n;
  if (lda < (ka > 1 ? ka : 1)) return arg_fail(9, "lda too small");
  if (ldb < (m > 1 ? m : 1)) return arg_fail(11, "ldb < max(1,m)");
  DM_INIT();
  T* ws = (T*)stream_workspace(stream_of(s), trsm_ws_elems<T>() * sizeof(T));
  if (!ws) return arg_fail(0, "trsm: workspace allocation failed");
  return trsm_full<T>(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, ws, nullptr,
                      stream_of(s));
}

extern "C" int dm_dtrsm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
             const double* A, int lda, double* B, int ldb, dm_stream_t s) {
  return trsm_api<double>(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, s);
}
extern "C" int dm_strsm(char side, char uplo, char transa, char diag, int m, int n, float alpha,
             const float* A, int lda, float* B, int ldb, dm_stream_t s) {
  return trsm_api<float>(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, s);
}

// dm_dtrsm with its workspace in the static per-GPU scratch instead of the
// per-stream workspace (which is allocated, and grown with a stream
// synchronization, on first use): this entry point allocates nothing.  The
// scratch need does not grow with m or n -- longer solves recurse through
// the same one-leaf buffer -- so no size is rejected for it.
extern "C" int dm_dtrsm_noalloc(char side, char uplo, char transa, char diag, int m, int n,
                                double alpha, const double* A, int lda, double* B, int ldb,
                                dm_stream_t s) {
  if (!is_left(side) && !is_right(side)) return arg_fail(1, "side must be 'L' or 'R'");
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(2, "uplo must be 'U' or 'L'");
  if (!is_trans(transa) && !is_notrans(transa)) return arg_fail(3, "transa must be 'N','T','C'");
  if (!is_unit(diag) && !is_nonunit(diag)) return arg_fail(4, "diag must be 'N' or 'U'");
  if (m < 0) return arg_fail(5, "m < 0");
  if (n < 0) return arg_fail(6, "n < 0");
  const int ka = is_left(side) ? m : n;
  if (lda < (ka > 1 ? ka : 1)) return arg_fail(9, "lda too small");
  if (ldb < (m > 1 ? m : 1)) return arg_fail(11, "ldb < max(1,m)");
  DM_INIT();
  double* ws = scratch_dev();
  if (!ws) {
    set_error("dm_dtrsm_noalloc: no static scratch address for this device");
    return DM_HIP_ERROR_BASE + (int)hipErrorInvalidSymbol;
  }
  return trsm_full<double>(side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, ws, nullptr,
                           stream_of(s));
}

template <typename T>
static int syrk_api(char uplo, char trans, int n, int k, T alpha, const T* A, int lda, T beta,
                    T* C, int ldc, dm_stream_t s) {
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(1, "uplo must be 'U' or 'L'");
  if (!is_trans(trans) && !is_notrans(trans)) return arg_fail(2, "trans must be 'N','T','C'");
  if (n < 0) return arg_fail(3, "n < 0");
  if (k < 0) return arg_fail(4, "k < 0");
  const int rows = is_notrans(trans) ? n : k;
  if (lda < (rows > 1 ? rows : 1)) return arg_fail(7, "lda too small");
  if (ldc < (n > 1 ? n : 1)) return arg_fail(10, "ldc < max(1,n)");
  DM_INIT();
  const bool tr = is_trans(trans);
  // trans 'T': C = alpha A^T A + beta C;  'N': C = alpha A A^T + beta C
  return gemm_single<T>(tr, !tr, n, n, k, alpha, A, lda, A, lda, beta, C, ldc,
                        is_upper(uplo) ? 1 : 2, nullptr, stream_of(s));
}

extern "C" int dm_dsyrk(char uplo, char trans, int n, int k, double alpha, const double* A, int lda,
             double beta, double* C, int ldc, dm_stream_t s) {
  return syrk_api<double>(uplo, trans, n, k, alpha, A, lda, beta, C, ldc, s);
}
extern "C" int dm_ssyrk(char uplo, char trans, int n, int k, float alpha, const float* A, int lda,
             float beta, float* C, int ldc, dm_stream_t s) {
  return syrk_api<float>(uplo, trans, n, k, alpha, A, lda, beta, C, ldc, s);
}

template <typename T>
static int gemm_api(char transa, char transb, int m, int n, int k, T alpha, const T* A, int lda,
                    const T* B, int ldb, T beta, T* C, int ldc, dm_stream_t s) {
  if (!is_trans(transa) && !is_notrans(transa)) return arg_fail(1, "transa must be 'N','T','C'");
  if (!is_trans(transb) && !is_notrans(transb)) return arg_fail(2, "transb must be 'N','T','C'");
  if (m < 0) return arg_fail(3, "m < 0");
  if (n < 0) return arg_fail(4, "n < 0");
  if (k < 0) return arg_fail(5, "k < 0");
  const bool ta = is_trans(transa), tb = is_trans(transb);
  const int ra = ta ? k : m, rb = tb ? n : k;
  if (lda < (ra > 1 ? ra : 1)) return arg_fail(8, "lda too small");
  if (ldb < (rb > 1 ? rb : 1)) return arg_fail(10, "ldb too small");
  if (ldc < (m > 1 ? m : 1)) return arg_fail(13, "ldc < max(1,m)");
  DM_INIT();
  return gemm_single<T>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0, nullptr,
                        stream_of(s));
}

extern "C" int dm_dgemm(char transa, char transb, int m, int n, int k, double alpha, const double* A,
             int lda, const double* B, int ldb, double beta, double* C, int ldc, dm_stream_t s) {
  return gemm_api<double>(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}
extern "C" int dm_sgemm(char transa, char transb, int m, int n, int k, float alpha, const float* A, int lda,
             const float* B, int ldb, float beta, float* C, int ldc, dm_stream_t s) {
  return gemm_api<float>(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

extern "C" int dm_sgemm_bf16x(int terms, char transa, char transb, int m, int n, int k, float alpha,
                              const float* A, int lda, const float* B, int ldb, float beta,
                              float* C, int ldc, dm_stream_t s) {
  if (terms != 3 && terms != 6) return arg_fail(1, "terms must be 3 or 6");
  if (!is_trans(transa)) return arg_fail(2, "transa must be 'T' (k-contiguous A)");
  if (!is_notrans(transb)) return arg_fail(3, "transb must be 'N' (k-contiguous B)");
  if (m < 0) return arg_fail(4, "m < 0");
  if (n < 0) return arg_fail(5, "n < 0");
  if (k < 0) return arg_fail(6, "k < 0");
  if (lda < (k > 1 ? k : 1) || (lda & 3)) return arg_fail(9, "lda < max(1,k) or not a multiple of 4");
  if (ldb < (k > 1 ? k : 1) || (ldb & 3)) return arg_fail(11, "ldb < max(1,k) or not a multiple of 4");
  if (ldc < (m > 1 ? m : 1)) return arg_fail(14, "ldc < max(1,m)");
  if (((uintptr_t)A & 15) || ((uintptr_t)B & 15)) return arg_fail(8, "A and B must be 16-byte aligned");
  DM_INIT();
  if (m == 0 || n == 0) return 0;
  GemmJob<float> j{};
  j.A = A; j.B = B; j.C = C;
  j.M = m; j.N = n; j.K = k;
  j.lda = lda; j.ldb = ldb; j.ldc = ldc;
  j.alpha = alpha; j.beta = beta; j.tri = 0;
  if (k == 0 || alpha == 0.0f) { j.alpha = 0.0f; j.K = 0; }
  const GemmCfgId cfg = gemm_split_cfg(terms, m, n, 1);
  const long total = gemm_prepare_jobs<float>(&j, 1, cfg);
  return gemm_split_launch(terms, cfg.bm, nullptr, 0, j, total, total, nullptr, stream_of(s));
}

extern "C" int dm_dtrsv(char uplo, char trans, char diag, int n, const double* A, int lda, double* x,
             int incx, dm_stream_t s) {
  if (!is_upper(uplo) && !is_lower(uplo)) return arg_fail(1, "uplo must be 'U' or 'L'");
  if (!is_trans(trans) && !is_notrans(trans)) return arg_fail(2, "trans must be 'N','T','C'");
  if (!is_unit(diag) && !is_nonunit(diag)) return arg_fail(3, "diag must be 'N' or 'U'");
  if (n < 0) return arg_fail(4, "n < 0");
  if (lda < (n > 1 ? n : 1)) return arg_fail(6, "lda < max(1,n)");
  if (incx <= 0) return arg_fail(8, "incx must be positive");
  DM_INIT();
  if (n == 0) return 0;
  double* ws = (double*)stream_workspace(stream_of(s), trsm_ws_elems<double>() * sizeof(double));
  if (!ws) return arg_fail(0, "trsv: workspace allocation failed");
  if (incx == 1)
    return trsm_full<double>('L', uplo, trans, diag, n, 1, 1.0, A, lda, x, n, ws, nullptr,
                             stream_of(s));
  // x with stride incx is the single row of a 1 x n matrix with ld = incx:
  // solve x^T op'(A) = b^T with op' the opposite transpose.
  return trsm_full<double>('R', uplo, is_trans(trans) ? 'N' : 'T', diag, 1, n, 1.0, A, lda, x,
                           incx, ws, nullptr, stream_of(s));
}

extern "C" int dm_dgemv(char trans, int m, int n, double alpha, const double* A, int lda, const double* x,
             int incx, double beta, double* y, int incy, dm_stream_t s) {
  if (!is_trans(trans) && !is_notrans(trans)) return arg_fail(1, "trans must be 'N','T','C'");
  if (m < 0) return arg_fail(2, "m < 0");
  if (n < 0) return arg_fail(3, "n < 0");
  if (lda < (m > 1 ? m : 1)) return arg_fail(6, "lda < max(1,m)");
  if (incx <= 0) return arg_fail(8, "incx must be positive");
  if (incy <= 0) return arg_fail(11, "incy must be positive");
  DM_INIT();
  return gemv_launch<double>(is_trans(trans), m, n, alpha, A, lda, x, incx, beta, y, incy,
                             nullptr, stream_of(s));
}

extern "C" int dm_fill_randn_d(double* x, size_t count, uint64_t seed, uint64_t offset, dm_stream_t s) {
  if (!x && count) return arg_fail(1, "x is NULL");
  DM_INIT();
  return randn_launch(x, count, seed, offset, stream_of(s));
}


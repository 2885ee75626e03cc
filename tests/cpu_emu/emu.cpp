// Fiber scheduler and HIP host API of the CPU emulation (see
// hip/hip_runtime.h).  Test infrastructure: built only into the emulated
// library used by the CPU bounds-checking tests.
#include <hip/hip_runtime.h>
#include <ucontext.h>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <vector>

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define EMU_ASAN 1
#include <sanitizer/common_interface_defs.h>
#endif
#endif

namespace emu {
namespace {

[[noreturn]] void die(const char* what) {
  fprintf(stderr, "emu: %s\n", what);
  fflush(stderr);
  abort();
}

enum State { RUNNABLE, AT_BARRIER, AT_WAVE, DONE };

struct Fiber {
  ucontext_t uc;
  dim3 tid;
  int linear = 0;
  State st = RUNNABLE;
  char* stack = nullptr;
};

constexpr size_t kStack = 256 * 1024;

struct WaveSlots {
  unsigned char buf[2][64][64];  // double-buffered by op parity
  size_t bytes[2] = {0, 0};
  int arrived[2] = {0, 0};
  long ops = 0;                  // completed wave ops of this wave
  std::vector<long> lane_ops;    // ops issued per lane
};

struct Run {
  std::vector<Fiber> f;
  ucontext_t sched;
  const void* sched_bottom = nullptr;
  size_t sched_size = 0;
  dim3 bid, bdim, gdim;
  std::function<void()>* body = nullptr;
  int n = 0, live = 0, at_barrier = 0;
  std::vector<WaveSlots> waves;
  std::vector<char*> stacks;  // reused across blocks
  unsigned char* dyn = nullptr;
  int cur = -1;
};

thread_local Run* g_run = nullptr;
std::recursive_mutex g_launch_mu;  // one kernel at a time, process-wide

Fiber& self() {
  if (!g_run || g_run->cur < 0) die("device call outside a kernel");
  return g_run->f[g_run->cur];
}

void to_scheduler() {
  Run* r = g_run;
  Fiber& me = r->f[r->cur];
#ifdef EMU_ASAN
  void* fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, r->sched_bottom, r->sched_size);
#endif
  swapcontext(&me.uc, &r->sched);
#ifdef EMU_ASAN
  __sanitizer_finish_switch_fiber(fake, &r->sched_bottom, &r->sched_size);
#endif
}

void trampoline() {
  Run* r = g_run;
#ifdef EMU_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &r->sched_bottom, &r->sched_size);
#endif
  (*r->body)();
  Fiber& me = r->f[r->cur];
  me.st = DONE;
  --r->live;
  // a barrier may now be complete for the remaining work-items
  if (r->live > 0 && r->at_barrier == r->live) {
    for (Fiber& o : r->f)
      if (o.st == AT_BARRIER) o.st = RUNNABLE;
    r->at_barrier = 0;
  }
#ifdef EMU_ASAN
  __sanitizer_start_switch_fiber(nullptr, r->sched_bottom, r->sched_size);
#endif
  setcontext(&r->sched);
}

void run_block(Run& r) {
  const int n = r.n;
  for (int i = 0; i < n; ++i) {
    Fiber& f = r.f[i];
    f.linear = i;
    f.tid = dim3(i % r.bdim.x, (i / r.bdim.x) % r.bdim.y, i / (r.bdim.x * r.bdim.y));
    f.st = RUNNABLE;
    f.stack = r.stacks[i];
    getcontext(&f.uc);
    f.uc.uc_stack.ss_sp = f.stack;
    f.uc.uc_stack.ss_size = kStack;
    f.uc.uc_link = nullptr;
    makecontext(&f.uc, (void (*)())trampoline, 0);
  }
  r.live = n;
  r.at_barrier = 0;
  for (WaveSlots& w : r.waves) {
    w.arrived[0] = w.arrived[1] = 0;
    w.ops = 0;
    std::fill(w.lane_ops.begin(), w.lane_ops.end(), 0);
  }
  while (r.live > 0) {
    bool progressed = false;
    for (int i = 0; i < n; ++i) {
      if (r.f[i].st != RUNNABLE) continue;
      progressed = true;
      r.cur = i;
#ifdef EMU_ASAN
      void* fake = nullptr;
      __sanitizer_start_switch_fiber(&fake, r.f[i].stack, kStack);
#endif
      swapcontext(&r.sched, &r.f[i].uc);
#ifdef EMU_ASAN
      __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
      r.cur = -1;
    }
    if (!progressed) {
      int cnt[4] = {0, 0, 0, 0};
      for (int i = 0; i < n; ++i) ++cnt[r.f[i].st];
      fprintf(stderr, "emu: block (%u,%u,%u): %d runnable, %d at barrier, %d at wave op, %d done\n",
              r.bid.x, r.bid.y, r.bid.z, cnt[RUNNABLE], cnt[AT_BARRIER], cnt[AT_WAVE], cnt[DONE]);
      for (size_t w = 0; w < r.waves.size(); ++w)
        fprintf(stderr, "emu:   wave %zu: ops %ld, arrived %d/%d\n", w, r.waves[w].ops,
                r.waves[w].arrived[0], r.waves[w].arrived[1]);
      die("deadlock: work-items wait at a barrier or wave op that not all reach");
    }
  }
}

}  // namespace

const dim3& thread_idx() { return self().tid; }
const dim3& block_idx() { return g_run->bid; }
const dim3& block_dim() { return g_run->bdim; }
const dim3& grid_dim() { return g_run->gdim; }
unsigned char* dynamic_shared() {
  if (!g_run->dyn) die("dynamic LDS used by a launch that requested none");
  return g_run->dyn;
}

void syncthreads() {
  Run* r = g_run;
  Fiber& me = self();
  me.st = AT_BARRIER;
  if (++r->at_barrier == r->live) {
    for (Fiber& o : r->f)
      if (o.st == AT_BARRIER) o.st = RUNNABLE;
    r->at_barrier = 0;
    me.st = RUNNABLE;
    return;
  }
  to_scheduler();
}

void wave_exchange(const void* in, size_t bytes, void* all) {
  Run* r = g_run;
  Fiber& me = self();
  if (bytes > 64) die("wave op operand too large");
  const int wv = me.linear / 64, lane = me.linear % 64;
  if ((wv + 1) * 64 > r->n) die("wave op in a partial wavefront");
  WaveSlots& w = r->waves[wv];
  const long op = w.lane_ops[lane]++;
  const int par = (int)(op & 1);
  if (op != w.ops) die("wave op: lanes of one wavefront diverged");
  if (bytes) std::memcpy(w.buf[par][lane], in, bytes);
  w.bytes[par] = bytes;
  if (++w.arrived[par] == 64) {
    w.arrived[par] = 0;
    ++w.ops;
    for (int l = 0; l < 64; ++l) {
      Fiber& o = r->f[wv * 64 + l];
      if (o.st == AT_WAVE) o.st = RUNNABLE;
    }
  } else {
    me.st = AT_WAVE;
    to_scheduler();
    if (w.ops <= op) die("wave op resumed early");
  }
  if (bytes)
    for (int l = 0; l < 64; ++l) std::memcpy((char*)all + l * bytes, w.buf[par][l], bytes);
}

// ------------------------------------------------------------ MFMA
// f64 16x16x4: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k;
// D[i][j] in lane j + 16*(i%4), item i/4.
d4 mfma_f64_16x16x4(double a, double b, d4 c) {
  struct In { double a, b; d4 c; } me{a, b, c}, all[64];
  wave_exchange(&me, sizeof(In), all);
  const int lane = self().linear % 64;
  d4 out;
  for (int r = 0; r < 4; ++r) {
    const int i = 4 * r + lane / 16, j = lane % 16;
    double acc = all[lane].c[r];
    for (int k = 0; k < 4; ++k) acc += all[i + 16 * k].a * all[j + 16 * k].b;
    out[r] = acc;
  }
  return out;
}

// f32 16x16x4: same inputs; D[i][j] in lane j + 16*(i/4), item i%4.
f4 mfma_f32_16x16x4(float a, float b, f4 c) {
  struct In { float a, b; f4 c; } me{a, b, c}, all[64];
  wave_exchange(&me, sizeof(In), all);
  const int lane = self().linear % 64;
  f4 out;
  for (int r = 0; r < 4; ++r) {
    const int i = 4 * (lane / 16) + r, j = lane % 16;
    float acc = all[lane].c[r];
    for (int k = 0; k < 4; ++k) acc += all[i + 16 * k].a * all[j + 16 * k].b;
    out[r] = acc;
  }
  return out;
}

// bf16 16x16x32: A[i][k] in lane i + 16*(k/8), element k%8; B[k][j] in
// lane j + 16*(k/8), element k%8; D as f32 16x16x4.
f4 mfma_f32_16x16x32_bf16(bf8 a, bf8 b, f4 c) {
  struct In { bf8 a, b; f4 c; } me{a, b, c}, all[64];
  wave_exchange(&me, sizeof(In), all);
  const int lane = self().linear % 64;
  f4 out;
  for (int r = 0; r < 4; ++r) {
    const int i = 4 * (lane / 16) + r, j = lane % 16;
    float acc = all[lane].c[r];
    for (int k = 0; k < 32; ++k)
      acc += (float)all[i + 16 * (k / 8)].a[k % 8] * (float)all[j + 16 * (k / 8)].b[k % 8];
    out[r] = acc;
  }
  return out;
}

// ------------------------------------------------------------ launches
namespace {
struct Capture {
  std::vector<std::function<void()>> ops;
};
thread_local Capture* g_capture = nullptr;
}  // namespace

static void run_launch(dim3 grid, dim3 block, size_t shmem, const std::function<void()>& body) {
  std::lock_guard<std::recursive_mutex> lk(g_launch_mu);
  if (g_run) die("nested kernel launch");
  const int n = (int)(block.x * block.y * block.z);
  if (n <= 0 || n > 1024) die("bad block size");
  Run r;
  r.n = n;
  r.f.resize(n);
  r.bdim = block;
  r.gdim = grid;
  r.waves.resize((n + 63) / 64);
  for (WaveSlots& w : r.waves) w.lane_ops.assign(64, 0);
  for (int i = 0; i < n; ++i) r.stacks.push_back((char*)std::malloc(kStack));
  std::function<void()> b = body;
  r.body = &b;
  g_run = &r;
  for (unsigned z = 0; z < grid.z; ++z)
    for (unsigned y = 0; y < grid.y; ++y)
      for (unsigned x = 0; x < grid.x; ++x) {
        r.bid = dim3(x, y, z);
        r.dyn = shmem ? (unsigned char*)std::malloc(shmem) : nullptr;
        run_block(r);
        std::free(r.dyn);
        r.dyn = nullptr;
      }
  g_run = nullptr;
  for (char* s : r.stacks) std::free(s);
}

void launch(dim3 grid, dim3 block, size_t shmem, hipStream_t, std::function<void()> body) {
  if ((size_t)grid.x * grid.y * grid.z == 0) return;
  if (g_capture) {
    g_capture->ops.push_back([=]() { run_launch(grid, block, shmem, body); });
    return;
  }
  run_launch(grid, block, shmem, body);
}

}  // namespace emu

// ------------------------------------------------------------ host API
struct ihipStream_t { int id; };
struct ihipEvent_t { double t = 0; };
struct ihipGraph { std::vector<std::function<void()>> ops; };
struct hipGraphExec { std::vector<std::function<void()>> ops; };

static double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

static void host_op(std::function<void()> f) {
  if (emu::g_capture) emu::g_capture->ops.push_back(std::move(f));
  else f();
}

hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
  // EMU_CUS lowers the CU count so that capped (looping) launches are exercised
  const char* e = std::getenv("EMU_CUS");
  *v = e ? std::atoi(e) : 256;
  return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  static int ids = 0;
  *s = new ihipStream_t{++ids};
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipGetSymbolAddress(void** p, const void* sym) {
  *p = const_cast<void*>(sym);
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) { *e = new ihipEvent_t; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  if (!e) return hipErrorInvalidValue;
  host_op([e]() { e->t = now_ms(); });
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = (float)(b->t - a->t);
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t e, unsigned) {
  return e ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  if (emu::g_capture) return hipErrorInvalidValue;
  emu::g_capture = new emu::Capture;
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  if (!emu::g_capture) return hipErrorInvalidValue;
  *g = new ihipGraph{std::move(emu::g_capture->ops)};
  delete emu::g_capture;
  emu::g_capture = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  *e = new hipGraphExec{g->ops};
  return hipSuccess;
}
hipError_t hipGraphGetNodes(hipGraph_t g, hipGraphNode_t*, size_t* n) {
  *n = g->ops.size();
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto& op : e->ops) op();
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
hipError_t hipMalloc(void** p, size_t n) {
  if (emu::g_capture) return hipErrorInvalidValue;  // not allowed while capturing
  void* q = nullptr;
  if (posix_memalign(&q, 256, n ? n : 1)) return hipErrorOutOfMemory;
  *p = q;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  if (emu::g_capture) return hipErrorInvalidValue;
  std::free(p);
  return hipSuccess;
}
hipError_t hipMallocAsync(void** p, size_t n, hipStream_t) { return hipMalloc(p, n); }
hipError_t hipFreeAsync(void* p, hipStream_t) { return hipFree(p); }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  if (n) std::memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  if (n) host_op([=]() { std::memmove(d, s, n); });
  return hipSuccess;
}
hipError_t hipMemcpy2DAsync(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h,
                            hipMemcpyKind, hipStream_t) {
  host_op([=]() {
    for (size_t r = 0; r < h; ++r) std::memmove((char*)d + r * dp, (const char*)s + r * sp, w);
  });
  return hipSuccess;
}
hipError_t hipMemset(void* d, int v, size_t n) {
  if (emu::g_capture) return hipErrorInvalidValue;
  std::memset(d, v, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
  if (emu::g_capture) return hipErrorInvalidValue;  // no memset nodes in captured graphs
  std::memset(d, v, n);
  return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "emulated HIP error"; }

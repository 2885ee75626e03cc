// CPU emulation of the HIP device model, for bounds checking the library's
// kernels under AddressSanitizer on the host (test infrastructure only).
//
// Every workgroup runs as a set of fibers (one per work-item) on the calling
// OS thread: __syncthreads is a fiber barrier (a barrier that not every live
// work-item reaches is reported as a deadlock), MFMA builtins and
// __shfl_down are wave-wide exchanges among the 64 lanes of a wavefront,
// using the lane layouts measured on gfx950.  __shared__ arrays become
// function statics (launches are serialized process-wide, so one copy is
// enough, and ASan puts redzones around them), the dynamic LDS buffer is an
// exact-size heap block, device memory is host memory (exact-size
// allocations, so ASan flags any access past an allocation), launches run
// synchronously, and stream capture records launches for replay.
// The lanes of a wave run one after another between sync points, so lanes
// that exchange data through LDS without a workgroup barrier must separate
// each write from the other lanes' reads (and each read from a later
// overwrite) by __builtin_amdgcn_wave_barrier(), which emits nothing on the
// hardware, where a wave executes in lockstep.
#pragma once
#include <algorithm>
#include <cmath>
using std::isfinite;  // a global device function under HIP
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>

typedef int hipError_t;
enum {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorNoDevice = 100,
  hipErrorInvalidDevice = 101,
  hipErrorInvalidSymbol = 701,
  hipErrorLaunchFailure = 719
};
typedef struct ihipStream_t* hipStream_t;
typedef struct ihipEvent_t* hipEvent_t;
typedef struct ihipGraph* hipGraph_t;
typedef struct hipGraphExec* hipGraphExec_t;
typedef struct hipGraphNode* hipGraphNode_t;
enum hipStreamCaptureMode { hipStreamCaptureModeThreadLocal = 0 };
enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
};
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 0 };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 0 };
#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

hipError_t hipGetDeviceCount(int*);
hipError_t hipSetDevice(int);
hipError_t hipGetDevice(int*);
hipError_t hipDeviceGetAttribute(int*, hipDeviceAttribute_t, int);
hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int);
hipError_t hipStreamCreateWithFlags(hipStream_t*, unsigned);
hipError_t hipStreamDestroy(hipStream_t);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipEventCreate(hipEvent_t*);
hipError_t hipEventCreateWithFlags(hipEvent_t*, unsigned);
hipError_t hipEventDestroy(hipEvent_t);
hipError_t hipEventRecord(hipEvent_t, hipStream_t);
hipError_t hipEventSynchronize(hipEvent_t);
hipError_t hipEventElapsedTime(float*, hipEvent_t, hipEvent_t);
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned);
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode);
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*);
hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, hipGraphNode_t*, char*, size_t);
hipError_t hipGraphGetNodes(hipGraph_t, hipGraphNode_t*, size_t*);
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t);
hipError_t hipGraphDestroy(hipGraph_t);
hipError_t hipGraphExecDestroy(hipGraphExec_t);
hipError_t hipMalloc(void**, size_t);
// a __device__ variable is an ordinary host object here: its "device
// address" is its host address
#define HIP_SYMBOL(x) (&(x))
#define __align__(n) __attribute__((aligned(n)))
hipError_t hipGetSymbolAddress(void**, const void*);
hipError_t hipFree(void*);
hipError_t hipMallocAsync(void**, size_t, hipStream_t);
hipError_t hipFreeAsync(void*, hipStream_t);
hipError_t hipMemcpy(void*, const void*, size_t, hipMemcpyKind);
hipError_t hipMemcpyAsync(void*, const void*, size_t, hipMemcpyKind, hipStream_t);
hipError_t hipMemcpy2DAsync(void*, size_t, const void*, size_t, size_t, size_t, hipMemcpyKind,
                            hipStream_t);
hipError_t hipMemset(void*, int, size_t);
hipError_t hipMemsetAsync(void*, int, size_t, hipStream_t);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t);

namespace emu {
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

const dim3& thread_idx();
const dim3& block_idx();
const dim3& block_dim();
const dim3& grid_dim();
void syncthreads();
// Wave-wide exchange: every lane of the current wavefront contributes
// `bytes` from `in`; on return `all` holds the 64 contributions (lane order).
void wave_exchange(const void* in, size_t bytes, void* all);
void launch(dim3 grid, dim3 block, size_t shmem, hipStream_t s, std::function<void()> body);
// Device-side bounds of the dynamic __shared__ buffer of the current launch.
unsigned char* dynamic_shared();

d4 mfma_f64_16x16x4(double a, double b, d4 c);
f4 mfma_f32_16x16x4(float a, float b, f4 c);
f4 mfma_f32_16x16x32_bf16(bf8 a, bf8 b, f4 c);

template <typename T>
inline T shfl_down(T v, unsigned d, int width) {
  T all[64];
  wave_exchange(&v, sizeof(T), all);
  const unsigned lane = thread_idx().x & 63;
  const unsigned w = (unsigned)width;
  const unsigned base = lane - lane % w;
  return (lane % w) + d < w ? all[base + (lane % w) + d] : v;
}
template <typename T>
inline T shfl(T v, int src, int width) {
  T all[64];
  wave_exchange(&v, sizeof(T), all);
  const unsigned lane = thread_idx().x & 63;
  const unsigned w = (unsigned)width;
  return all[lane - lane % w + (unsigned)src % w];
}
inline int readlane(int v, int l) {
  int all[64];
  wave_exchange(&v, sizeof(int), all);
  return all[l];
}
}  // namespace emu

#define threadIdx (::emu::thread_idx())
#define blockIdx (::emu::block_idx())
#define blockDim (::emu::block_dim())
#define gridDim (::emu::grid_dim())
#define __syncthreads() ::emu::syncthreads()
#define __global__
#define __device__
#define __host__
#define __shared__ static
#define DM_OPAQUE(x) asm volatile("" : "+r"(x))
#define DM_DYNAMIC_LDS(name) unsigned char* name = ::emu::dynamic_shared()
#define __forceinline__ inline
#define __launch_bounds__(...)
// Reciprocal square root estimates: the exact value perturbed by a relative
// 2^-22 (FP64) / 2^-20 (FP32), so the callers' Newton refinement is exercised.
#define __builtin_amdgcn_rsq(x) ((1.0 / std::sqrt((double)(x))) * (1.0 + 0x1p-22))
#define __builtin_amdgcn_rsqf(x) ((1.0f / std::sqrt((float)(x))) * (1.0f + 0x1p-20f))
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) ::emu::mfma_f64_16x16x4((a), (b), (c))
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, x, y, z) ::emu::mfma_f32_16x16x4((a), (b), (c))
#define __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, x, y, z) \
  ::emu::mfma_f32_16x16x32_bf16((a), (b), (c))
#define __shfl_down(v, d, w) ::emu::shfl_down((v), (d), (w))
#define __shfl(v, l, w) ::emu::shfl((v), (l), (w))
#define __builtin_amdgcn_wave_barrier() ::emu::wave_exchange(nullptr, 0, nullptr)
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define __builtin_amdgcn_sched_barrier(mask) ((void)0)
#define __builtin_amdgcn_s_setprio(p) ((void)0)
#define DM_GLOBAL_AS  /* host memory: no device address spaces */
#define __builtin_amdgcn_readlane(v, l) ::emu::readlane((v), (l))
#define hipLaunchKernelGGL(K, G, B, SH, ST, ...) \
  ::emu::launch(dim3(G), dim3(B), (size_t)(SH), (ST), [=]() { K(__VA_ARGS__); })

inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline long min(long a, long b) { return a < b ? a : b; }
inline long max(long a, long b) { return a > b ? a : b; }
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }

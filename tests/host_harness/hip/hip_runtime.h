// Host-only stand-in for the HIP runtime header, for CPU tests of the
// planner (no device code in the translation units that include it).
#pragma once
#include <cstddef>
#include <cstdint>
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorNoDevice = 100 };
typedef struct ihipStream_t* hipStream_t;
typedef struct ihipEvent_t* hipEvent_t;
typedef struct ihipGraph* hipGraph_t;
typedef struct hipGraphExec* hipGraphExec_t;
typedef struct hipGraphNode* hipGraphNode_t;
enum hipStreamCaptureMode { hipStreamCaptureModeThreadLocal = 0 };
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u
hipError_t hipStreamCreateWithFlags(hipStream_t*, unsigned);
hipError_t hipStreamDestroy(hipStream_t);
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 0 };
hipError_t hipGetDevice(int*);
hipError_t hipDeviceGetAttribute(int*, hipDeviceAttribute_t, int);
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
hipError_t hipFree(void*);
hipError_t hipMemcpy(void*, const void*, size_t, hipMemcpyKind);
hipError_t hipMemset(void*, int, size_t);
hipError_t hipMemcpyAsync(void*, const void*, size_t, hipMemcpyKind, hipStream_t);
hipError_t hipMemsetAsync(void*, int, size_t, hipStream_t);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t);
// device-function qualifiers of the shared internal header (no device code
// is compiled here)
#define __device__
#define __forceinline__ inline
#define __builtin_amdgcn_s_setprio(p) ((void)0)
#define DM_GLOBAL_AS

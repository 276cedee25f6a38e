// Host stand-in for the parts of the HIP runtime API that the afivo_amd host
// code (afivo_amd/csrc/afg_context.cpp) uses, for the CPU emulation build of
// the library in tests/emul.  TEST INFRASTRUCTURE ONLY: the product library
// is built with hipcc against the real runtime.
#pragma once

#include <cstddef>
#include <cstdint>

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorMemoryAllocation = 2,
       hipErrorNoDevice = 100, hipErrorLaunchFailure = 719, hipErrorNotSupported = 801 };
typedef struct EmulStream *hipStream_t;
typedef struct EmulEvent *hipEvent_t;
typedef struct EmulGraph *hipGraph_t;
typedef struct EmulGraph *hipGraphExec_t;
typedef struct EmulNode *hipGraphNode_t;
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
enum hipDeviceAttribute_t { hipDeviceAttributeMaxSharedMemoryPerBlock = 1,
                            hipDeviceAttributeSharedMemPerBlockOptin = 2 };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 1 };
#define hipStreamNonBlocking 1u
#define hipHostMallocDefault 0u
#define hipEventRecordExternal 1u

hipError_t hipGetDevice(int *dev);
hipError_t hipSetDevice(int dev);
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int dev);
const char *hipGetErrorString(hipError_t e);
hipError_t hipGetLastError();
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipMalloc(void **p, size_t bytes);
template <typename T>
inline hipError_t hipMalloc(T **p, size_t bytes) { return hipMalloc((void **)p, bytes); }
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k, hipStream_t s);
hipError_t hipMemset(void *p, int v, size_t bytes);
hipError_t hipMemsetAsync(void *p, int v, size_t bytes, hipStream_t s);
hipError_t hipEventCreate(hipEvent_t *e);
#define hipEventDisableTiming 2u
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecordWithFlags(hipEvent_t e, hipStream_t s, unsigned flags);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *n, char *log, size_t sz);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);

// Host implementation of the HIP runtime subset used by afivo_amd's host code,
// for the CPU emulation build of the library.  TEST INFRASTRUCTURE ONLY.
//
// Device memory is host memory with PROT_NONE guard pages right after the end
// of every allocation, so a kernel body that reads or writes past an array
// faults at once.  Streams run work immediately; stream capture records the
// submitted work into a graph that a launch replays.
#include <sys/mman.h>
#include <unistd.h>

#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "emul.h"

struct EmulGraph {
    std::vector<std::function<void()>> work;
};
struct EmulStream {
    bool capturing = false;
    EmulGraph *graph = nullptr;
};
struct EmulEvent {
    long stamp = 0;
};

namespace {
std::mutex g_mu;
std::map<void *, std::pair<void *, size_t>> g_maps;   // user ptr -> (mapping, length)
long g_clock = 0;
}  // namespace

void *emul_alloc(size_t bytes)
{
    const size_t page = (size_t)sysconf(_SC_PAGESIZE);
    const size_t data = (bytes + page - 1) / page * page;
    const size_t len = data + 2 * page;
    char *m = (char *)mmap(nullptr, len, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (m == MAP_FAILED) return nullptr;
    mprotect(m, page, PROT_NONE);                      // guard before
    mprotect(m + page + data, page, PROT_NONE);        // guard after
    // end of the buffer flush with the guard after it (8-byte aligned)
    char *p = m + page + (data - (bytes + 7) / 8 * 8);
    std::lock_guard<std::mutex> lk(g_mu);
    g_maps[p] = {m, len};
    return p;
}

void emul_free(void *p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_maps.find(p);
    if (it == g_maps.end()) return;
    munmap(it->second.first, it->second.second);
    g_maps.erase(it);
}

namespace {
thread_local hipError_t g_last_error = hipSuccess;
int g_fail_launches = 0;   // the next this many submissions fail as a launch would
}  // namespace

// Test hook: make the next n submitted kernels fail to launch (they do not
// run, and hipGetLastError reports hipErrorLaunchFailure), as a launch with
// a bad configuration does on the device.
extern "C" void emul_fail_launches(int n) { g_fail_launches = n; }

hipError_t hipGetLastError()
{
    const hipError_t e = g_last_error;
    g_last_error = hipSuccess;
    return e;
}

void emul_submit(hipStream_t s, std::function<void()> work)
{
    if (g_fail_launches > 0) {
        g_fail_launches--;
        g_last_error = hipErrorLaunchFailure;
        return;
    }
    if (s && s->capturing)
        s->graph->work.push_back(std::move(work));
    else
        work();
}

hipError_t hipGetDevice(int *dev)
{
    *dev = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int)
{
    *v = a == hipDeviceAttributeSharedMemPerBlockOptin ? 163840 : 65536;
    return hipSuccess;
}
const char *hipGetErrorString(hipError_t e)
{
    return e == hipSuccess ? "no error" : e == hipErrorLaunchFailure ? "unspecified launch failure" : "emulated error";
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned)
{
    *s = new EmulStream;
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipMalloc(void **p, size_t bytes)
{
    *p = emul_alloc(bytes ? bytes : 8);
    return *p ? hipSuccess : hipErrorMemoryAllocation;
}
hipError_t hipFree(void *p)
{
    emul_free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned)
{
    *p = std::malloc(bytes ? bytes : 8);
    return *p ? hipSuccess : hipErrorMemoryAllocation;
}
hipError_t hipHostFree(void *p)
{
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind)
{
    std::memmove(dst, src, bytes);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind, hipStream_t s)
{
    if (s && s->capturing) return hipErrorNotSupported;
    std::memmove(dst, src, bytes);
    return hipSuccess;
}
hipError_t hipMemset(void *p, int v, size_t bytes)
{
    std::memset(p, v, bytes);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *p, int v, size_t bytes, hipStream_t s)
{
    if (s && s->capturing) return hipErrorNotSupported;   // no memset nodes in graphs
    std::memset(p, v, bytes);
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e)
{
    *e = new EmulEvent;
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e)
{
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecordWithFlags(hipEvent_t e, hipStream_t s, unsigned)
{
    emul_submit(s, [e] { e->stamp = ++g_clock; });
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) { return hipEventRecordWithFlags(e, s, 0); }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
    *ms = 0.001f * (float)(b->stamp - a->stamp);
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode)
{
    if (s->capturing) return hipErrorInvalidValue;
    s->capturing = true;
    s->graph = new EmulGraph;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g)
{
    if (!s->capturing) return hipErrorInvalidValue;
    s->capturing = false;
    *g = s->graph;
    s->graph = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t)
{
    *e = new EmulGraph(*g);
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete g;
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e)
{
    delete e;
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s)
{
    if (s && s->capturing) return hipErrorNotSupported;
    for (auto &w : e->work) w();
    return hipSuccess;
}

// Internal interface of the CPU emulation build (tests/emul).
#pragma once

#include <functional>

#include "hip/hip_runtime.h"

// Run `work` in stream order: immediately, or recorded into the graph being
// captured on the stream.
void emul_submit(hipStream_t s, std::function<void()> work);
// Host buffer with guard pages after its end (out-of-bounds accesses fault).
void *emul_alloc(size_t bytes);
void emul_free(void *p);

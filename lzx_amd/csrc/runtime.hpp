// Shared host-side runtime pieces of liblzx_gpu: status handling, HIP error
// capture and the per-device bookkeeping used by every module.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lzx_gpu.h"

namespace lzx {

// Records the HIP error text for lzx_last_device_error() and maps it to a
// status code.
lzx_status device_fail(hipError_t e, const char *what);

// Records a message for lzx_last_device_error() without a HIP error (a
// refusal whose reason a caller may want to read).
void note_error(const char *what);

// Stream-ordered device scratch from the library's per-device pool (freed
// memory stays cached in the pool for the next call).
hipError_t scratch_alloc(void **p, size_t bytes, hipStream_t s);
hipError_t scratch_free(void *p, hipStream_t s);

// A pinned host buffer from the library's reuse list.
struct HostStage {
    void *p = nullptr;
    size_t cap = 0;
};
hipError_t stage_acquire(HostStage &h, size_t bytes);
void stage_release(HostStage &h);

} // namespace lzx

#define LZX_HIP(call)                                                                                        \
    do {                                                                                                     \
        hipError_t lzx_e_ = (call);                                                                          \
        if (lzx_e_ != hipSuccess) return ::lzx::device_fail(lzx_e_, #call);                                  \
    } while (0)

#define LZX_LAUNCH_CHECK(what)                                                                               \
    do {                                                                                                     \
        hipError_t lzx_e_ = hipGetLastError();                                                               \
        if (lzx_e_ != hipSuccess) return ::lzx::device_fail(lzx_e_, what);                                   \
    } while (0)

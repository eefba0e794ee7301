// Shared host-side runtime pieces of liblzx_gpu: status handling, HIP error
// capture and the per-device bookkeeping used by every module.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <string>

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

// Tuned choices (a u32 per text key) shared by every module: kept for the
// process and, when LZX_AUTOTUNE_CACHE names a file, loaded from and saved to
// it (one "key<TAB>value" line each).
bool autotune_get(const std::string &key, uint32_t &val);
void autotune_put(const std::string &key, uint32_t val);

// A pinned host buffer from the library's reuse list.
struct HostStage {
    void *p = nullptr;
    size_t cap = 0;
};
hipError_t stage_acquire(HostStage &h, size_t bytes);
void stage_release(HostStage &h);

// A synchronous call on host slices (the reference's &mut [u32] / &[RingElement]
// signatures): in_bytes of `in` are staged through a pinned buffer into
// dev_bytes of pool scratch on the null stream of `device`, run(dev, stream)
// is enqueued, out_bytes from dev + out_off are read back into `out`, with one
// synchronisation.  Nothing is allocated once the pool and the pinned reuse
// list hold buffers of this size.
lzx_status host_slice_call(int device, const void *in, size_t in_bytes, void *out, size_t out_off, size_t out_bytes,
                           size_t dev_bytes, const std::function<lzx_status(void *, hipStream_t)> &run);

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

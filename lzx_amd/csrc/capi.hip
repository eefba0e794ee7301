// Status strings, HIP error capture and the library's scratch allocators for
// the C ABI (include/lzx_gpu.h).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "runtime.hpp"

namespace {
thread_local std::string g_last_error;

constexpr int MAX_DEVICES = 64;
std::mutex g_alloc_mu;
hipMemPool_t g_pools[MAX_DEVICES];            // one stream-ordered pool per device
std::vector<lzx::HostStage> g_stages;         // idle pinned staging buffers
} // namespace

namespace lzx {

lzx_status device_fail(hipError_t e, const char *what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    if (e == hipErrorOutOfMemory) return LZX_DEVICE_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return LZX_NO_DEVICE;
    return LZX_DEVICE_ERROR;
}

void note_error(const char *what) { g_last_error = what; }

// Stream-ordered scratch from a per-device pool that keeps freed memory for
// reuse (release threshold = unlimited), so a call that needs the same
// working set as the last one maps no new memory.  The pool holds the
// largest working set seen until the process ends.
hipError_t scratch_alloc(void **p, size_t bytes, hipStream_t s)
{
    hipDevice_t dev = 0;
    hipError_t e = hipStreamGetDevice(s, &dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipMallocAsync(p, bytes, s);
    hipMemPool_t pool;
    {
        std::lock_guard<std::mutex> g(g_alloc_mu);
        if (!g_pools[dev]) {
            hipMemPoolProps props{};
            props.allocType = hipMemAllocationTypePinned;
            props.location.type = hipMemLocationTypeDevice;
            props.location.id = dev;
            hipMemPool_t np = nullptr;
            e = hipMemPoolCreate(&np, &props);
            if (e != hipSuccess) return e;
            uint64_t keep = UINT64_MAX;
            e = hipMemPoolSetAttribute(np, hipMemPoolAttrReleaseThreshold, &keep);
            if (e != hipSuccess) {
                (void)hipMemPoolDestroy(np);
                return e;
            }
            g_pools[dev] = np;
        }
        pool = g_pools[dev];
    }
    return hipMallocFromPoolAsync(p, bytes, pool, s);
}

hipError_t scratch_free(void *p, hipStream_t s) { return p ? hipFreeAsync(p, s) : hipSuccess; }

// Pinned host staging buffers, reused across calls and threads: a call takes
// an idle buffer of sufficient size (or allocates one) and hands it back.
// ---- tuning cache: process-wide, optionally backed by a text file ----------

namespace {
std::mutex g_tune_mu;
std::map<std::string, uint32_t> g_tune;
bool g_tune_loaded = false;

void tune_load_locked()
{
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char *path = getenv("LZX_AUTOTUNE_CACHE");
    if (!path || !*path) return;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        const size_t tab = line.rfind('\t');
        if (tab == std::string::npos) continue;
        g_tune[line.substr(0, tab)] = (uint32_t)strtoul(line.c_str() + tab + 1, nullptr, 10);
    }
}
} // namespace

bool autotune_get(const std::string &key, uint32_t &val)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    val = it->second;
    return true;
}

void autotune_put(const std::string &key, uint32_t val)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    g_tune[key] = val;
    const char *path = getenv("LZX_AUTOTUNE_CACHE");
    if (!path || !*path) return;
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream f(tmp, std::ios::trunc);
        for (const auto &kv : g_tune) f << kv.first << '\t' << kv.second << '\n';
    }
    std::rename(tmp.c_str(), path);
}

hipError_t stage_acquire(HostStage &h, size_t bytes)
{
    {
        std::lock_guard<std::mutex> g(g_alloc_mu);
        for (size_t i = 0; i < g_stages.size(); i++)
            if (g_stages[i].cap >= bytes) {
                h = g_stages[i];
                g_stages[i] = g_stages.back();
                g_stages.pop_back();
                return hipSuccess;
            }
    }
    const size_t cap = bytes < 65536 ? 65536 : bytes;
    h.cap = 0;
    hipError_t e = hipHostMalloc(&h.p, cap, 0);
    if (e == hipSuccess) h.cap = cap;
    return e;
}

void stage_release(HostStage &h)
{
    if (!h.p) return;
    std::lock_guard<std::mutex> g(g_alloc_mu);
    g_stages.push_back(h);
    h = HostStage{};
}

lzx_status host_slice_call(int device, const void *in, size_t in_bytes, void *out, size_t out_off, size_t out_bytes,
                           size_t dev_bytes, const std::function<lzx_status(void *, hipStream_t)> &run)
{
    LZX_HIP(hipSetDevice(device));
    hipStream_t s = nullptr;
    HostStage h;
    const size_t stage_bytes = in_bytes > out_bytes ? in_bytes : out_bytes;
    LZX_HIP(stage_acquire(h, stage_bytes ? stage_bytes : 1));
    void *d = nullptr;
    hipError_t e = scratch_alloc(&d, dev_bytes ? dev_bytes : 1, s);
    if (e != hipSuccess) {
        stage_release(h);
        return device_fail(e, "host slice scratch");
    }
    lzx_status st = LZX_OK;
    if (in_bytes) {
        memcpy(h.p, in, in_bytes);
        e = hipMemcpyAsync(d, h.p, in_bytes, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) st = device_fail(e, "host slice upload");
    }
    if (st == LZX_OK) st = run(d, s);
    if (st == LZX_OK && out_bytes) {
        e = hipMemcpyAsync(h.p, (uint8_t *)d + out_off, out_bytes, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) st = device_fail(e, "host slice readback");
    }
    (void)scratch_free(d, s);
    e = hipStreamSynchronize(s);
    if (st == LZX_OK && e != hipSuccess) st = device_fail(e, "host slice synchronize");
    if (st == LZX_OK && out_bytes) memcpy(out, h.p, out_bytes);
    stage_release(h);
    return st;
}

} // namespace lzx

extern "C" const char *lzx_last_device_error(void) { return g_last_error.c_str(); }

extern "C" const char *lzx_status_string(lzx_status s)
{
    switch (s) {
    case LZX_OK: return "Ok";
    case LZX_NTT_LENGTH_MISMATCH: return "NttError::LengthMismatch";
    case LZX_NTT_ADICITY_TOO_SMALL: return "NttError::AdicityTooSmall";
    case LZX_RING_LENGTH_MISMATCH: return "RingError::LengthMismatch";
    case LZX_AJTAI_DIMENSION_MISMATCH: return "AjtaiError::DimensionMismatch";
    case LZX_AJTAI_NORM_EXCEEDED: return "AjtaiError::NormExceeded";
    case LZX_AJTAI_COMMITMENT_MISMATCH: return "AjtaiError::CommitmentMismatch";
    case LZX_SUMCHECK_BAD_ROUND_SHAPE: return "SumcheckError::BadRoundShape";
    case LZX_SUMCHECK_ROUND_CHECK_FAILED: return "SumcheckError::RoundCheckFailed";
    case LZX_SUMCHECK_FINAL_CHECK_FAILED: return "SumcheckError::FinalCheckFailed";
    case LZX_SUMCHECK_CLAIM_MISMATCH: return "SumcheckError::ClaimMismatch";
    case LZX_SUMCHECK_VIRTUAL_POLY: return "SumcheckError::VirtualPoly";
    case LZX_ZEROCHECK_FINAL_CLAIM_NONZERO: return "ZerocheckError::FinalClaimNonZero";
    case LZX_BATCH_VARIABLE_COUNT_MISMATCH: return "BatchError::VariableCountMismatch";
    case LZX_BATCH_EMPTY: return "BatchError::EmptyBatch";
    case LZX_BATCH_CHALLENGE_COUNT: return "BatchError::ChallengeCount";
    case LZX_SHORT_CHALLENGE_REJECTION_BUDGET: return "ShortChallengeError::RejectionBudgetExceeded";
    case LZX_SHORT_CHALLENGE_INVALID_PARAMETERS: return "ShortChallengeError::InvalidParameters";
    case LZX_SHORT_CHALLENGE_GAMMA_CAP_EXCEEDED: return "ShortChallengeError::GammaCapExceeded";
    case LZX_MOD50_LOG_N_OUT_OF_RANGE: return "Mod50Error::LogNOutOfRange";
    case LZX_MOD50_ADICITY_TOO_SMALL: return "Mod50Error::AdicityTooSmall";
    case LZX_MOD50_GENERATOR_INVALID: return "Mod50Error::GeneratorInvalid";
    case LZX_MOD50_LENGTH_MISMATCH: return "Mod50Error::LengthMismatch";
    case LZX_PACKING_LENGTH_MISMATCH: return "PackingError::LengthMismatch";
    case LZX_PACKING_LIMB_OUT_OF_RANGE: return "PackingError::LimbOutOfRange";
    case LZX_PACKING_MODULUS_TOO_LARGE: return "PackingError::ModulusTooLarge";
    case LZX_PACKING_VALUE_OUT_OF_RANGE: return "PackingError::ValueOutOfRange";
    case LZX_PACKING_NOT_COPRIME: return "PackingError::NotCoprime";
    case LZX_DECOMP_DIGIT_COUNT_MISMATCH: return "DecompositionError::DigitCountMismatch";
    case LZX_DECOMP_DIGIT_OUT_OF_RANGE: return "DecompositionError::DigitOutOfRange";
    case LZX_DECOMP_VALUE_OUT_OF_RANGE: return "DecompositionError::ValueOutOfRange";
    case LZX_DECOMP_VALUE_TOO_LARGE: return "DecompositionError::ValueTooLarge";
    case LZX_TRANSCRIPT_REJECTION_BUDGET: return "TranscriptError::RejectionBudgetExceeded";
    case LZX_TRANSCRIPT_MESSAGE_TOO_LARGE: return "TranscriptError::MessageTooLarge";
    case LZX_INVALID_ARGUMENT: return "InvalidArgument";
    case LZX_UNSUPPORTED: return "Unsupported";
    case LZX_DEVICE_ERROR: return "DeviceError";
    case LZX_DEVICE_OUT_OF_MEMORY: return "DeviceOutOfMemory";
    case LZX_NO_DEVICE: return "NoDevice";
    default: return "Unknown";
    }
}

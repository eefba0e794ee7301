// Status strings and HIP error capture for the C ABI (include/lzx_gpu.h).
#include <string>

#include "runtime.hpp"

namespace {
thread_local std::string g_last_error;
}

namespace lzx {

lzx_status device_fail(hipError_t e, const char *what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    if (e == hipErrorOutOfMemory) return LZX_DEVICE_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return LZX_NO_DEVICE;
    return LZX_DEVICE_ERROR;
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

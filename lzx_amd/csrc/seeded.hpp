// The Fiat-Shamir transcript object and its host-side helpers, shared by the
// transcript C ABI and the sum-check prover (which runs the same transcript
// on the device between rounds and hands the state back).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "keccak.hpp"
#include "runtime.hpp"

struct lzx_transcript {
    lzx::Sponge sponge;
    uint64_t query_count;
    uint64_t query_budget;
};

namespace lzx {

// append_field_slice / challenge_field with C-string labels.
lzx_status transcript_append_fields(lzx_transcript *t, const char *label, const uint64_t *v, size_t n);
lzx_status transcript_challenge(lzx_transcript *t, const char *label, uint64_t *out);

} // namespace lzx

// Host-side seeded sampling and transcript helpers shared across modules.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "runtime.hpp"

namespace lzx {

// RingConfig::uniform_from_seed (ring.rs:102-180) for one ring element.
void uniform_from_seed(uint32_t q, uint32_t log_n, const uint8_t *domain, size_t dlen, const uint8_t *seed,
                       size_t slen, uint64_t index, uint32_t *coeffs);

// append_field_slice / challenge_field with C-string labels.
lzx_status transcript_append_fields(lzx_transcript *t, const char *label, const uint64_t *v, size_t n);
lzx_status transcript_challenge(lzx_transcript *t, const char *label, uint64_t *out);

} // namespace lzx

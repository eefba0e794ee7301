// Host-side transcript helpers shared across modules.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "runtime.hpp"

namespace lzx {

// append_field_slice / challenge_field with C-string labels.
lzx_status transcript_append_fields(lzx_transcript *t, const char *label, const uint64_t *v, size_t n);
lzx_status transcript_challenge(lzx_transcript *t, const char *label, uint64_t *out);

} // namespace lzx

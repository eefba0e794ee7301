// Fiat-Shamir transcript and seeded samplers (host side).
//
// Transcript restates crates/lattice-core/src/transcript.rs:36-238:
//   new            absorb "LZX1" || u32le|protocol| || protocol       (:46-59)
//   append_message absorb u32le|label| || label || u32le|msg| || msg  (:67-78)
//   challenge_fields  query_count++ (budget check), absorb "CHAL" ||
//                  u32le|label| || label || u32le(n), then squeeze from a COPY
//                  of the raw state: stream block b is available after b+1
//                  permutations, no padding is applied; 16-byte chunks, the
//                  first 8 bytes LE is a candidate accepted iff < p (:109-205)
#include <stdlib.h>

#include <vector>

#include "keccak.hpp"
#include "runtime.hpp"
#include "seeded.hpp"

namespace lzx {

static void squeeze_copy(const lzx_transcript *t, size_t *offset, uint8_t *out, size_t n)
{
    uint64_t st[25];
    for (int i = 0; i < 25; i++) st[i] = t->sponge.st[i];
    const size_t block = *offset / 136;
    for (size_t i = 0; i <= block; i++) keccak_f1600(st);
    size_t pos = *offset % 136, done = 0;
    while (done < n) {
        const size_t take = (136 - pos) < (n - done) ? (136 - pos) : (n - done);
        for (size_t i = 0; i < take; i++) out[done + i] = (uint8_t)(st[(pos + i) >> 3] >> (8 * ((pos + i) & 7)));
        done += take;
        pos += take;
        if (pos == 136) {
            keccak_f1600(st);
            pos = 0;
        }
    }
    *offset += n;
}

lzx_status transcript_append_fields(lzx_transcript *t, const char *label, const uint64_t *v, size_t n)
{
    std::vector<uint8_t> bytes(n * 8);
    for (size_t i = 0; i < n; i++)
        for (int b = 0; b < 8; b++) bytes[8 * i + b] = (uint8_t)(v[i] >> (8 * b));
    return lzx_transcript_append_message(t, (const uint8_t *)label, strlen(label), bytes.data(), bytes.size());
}

lzx_status transcript_challenge(lzx_transcript *t, const char *label, uint64_t *out)
{
    return lzx_transcript_challenge_fields(t, (const uint8_t *)label, strlen(label), 1, out);
}


} // namespace lzx

using namespace lzx;

extern "C" lzx_status lzx_transcript_new(const uint8_t *protocol, size_t len, uint64_t query_budget,
                                         lzx_transcript **out)
{
    if (!out || (!protocol && len)) return LZX_INVALID_ARGUMENT;
    lzx_transcript *t = new lzx_transcript();
    t->sponge.init(0x1f);
    const uint8_t tag[4] = {'L', 'Z', 'X', '1'};
    t->sponge.absorb(tag, 4);
    t->sponge.absorb_u32le((uint32_t)len);
    if (len) t->sponge.absorb(protocol, len);
    t->query_count = 0;
    t->query_budget = query_budget;
    *out = t;
    return LZX_OK;
}

extern "C" lzx_status lzx_transcript_clone(const lzx_transcript *t, lzx_transcript **out)
{
    if (!t || !out) return LZX_INVALID_ARGUMENT;
    *out = new lzx_transcript(*t);
    return LZX_OK;
}

extern "C" void lzx_transcript_destroy(lzx_transcript *t) { delete t; }

extern "C" lzx_status lzx_transcript_append_message(lzx_transcript *t, const uint8_t *label, size_t label_len,
                                                    const uint8_t *msg, size_t msg_len)
{
    if (!t || (!label && label_len) || (!msg && msg_len)) return LZX_INVALID_ARGUMENT;
    if (msg_len > ((size_t)1 << 26)) return LZX_TRANSCRIPT_MESSAGE_TOO_LARGE;
    t->sponge.absorb_u32le((uint32_t)label_len);
    if (label_len) t->sponge.absorb(label, label_len);
    t->sponge.absorb_u32le((uint32_t)msg_len);
    if (msg_len) t->sponge.absorb(msg, msg_len);
    return LZX_OK;
}

extern "C" lzx_status lzx_transcript_append_field_slice(lzx_transcript *t, const uint8_t *label, size_t label_len,
                                                        const uint64_t *values, size_t n)
{
    if (!t || (!values && n)) return LZX_INVALID_ARGUMENT;
    std::vector<uint8_t> bytes(n * 8);
    for (size_t i = 0; i < n; i++)
        for (int b = 0; b < 8; b++) bytes[8 * i + b] = (uint8_t)(values[i] >> (8 * b));
    return lzx_transcript_append_message(t, label, label_len, bytes.data(), bytes.size());
}

extern "C" lzx_status lzx_transcript_challenge_fields(lzx_transcript *t, const uint8_t *label, size_t label_len,
                                                      size_t n, uint64_t *out)
{
    if (!t || (!label && label_len) || (!out && n)) return LZX_INVALID_ARGUMENT;
    t->query_count++;
    if (t->query_count > t->query_budget) return LZX_TRANSCRIPT_REJECTION_BUDGET;
    const uint8_t tag[4] = {'C', 'H', 'A', 'L'};
    t->sponge.absorb(tag, 4);
    t->sponge.absorb_u32le((uint32_t)label_len);
    if (label_len) t->sponge.absorb(label, label_len);
    t->sponge.absorb_u32le((uint32_t)n);
    size_t got = 0, offset = 0;
    std::vector<uint8_t> bytes;
    while (got < n) {
        size_t need = (n - got) * 16;
        if (need < 32) need = 32;
        bytes.assign(need, 0);
        squeeze_copy(t, &offset, bytes.data(), need);
        for (size_t c = 0; c < need && got < n; c += 16) {
            uint64_t cand = 0;
            for (int b = 7; b >= 0; b--) cand = (cand << 8) | (c + b < need ? bytes[c + b] : 0);
            if (cand < GL_P) out[got++] = cand;
        }
    }
    return LZX_OK;
}

extern "C" uint64_t lzx_transcript_query_count(const lzx_transcript *t) { return t ? t->query_count : 0; }

// Short ring-element challenges (crates/lattice-core/src/short_challenge.rs:47-500)
// and the extension-ring challenge sampler (crates/lattice-ring/src/extension.rs:180-216).
//
// Host code: a challenge is a few hundred coefficients drawn from one SHAKE-256
// stream per sample (the library's Keccak sponge), so there is nothing for the
// device to do; the device consumes the result as an ordinary ring element.
#include <cstring>
#include <vector>

#include "keccak.hpp"
#include "runtime.hpp"

namespace lzx {

// The stream of one sample (short_challenge.rs:343-352):
// SHAKE256("SHORTCHAL" || u32le|label| || label || u32le|seed| || seed || u64le(0)).
static std::vector<uint8_t> sc_stream(const char *label, const uint8_t *seed, size_t seed_len, size_t bytes)
{
    Sponge sp;
    sp.init(0x1f);
    sp.absorb((const uint8_t *)"SHORTCHAL", 9);
    const size_t ll = strlen(label);
    sp.absorb_u32le((uint32_t)ll);
    sp.absorb((const uint8_t *)label, ll);
    sp.absorb_u32le((uint32_t)seed_len);
    if (seed_len) sp.absorb(seed, seed_len);
    const uint8_t attempt[8] = {0};
    sp.absorb(attempt, 8);
    std::vector<uint8_t> out(bytes);
    sp.finalize(out.data(), bytes);
    return out;
}

// UniformCursor (short_challenge.rs:355-410): little-endian words off the
// stream, unbiased indices by rejection; false when the stream is exhausted.
struct ScCursor {
    const std::vector<uint8_t> &s;
    size_t pos = 0;
    bool word32(uint32_t &v)
    {
        if (pos + 4 > s.size()) return false;
        v = (uint32_t)s[pos] | (uint32_t)s[pos + 1] << 8 | (uint32_t)s[pos + 2] << 16 | (uint32_t)s[pos + 3] << 24;
        pos += 4;
        return true;
    }
    bool word64(uint64_t &v)
    {
        if (pos + 8 > s.size()) return false;
        v = 0;
        for (int i = 7; i >= 0; i--) v = v << 8 | s[pos + i];
        pos += 8;
        return true;
    }
    bool index(uint64_t range, uint64_t &out)
    {
        if (range == 0) return false;
        if (range == 1) return out = 0, true; // no bytes drawn
        if (range > UINT32_MAX) {
            const uint64_t lim = UINT64_MAX / range * range;
            for (uint64_t raw; word64(raw);)
                if (raw < lim) return out = raw % range, true;
            return false;
        }
        const uint64_t lim = (1ULL << 32) / range * range;
        for (uint32_t raw; word32(raw);)
            if (raw < lim) return out = raw % range, true;
        return false;
    }
};

// min(ceil(sqrt(x)), 2^64 - 1) for x < 2^128 (short_challenge.rs:281-340):
// integer Newton from above for floor(sqrt(x)), then the ceiling.
static uint64_t ceil_sqrt128(unsigned __int128 x)
{
    if (x == 0) return 0;
    unsigned __int128 g = (unsigned __int128)1 << 64; // >= floor(sqrt(x)) for every x < 2^128
    for (;;) {
        const unsigned __int128 nx = (g + x / g) >> 1;
        if (nx >= g) break;
        g = nx;
    }
    // g = floor(sqrt(x)) (g <= 2^64 - 1 here, so g * g does not overflow)
    if (g * g == x) return (uint64_t)g;
    return g == UINT64_MAX ? UINT64_MAX : (uint64_t)g + 1;
}

static bool sc_valid(const lzx_short_challenge_spec &s)
{
    if (s.n == 0 || s.n > ((size_t)1 << 20)) return false; // short_challenge.rs:97-124
    switch (s.family) {
    case LZX_SC_FIXED_WEIGHT: return s.weight != 0 && s.weight <= s.n && s.amplitude > 0;
    case LZX_SC_BIASED_TERNARY: return s.p_nonzero_permille != 0 && s.p_nonzero_permille <= 1000;
    case LZX_SC_SMALL_SET: return s.values && s.nvalues != 0 && s.nvalues <= 1024;
    case LZX_SC_FIXED_WEIGHT_SMALL_SET:
        return s.weight != 0 && s.weight <= s.n && s.values && s.nvalues != 0 && s.nvalues <= 1024;
    default: return false;
    }
}

static lzx_status sc_sample(const lzx_short_challenge_spec &sp, const uint8_t *seed, size_t seed_len, int64_t *c,
                            uint64_t *gamma)
{
    if (!sc_valid(sp)) return LZX_SHORT_CHALLENGE_INVALID_PARAMETERS;
    const size_t n = sp.n;
    for (size_t i = 0; i < n; i++) c[i] = 0;
    bool ok = true;
    if (sp.family == LZX_SC_FIXED_WEIGHT || sp.family == LZX_SC_FIXED_WEIGHT_SMALL_SET) {
        // partial Fisher-Yates: slot k draws a position among the n - k not yet taken (:392-456)
        const bool fw = sp.family == LZX_SC_FIXED_WEIGHT;
        const std::vector<uint8_t> s = sc_stream(fw ? "fw" : "fwset", seed, seed_len, sp.weight * (fw ? 24 : 32) + 32);
        ScCursor cur{s};
        std::vector<uint32_t> free_pos(n);
        for (size_t i = 0; i < n; i++) free_pos[i] = (uint32_t)i;
        for (size_t k = 0; k < sp.weight && ok; k++) {
            uint64_t pick = 0, v = 0;
            ok = cur.index(n - k, pick);
            if (!ok) break;
            const uint32_t at = free_pos[pick];
            std::swap(free_pos[pick], free_pos[n - 1 - k]);
            ok = cur.index(fw ? 2 : sp.nvalues, v);
            if (ok) c[at] = fw ? (v == 0 ? sp.amplitude : -sp.amplitude) : sp.values[v];
        }
    } else if (sp.family == LZX_SC_BIASED_TERNARY) {
        // a 10-bit coin below the permille threshold makes the slot non-zero, a
        // second word's low bit its sign (:458-481)
        const std::vector<uint8_t> s = sc_stream("bias3", seed, seed_len, n * 6 + 32);
        ScCursor cur{s};
        for (size_t i = 0; i < n && ok; i++) {
            uint32_t coin = 0, sgn = 0;
            ok = cur.word32(coin);
            if (!ok || (coin & 0x3FF) >= sp.p_nonzero_permille) continue;
            ok = cur.word32(sgn);
            if (ok) c[i] = (sgn & 1) ? -1 : 1;
        }
    } else {
        const std::vector<uint8_t> s = sc_stream("smallset", seed, seed_len, n * 8 + 32);
        ScCursor cur{s};
        for (size_t i = 0; i < n && ok; i++) {
            uint64_t v = 0;
            ok = cur.index(sp.nvalues, v);
            if (ok) c[i] = sp.values[v];
        }
    }
    if (!ok) return LZX_SHORT_CHALLENGE_REJECTION_BUDGET;
    unsigned __int128 l2 = 0; // certified_gamma (:271-277)
    for (size_t i = 0; i < n; i++) l2 += (unsigned __int128)((__int128)c[i] * c[i]);
    *gamma = ceil_sqrt128(l2);
    return LZX_OK;
}

} // namespace lzx

using namespace lzx;

extern "C" lzx_status lzx_short_challenge_sample(const lzx_short_challenge_spec *spec, const uint8_t *seed,
                                                 size_t seed_len, int64_t *coeffs_out, uint64_t *gamma_out)
{
    if (!spec || (seed_len && !seed) || !gamma_out || (spec->n && !coeffs_out)) return LZX_INVALID_ARGUMENT;
    return sc_sample(*spec, seed, seed_len, coeffs_out, gamma_out);
}

extern "C" lzx_status lzx_short_challenge_sample_with_gamma_cap(const lzx_short_challenge_spec *spec,
                                                                const uint8_t *seed, size_t seed_len, uint64_t cap,
                                                                size_t retries, int64_t *coeffs_out,
                                                                uint64_t *gamma_out)
{
    if (!spec || (seed_len && !seed) || !gamma_out || (spec->n && !coeffs_out)) return LZX_INVALID_ARGUMENT;
    if (cap == 0) return LZX_SHORT_CHALLENGE_INVALID_PARAMETERS; // short_challenge.rs:214-216
    std::vector<uint8_t> dom(seed_len + 8);
    if (seed_len) memcpy(dom.data(), seed, seed_len);
    for (size_t attempt = 0; attempt <= retries; attempt++) { // seed || u64le(attempt) per try (:218-231)
        for (int i = 0; i < 8; i++) dom[seed_len + i] = (uint8_t)((uint64_t)attempt >> (8 * i));
        const lzx_status st = sc_sample(*spec, dom.data(), dom.size(), coeffs_out, gamma_out);
        if (st != LZX_OK) return st;
        if (*gamma_out <= cap) return LZX_OK;
    }
    return LZX_SHORT_CHALLENGE_GAMMA_CAP_EXCEEDED; // *gamma_out holds the last Γ
}

extern "C" lzx_status lzx_rq2_sample_challenge(uint32_t q, uint32_t log_n, const lzx_short_challenge_spec *spec,
                                               const uint8_t *seed, size_t seed_len, uint32_t *out_host,
                                               uint64_t *g0, uint64_t *g1)
{
    if (!spec || (seed_len && !seed) || !out_host || !g0 || !g1 || q < 2 || log_n > 20) return LZX_INVALID_ARGUMENT;
    const size_t n = (size_t)1 << log_n;
    if (spec->n != 2 * n) return LZX_SHORT_CHALLENGE_INVALID_PARAMETERS; // extension.rs:185-187
    std::vector<int64_t> c(2 * n);
    uint64_t gam = 0;
    const lzx_status st = sc_sample(*spec, seed, seed_len, c.data(), &gam);
    if (st != LZX_OK) return st;
    unsigned __int128 l0 = 0, l1 = 0;
    for (size_t i = 0; i < n; i++) { // even stream positions -> c0, odd -> c1 (RingElement::from_signed)
        const int64_t a = c[2 * i], b = c[2 * i + 1];
        const int64_t ra = a % (int64_t)q, rb = b % (int64_t)q;
        out_host[i] = (uint32_t)(ra < 0 ? ra + (int64_t)q : ra);
        out_host[n + i] = (uint32_t)(rb < 0 ? rb + (int64_t)q : rb);
        l0 += (unsigned __int128)((__int128)a * a);
        l1 += (unsigned __int128)((__int128)b * b);
    }
    *g0 = ceil_sqrt128(l0);
    *g1 = ceil_sqrt128(l1);
    return LZX_OK;
}

extern "C" uint64_t lzx_ceil_sqrt_u128(uint64_t lo, uint64_t hi)
{
    return ceil_sqrt128((unsigned __int128)hi << 64 | lo);
}

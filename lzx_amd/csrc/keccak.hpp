// Keccak-f[1600], the SHAKE-256 sponge and the LZX XOF / seeded samplers.
//
// Restates crates/lattice-core/src/keccak.rs (FIPS 202 permutation, sponge
// with caller-supplied domain byte), transcript.rs:229-237 (xof framing) and
// crates/lattice-ring/src/ring.rs:102-180 (uniform_from_seed).  Written as
// host/device inline code so key derivation can run on either side.
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "modarith.hpp"

namespace lzx {

LZX_HD uint64_t rotl(uint64_t x, unsigned r) { return r ? (x << r) | (x >> (64 - r)) : x; }

LZX_HD void keccak_f1600(uint64_t s[25])
{
    constexpr uint64_t RC[24] = {
        0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
        0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
        0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
        0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
        0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
        0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};
    // rho offset and pi destination for lane (x + 5y): B[y + 5((2x+3y) mod 5)]
    constexpr unsigned RHO[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43,
                                  25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
    for (int round = 0; round < 24; round++) {
        uint64_t c0 = s[0] ^ s[5] ^ s[10] ^ s[15] ^ s[20];
        uint64_t c1 = s[1] ^ s[6] ^ s[11] ^ s[16] ^ s[21];
        uint64_t c2 = s[2] ^ s[7] ^ s[12] ^ s[17] ^ s[22];
        uint64_t c3 = s[3] ^ s[8] ^ s[13] ^ s[18] ^ s[23];
        uint64_t c4 = s[4] ^ s[9] ^ s[14] ^ s[19] ^ s[24];
        const uint64_t d[5] = {c4 ^ rotl(c1, 1), c0 ^ rotl(c2, 1), c1 ^ rotl(c3, 1), c2 ^ rotl(c4, 1),
                               c3 ^ rotl(c0, 1)};
        uint64_t b[25];
        for (int y = 0; y < 5; y++)
            for (int x = 0; x < 5; x++) {
                const int src = x + 5 * y;
                b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl(s[src] ^ d[x], RHO[src]);
            }
        for (int y = 0; y < 5; y++)
            for (int x = 0; x < 5; x++)
                s[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
        s[0] ^= RC[round];
    }
}

// ---- Keccak-f[1600] across the lanes of one wave, bit-interleaved ----------
//
// Lane L < 50 holds one half of state word w = L / 2 = x + 5y in the
// bit-interleaved form: h = L % 2 selects the even (h = 0) or odd (h = 1) bits
// of the 64-bit word, packed into 32 bits.  A 64-bit rotation by r then moves
// each half by a 32-bit rotation (r even: same half, r/2; r odd: the other
// half, (r+1)/2 into the even half and (r-1)/2 into the odd half), so every
// exchange and every ALU operation is 32-bit.  A round is two exchanges:
// theta gathers the halves of columns x-1 (half h) and x+1 (half 1-h, since
// rotl64(C, 1) swaps the halves), and rho + pi + chi gathers the three
// post-theta halves that land at (x, y), (x+1, y), (x+2, y) under pi
// (B[y][2x+3y] = rotl64(A[x][y], RHO)) and rotates them.  Lanes 50..63
// compute values that are never used.

// even bits of v in the low 32 bits, odd bits in the high 32 bits
LZX_HD constexpr uint64_t bit_unshuffle64(uint64_t x)
{
    uint64_t t = (x ^ (x >> 1)) & 0x2222222222222222ULL;
    x ^= t ^ (t << 1);
    t = (x ^ (x >> 2)) & 0x0C0C0C0C0C0C0C0CULL;
    x ^= t ^ (t << 2);
    t = (x ^ (x >> 4)) & 0x00F000F000F000F0ULL;
    x ^= t ^ (t << 4);
    t = (x ^ (x >> 8)) & 0x0000FF000000FF00ULL;
    x ^= t ^ (t << 8);
    t = (x ^ (x >> 16)) & 0x00000000FFFF0000ULL;
    x ^= t ^ (t << 16);
    return x;
}

// inverse of bit_unshuffle64
LZX_HD constexpr uint64_t bit_shuffle64(uint64_t x)
{
    uint64_t t = (x ^ (x >> 16)) & 0x00000000FFFF0000ULL;
    x ^= t ^ (t << 16);
    t = (x ^ (x >> 8)) & 0x0000FF000000FF00ULL;
    x ^= t ^ (t << 8);
    t = (x ^ (x >> 4)) & 0x00F000F000F000F0ULL;
    x ^= t ^ (t << 4);
    t = (x ^ (x >> 2)) & 0x0C0C0C0C0C0C0C0CULL;
    x ^= t ^ (t << 2);
    t = (x ^ (x >> 1)) & 0x2222222222222222ULL;
    x ^= t ^ (t << 1);
    return x;
}

struct KLanePlan {
    uint8_t cm[5];  // lanes of column x-1, half h
    uint8_t cp[5];  // lanes of column x+1, half 1-h
    uint8_t rot_d;  // rotation of the gathered x+1 half (1 into the even half, 0 into the odd)
    uint8_t src[3]; // lanes of the post-theta halves landing at (x, y), (x+1, y), (x+2, y)
    uint8_t rot[3]; // their left rotations
};

LZX_HD KLanePlan klane_plan(uint32_t lane)
{
    constexpr uint8_t RHO[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43,
                                 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
    const uint32_t L = lane < 50 ? lane : 0;
    const uint32_t w = L >> 1, h = L & 1, x = w % 5, y = w / 5;
    KLanePlan p{};
    for (uint32_t k = 0; k < 5; k++) {
        p.cm[k] = (uint8_t)(2 * ((x + 4) % 5 + 5 * k) + h);
        p.cp[k] = (uint8_t)(2 * ((x + 1) % 5 + 5 * k) + (1 - h));
    }
    p.rot_d = h == 0 ? 1 : 0;
    for (uint32_t j = 0; j < 3; j++) {
        const uint32_t X = (x + j) % 5, Y = y;
        // pi: position (X, Y) receives word (x', y') with y' = X, x' = 3 (Y - 3X) mod 5
        const uint32_t sw = (3 * (Y + 15 - 3 * X)) % 5 + 5 * X;
        const uint32_t r = RHO[sw];
        if ((r & 1) == 0) {
            p.src[j] = (uint8_t)(2 * sw + h);
            p.rot[j] = (uint8_t)(r / 2);
        } else {
            p.src[j] = (uint8_t)(2 * sw + (1 - h));
            p.rot[j] = (uint8_t)(h == 0 ? (r + 1) / 2 : (r - 1) / 2);
        }
    }
    return p;
}

LZX_HD uint32_t rotl32(uint32_t x, uint32_t s) { return s ? (x << s) | (x >> (32 - s)) : x; }

// Round constants in the interleaved form: [round] = (even half) | (odd half) << 32.
struct KRoundConsts {
    uint64_t v[24];
};
LZX_HD constexpr KRoundConsts keccak_rc_interleaved()
{
    constexpr uint64_t RC[24] = {
        0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
        0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
        0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
        0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
        0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
        0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};
    KRoundConsts k{};
    for (int i = 0; i < 24; i++) k.v[i] = bit_unshuffle64(RC[i]);
    return k;
}

// One Keccak round for a lane, given the gathered halves (shared by the
// device permutation and the host emulation used to test the plan).
LZX_HD uint32_t klane_theta(uint32_t a, const uint32_t (&gm)[5], const uint32_t (&gp)[5], const KLanePlan &p)
{
    const uint32_t dm = gm[0] ^ gm[1] ^ gm[2] ^ gm[3] ^ gm[4];
    const uint32_t dp = gp[0] ^ gp[1] ^ gp[2] ^ gp[3] ^ gp[4];
    return a ^ dm ^ rotl32(dp, p.rot_d);
}

LZX_HD uint32_t klane_chi(const uint32_t (&b)[3], const KLanePlan &p)
{
    const uint32_t b0 = rotl32(b[0], p.rot[0]), b1 = rotl32(b[1], p.rot[1]), b2 = rotl32(b[2], p.rot[2]);
    return b0 ^ (~b1 & b2);
}

__device__ __forceinline__ uint32_t lane_get(uint32_t v, uint32_t src)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}

__device__ __forceinline__ uint32_t keccak_f1600_lanes(uint32_t a, uint32_t lane)
{
    constexpr KRoundConsts RCI = keccak_rc_interleaved();
    const KLanePlan p = klane_plan(lane);
    // iota touches word 0: its even half on lane 0, its odd half on lane 1
    const uint32_t m0 = lane == 0 ? ~0u : 0u, m1 = lane == 1 ? ~0u : 0u;
#pragma unroll
    for (int round = 0; round < 24; round++) {
        uint32_t gm[5], gp[5];
#pragma unroll
        for (int k = 0; k < 5; k++) {
            gm[k] = lane_get(a, p.cm[k]);
            gp[k] = lane_get(a, p.cp[k]);
        }
        a = klane_theta(a, gm, gp, p);
        uint32_t b[3];
#pragma unroll
        for (int j = 0; j < 3; j++) b[j] = lane_get(a, p.src[j]);
        a = klane_chi(b, p) ^ ((m0 & (uint32_t)RCI.v[round]) | (m1 & (uint32_t)(RCI.v[round] >> 32)));
    }
    return a;
}

// Sponge with rate 136 (keccak.rs:241-356).  suffix 0x1F = SHAKE-256,
// 0x06 = SHA3-256.
struct Sponge {
    uint64_t st[25];
    uint32_t pos;
    uint8_t suffix;

    LZX_HD void init(uint8_t sfx)
    {
        for (int i = 0; i < 25; i++) st[i] = 0;
        pos = 0;
        suffix = sfx;
    }
    LZX_HD void absorb(const uint8_t *data, size_t len)
    {
        for (size_t i = 0; i < len; i++) {
            st[pos >> 3] ^= (uint64_t)data[i] << (8 * (pos & 7));
            if (++pos == 136) {
                keccak_f1600(st);
                pos = 0;
            }
        }
    }
    LZX_HD void absorb_u32le(uint32_t v)
    {
        uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
        absorb(b, 4);
    }
    // pad10*1 with the domain byte, then squeeze out_len bytes (keccak.rs:334-355).
    LZX_HD void finalize(uint8_t *out, size_t out_len)
    {
        st[pos >> 3] ^= (uint64_t)suffix << (8 * (pos & 7));
        st[135 >> 3] ^= 0x80ULL << (8 * (135 & 7));
        keccak_f1600(st);
        size_t done = 0;
        while (done < out_len) {
            size_t take = out_len - done < 136 ? out_len - done : 136;
            for (size_t i = 0; i < take; i++) out[done + i] = (uint8_t)(st[i >> 3] >> (8 * (i & 7)));
            done += take;
            if (done < out_len) keccak_f1600(st);
        }
    }
};

// Transcript::xof (transcript.rs:229-237) absorb phase:
// "LZXX" || u32le|domain| || domain || u32le|msg| || msg.
LZX_HD void xof_begin(Sponge &sp, const uint8_t *domain, size_t dlen, size_t mlen)
{
    sp.init(0x1f);
    const uint8_t tag[4] = {'L', 'Z', 'X', 'X'};
    sp.absorb(tag, 4);
    sp.absorb_u32le((uint32_t)dlen);
    sp.absorb(domain, dlen);
    sp.absorb_u32le((uint32_t)mlen);
}

} // namespace lzx

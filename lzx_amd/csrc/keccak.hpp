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

// One negacyclic NTT held in a lane's VGPRs (the butterflies the batched NTT
// and the fused extension-ring kernels share); tables as laid out by TabLayout.
#pragma once

#include "ntt.hpp"

namespace lzx {

__host__ __device__ constexpr uint32_t brv_bits(uint32_t x, uint32_t bits)
{
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

__device__ __forceinline__ uint32_t brv_dev(uint32_t x, uint32_t bits)
{
    return bits == 0 ? 0 : (__brev(x) >> (32 - bits));
}

// Forward (merged psi twiddles, CT from natural order, output permuted to
// natural-order evaluations) and inverse (GS, n^{-1} folded into the last
// level) of one polynomial held in VGPRs.  a[] is the working order: the
// forward reads coefficients into a[c]; the inverse reads evaluation k into
// a[brv(k)].
template <int LOG_N, int AR, int MODE>
__device__ __forceinline__ void ntt_regs(uint32_t (&a)[1u << LOG_N], const uint32_t *__restrict__ table)
{
    // the tables are read-only for a kernel's life: reading them through the
    // constant address space makes the wave-uniform twiddles scalar loads
    // (counted by lgkmcnt, so they never wait behind the tile's DMA loads)
    const __attribute__((address_space(4))) uint32_t *tab = (const __attribute__((address_space(4))) uint32_t *)table;
    constexpr uint32_t N = 1u << LOG_N;
    const QConst Q{tab[0], tab[1], tab[5], tab[6]};
    if (MODE == NTT_FWD) {
#pragma unroll
        for (uint32_t len = N / 2; len >= 1; len >>= 1) {
#pragma unroll
            for (uint32_t s = 0; s < N; s += 2 * len) {
                const uint32_t k = N / (2 * len) + s / (2 * len);
                const uint32_t w = tab[TabLayout::fwd_w(N) + k];
                const uint32_t wp = tab[TabLayout::fwd_wp(N) + k];
#pragma unroll
                for (uint32_t j = s; j < s + len; j++) {
                    const uint32_t t = mulq<AR>(a[j + len], w, wp, Q);
                    const uint32_t u = a[j];
                    a[j] = addq<AR>(u, t, Q);
                    a[j + len] = subq<AR>(u, t, Q);
                }
            }
        }
    } else {
#pragma unroll
        for (uint32_t len = 1; len <= N / 2; len <<= 1) {
#pragma unroll
            for (uint32_t s = 0; s < N; s += 2 * len) {
                const uint32_t k = N / (2 * len) + s / (2 * len);
                if (len == N / 2) {
                    const uint32_t nw = tab[8], nwp = tab[9], lw = tab[10], lwp = tab[11];
#pragma unroll
                    for (uint32_t j = s; j < s + len; j++) {
                        const uint32_t u = a[j], v = a[j + len];
                        a[j] = mulq<AR>(addq<AR>(u, v, Q), nw, nwp, Q);
                        a[j + len] = mulq<AR>(subq<AR>(u, v, Q), lw, lwp, Q);
                    }
                } else {
                    const uint32_t w = tab[TabLayout::inv_w(N) + k];
                    const uint32_t wp = tab[TabLayout::inv_wp(N) + k];
#pragma unroll
                    for (uint32_t j = s; j < s + len; j++) {
                        const uint32_t u = a[j], v = a[j + len];
                        a[j] = addq<AR>(u, v, Q);
                        a[j + len] = mulq<AR>(subq<AR>(u, v, Q), w, wp, Q);
                    }
                }
            }
        }
    }
}

// ---- tiles of 64 polynomials, shared by the one-wave register kernels --------

template <uint32_t C>
__device__ __forceinline__ uint32_t tile_slot(uint32_t p, uint32_t c)
{
    return C * p + ((c + p) & (C - 1));
}

// Chunk c (16 bytes) of polynomial p of a 64-polynomial tile lives at LDS
// slot tile_slot(p, c): a lane reading its own polynomial's chunks, and a
// wave writing 64 consecutive slots, are both bank-conflict free.

// For blocks of one wave: lanes see each other's LDS writes in issue order (a
// wave's LDS operations execute in order), so a wave-scope fence and a
// code-motion barrier order them.  A workgroup barrier's release fence would
// also wait for the wave's in-flight global stores.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Chunk i of a lane's prefetch: the 16-byte piece of LDS slot s = 64 i + lane,
// i.e. chunk ((s mod C) - p) mod C of polynomial p = s / C (the rotated layout
// of tile_slot), read from HBM with the same coalesced addressing.  Slots past
// the tile's last polynomial reload polynomial 0 (in bounds; never stored), so
// every load is unconditional and the array stays in VGPRs.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4))); // promotable to VGPRs across loop iterations

template <uint32_t C>
__device__ __forceinline__ void tile_fetch(const uint32_t *__restrict__ g, uint32_t np, u32x4 (&r)[C], uint32_t lane)
{
#pragma unroll
    for (uint32_t i = 0; i < C; i++) {
        const uint32_t s = 64 * i + lane, p = s / C, c = ((s % C) - p) & (C - 1);
        const uint32_t pp = p < np ? p : 0;
        r[i] = *reinterpret_cast<const u32x4 *>(g + (pp * C + c) * 4);
    }
}

} // namespace lzx

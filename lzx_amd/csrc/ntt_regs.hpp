// One negacyclic NTT held in a lane's VGPRs (the butterflies the batched NTT
// and the fused extension-ring kernels share); tables as laid out by TabLayout.
#pragma once

#include <type_traits>

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

// ---- lazily reduced u32 transforms (q < 2^31) ---------------------------------
//
// Shoup's product a w - mulhi(a, w') q lies in [0, 2q) for every 32-bit a, so
// values need not be canonical between levels.  Two classes:
//   HV = false, q < 2^31: values in [0, 2q) (2q < 2^32).  A CT butterfly
//     reduces its X input and the product to [0, q) (one subtract and one
//     unsigned min each) and emits X + T and X - T + q, both in [0, 2q): 3
//     multiplies + 8 ALU operations against 3 + 9 canonical;
//   HV = true, q < 2^30 (4q < 2^32, Harvey's range): values in [0, 4q); X is
//     reduced to [0, 2q), the product is used as it comes in [0, 2q), and the
//     outputs X + T and X - T + 2q are in [0, 4q): 3 multiplies + 6.
// The inverse (GS) keeps its inputs in [0, 2q): for HV it forms U + V in
// [0, 4q) and reduces it once, for !HV it reduces both inputs to [0, q).
// The first level's inputs are canonical and skip their reductions; every
// output leaves canonical.

// x in [0, 2c) -> x mod c: the unsigned min picks x - c exactly when x >= c
__device__ __forceinline__ uint32_t red_once(uint32_t x, uint32_t c)
{
    const uint32_t t = x - c;
    return t < x ? t : x;
}

// a w mod q up to one multiple of q: in [0, 2q) for every 32-bit a (Shoup)
__device__ __forceinline__ uint32_t shoup_lazy(uint32_t a, uint32_t w, uint32_t wp, uint32_t q)
{
    return a * w - __umulhi(a, wp) * q;
}

template <int LOG_N, bool HV, int MODE>
__device__ __forceinline__ void ntt_regs_lazy(uint32_t (&a)[1u << LOG_N], const uint32_t *__restrict__ table)
{
    const __attribute__((address_space(4))) uint32_t *tab = (const __attribute__((address_space(4))) uint32_t *)table;
    constexpr uint32_t N = 1u << LOG_N;
    const uint32_t q = tab[0], q2 = 2 * q;
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
                    const bool first = len == N / 2;
                    if (HV) {
                        const uint32_t x = first ? a[j] : red_once(a[j], q2);
                        const uint32_t t = shoup_lazy(a[j + len], w, wp, q);
                        a[j] = x + t;
                        a[j + len] = x - t + q2;
                    } else {
                        const uint32_t x = first ? a[j] : red_once(a[j], q);
                        const uint32_t t = red_once(shoup_lazy(a[j + len], w, wp, q), q);
                        a[j] = x + t;
                        a[j + len] = x - t + q;
                    }
                }
            }
        }
#pragma unroll
        for (uint32_t i = 0; i < N; i++) a[i] = HV ? red_once(red_once(a[i], q2), q) : red_once(a[i], q);
    } else {
#pragma unroll
        for (uint32_t len = 1; len <= N / 2; len <<= 1) {
#pragma unroll
            for (uint32_t s = 0; s < N; s += 2 * len) {
                const uint32_t k = N / (2 * len) + s / (2 * len);
                const bool first = len == 1, last = len == N / 2;
                const uint32_t w = last ? tab[8] : tab[TabLayout::inv_w(N) + k];
                const uint32_t wp = last ? tab[9] : tab[TabLayout::inv_wp(N) + k];
                const uint32_t lw = tab[10], lwp = tab[11];
#pragma unroll
                for (uint32_t j = s; j < s + len; j++) {
                    uint32_t u = a[j], v = a[j + len];
                    if (!HV && !first) {
                        u = red_once(u, q);
                        v = red_once(v, q);
                    }
                    const uint32_t d = HV ? u - v + q2 : u - v + q;
                    if (last) {
                        a[j] = red_once(shoup_lazy(u + v, w, wp, q), q);
                        a[j + len] = red_once(shoup_lazy(d, lw, lwp, q), q);
                    } else {
                        a[j] = HV ? red_once(u + v, q2) : u + v;
                        a[j + len] = shoup_lazy(d, w, wp, q);
                    }
                }
            }
        }
    }
}

// ---- FP64 register transforms (any q < 2^32) ----------------------------------
//
// Residues are integer doubles.  For |a| < 2^40 and a constant 0 <= w < q,
// a w = hi + lo exactly (hi = fl(a w), lo = fma(a, w, -hi)); t = rint(a fl(w/q))
// is within 1/2 + 2^-11 of a w / q, so r = fma(-t, q, hi) + lo = a w - t q
// exactly, |r| <= q/2 + 2^-11 q: six FP64 operations in place of Montgomery's
// four 32-bit multiplies and its 33-bit corrections.  Sums and differences are
// left unreduced: the forward adds at most (1/2 + 2^-11) q per level to |x|
// (about 4q after six levels from input below q), the inverse at most doubles
// it (32q before its last level, which multiplies both outputs).  Every value
// stays below 2^38 (2^40 for any u32 input), so nothing rounds.  (No FMA contraction: each step is
// written as the rounding it needs.)

__device__ __forceinline__ double f64_mulc(double a, double w, double wq, double q)
{
#pragma clang fp contract(off)
    const double hi = a * w, lo = __fma_rn(a, w, -hi);
    const double t = __builtin_rint(a * wq);
    return __fma_rn(-t, q, hi) + lo;
}

// a b mod q up to a multiple of q, for integers with |a b| < 2^100: the
// quotient estimate rint(fl(a b) fl(1/q)) is within 1/2 + |a b| 1.5 2^-52 / q
// of a b / q, so |result| <= q/2 + |a b| 1.5 2^-52 and neither step rounds
__device__ __forceinline__ double f64_mul(double a, double b, double q, double qinv)
{
#pragma clang fp contract(off)
    const double hi = a * b, lo = __fma_rn(a, b, -hi);
    const double t = __builtin_rint(hi * qinv);
    return __fma_rn(-t, q, hi) + lo;
}

// x mod q in [0, q) for integer |x| < 2^50: floor(x fl(1/q)) is exact unless
// x = k q, where it may give k - 1 and leave exactly q
__device__ __forceinline__ uint32_t f64_canonical(double x, double q, double qinv)
{
#pragma clang fp contract(off)
    const double r = __fma_rn(-__builtin_floor(x * qinv), q, x);
    return (uint32_t)(r >= q ? 0.0 : r);
}

// One level of length LEN over the blocks in [S0, S1) (every index a template
// constant, so the array stays in VGPRs).
typedef const __attribute__((address_space(4))) double *f64_tab_t;

template <uint32_t N, uint32_t LEN, uint32_t S0, uint32_t S1>
__device__ __forceinline__ void f64_fwd_level(double (&a)[N], f64_tab_t dt, double q)
{
#pragma unroll
    for (uint32_t s = S0; s < S1; s += 2 * LEN) {
        const uint32_t k = N / (2 * LEN) + s / (2 * LEN);
        const double w = dt[TabLayout::dbl_fwd(N) + 2 * k], wq = dt[TabLayout::dbl_fwd(N) + 2 * k + 1];
#pragma unroll
        for (uint32_t j = s; j < s + LEN; j++) {
            const double t = f64_mulc(a[j + LEN], w, wq, q);
            const double u = a[j];
            a[j] = u + t;
            a[j + LEN] = u - t;
        }
    }
}

// levels LEN, LEN/2, ..., 1 over [S0, S1)
template <uint32_t N, uint32_t LEN, uint32_t S0, uint32_t S1>
__device__ __forceinline__ void f64_fwd_levels(double (&a)[N], f64_tab_t dt, double q)
{
    if constexpr (LEN >= 1) {
        f64_fwd_level<N, LEN, S0, S1>(a, dt, q);
        f64_fwd_levels<N, LEN / 2, S0, S1>(a, dt, q);
    }
}

template <uint32_t N, uint32_t LEN, uint32_t S0, uint32_t S1>
__device__ __forceinline__ void f64_inv_level(double (&a)[N], f64_tab_t dt, double q)
{
#pragma unroll
    for (uint32_t s = S0; s < S1; s += 2 * LEN) {
        const uint32_t k = N / (2 * LEN) + s / (2 * LEN);
        if constexpr (LEN == N / 2) {
            const double nw = dt[2], nwq = dt[3], lw = dt[4], lwq = dt[5];
#pragma unroll
            for (uint32_t j = s; j < s + LEN; j++) {
                const double u = a[j], v = a[j + LEN];
                a[j] = f64_mulc(u + v, nw, nwq, q);
                a[j + LEN] = f64_mulc(u - v, lw, lwq, q);
            }
        } else {
            const double w = dt[TabLayout::dbl_inv(N) + 2 * k], wq = dt[TabLayout::dbl_inv(N) + 2 * k + 1];
#pragma unroll
            for (uint32_t j = s; j < s + LEN; j++) {
                const double u = a[j], v = a[j + LEN];
                a[j] = u + v;
                a[j + LEN] = f64_mulc(u - v, w, wq, q);
            }
        }
    }
}

// levels LEN, 2 LEN, ..., TOP over [S0, S1)
template <uint32_t N, uint32_t LEN, uint32_t TOP, uint32_t S0, uint32_t S1>
__device__ __forceinline__ void f64_inv_levels(double (&a)[N], f64_tab_t dt, double q)
{
    if constexpr (LEN >= 1 && LEN <= TOP) {
        f64_inv_level<N, LEN, S0, S1>(a, dt, q);
        f64_inv_levels<N, 2 * LEN, TOP, S0, S1>(a, dt, q);
    }
}

// the same transforms as ntt_regs, on doubles; the outputs are lazy (reduce
// them with f64_canonical).  After the forward's first level (before the
// inverse's last) the two halves are independent transforms: they run one
// after the other, with the scheduler kept from interleaving them, so only
// one half's working set is live beside the other half's values.
template <int LOG_N, int MODE>
__device__ __forceinline__ void ntt_regs_f64(double (&a)[1u << LOG_N], const uint32_t *__restrict__ table)
{
    constexpr uint32_t N = 1u << LOG_N;
    const f64_tab_t dt = (f64_tab_t)(table + TabLayout::dbl(N));
    const double q = dt[0];
    if (MODE == NTT_FWD) {
        f64_fwd_level<N, N / 2, 0, N>(a, dt, q);
        if constexpr (N >= 8) __builtin_amdgcn_sched_barrier(0);
        f64_fwd_levels<N, N / 4, 0, N / 2>(a, dt, q);
        if constexpr (N >= 8) __builtin_amdgcn_sched_barrier(0);
        f64_fwd_levels<N, N / 4, N / 2, N>(a, dt, q);
    } else {
        f64_inv_levels<N, 1, N / 4, 0, N / 2>(a, dt, q);
        if constexpr (N >= 8) __builtin_amdgcn_sched_barrier(0);
        f64_inv_levels<N, 1, N / 4, N / 2, N>(a, dt, q);
        if constexpr (N >= 8) __builtin_amdgcn_sched_barrier(0);
        f64_inv_level<N, N / 2, 0, N>(a, dt, q);
    }
}

// value type of a register transform's working array
template <int AR>
using ntt_val_t = typename std::conditional<AR == ARITH_F64, double, uint32_t>::type;

template <int LOG_N, int AR, int MODE>
__device__ __forceinline__ void ntt_regs_any(ntt_val_t<AR> (&a)[1u << LOG_N], const uint32_t *__restrict__ table)
{
    if constexpr (AR == ARITH_F64)
        ntt_regs_f64<LOG_N, MODE>(a, table);
    else
        ntt_regs<LOG_N, AR, MODE>(a, table);
}

// a working value -> canonical u32
template <int AR>
__device__ __forceinline__ uint32_t ntt_out(ntt_val_t<AR> v, const uint32_t *__restrict__ table, uint32_t n)
{
    if constexpr (AR == ARITH_F64) {
        const __attribute__((address_space(4))) double *dt =
            (const __attribute__((address_space(4))) double *)(table + TabLayout::dbl(n));
        return f64_canonical(v, dt[0], dt[1]);
    } else {
        return v;
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

// 16-byte store at lane byte offset off from the wave-uniform tile base g,
// through a buffer resource over the tile's bytes with system-scope cache
// policy (sc0 sc1, operand 0x11 on gfx950): the store writes through L2, so a
// launch ends with no dirty lines for its release to write back.
__device__ __forceinline__ void tile_store_wt(void *g, uint32_t bytes, uint32_t off, u32x4 v)
{
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(g, 0, (int)bytes, 0x00020000);
    __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, 0x11);
}

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

// Batched negacyclic NTT over Z_q[X]/(X^n+1) for gfx950.
//
// Reference semantics (crates/lattice-ring/src/ntt.rs):
//   forward  (ntt.rs:118-182): out[k] = sum_i a_i psi^(i(2k+1)), natural order,
//            psi = g^((q-1) >> (log_n+1)) (modulus.rs:149-161);
//   inverse  (ntt.rs:187-250): the exact inverse;
//   forward_partial / complete_partial (ntt.rs:257-316): the reference's own
//            level structure (prescale by psi^i, bit-reverse, CT levels with
//            omega = psi^2), whose intermediate state is part of the contract.
//
// Kernels
//   ntt_pipe_kernel n in [4, 64] (default): one polynomial per lane, all
//       log n levels in VGPRs (no cross-lane traffic), twiddles wave-uniform
//       (scalar loads from the constant address space); persistent waves
//       stream 64-polynomial tiles, the next tile's loads in flight in VGPRs
//       while the current one is transformed (see below).  psi is folded into
//       the twiddles; the forward permutes to natural order on the way back to
//       LDS and the inverse folds n^{-1} into its last level.
//   ntt_reg_kernel  the same butterflies with a plain LDS-staged tile per
//       block, for buffers that are not 16-byte aligned.
//   ntt_lds_kernel  any n up to 2^13 and the incomplete transforms: a tile of
//       polynomials in LDS, one butterfly per thread per level.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>

#include "ntt.hpp"
#include "ntt_mfma.hpp"
#include "ntt_regs.hpp"

namespace lzx {

// LDS-staged register kernel (fallback for buffers that are not 16-byte
// aligned): one 64-polynomial tile per block, padded LDS rows.
template <int LOG_N, int AR, int MODE>
__global__ __launch_bounds__(64) void ntt_reg_kernel(uint32_t *__restrict__ data, TabSet tabs, uint32_t polys,
                                                     uint32_t blocks_per_plan)
{
    constexpr uint32_t N = 1u << LOG_N;
    constexpr uint32_t S = N + 1;
    __shared__ uint32_t tile[64 * S];
    const uint32_t plan = blockIdx.x / blocks_per_plan;
    const uint32_t blk = blockIdx.x - plan * blocks_per_plan;
    const uint32_t *__restrict__ tab = tabs.p[plan];
    const uint32_t p0 = blk * 64;
    const uint32_t np = min(64u, polys - p0);
    const uint32_t tid = threadIdx.x;
    uint32_t *__restrict__ g = data + ((size_t)plan * polys + p0) * N;
    const uint32_t nw = np * N;
    for (uint32_t i = tid; i < nw; i += 64) tile[(i / N) * S + (i % N)] = g[i];
    __syncthreads();
    ntt_val_t<AR> a[N];
    uint32_t *row = tile + tid * S;
#pragma unroll
    for (uint32_t c = 0; c < N; c++) a[MODE == NTT_FWD ? c : brv_bits(c, LOG_N)] = row[c];
    ntt_regs_any<LOG_N, AR, MODE>(a, tab);
#pragma unroll
    for (uint32_t c = 0; c < N; c++) row[c] = ntt_out<AR>(a[MODE == NTT_FWD ? brv_bits(c, LOG_N) : c], tab, N);
    __syncthreads();
    for (uint32_t i = tid; i < nw; i += 64) g[i] = tile[(i / N) * S + (i % N)];
}

// Persistent register kernel (default for n in [4, 64]).  A tile is 64
// polynomials of C = n/4 16-byte chunks.  Each lane loads C chunks of the next
// tile with coalesced 16-byte loads (one instruction moves 1 KiB of
// consecutive polynomials) into VGPRs while the wave runs the butterflies of
// the current tile; at the top of the next iteration the chunks go to LDS,
// where chunk c of polynomial p lives at slot C*p + ((c + p) mod C), so the
// transposed per-lane reads (lane = polynomial) and writes are bank-conflict
// free.  Results go back through the same slots and leave as coalesced
// 16-byte stores.  One 16 KiB tile per wave: at n = 64 (Shoup) two waves per
// SIMD are resident, and the grid takes exactly the resident waves.
// Prefetching into VGPRs (not global_load_lds) lets the compiler wait only for
// the registers it stages; an LDS DMA would make it wait for every DMA before
// any LDS access, the next tile's included.
template <int LOG_N, int AR, int MODE>
__global__ __launch_bounds__(64) void ntt_pipe_kernel(uint32_t *__restrict__ data, TabSet tabs, uint32_t polys,
                                                      uint32_t blocks_per_plan, uint32_t total)
{
    constexpr uint32_t N = 1u << LOG_N, C = N / 4;
    __shared__ __attribute__((aligned(16))) uint32_t T[64 * N];
    const uint32_t lane = threadIdx.x;
    // each wave owns a contiguous range of tiles, so it stays on one modulus
    // (one twiddle table in the scalar cache) for most of its tiles
    const uint32_t per = (total + gridDim.x - 1) / gridDim.x;
    uint32_t b = blockIdx.x * per;
    const uint32_t b_end = min(total, b + per);
    if (b >= b_end) return;
    auto locate = [&](uint32_t bb, uint32_t &np, uint32_t &plan) {
        plan = bb / blocks_per_plan;
        const uint32_t p0 = (bb - plan * blocks_per_plan) * 64;
        np = min(64u, polys - p0);
        return data + ((size_t)plan * polys + p0) * N;
    };
    // PF: the next tile's loads are in flight in VGPRs while this one is
    // transformed.  The FP64 class holds a polynomial in 2n VGPRs, so at n = 64
    // a prefetch buffer would leave one wave per SIMD; it loads each tile at
    // the top of its iteration instead and relies on two resident waves.
    constexpr bool PF = AR != ARITH_F64;
    uint32_t np, plan;
    u32x4 nx[C]; // the tile being staged (PF: the next tile, in flight)
    // The compiler's waits before staging nx count the vector memory ops issued
    // after each load; on the steady path those include the previous tile's C
    // stores, so the stores stay in flight.  Every other way into the loop
    // header (this prologue, a partial tile) ends with nothing pending, so it
    // cannot lower those counts.
    constexpr int WAIT_VM0 = 0x0F70; // s_waitcnt vmcnt(0), lgkmcnt and expcnt free
    if (PF) {
        tile_fetch<C>(locate(b, np, plan), np, nx, lane);
        __builtin_amdgcn_s_waitcnt(WAIT_VM0);
    }
    for (; b < b_end; b++) {
        uint32_t *g = locate(b, np, plan);
        const uint32_t *__restrict__ tab = tabs.p[plan];
        if (!PF) tile_fetch<C>(g, np, nx, lane);
#pragma unroll
        for (uint32_t i = 0; i < C; i++) *reinterpret_cast<u32x4 *>(T + 4 * (64 * i + lane)) = nx[i];
        wave_sync();
        ntt_val_t<AR> a[N];
#pragma unroll
        for (uint32_t c4 = 0; c4 < C; c4++) {
            const uint4 v = *reinterpret_cast<const uint4 *>(T + 4 * tile_slot<C>(lane, c4));
            const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (uint32_t e = 0; e < 4; e++) a[MODE == NTT_FWD ? 4 * c4 + e : brv_bits(4 * c4 + e, LOG_N)] = vv[e];
        }
        if (PF && b + 1 < b_end) {
            uint32_t npn, pn;
            const uint32_t *gn = locate(b + 1, npn, pn);
            tile_fetch<C>(gn, npn, nx, lane);
        }
        ntt_regs_any<LOG_N, AR, MODE>(a, tab);
        wave_sync(); // every lane has read its row before the results overwrite the tile
#pragma unroll
        for (uint32_t c4 = 0; c4 < C; c4++) {
            uint32_t vv[4];
#pragma unroll
            for (uint32_t e = 0; e < 4; e++)
                vv[e] = ntt_out<AR>(a[MODE == NTT_FWD ? brv_bits(4 * c4 + e, LOG_N) : 4 * c4 + e], tab, N);
            *reinterpret_cast<uint4 *>(T + 4 * tile_slot<C>(lane, c4)) = make_uint4(vv[0], vv[1], vv[2], vv[3]);
        }
        wave_sync();
        if (np == 64) {
#pragma unroll
            for (uint32_t i = 0; i < C; i++) {
                const uint32_t s = 64 * i + lane, p = s / C, c = ((s % C) - p) & (C - 1);
                tile_store_wt(g, 64 * N * 4, (p * C + c) * 16, *reinterpret_cast<const u32x4 *>(T + 4 * s));
            }
        } else {
#pragma unroll
            for (uint32_t i = 0; i < C; i++) {
                const uint32_t s = 64 * i + lane, p = s / C, c = ((s % C) - p) & (C - 1);
                if (p < np)
                    tile_store_wt(g, np * N * 4, (p * C + c) * 16, *reinterpret_cast<const u32x4 *>(T + 4 * s));
            }
            __builtin_amdgcn_s_waitcnt(WAIT_VM0);
        }
        wave_sync(); // the stores have read the tile before the next one is staged
    }
}

// Persistent register kernel of the u32 classes below 2^31 (the default for
// n in [4, 64] at those primes): the lazily reduced butterflies of
// ntt_regs_lazy, Harvey's [0, 4q) range where the plan's flag (q < 2^30)
// allows it.  A wave runs its tiles in one loop per class and switches loops
// where its tile range crosses into a plan of the other class, so a CRT batch
// whose primes straddle 2^30 is still one launch (each loop keeps its own
// registers: one loop holding both bodies spilled).  The tile goes through
// LDS in padded rows of n + 4 words: polynomial p's chunk c sits at word
// (n + 4) p + 4 c.  The coalesced staging writes (16 lanes per polynomial row)
// and the per-lane row reads both touch distinct banks in every 16-lane group,
// and every global and LDS address is a lane constant plus an immediate, so a
// full tile spends no VALU on addressing.  The next tile's loads are in flight
// in VGPRs while the current one is transformed.  The last tile of a plan may
// be partial: its loads clamp to polynomial 0 and its stores are masked.
typedef __attribute__((address_space(1))) u32x4 g_u32x4;
typedef __attribute__((address_space(1))) uint8_t g_u8;

template <uint32_t N>
struct LazyTile {
    static constexpr uint32_t C = N / 4, ROW = N + 4, PPI = 64 / C, GROUPS = (C + 3) / 4;
    uint32_t *data;
    const TabSet *tabs;
    uint32_t polys, blocks_per_plan;
    uint32_t lane, stage, row;
    uint32_t voff[GROUPS]; // lane byte offsets of each group of four 1 KiB loads

    // a tile's base, its polynomial count and plan (wave-uniform: readfirstlane
    // keeps them in SGPRs, since the division runs on the VALU)
    __device__ __forceinline__ g_u8 *locate(uint32_t bb, uint32_t &np, uint32_t &plan) const
    {
        plan = (uint32_t)__builtin_amdgcn_readfirstlane(bb / blocks_per_plan);
        const uint32_t p0 = (bb - plan * blocks_per_plan) * 64;
        np = min(64u, polys - p0);
        const uint64_t g = (uint64_t)(data + ((size_t)plan * polys + p0) * N);
        // (readfirstlane returns a signed int: zero-extend it, never sign-extend)
        const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)g);
        const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(g >> 32));
        return (g_u8 *)((hi << 32) | lo);
    }
    __device__ __forceinline__ g_u32x4 *full(g_u8 *g, uint32_t i) const
    {
        return (g_u32x4 *)(g + voff[i / 4] + 1024 * (i % 4));
    }
    __device__ __forceinline__ g_u32x4 *part(g_u8 *g, uint32_t i, uint32_t p) const
    {
        return (g_u32x4 *)(g + 16 * (C * p + lane % C));
    }
    __device__ __forceinline__ void fetch(g_u8 *g, uint32_t np, u32x4 (&r)[C]) const
    {
        if (np == 64) {
#pragma unroll
            for (uint32_t i = 0; i < C; i++) r[i] = *full(g, i);
        } else {
#pragma unroll
            for (uint32_t i = 0; i < C; i++) {
                const uint32_t p = PPI * i + lane / C;
                r[i] = *part(g, i, p < np ? p : 0);
            }
        }
    }
};

constexpr int WAIT_VM0 = 0x0F70; // s_waitcnt vmcnt(0), lgkmcnt and expcnt free

// Tiles [b, b_end) of the wave while their plan's class is HV; nx holds tile
// b's chunks on entry (and the next tile's on return).
template <int LOG_N, int MODE, bool HV>
__device__ __forceinline__ void lazy_run(const LazyTile<1u << LOG_N> &L, uint32_t *__restrict__ T, uint32_t &b,
                                         uint32_t b_end, u32x4 (&nx)[(1u << LOG_N) / 4])
{
    constexpr uint32_t N = 1u << LOG_N, C = N / 4, ROW = N + 4, PPI = 64 / C;
    for (; b < b_end; b++) {
        uint32_t np, plan;
        g_u8 *g = L.locate(b, np, plan);
        const uint32_t *__restrict__ tab = L.tabs->p[plan];
        if ((((const __attribute__((address_space(4))) uint32_t *)tab)[12] != 0) != HV) return;
#pragma unroll
        for (uint32_t i = 0; i < C; i++) *reinterpret_cast<u32x4 *>(T + L.stage + ROW * PPI * i) = nx[i];
        wave_sync();
        uint32_t a[N];
#pragma unroll
        for (uint32_t c4 = 0; c4 < C; c4++) {
            const uint4 v = *reinterpret_cast<const uint4 *>(T + L.row + 4 * c4);
            const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (uint32_t e = 0; e < 4; e++) a[MODE == NTT_FWD ? 4 * c4 + e : brv_bits(4 * c4 + e, LOG_N)] = vv[e];
        }
        if (b + 1 < b_end) {
            uint32_t npn, pn;
            g_u8 *gn = L.locate(b + 1, npn, pn);
            L.fetch(gn, npn, nx);
        }
        ntt_regs_lazy<LOG_N, HV, MODE>(a, tab);
        wave_sync(); // every lane has read its row before the results overwrite the tile
#pragma unroll
        for (uint32_t c4 = 0; c4 < C; c4++) {
            uint32_t vv[4];
#pragma unroll
            for (uint32_t e = 0; e < 4; e++) vv[e] = a[MODE == NTT_FWD ? brv_bits(4 * c4 + e, LOG_N) : 4 * c4 + e];
            *reinterpret_cast<uint4 *>(T + L.row + 4 * c4) = make_uint4(vv[0], vv[1], vv[2], vv[3]);
        }
        wave_sync();
        // The next tile's loads were issued before the butterflies and have
        // landed by now.  Waiting for them here, before this tile's stores are
        // issued, leaves only stores in flight at the loop head, so staging the
        // next tile does not wait for them: loads and stores share one counter
        // and may complete out of order with respect to each other, so a wait
        // for loads issued before stores also waits for the stores.
        __builtin_amdgcn_s_waitcnt(WAIT_VM0);
        if (np == 64) {
#pragma unroll
            for (uint32_t i = 0; i < C; i++)
                tile_store_wt((uint8_t *)g, 64 * N * 4, L.voff[i / 4] + 1024 * (i % 4),
                              *reinterpret_cast<const u32x4 *>(T + L.stage + ROW * PPI * i));
        } else {
#pragma unroll
            for (uint32_t i = 0; i < C; i++) {
                const uint32_t p = PPI * i + L.lane / C;
                if (p < np)
                    tile_store_wt((uint8_t *)g, np * N * 4, 16 * (C * p + L.lane % C),
                                  *reinterpret_cast<const u32x4 *>(T + L.stage + ROW * PPI * i));
            }
            __builtin_amdgcn_s_waitcnt(WAIT_VM0);
        }
        wave_sync(); // the stores have read the tile before the next one is staged
    }
}

// skew: the second half of the grid (the second resident wave on each SIMD
// under round-robin dispatch) starts skew x 1024 cycles late, so the first
// tiles of the first half arrive in half the time and their butterflies
// overlap the second half's loads instead of every wave waiting on one burst.
template <int LOG_N, int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void ntt_lazy_kernel(
    uint32_t *__restrict__ data, TabSet tabs, uint32_t polys, uint32_t blocks_per_plan, uint32_t total,
    uint32_t skew)
{
    constexpr uint32_t N = 1u << LOG_N, C = N / 4;
    __shared__ __attribute__((aligned(16))) uint32_t T[64 * (N + 4)];
    LazyTile<N> L;
    L.data = data;
    L.tabs = &tabs;
    L.polys = polys;
    L.blocks_per_plan = blocks_per_plan;
    L.lane = threadIdx.x;
    L.stage = (N + 4) * (L.lane / C) + 4 * (L.lane % C); // load i's chunk: LDS word stage + ROW PPI i
    L.row = (N + 4) * L.lane;                           // this lane's polynomial
#pragma unroll
    for (uint32_t k = 0; k < LazyTile<N>::GROUPS; k++) {
        // hidden from the optimiser, which would otherwise fold the group offset
        // into a per-lane 64-bit address and an immediate too large to encode
        L.voff[k] = 16 * L.lane + 4096 * k;
        asm volatile("" : "+v"(L.voff[k]));
    }
    const uint32_t per = (total + gridDim.x - 1) / gridDim.x;
    uint32_t b = blockIdx.x * per;
    const uint32_t b_end = min(total, b + per);
    if (b >= b_end) return;
    if (blockIdx.x >= gridDim.x / 2)
        for (uint32_t z = 0; z < skew; z++) __builtin_amdgcn_s_sleep(16);
    u32x4 nx[C];
    {
        uint32_t np, plan;
        L.fetch(L.locate(b, np, plan), np, nx);
    }
    __builtin_amdgcn_s_waitcnt(WAIT_VM0);
    while (b < b_end) {
        lazy_run<LOG_N, MODE, false>(L, T, b, b_end, nx);
        lazy_run<LOG_N, MODE, true>(L, T, b, b_end, nx);
    }
}

// Forward NTT of small-digit polynomials (i8 in, u32 NTT-domain out): the
// witness transform of the digit-input Ajtai commit with the digit -> residue
// conversion fused into the load (ring.rs:232-243 from_signed).  One
// 64-polynomial tile per block: each lane turns one coalesced 16-byte digit
// chunk (16 coefficients) into four residue chunks in the conflict-free LDS
// slots, then runs the register butterflies of one polynomial; results leave
// through the same slots as coalesced 16-byte stores.  n >= 16.
template <int LOG_N, int AR>
__global__ __launch_bounds__(64) void ntt_digits_kernel(const int8_t *__restrict__ digits, uint32_t *__restrict__ out,
                                                        const uint32_t *__restrict__ tab, uint32_t npolys)
{
    constexpr uint32_t N = 1u << LOG_N, C = N / 4, D = N / 16; // 16-byte output / digit chunks per polynomial
    __shared__ __attribute__((aligned(16))) uint32_t T[64 * N];
    const uint32_t lane = threadIdx.x;
    const uint32_t p0 = blockIdx.x * 64, np = min(64u, npolys - p0);
    const uint32_t q = tab[0];
    const int8_t *__restrict__ dg = digits + (size_t)p0 * N;
    uint32_t *__restrict__ g = out + (size_t)p0 * N;
#pragma unroll
    for (uint32_t i = 0; i < D; i++) {
        const uint32_t d = 64 * i + lane, p = d / D, c0 = 4 * (d % D);
        if (p < np) {
            const uint4 v = *reinterpret_cast<const uint4 *>(dg + 16 * (size_t)d);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                uint32_t r[4];
#pragma unroll
                for (uint32_t e = 0; e < 4; e++) {
                    const int x = (int8_t)(w[j] >> (8 * e));
                    r[e] = x < 0 ? q - (uint32_t)(-x) : (uint32_t)x;
                }
                *reinterpret_cast<uint4 *>(T + 4 * tile_slot<C>(p, c0 + j)) = make_uint4(r[0], r[1], r[2], r[3]);
            }
        }
    }
    __syncthreads();
    ntt_val_t<AR> a[N];
#pragma unroll
    for (uint32_t c4 = 0; c4 < C; c4++) {
        const uint4 v = *reinterpret_cast<const uint4 *>(T + 4 * tile_slot<C>(lane, c4));
        a[4 * c4] = v.x, a[4 * c4 + 1] = v.y, a[4 * c4 + 2] = v.z, a[4 * c4 + 3] = v.w;
    }
    ntt_regs_any<LOG_N, AR, NTT_FWD>(a, tab);
#pragma unroll
    for (uint32_t c4 = 0; c4 < C; c4++) {
        uint32_t vv[4];
#pragma unroll
        for (uint32_t e = 0; e < 4; e++) vv[e] = ntt_out<AR>(a[brv_bits(4 * c4 + e, LOG_N)], tab, N);
        *reinterpret_cast<uint4 *>(T + 4 * tile_slot<C>(lane, c4)) = make_uint4(vv[0], vv[1], vv[2], vv[3]);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < C; i++) {
        const uint32_t s = 64 * i + lane, p = s / C, c = ((s % C) - p) & (C - 1);
        if (p < np) *reinterpret_cast<uint4 *>(g + (p * C + c) * 4) = *reinterpret_cast<const uint4 *>(T + 4 * s);
    }
}

template <int AR, int MODE>
__global__ __launch_bounds__(256) void ntt_lds_kernel(uint32_t *__restrict__ data, TabSet tabs, uint32_t polys,
                                                      uint32_t blocks_per_plan, uint32_t log_n, uint32_t ppb,
                                                      uint32_t levels)
{
    extern __shared__ uint32_t tile[];
    const uint32_t n = 1u << log_n;
    const uint32_t half = n >> 1;
    const uint32_t plan = blockIdx.x / blocks_per_plan;
    const uint32_t blk = blockIdx.x - plan * blocks_per_plan;
    const uint32_t *__restrict__ tab = tabs.p[plan];
    const uint32_t p0 = blk * ppb;
    const uint32_t np = min(ppb, polys - p0);
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    uint32_t *__restrict__ g = data + ((size_t)plan * polys + p0) * n;
    const uint32_t nwords = np * n;
    const QConst Q{tab[0], tab[1], tab[5], tab[6]};

    for (uint32_t i = tid; i < nwords; i += nt) {
        const uint32_t p = i >> log_n, c = i & (n - 1);
        const uint32_t x = g[i];
        if (MODE == NTT_INV) {
            tile[p * n + brv_dev(c, log_n)] = x;
        } else if (MODE == NTT_PARTIAL) {
            // ntt.rs:262-265: pre-scale by psi^c, then bit-reverse
            tile[p * n + brv_dev(c, log_n)] =
                mulq<AR>(x, tab[TabLayout::psi_w(n) + c], tab[TabLayout::psi_wp(n) + c], Q);
        } else {
            tile[p * n + c] = x;
        }
    }
    __syncthreads();

    const uint32_t nbf = np * half;
    if (MODE == NTT_FWD) {
        for (uint32_t len = half; len >= 1; len >>= 1) {
            for (uint32_t b = tid; b < nbf; b += nt) {
                const uint32_t p = b >> (log_n - 1), bi = b & (half - 1);
                const uint32_t grp = bi / len, j = bi % len;
                const uint32_t k = n / (2 * len) + grp;
                const uint32_t i0 = p * n + grp * 2 * len + j, i1 = i0 + len;
                const uint32_t t = mulq<AR>(tile[i1], tab[TabLayout::fwd_w(n) + k],
                                              tab[TabLayout::fwd_wp(n) + k], Q);
                const uint32_t u = tile[i0];
                tile[i0] = addq<AR>(u, t, Q);
                tile[i1] = subq<AR>(u, t, Q);
            }
            __syncthreads();
        }
    } else if (MODE == NTT_INV) {
        for (uint32_t len = 1; len <= half; len <<= 1) {
            for (uint32_t b = tid; b < nbf; b += nt) {
                const uint32_t p = b >> (log_n - 1), bi = b & (half - 1);
                const uint32_t grp = bi / len, j = bi % len;
                const uint32_t k = n / (2 * len) + grp;
                const uint32_t i0 = p * n + grp * 2 * len + j, i1 = i0 + len;
                const uint32_t u = tile[i0], v = tile[i1];
                if (len == half) {
                    tile[i0] = mulq<AR>(addq<AR>(u, v, Q), tab[8], tab[9], Q);
                    tile[i1] = mulq<AR>(subq<AR>(u, v, Q), tab[10], tab[11], Q);
                } else {
                    tile[i0] = addq<AR>(u, v, Q);
                    tile[i1] = mulq<AR>(subq<AR>(u, v, Q), tab[TabLayout::inv_w(n) + k],
                                          tab[TabLayout::inv_wp(n) + k], Q);
                }
            }
            __syncthreads();
        }
    } else {
        // ntt.rs:267-284 (partial: len = 2 .. 2^levels) and ntt.rs:299-314
        // (complete: len = 2^(levels+1) .. n), twiddle omega^((n/len) j).
        uint32_t len = MODE == NTT_PARTIAL ? 2u : (1u << (levels + 1));
        const uint32_t max_levels = MODE == NTT_PARTIAL ? levels : 64u;
        for (uint32_t applied = 0; len <= n && applied < max_levels; len <<= 1, applied++) {
            const uint32_t hl = len >> 1, stride = n / len;
            for (uint32_t b = tid; b < nbf; b += nt) {
                const uint32_t p = b >> (log_n - 1), bi = b & (half - 1);
                const uint32_t grp = bi / hl, j = bi % hl;
                const uint32_t i0 = p * n + grp * len + j, i1 = i0 + hl;
                const uint32_t e = stride * j;
                const uint32_t u = tile[i0];
                const uint32_t v = mulq<AR>(tile[i1], tab[TabLayout::omg_w(n) + e], tab[TabLayout::omg_wp(n) + e], Q);
                tile[i0] = addq<AR>(u, v, Q);
                tile[i1] = subq<AR>(u, v, Q);
            }
            __syncthreads();
        }
    }

    for (uint32_t i = tid; i < nwords; i += nt) {
        const uint32_t p = i >> log_n, c = i & (n - 1);
        g[i] = MODE == NTT_FWD ? tile[p * n + brv_dev(c, log_n)] : tile[p * n + c];
    }
}

// out[i] = out[i] * b[i] mod q (general residues), RingElement::mul's
// pointwise step (ring.rs:329-331).
__global__ void pointwise_mul_kernel(uint32_t *__restrict__ out, const uint32_t *__restrict__ b, size_t len,
                                     uint32_t q, uint64_t barrett)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x)
        out[i] = reduce_u64((uint64_t)out[i] * b[i], q, barrett);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// Resident one-wave blocks per CU of a persistent pipe kernel (its VGPR and
// LDS use), measured once per instantiation.
template <int LOG_N, int AR, int MODE>
static uint32_t pipe_blocks_per_cu()
{
    static const uint32_t nb = [] {
        int v = 0;
        const hipError_t e =
            AR == ARITH_SHOUP
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, ntt_lazy_kernel<LOG_N, MODE>, 64, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, ntt_pipe_kernel<LOG_N, AR, MODE>, 64, 0);
        return (uint32_t)(e == hipSuccess && v > 0 ? v : 4);
    }();
    return nb;
}

// Start skew of the lazy kernel's second half (units of 1024 cycles): 6 by
// measurement on the 4-prime x 2^16 x n = 64 batch (forward 29.5 -> 29.0 us,
// inverse 31.4 -> 29.0 us; 8 and above lose again); LZX_NTT_SKEW overrides.
static uint32_t ntt_skew()
{
    static const uint32_t v = [] {
        const char *e = getenv("LZX_NTT_SKEW");
        return e ? (uint32_t)strtoul(e, nullptr, 10) : 6u;
    }();
    return v;
}

// grid: for the persistent pipe kernel, the number of CUs (the launch takes
// as many waves as stay resident); otherwise one block per tile.
template <int LOG_N, int AR>
static void launch_reg(NttMode mode, bool dma, dim3 grid, uint32_t *data, const TabSet &ts, uint32_t polys,
                       uint32_t bpp, uint32_t total, hipStream_t stream)
{
    if (dma && AR == ARITH_SHOUP) {
        const uint32_t cus = grid.x;
        if (mode == NTT_FWD) {
            const dim3 g(std::min<uint32_t>(total, cus * pipe_blocks_per_cu<LOG_N, AR, NTT_FWD>()));
            hipLaunchKernelGGL((ntt_lazy_kernel<LOG_N, NTT_FWD>), g, dim3(64), 0, stream, data, ts, polys, bpp, total,
                               ntt_skew());
        } else {
            const dim3 g(std::min<uint32_t>(total, cus * pipe_blocks_per_cu<LOG_N, AR, NTT_INV>()));
            hipLaunchKernelGGL((ntt_lazy_kernel<LOG_N, NTT_INV>), g, dim3(64), 0, stream, data, ts, polys, bpp, total,
                               ntt_skew());
        }
    } else if (dma) {
        const uint32_t cus = grid.x;
        if (mode == NTT_FWD) {
            const dim3 g(std::min<uint32_t>(total, cus * pipe_blocks_per_cu<LOG_N, AR, NTT_FWD>()));
            hipLaunchKernelGGL((ntt_pipe_kernel<LOG_N, AR, NTT_FWD>), g, dim3(64), 0, stream, data, ts, polys, bpp,
                               total);
        } else {
            const dim3 g(std::min<uint32_t>(total, cus * pipe_blocks_per_cu<LOG_N, AR, NTT_INV>()));
            hipLaunchKernelGGL((ntt_pipe_kernel<LOG_N, AR, NTT_INV>), g, dim3(64), 0, stream, data, ts, polys, bpp,
                               total);
        }
    } else {
        if (mode == NTT_FWD)
            hipLaunchKernelGGL((ntt_reg_kernel<LOG_N, AR, NTT_FWD>), grid, dim3(64), 0, stream, data, ts, polys, bpp);
        else
            hipLaunchKernelGGL((ntt_reg_kernel<LOG_N, AR, NTT_INV>), grid, dim3(64), 0, stream, data, ts, polys, bpp);
    }
}

template <int AR>
static void dispatch_reg_n(uint32_t log_n, NttMode mode, bool dma, dim3 grid, uint32_t *data, const TabSet &ts,
                           uint32_t polys, uint32_t bpp, uint32_t total, hipStream_t stream)
{
    switch (log_n) {
    case 2: launch_reg<2, AR>(mode, dma, grid, data, ts, polys, bpp, total, stream); break;
    case 3: launch_reg<3, AR>(mode, dma, grid, data, ts, polys, bpp, total, stream); break;
    case 4: launch_reg<4, AR>(mode, dma, grid, data, ts, polys, bpp, total, stream); break;
    case 5: launch_reg<5, AR>(mode, dma, grid, data, ts, polys, bpp, total, stream); break;
    default: launch_reg<6, AR>(mode, dma, grid, data, ts, polys, bpp, total, stream); break;
    }
}

static void dispatch_reg(int arith, uint32_t log_n, NttMode mode, bool dma, dim3 grid, uint32_t *data,
                         const TabSet &ts, uint32_t polys, uint32_t bpp, uint32_t total, hipStream_t stream)
{
    // q > 2^31 runs the register kernels in exact FP64 (8 operations of ~4.3
    // clocks per butterfly against Montgomery's 4 multiplies of ~4.1 clocks and
    // 33-bit add/sub): 16% faster per pass at n = 64.  Below 2^31 a Shoup
    // butterfly (3 multiplies, 9 single-rate operations) costs as much as the
    // FP64 one and needs half the registers.
    if (arith == ARITH_MONT)
        dispatch_reg_n<ARITH_F64>(log_n, mode, dma, grid, data, ts, polys, bpp, total, stream);
    else
        dispatch_reg_n<ARITH_SHOUP>(log_n, mode, dma, grid, data, ts, polys, bpp, total, stream);
}

template <int AR>
static void dispatch_lds_mode(NttMode mode, dim3 grid, size_t shmem, uint32_t *data, const TabSet &ts,
                              uint32_t polys, uint32_t bpp, uint32_t log_n, uint32_t ppb, uint32_t levels,
                              hipStream_t stream)
{
    switch (mode) {
    case NTT_FWD:
        hipLaunchKernelGGL((ntt_lds_kernel<AR, NTT_FWD>), grid, dim3(256), shmem, stream, data, ts, polys, bpp, log_n,
                           ppb, levels);
        break;
    case NTT_INV:
        hipLaunchKernelGGL((ntt_lds_kernel<AR, NTT_INV>), grid, dim3(256), shmem, stream, data, ts, polys, bpp, log_n,
                           ppb, levels);
        break;
    case NTT_PARTIAL:
        hipLaunchKernelGGL((ntt_lds_kernel<AR, NTT_PARTIAL>), grid, dim3(256), shmem, stream, data, ts, polys, bpp,
                           log_n, ppb, levels);
        break;
    default:
        hipLaunchKernelGGL((ntt_lds_kernel<AR, NTT_COMPLETE>), grid, dim3(256), shmem, stream, data, ts, polys, bpp,
                           log_n, ppb, levels);
        break;
    }
}

static void dispatch_lds(int arith, NttMode mode, dim3 grid, size_t shmem, uint32_t *data, const TabSet &ts,
                         uint32_t polys, uint32_t bpp, uint32_t log_n, uint32_t ppb, uint32_t levels,
                         hipStream_t stream)
{
    if (arith == ARITH_MONT)
        dispatch_lds_mode<ARITH_MONT>(mode, grid, shmem, data, ts, polys, bpp, log_n, ppb, levels, stream);
    else
        dispatch_lds_mode<ARITH_SHOUP>(mode, grid, shmem, data, ts, polys, bpp, log_n, ppb, levels, stream);
}

// ---- the kernel class of an MFMA-capable launch, per shape -------------
//
// Both classes give the same transform.  A launch whose plans all hold MFMA
// tables runs the class their creation asked for (LZX_NTT_KERNEL), or, for
// the default AUTO, the class kept in the tuning cache for (device, moduli,
// polynomials per plan, direction).  With none kept yet, the call times both
// classes on a pool scratch copy of its data (the transforms are in place):
// one untimed launch each, then five launches of each class alternating, and
// keeps the class with the faster best launch; then it runs that class on the
// caller's data.  Under stream capture it takes MFMA unmeasured.
static std::string ntt_path_key(const lzx_ntt_plan *const *plans, size_t cnt, size_t polys, NttMode mode)
{
    std::string k = "ntt_path dev=" + std::to_string(plans[0]->device) + " n=" + std::to_string(plans[0]->n) +
                    " polys=" + std::to_string(polys) + (mode == NTT_FWD ? " fwd q=" : " inv q=");
    for (size_t c = 0; c < cnt; c++) k += std::to_string(plans[c]->q) + (c + 1 < cnt ? "," : "");
    return k;
}

static void launch_register(const lzx_ntt_plan *const *plans, size_t cnt, const TabSet &ts, uint32_t *chunk,
                            size_t polys, NttMode mode, hipStream_t stream)
{
    const uint32_t bpp = (uint32_t)((polys + 63) / 64);
    const uint32_t total = (uint32_t)(bpp * cnt);
    const bool dma = plans[0]->stage_dma;
    const dim3 grid(dma ? (uint32_t)plans[0]->cus : total);
    dispatch_reg(plans[0]->arith, plans[0]->log_n, mode, dma, grid, chunk, ts, (uint32_t)polys, bpp, total, stream);
}

static lzx_status choose_class(const lzx_ntt_plan *const *plans, size_t cnt, const TabSet &ts, uint32_t *chunk,
                               size_t polys, NttMode mode, hipStream_t stream, uint32_t &path)
{
    if (plans[0]->kernel_pref != LZX_NTT_PATH_AUTO) {
        path = plans[0]->kernel_pref;
        return LZX_OK;
    }
    const std::string key = ntt_path_key(plans, cnt, polys, mode);
    uint32_t kept = 0;
    if (autotune_get(key, kept) && (kept == LZX_NTT_PATH_MFMA || kept == LZX_NTT_PATH_REGISTER)) {
        path = kept;
        return LZX_OK;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    LZX_HIP(hipStreamIsCapturing(stream, &cs));
    path = LZX_NTT_PATH_MFMA;
    if (cs != hipStreamCaptureStatusNone) return LZX_OK;
    const size_t bytes = cnt * polys * plans[0]->n * sizeof(uint32_t);
    uint32_t *tmp = nullptr;
    LZX_HIP(scratch_alloc((void **)&tmp, bytes, stream));
    // the two classes alternate launch by launch, each between its own events,
    // and each class's fastest launch counts: clocks that are still ramping or
    // a neighbour's burst then bias neither class
    constexpr int PAIRS = 5, NEV = 2 * PAIRS + 1;
    hipEvent_t ev[NEV] = {};
    lzx_status st = LZX_OK;
    for (int k = 0; k < NEV && st == LZX_OK; k++)
        if (hipEventCreate(&ev[k]) != hipSuccess) st = device_fail(hipGetLastError(), "ntt path events");
    if (st == LZX_OK && hipMemcpyAsync(tmp, chunk, bytes, hipMemcpyDeviceToDevice, stream) != hipSuccess)
        st = device_fail(hipGetLastError(), "ntt path copy");
    auto run = [&](uint32_t cls) {
        if (cls == LZX_NTT_PATH_MFMA)
            ntt_mfma_launch(plans, cnt, tmp, (uint32_t)polys, mode, stream);
        else
            launch_register(plans, cnt, ts, tmp, polys, mode, stream);
    };
    float t_mf = 1e30f, t_reg = 1e30f;
    if (st == LZX_OK) {
        run(LZX_NTT_PATH_MFMA); // untimed: first-launch costs
        run(LZX_NTT_PATH_REGISTER);
        if (hipEventRecord(ev[0], stream) != hipSuccess) st = device_fail(hipGetLastError(), "event");
    }
    for (int k = 0; k < 2 * PAIRS && st == LZX_OK; k++) {
        run(k % 2 == 0 ? LZX_NTT_PATH_MFMA : LZX_NTT_PATH_REGISTER);
        if (hipEventRecord(ev[k + 1], stream) != hipSuccess) st = device_fail(hipGetLastError(), "event");
    }
    if (st == LZX_OK) {
        hipError_t e = hipEventSynchronize(ev[NEV - 1]);
        for (int k = 0; k < 2 * PAIRS && e == hipSuccess; k++) {
            float t = 0;
            e = hipEventElapsedTime(&t, ev[k], ev[k + 1]);
            float &best = k % 2 == 0 ? t_mf : t_reg;
            if (e == hipSuccess) best = std::min(best, t);
        }
        if (e != hipSuccess) st = device_fail(e, "ntt path timing");
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    (void)scratch_free(tmp, stream);
    if (st != LZX_OK) return st;
    path = t_mf <= t_reg ? LZX_NTT_PATH_MFMA : LZX_NTT_PATH_REGISTER;
    autotune_put(key, path);
    return LZX_OK;
}

static bool mfma_capable(const lzx_ntt_plan *const *plans, size_t cnt, NttMode mode, bool aligned)
{
    if (!(mode == NTT_FWD || mode == NTT_INV) || !aligned || plans[0]->log_n != 6) return false;
    for (size_t c = 0; c < cnt; c++)
        if (!plans[c]->mfma_dev) return false;
    return true;
}

lzx_status ntt_launch(const lzx_ntt_plan *const *plans, size_t nplans, uint32_t *data, size_t polys, NttMode mode,
                      uint32_t levels, hipStream_t stream)
{
    if (nplans == 0 || polys == 0) return LZX_OK;
    if (!data) return LZX_INVALID_ARGUMENT;
    const uint32_t log_n = plans[0]->log_n, n = plans[0]->n;
    if (polys > 0xFFFFFFFFull / 2) return LZX_UNSUPPORTED;
    const bool aligned = ((uintptr_t)data % 16) == 0;
    size_t i = 0;
    while (i < nplans) {
        // a launch covers consecutive plans of one arithmetic class
        const int arith = plans[i]->arith;
        size_t cnt = 0;
        TabSet ts{};
        while (i + cnt < nplans && cnt < (size_t)MAX_PLANS_PER_LAUNCH && plans[i + cnt]->arith == arith) {
            if (plans[i + cnt]->n != n) return LZX_INVALID_ARGUMENT;
            ts.p[cnt] = plans[i + cnt]->dev_tab;
            cnt++;
        }
        uint32_t *chunk = data + i * polys * n;
        const bool use_reg = (mode == NTT_FWD || mode == NTT_INV) && log_n >= 2 && log_n <= 6 && aligned;
        uint32_t path = LZX_NTT_PATH_REGISTER;
        if (mfma_capable(plans + i, cnt, mode, aligned)) {
            const lzx_status st = choose_class(plans + i, cnt, ts, chunk, polys, mode, stream, path);
            if (st != LZX_OK) return st;
        }
        if (path == LZX_NTT_PATH_MFMA) {
            ntt_mfma_launch(plans + i, cnt, chunk, (uint32_t)polys, mode, stream);
        } else if (use_reg) {
            launch_register(plans + i, cnt, ts, chunk, polys, mode, stream);
        } else {
            const uint32_t ppb = std::max<uint32_t>(1u, 2048u >> log_n);
            const uint32_t bpp = (uint32_t)((polys + ppb - 1) / ppb);
            const dim3 grid((uint32_t)(bpp * cnt));
            const size_t shmem = (size_t)ppb * n * sizeof(uint32_t);
            dispatch_lds(arith, mode, grid, shmem, chunk, ts, (uint32_t)polys, bpp, log_n, ppb, levels, stream);
        }
        LZX_LAUNCH_CHECK("ntt kernel launch");
        i += cnt;
    }
    return LZX_OK;
}

lzx_status ntt_batch_path(const lzx_ntt_plan *const *plans, size_t nplans, size_t polys, NttMode mode, uint32_t &path)
{
    // the first launch's plans (a call covers up to MAX_PLANS_PER_LAUNCH per launch)
    const size_t cnt = std::min<size_t>(MAX_PLANS_PER_LAUNCH, nplans);
    path = plans[0]->log_n <= 6 ? LZX_NTT_PATH_REGISTER : LZX_NTT_PATH_LDS;
    if (!mfma_capable(plans, cnt, mode, true)) return LZX_OK;
    if (plans[0]->kernel_pref != LZX_NTT_PATH_AUTO) {
        path = plans[0]->kernel_pref;
        return LZX_OK;
    }
    uint32_t kept = 0;
    path = autotune_get(ntt_path_key(plans, cnt, polys, mode), kept) ? kept : LZX_NTT_PATH_AUTO;
    return LZX_OK;
}

// Host table construction restating NttTables::new (ntt.rs:46-106).
static void build_tables(lzx_ntt_plan *p)
{
    const uint32_t q = p->q, n = p->n, log_n = p->log_n;
    std::vector<uint32_t> &t = p->host_tab;
    t.assign(TabLayout::words(n), 0);
    auto put = [&](uint32_t off_w, uint32_t off_wp, uint32_t idx, uint32_t w) {
        if (p->bigq) {
            t[off_w + idx] = to_mont(w, q);
            t[off_wp + idx] = 0;
        } else {
            t[off_w + idx] = w;
            t[off_wp + idx] = shoup_companion(w, q);
        }
    };
    t[0] = q;
    t[1] = p->mod.qinv;
    t[2] = n;
    t[3] = log_n;
    t[4] = p->bigq ? 1 : 0;
    t[5] = p->prime_c;
    t[6] = p->prime_k;
    t[7] = (uint32_t)p->arith;
    t[12] = q < (1u << 30) ? 1 : 0; // Harvey's lazy range [0, 4q) fits 32 bits
    put(8, 9, 0, p->n_inv);
    const uint32_t z1inv = host_pow(p->psi_inv, brv_bits(1, log_n), q);
    put(10, 11, 0, (uint32_t)((uint64_t)z1inv * p->n_inv % q));
    for (uint32_t k = 1; k < n; k++) {
        const uint32_t e = brv_bits(k, log_n);
        put(TabLayout::fwd_w(n), TabLayout::fwd_wp(n), k, host_pow(p->psi, e, q));
        put(TabLayout::inv_w(n), TabLayout::inv_wp(n), k, host_pow(p->psi_inv, e, q));
    }
    // FP64 constants: each w (< 2^32, exact) beside fl(w / q)
    const double qd = (double)q;
    auto dset = [&](uint32_t off, double v) { memcpy(&t[TabLayout::dbl(n) + 2 * off], &v, sizeof v); };
    auto dput = [&](uint32_t off, uint64_t w) { dset(off, (double)w), dset(off + 1, (double)w / qd); };
    dset(0, qd);
    dset(1, 1.0 / qd);
    dput(2, p->n_inv);
    dput(4, (uint64_t)z1inv * p->n_inv % q);
    for (uint32_t k = 1; k < n; k++) {
        const uint32_t e = brv_bits(k, log_n);
        dput(TabLayout::dbl_fwd(n) + 2 * k, host_pow(p->psi, e, q));
        dput(TabLayout::dbl_inv(n) + 2 * k, host_pow(p->psi_inv, e, q));
    }
    uint64_t pw = 1;
    for (uint32_t i = 0; i < n; i++) {
        put(TabLayout::psi_w(n), TabLayout::psi_wp(n), i, (uint32_t)pw);
        pw = pw * p->psi % q;
    }
    const uint32_t omega = (uint32_t)((uint64_t)p->psi * p->psi % q);
    uint64_t ow = 1;
    for (uint32_t e = 0; e < (n > 1 ? n / 2 : 1); e++) {
        put(TabLayout::omg_w(n), TabLayout::omg_wp(n), e, (uint32_t)ow);
        ow = ow * omega % q;
    }
}

lzx_status ntt_forward_digits(const lzx_ntt_plan *plan, const int8_t *digits, uint32_t *out, size_t polys,
                              hipStream_t stream)
{
    if (polys == 0) return LZX_OK;
    if (plan->log_n < 4 || plan->log_n > 6 || ((uintptr_t)digits % 16) || ((uintptr_t)out % 16))
        return LZX_UNSUPPORTED;
    if (polys > 0xFFFFFFFFull - 63) return LZX_UNSUPPORTED;
    const dim3 grid((uint32_t)((polys + 63) / 64));
    const uint32_t np = (uint32_t)polys;
#define LZX_DIGITS_LAUNCH(LN, AR)                                                                                   \
    hipLaunchKernelGGL((ntt_digits_kernel<LN, AR>), grid, dim3(64), 0, stream, digits, out, plan->dev_tab, np)
    const bool big = plan->arith == ARITH_MONT; // q > 2^31: the FP64 register transform
    switch (plan->log_n) {
    case 4: if (big) LZX_DIGITS_LAUNCH(4, ARITH_F64); else LZX_DIGITS_LAUNCH(4, ARITH_SHOUP); break;
    case 5: if (big) LZX_DIGITS_LAUNCH(5, ARITH_F64); else LZX_DIGITS_LAUNCH(5, ARITH_SHOUP); break;
    default: if (big) LZX_DIGITS_LAUNCH(6, ARITH_F64); else LZX_DIGITS_LAUNCH(6, ARITH_SHOUP); break;
    }
#undef LZX_DIGITS_LAUNCH
    LZX_LAUNCH_CHECK("ntt_digits_kernel");
    return LZX_OK;
}

static lzx_status run_host(const lzx_ntt_plan *plan, uint32_t *a_host, size_t len, NttMode mode, uint32_t levels)
{
    if (!plan || (!a_host && len)) return LZX_INVALID_ARGUMENT;
    if (len % plan->n) return LZX_NTT_LENGTH_MISMATCH;
    if (len == 0) return LZX_OK;
    const size_t bytes = len * sizeof(uint32_t);
    return host_slice_call(plan->device, a_host, bytes, a_host, 0, bytes, bytes, [&](void *d, hipStream_t s) {
        return ntt_launch(&plan, 1, (uint32_t *)d, len / plan->n, mode, levels, s);
    });
}

} // namespace lzx

using namespace lzx;

extern "C" lzx_status lzx_ntt_plan_create(uint32_t q, uint32_t two_adicity, uint32_t generator, uint32_t log_n,
                                          lzx_ntt_plan **out)
{
    if (!out) return LZX_INVALID_ARGUMENT;
    *out = nullptr;
    if (log_n + 1 > two_adicity) return LZX_NTT_ADICITY_TOO_SMALL; // ntt.rs:47-50
    if (q < 3 || (q & 1) == 0) return LZX_INVALID_ARGUMENT;
    if (log_n > NTT_MAX_LOG_N) return LZX_UNSUPPORTED;
    lzx_ntt_plan *p = new lzx_ntt_plan();
    p->q = q;
    p->two_adicity = two_adicity;
    p->generator = generator;
    p->log_n = log_n;
    p->n = 1u << log_n;
    p->bigq = q >= (1u << 31);
    p->prime_k = (uint32_t)__builtin_ctz(q - 1);
    p->prime_c = (q - 1) >> p->prime_k;
    p->arith = p->bigq ? ARITH_MONT : ARITH_SHOUP;
    p->mod.q = q;
    p->mod.qinv = host_inv32(q);
    p->mod.barrett = UINT64_MAX / q;
    // modulus.rs:149-161 and ntt.rs:52-58
    p->psi = host_pow(generator, ((uint64_t)q - 1) >> (log_n + 1), q);
    p->psi_inv = host_pow(p->psi, (uint64_t)q - 2, q);
    p->n_inv = host_pow(p->n, (uint64_t)q - 2, q);
    {
        // LZX_NTT_STAGE=lds selects the LDS-staged fallback kernel everywhere
        const char *stage = getenv("LZX_NTT_STAGE");
        p->stage_dma = !(stage && strcmp(stage, "lds") == 0);
    }
    build_tables(p);
    p->dev_tab = nullptr;
    hipError_t e = hipGetDevice(&p->device);
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device);
    p->cus = std::max(cus, 1);
    if (e == hipSuccess) e = hipMalloc(&p->dev_tab, p->host_tab.size() * sizeof(uint32_t));
    if (e == hipSuccess)
        e = hipMemcpy(p->dev_tab, p->host_tab.data(), p->host_tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    p->mfma_dev = nullptr;
    // LZX_NTT_MFMA=0: no MFMA tables (the butterfly kernels only);
    // LZX_NTT_KERNEL=mfma|register: that class for every MFMA-capable launch
    const char *mf = getenv("LZX_NTT_MFMA");
    const char *kp = getenv("LZX_NTT_KERNEL");
    p->kernel_pref = kp && strcmp(kp, "mfma") == 0       ? LZX_NTT_PATH_MFMA
                     : kp && strcmp(kp, "register") == 0 ? LZX_NTT_PATH_REGISTER
                                                         : LZX_NTT_PATH_AUTO;
    if (e == hipSuccess && ntt_mfma_supported(p) && !(mf && strcmp(mf, "0") == 0)) {
        const std::vector<uint8_t> mt = ntt_mfma_tables(p);
        if (!mt.empty()) {
            e = hipMalloc(&p->mfma_dev, mt.size());
            if (e == hipSuccess) e = hipMemcpy(p->mfma_dev, mt.data(), mt.size(), hipMemcpyHostToDevice);
        }
    }
    if (e != hipSuccess) {
        lzx_status st = device_fail(e, "ntt plan upload");
        if (p->dev_tab) (void)hipFree(p->dev_tab);
        if (p->mfma_dev) (void)hipFree(p->mfma_dev);
        delete p;
        return st;
    }
    *out = p;
    return LZX_OK;
}

extern "C" void lzx_ntt_plan_destroy(lzx_ntt_plan *plan)
{
    if (!plan) return;
    if (plan->dev_tab) (void)hipFree(plan->dev_tab);
    if (plan->mfma_dev) (void)hipFree(plan->mfma_dev);
    delete plan;
}

extern "C" lzx_status lzx_ntt_batch_path(const lzx_ntt_plan *const *plans, size_t nplans, size_t polys_per_plan,
                                         uint32_t inverse, uint32_t *path)
{
    if (!plans || !path || nplans == 0) return LZX_INVALID_ARGUMENT;
    for (size_t i = 0; i < nplans; i++)
        if (!plans[i] || plans[i]->n != plans[0]->n) return LZX_INVALID_ARGUMENT;
    return ntt_batch_path(plans, nplans, polys_per_plan, inverse ? NTT_INV : NTT_FWD, *path);
}

extern "C" lzx_status lzx_ntt_plan_constants(const lzx_ntt_plan *plan, uint32_t *psi, uint32_t *n_inv)
{
    if (!plan) return LZX_INVALID_ARGUMENT;
    if (psi) *psi = plan->psi;
    if (n_inv) *n_inv = plan->n_inv;
    return LZX_OK;
}

static lzx_status batch(const lzx_ntt_plan *plan, uint32_t *data, size_t len, NttMode mode, uint32_t levels,
                        void *stream)
{
    if (!plan) return LZX_INVALID_ARGUMENT;
    if (len % plan->n) return LZX_NTT_LENGTH_MISMATCH;
    return ntt_launch(&plan, 1, data, len / plan->n, mode, levels, (hipStream_t)stream);
}

extern "C" lzx_status lzx_ntt_forward_batch(const lzx_ntt_plan *plan, uint32_t *data_dev, size_t len, void *stream)
{
    return batch(plan, data_dev, len, NTT_FWD, 0, stream);
}

extern "C" lzx_status lzx_ntt_inverse_batch(const lzx_ntt_plan *plan, uint32_t *data_dev, size_t len, void *stream)
{
    return batch(plan, data_dev, len, NTT_INV, 0, stream);
}

extern "C" lzx_status lzx_ntt_forward_partial_batch(const lzx_ntt_plan *plan, uint32_t *data_dev, size_t len,
                                                    uint32_t levels, void *stream)
{
    if (plan && levels > plan->log_n) return LZX_NTT_LENGTH_MISMATCH; // ntt.rs:258
    return batch(plan, data_dev, len, NTT_PARTIAL, levels, stream);
}

extern "C" lzx_status lzx_ntt_complete_partial_batch(const lzx_ntt_plan *plan, uint32_t *data_dev, size_t len,
                                                     uint32_t already, void *stream)
{
    if (plan && already > plan->log_n) return LZX_NTT_LENGTH_MISMATCH; // ntt.rs:292
    return batch(plan, data_dev, len, NTT_COMPLETE, already, stream);
}

static lzx_status multi(const lzx_ntt_plan *const *plans, size_t nplans, uint32_t *data, size_t polys, NttMode mode,
                        void *stream)
{
    if (!plans) return LZX_INVALID_ARGUMENT;
    for (size_t i = 0; i < nplans; i++)
        if (!plans[i] || plans[i]->n != plans[0]->n) return LZX_INVALID_ARGUMENT;
    return ntt_launch(plans, nplans, data, polys, mode, 0, (hipStream_t)stream);
}

extern "C" lzx_status lzx_ntt_forward_multi(const lzx_ntt_plan *const *plans, size_t nplans, uint32_t *data_dev,
                                            size_t polys_per_plan, void *stream)
{
    return multi(plans, nplans, data_dev, polys_per_plan, NTT_FWD, stream);
}

extern "C" lzx_status lzx_ntt_inverse_multi(const lzx_ntt_plan *const *plans, size_t nplans, uint32_t *data_dev,
                                            size_t polys_per_plan, void *stream)
{
    return multi(plans, nplans, data_dev, polys_per_plan, NTT_INV, stream);
}

extern "C" lzx_status lzx_ntt_forward_host(const lzx_ntt_plan *plan, uint32_t *a_host, size_t len)
{
    return run_host(plan, a_host, len, NTT_FWD, 0);
}

extern "C" lzx_status lzx_ntt_inverse_host(const lzx_ntt_plan *plan, uint32_t *a_host, size_t len)
{
    return run_host(plan, a_host, len, NTT_INV, 0);
}

extern "C" lzx_status lzx_ring_mul_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, const uint32_t *b_dev,
                                         uint32_t *out_dev, size_t len, void *stream)
{
    if (!plan) return LZX_INVALID_ARGUMENT;
    if (len % plan->n) return LZX_RING_LENGTH_MISMATCH;
    if (len == 0) return LZX_OK; // an empty batch may pass null pointers
    if (!a_dev || !b_dev || !out_dev) return LZX_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)stream;
    uint32_t *tmp = nullptr;
    LZX_HIP(scratch_alloc((void **)&tmp, len * sizeof(uint32_t), s));
    lzx_status st = LZX_OK;
    hipError_t e = hipMemcpyAsync(tmp, b_dev, len * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && out_dev != a_dev)
        e = hipMemcpyAsync(out_dev, a_dev, len * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) st = device_fail(e, "ring mul staging copy");
    const size_t polys = len / plan->n;
    if (st == LZX_OK) st = ntt_launch(&plan, 1, out_dev, polys, NTT_FWD, 0, s);
    if (st == LZX_OK) st = ntt_launch(&plan, 1, tmp, polys, NTT_FWD, 0, s);
    if (st == LZX_OK) {
        const uint32_t blocks = (uint32_t)std::min<size_t>((len + 255) / 256, 65536);
        hipLaunchKernelGGL(pointwise_mul_kernel, dim3(blocks), dim3(256), 0, s, out_dev, tmp, len, plan->q,
                           plan->mod.barrett);
        hipError_t le = hipGetLastError();
        if (le != hipSuccess) st = device_fail(le, "pointwise_mul_kernel");
    }
    if (st == LZX_OK) st = ntt_launch(&plan, 1, out_dev, polys, NTT_INV, 0, s);
    (void)scratch_free(tmp, s);
    return st;
}

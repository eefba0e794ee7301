// Batched arithmetic in R_q[Y]/(Y^2+1) (crates/lattice-ring/src/extension.rs:27-176).
//
// An element c0 + c1 Y is stored as [c0 | c1], n canonical u32 coefficients
// each, so a batch is [count][2][n] (the reference's to_bytes order).
//
// rq2_kernel (n = 4 .. 64): one 64-lane wave owns a tile of 32 elements, lane
// 2e + h holding half h of element e, so row l of a 64-polynomial tile is
// exactly lane l's polynomial. Each operand arrives as coalesced 16-byte
// chunks (b's loads in flight while a is transformed) and is transposed
// through one LDS tile in the rotated, bank-conflict-free slot layout of the
// batched NTT (16 KiB at n = 64: two waves per SIMD). Each lane transforms its
// halves in VGPRs (the shared register NTT), combines slot-wise with its
// partner lane (lane ^ 1, one DPP move per value) and inverts its result half:
//   mul      c0 = a0 b0 - a1 b1 (even lane), c1 = a0 b1 + a1 b0 (odd lane),
//            two slot products per lane (in the NTT domain the Karatsuba
//            saving is one slot product; the result is the same ring element);
//   norm     N = a0^2 + a1^2 (extension.rs:142-147), written by the even lane;
//   inverse  conj(z) N(z)^{-1} (extension.rs:151-167) with every slot of N
//            inverted by one exponentiation (prefix products, Montgomery's
//            batch inversion); an element with a zero slot has no inverse:
//            its flag is 0 and its output is zero.
// Nothing but the inputs and the outputs touches HBM.  For q > 2^31 the same
// pass runs in exact FP64 (rq2_f64_kernel), in place of Montgomery.
//
// Other n (1, 2 and 128 .. 8192): the tuned batched NTT launches on scratch
// copies, rq2_slot_kernel for the slot-wise step, and the inverse launch.
#include <algorithm>

#include "ntt.hpp"
#include "ntt_regs.hpp"
#include "runtime.hpp"

namespace lzx {

enum Rq2Mode { RQ2_MUL = 0, RQ2_NORM = 1, RQ2_INV = 2 };
enum Rq2Linear { RQ2_ADD = 0, RQ2_SUB = 1, RQ2_NEG = 2, RQ2_CONJ = 3 };
constexpr uint32_t RQ2_TILE = 32;        // elements per wave tile
constexpr uint32_t RQ2_MAX_BLOCKS = 8192; // grid-stride beyond this

// the partner lane's value (lane ^ 1): DPP quad_perm [1, 0, 3, 2]
__device__ __forceinline__ uint32_t rq2_swap(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);
}

__device__ __forceinline__ uint32_t rq2_mul(uint32_t a, uint32_t b, uint32_t q, uint64_t barrett)
{
    return reduce_u64((uint64_t)a * b, q, barrett);
}

__device__ __forceinline__ uint32_t rq2_pow(uint32_t a, uint32_t e, uint32_t q, uint64_t barrett)
{
    uint32_t acc = 1;
    for (; e; e >>= 1) {
        if (e & 1) acc = rq2_mul(acc, a, q, barrett);
        a = rq2_mul(a, a, q, barrett);
    }
    return acc;
}

// out may alias a (in-place use): each tile is read whole before it is written
template <int LOG_N, int AR, int MODE>
__global__ __launch_bounds__(64) void rq2_kernel(const uint32_t *a, const uint32_t *b, uint32_t *out, uint8_t *ok,
                                                 const uint32_t *__restrict__ tab, uint64_t barrett, size_t count)
{
    constexpr uint32_t N = 1u << LOG_N, C = N / 4;
    __shared__ __attribute__((aligned(16))) uint32_t T[64 * N]; // one tile: row l = lane l's polynomial
    const QConst Q{tab[0], tab[1], tab[5], tab[6]};
    const uint32_t q = Q.q;
    const uint32_t lane = threadIdx.x, odd = lane & 1;
    const size_t tiles = (count + RQ2_TILE - 1) / RQ2_TILE;
    u32x4 r[C];
    for (size_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const size_t e0 = t * RQ2_TILE;
        const uint32_t ne = count - e0 < RQ2_TILE ? (uint32_t)(count - e0) : RQ2_TILE;
        const uint32_t rows = 2 * ne;
        tile_fetch<C>(a + e0 * 2 * N, rows, r, lane);
#pragma unroll
        for (uint32_t i = 0; i < C; i++) *reinterpret_cast<u32x4 *>(T + 4 * (64 * i + lane)) = r[i];
        wave_sync();
        // every lane computes (rows past 2 ne hold stale values and are not stored),
        // so the partner exchange always sees both lanes of a pair
        uint32_t x[N];
#pragma unroll
        for (uint32_t c4 = 0; c4 < C; c4++) {
            const u32x4 v = *reinterpret_cast<const u32x4 *>(T + 4 * tile_slot<C>(lane, c4));
            x[4 * c4] = v.x, x[4 * c4 + 1] = v.y, x[4 * c4 + 2] = v.z, x[4 * c4 + 3] = v.w;
        }
        if (MODE == RQ2_MUL) tile_fetch<C>(b + e0 * 2 * N, rows, r, lane); // in flight during a's transform
        ntt_regs<LOG_N, AR, NTT_FWD>(x, tab); // x[j]: evaluation brv(j), the order the inverse reads
        bool good = true;
        if (MODE == RQ2_MUL) {
            wave_sync(); // every lane has read its row of a: b goes through the same tile
#pragma unroll
            for (uint32_t i = 0; i < C; i++) *reinterpret_cast<u32x4 *>(T + 4 * (64 * i + lane)) = r[i];
            wave_sync();
            uint32_t y[N];
#pragma unroll
            for (uint32_t c4 = 0; c4 < C; c4++) {
                const u32x4 v = *reinterpret_cast<const u32x4 *>(T + 4 * tile_slot<C>(lane, c4));
                y[4 * c4] = v.x, y[4 * c4 + 1] = v.y, y[4 * c4 + 2] = v.z, y[4 * c4 + 3] = v.w;
            }
            ntt_regs<LOG_N, AR, NTT_FWD>(y, tab);
#pragma unroll
            for (uint32_t j = 0; j < N; j++) {
                const uint32_t pa = rq2_swap(x[j]), pb = rq2_swap(y[j]);
                // even: x y - pa pb = a0 b0 - a1 b1; odd: pa y + x pb = a0 b1 + a1 b0
                const uint32_t p = rq2_mul(odd ? pa : x[j], y[j], q, barrett);
                const uint32_t rr = rq2_mul(odd ? x[j] : pa, pb, q, barrett);
                x[j] = odd ? addq<AR>(p, rr, Q) : subq<AR>(p, rr, Q);
            }
        } else if (MODE == RQ2_NORM) {
#pragma unroll
            for (uint32_t j = 0; j < N; j++) {
                const uint32_t s = rq2_mul(x[j], x[j], q, barrett);
                x[j] = addq<AR>(s, rq2_swap(s), Q);
            }
        } else {
            // pre[j] = N_0 ... N_{j-1}, kept in this lane's own row of the tile (read already)
            uint32_t acc = 1;
#pragma unroll
            for (uint32_t j = 0; j < N; j++) {
                const uint32_t s = rq2_mul(x[j], x[j], q, barrett);
                T[4 * tile_slot<C>(lane, j / 4) + (j % 4)] = acc;
                acc = rq2_mul(acc, addq<AR>(s, rq2_swap(s), Q), q, barrett);
            }
            good = acc != 0; // q is prime: the product vanishes iff a slot does
            uint32_t inv = rq2_pow(acc, q - 2, q, barrett);
#pragma unroll
            for (uint32_t jj = 0; jj < N; jj++) {
                const uint32_t j = N - 1 - jj;
                const uint32_t s = rq2_mul(x[j], x[j], q, barrett);
                const uint32_t nj = addq<AR>(s, rq2_swap(s), Q);
                const uint32_t ninv = rq2_mul(inv, T[4 * tile_slot<C>(lane, j / 4) + (j % 4)], q, barrett);
                inv = rq2_mul(inv, nj, q, barrett);
                const uint32_t v = odd ? (x[j] ? q - x[j] : 0) : x[j]; // conj: -c1 on the odd lane
                x[j] = rq2_mul(v, ninv, q, barrett);
            }
        }
        ntt_regs<LOG_N, AR, NTT_INV>(x, tab);
        // results back through the tile: row l for the product and inverse, row l/2
        // (the element) for the norm, written by the even lane
        wave_sync();
        const uint32_t orow = MODE == RQ2_NORM ? lane >> 1 : lane;
        if (MODE != RQ2_NORM || !odd) {
#pragma unroll
            for (uint32_t c4 = 0; c4 < C; c4++)
                *reinterpret_cast<u32x4 *>(T + 4 * tile_slot<C>(orow, c4)) =
                    u32x4{x[4 * c4], x[4 * c4 + 1], x[4 * c4 + 2], x[4 * c4 + 3]};
        }
        if (MODE == RQ2_INV && lane < rows && !odd) ok[e0 + (lane >> 1)] = good ? 1 : 0;
        wave_sync();
        const uint32_t orows = MODE == RQ2_NORM ? ne : rows;
        uint32_t *go = out + e0 * (MODE == RQ2_NORM ? N : 2 * N);
#pragma unroll
        for (uint32_t i = 0; i < C; i++) {
            const uint32_t s = 64 * i + lane, p = s / C, c = ((s % C) - p) & (C - 1);
            if (p < orows) *reinterpret_cast<u32x4 *>(go + (p * C + c) * 4) = *reinterpret_cast<const u32x4 *>(T + 4 * s);
        }
        wave_sync(); // the stores have read the tile before the next one is staged
    }
}

// The fused pass for q > 2^31 in exact FP64 (the register NTT's ARITH_F64
// class: ntt_regs_f64, f64_mul).  Same tiles, lanes and slot algebra as
// rq2_kernel; values are integer doubles, reduced lazily.  For the product,
// a's transform is parked as canonical u32 while b is staged and transformed,
// so at most one polynomial is held as doubles.  The twiddle pointer is made
// opaque per tile so the compiler reloads each transform's constants from the
// scalar cache instead of keeping all of them in SGPRs across the loop.
__device__ __forceinline__ double rq2_swap_d(double v)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = rq2_swap((uint32_t)u), hi = rq2_swap((uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// The grid is persistent (the resident blocks, each taking every gridDim.x-th
// tile).  Norm and inverse (PF) load a block's next tile of a into r right
// after the current one is staged, so it arrives while the current tile is
// transformed, and stage it into the LDS tile before the current results are
// stored, so the wait for it has no stores in flight.  Past the last tile the
// prefetch reads row 0 of the current tile (in bounds, never used), so every
// load is unconditional.  Tile t + gridDim.x belongs to this block alone, so
// the early read is safe wherever the host allows aliasing (out == a exactly).
// The product loads a at the top of each tile: its b is already in flight
// during a's transform, and a prefetch of the next a kept r live beside a's
// parked transform, whose extra register moves cost more than it hid.
template <int LOG_N, int MODE>
__global__ __launch_bounds__(64) void rq2_f64_kernel(const uint32_t *a, const uint32_t *b, uint32_t *out, uint8_t *ok,
                                                     const uint32_t *__restrict__ tab, size_t count)
{
    constexpr uint32_t N = 1u << LOG_N, C = N / 4;
    __shared__ __attribute__((aligned(16))) uint32_t T[64 * N]; // one tile: row l = lane l's polynomial
    const uint32_t lane = threadIdx.x, odd = lane & 1;
    const size_t tiles = (count + RQ2_TILE - 1) / RQ2_TILE;
    auto rows_of = [&](size_t tt) -> uint32_t {
        const size_t e = tt * RQ2_TILE;
        return 2 * (count - e < RQ2_TILE ? (uint32_t)(count - e) : RQ2_TILE);
    };
    // tile tn's a, or row 0 of tile t when tn is past the end
    auto fetch_next_a = [&](size_t t, size_t tn, u32x4(&r)[C]) {
        const bool in = tn < tiles;
        tile_fetch<C>(a + (in ? tn : t) * RQ2_TILE * 2 * N, in ? rows_of(tn) : 0, r, lane);
    };
    constexpr bool PF = MODE != RQ2_MUL;
    u32x4 r[C];
    if (PF && blockIdx.x < tiles) {
        tile_fetch<C>(a + (size_t)blockIdx.x * RQ2_TILE * 2 * N, rows_of(blockIdx.x), r, lane);
#pragma unroll
        for (uint32_t i = 0; i < C; i++) *reinterpret_cast<u32x4 *>(T + 4 * (64 * i + lane)) = r[i];
    }
    for (size_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint32_t *tb = tab;
        asm volatile("" : "+s"(tb));
        const __attribute__((address_space(4))) double *dt =
            (const __attribute__((address_space(4))) double *)(tb + TabLayout::dbl(N));
        const double q = dt[0], qinv = dt[1];
        const size_t e0 = t * RQ2_TILE, tn = t + gridDim.x;
        const uint32_t ne = count - e0 < RQ2_TILE ? (uint32_t)(count - e0) : RQ2_TILE;
        const uint32_t rows = 2 * ne;
        if constexpr (!PF) {
            tile_fetch<C>(a + e0 * 2 * N, rows, r, lane);
#pragma unroll
            for (uint32_t i = 0; i < C; i++) *reinterpret_cast<u32x4 *>(T + 4 * (64 * i + lane)) = r[i];
        }
        wave_sync(); // a's tile is staged
        // every lane computes (rows past 2 ne hold stale values and are not stored),
        // so the partner exchange always sees both lanes of a pair
        double x[N];
#pragma unroll
        for (uint32_t c4 = 0; c4 < C; c4++) {
            const u32x4 v = *reinterpret_cast<const u32x4 *>(T + 4 * tile_slot<C>(lane, c4));
            x[4 * c4] = v.x, x[4 * c4 + 1] = v.y, x[4 * c4 + 2] = v.z, x[4 * c4 + 3] = v.w;
        }
        if (MODE == RQ2_MUL)
            tile_fetch<C>(b + e0 * 2 * N, rows, r, lane); // in flight during a's transform
        else
            fetch_next_a(t, tn, r);
        ntt_regs_f64<LOG_N, NTT_FWD>(x, tb); // x[j]: evaluation brv(j), the order the inverse reads
        bool good = true;
        if (MODE == RQ2_MUL) {
            uint32_t xa[N]; // a's transform, parked
#pragma unroll
            for (uint32_t j = 0; j < N; j++) xa[j] = f64_canonical(x[j], q, qinv);
            wave_sync(); // every lane has read its row of a: b goes through the same tile
#pragma unroll
            for (uint32_t i = 0; i < C; i++) *reinterpret_cast<u32x4 *>(T + 4 * (64 * i + lane)) = r[i];
            wave_sync();
#pragma unroll
            for (uint32_t c4 = 0; c4 < C; c4++) {
                const u32x4 v = *reinterpret_cast<const u32x4 *>(T + 4 * tile_slot<C>(lane, c4));
                x[4 * c4] = v.x, x[4 * c4 + 1] = v.y, x[4 * c4 + 2] = v.z, x[4 * c4 + 3] = v.w;
            }
            ntt_regs_f64<LOG_N, NTT_FWD>(x, tb); // b's transform
#pragma unroll
            for (uint32_t j = 0; j < N; j++) {
                const double y = x[j], pb = rq2_swap_d(y);
                const double xs = xa[j], pa = rq2_swap(xa[j]);
                // even: x y - pa pb = a0 b0 - a1 b1; odd: pa y + x pb = a0 b1 + a1 b0
                const double p = f64_mul(odd ? pa : xs, y, q, qinv);
                const double rr = f64_mul(odd ? xs : pa, pb, q, qinv);
                x[j] = odd ? p + rr : p - rr;
            }
        } else if (MODE == RQ2_NORM) {
#pragma unroll
            for (uint32_t j = 0; j < N; j++) {
                const double s = f64_mul(x[j], x[j], q, qinv);
                x[j] = s + rq2_swap_d(s);
            }
        } else {
            // pre[j] = N_0 ... N_{j-1} (canonical), kept in this lane's own row of the tile (read already)
            double acc = 1.0;
#pragma unroll
            for (uint32_t j = 0; j < N; j++) {
                const double s = f64_mul(x[j], x[j], q, qinv);
                T[4 * tile_slot<C>(lane, j / 4) + (j % 4)] = f64_canonical(acc, q, qinv);
                acc = f64_mul(acc, s + rq2_swap_d(s), q, qinv);
            }
            const uint32_t prod = f64_canonical(acc, q, qinv);
            good = prod != 0; // q is prime: the product vanishes iff a slot does
            double inv = 1.0, base = prod;
            for (uint32_t e = (uint32_t)q - 2; e; e >>= 1) { // Fermat: prod^(q-2)
                if (e & 1) inv = f64_mul(inv, base, q, qinv);
                base = f64_mul(base, base, q, qinv);
            }
#pragma unroll
            for (uint32_t jj = 0; jj < N; jj++) {
                const uint32_t j = N - 1 - jj;
                const double s = f64_mul(x[j], x[j], q, qinv);
                const double ninv = f64_mul(inv, (double)T[4 * tile_slot<C>(lane, j / 4) + (j % 4)], q, qinv);
                inv = f64_mul(inv, s + rq2_swap_d(s), q, qinv);
                x[j] = f64_mul(odd ? -x[j] : x[j], ninv, q, qinv); // conj: -c1 on the odd lane
            }
        }
        ntt_regs_f64<LOG_N, NTT_INV>(x, tb);
        // results back through the tile: row l for the product and inverse, row l/2
        // (the element) for the norm, written by the even lane
        wave_sync();
        const uint32_t orow = MODE == RQ2_NORM ? lane >> 1 : lane;
        if (MODE != RQ2_NORM || !odd) {
#pragma unroll
            for (uint32_t c4 = 0; c4 < C; c4++)
                *reinterpret_cast<u32x4 *>(T + 4 * tile_slot<C>(orow, c4)) =
                    u32x4{f64_canonical(x[4 * c4], q, qinv), f64_canonical(x[4 * c4 + 1], q, qinv),
                          f64_canonical(x[4 * c4 + 2], q, qinv), f64_canonical(x[4 * c4 + 3], q, qinv)};
        }
        if (MODE == RQ2_INV && lane < rows && !odd) ok[e0 + (lane >> 1)] = good ? 1 : 0;
        wave_sync();
        const uint32_t orows = MODE == RQ2_NORM ? ne : rows;
        uint32_t *go = out + e0 * (MODE == RQ2_NORM ? N : 2 * N);
        u32x4 v[C];
#pragma unroll
        for (uint32_t i = 0; i < C; i++) v[i] = *reinterpret_cast<const u32x4 *>(T + 4 * (64 * i + lane));
        if constexpr (PF) {
            wave_sync(); // every lane has read its results: the next tile's a goes into T
#pragma unroll
            for (uint32_t i = 0; i < C; i++) *reinterpret_cast<u32x4 *>(T + 4 * (64 * i + lane)) = r[i];
        }
#pragma unroll
        for (uint32_t i = 0; i < C; i++) {
            const uint32_t s = 64 * i + lane, p = s / C, c = ((s % C) - p) & (C - 1);
            if (p < orows) *reinterpret_cast<u32x4 *>(go + (p * C + c) * 4) = v[i];
        }
        if constexpr (!PF) wave_sync(); // the results are read before the next tile is staged
    }
}

// Slot-wise step of the composed path on NTT-domain batches: A, B are
// [count][2][n]; MUL writes C = A B to out ([count][2][n]); NORM writes
// A0^2 + A1^2 to out ([count][n]); INV_FLAG clears ok[e] when a slot of N
// vanishes; INV writes ok[e] ? (A0, -A1) N^{-1} : 0 to out.
enum Rq2Slot { SLOT_MUL = 0, SLOT_NORM = 1, SLOT_INV_FLAG = 2, SLOT_INV = 3 };

__global__ void rq2_slot_kernel(int mode, const uint32_t *__restrict__ A, const uint32_t *__restrict__ B,
                                uint32_t *__restrict__ out, uint8_t *__restrict__ ok, uint32_t log_n, size_t count,
                                uint32_t q, uint64_t barrett)
{
    const uint32_t n = 1u << log_n;
    const size_t total = count << log_n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t e = i >> log_n, j = i & (n - 1);
        const uint32_t *ae = A + e * 2 * n;
        const uint32_t a0 = ae[j], a1 = ae[n + j];
        if (mode == SLOT_MUL) {
            const uint32_t *be = B + e * 2 * n;
            const uint32_t b0 = be[j], b1 = be[n + j];
            const uint32_t p = rq2_mul(a0, b0, q, barrett), r = rq2_mul(a1, b1, q, barrett);
            const uint32_t u = rq2_mul(a0, b1, q, barrett), v = rq2_mul(a1, b0, q, barrett);
            out[e * 2 * n + j] = p >= r ? p - r : p + (q - r);
            const uint64_t s = (uint64_t)u + v;
            out[e * 2 * n + n + j] = (uint32_t)(s >= q ? s - q : s);
            continue;
        }
        const uint64_t s = (uint64_t)rq2_mul(a0, a0, q, barrett) + rq2_mul(a1, a1, q, barrett);
        const uint32_t nj = (uint32_t)(s >= q ? s - q : s);
        if (mode == SLOT_NORM) {
            out[i] = nj;
        } else if (mode == SLOT_INV_FLAG) {
            if (nj == 0) ok[e] = 0; // every writer stores 0
        } else {
            const uint32_t ninv = ok[e] ? rq2_pow(nj, q - 2, q, barrett) : 0;
            out[e * 2 * n + j] = rq2_mul(a0, ninv, q, barrett);
            out[e * 2 * n + n + j] = rq2_mul(a1 ? q - a1 : 0, ninv, q, barrett);
        }
    }
}

__global__ void rq2_fill_u8_kernel(uint8_t *__restrict__ p, size_t n, uint8_t v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// add / sub / neg / conjugate (extension.rs:91-139), coefficient-wise on [count][2][n]
__global__ void rq2_linear_kernel(int op, const uint32_t *a, const uint32_t *b, uint32_t *out, uint32_t log_n,
                                  size_t count, uint32_t q)
{
    const size_t total = count << (log_n + 1);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t x = a[i];
        uint32_t r;
        if (op == RQ2_ADD) {
            const uint64_t s = (uint64_t)x + b[i];
            r = (uint32_t)(s >= q ? s - q : s);
        } else if (op == RQ2_SUB) {
            const uint32_t y = b[i];
            r = x >= y ? x - y : x + (q - y);
        } else {
            const bool neg = op == RQ2_NEG || ((i >> log_n) & 1); // conjugate negates c1 only
            r = neg && x ? q - x : x;
        }
        out[i] = r;
    }
}

static uint32_t rq2_grid(size_t work, uint32_t block)
{
    return (uint32_t)std::max<size_t>(1, std::min<size_t>((work + block - 1) / block, 65536));
}

template <int LOG_N, int AR>
static void launch_fused(int mode, const uint32_t *a, const uint32_t *b, uint32_t *out, uint8_t *ok,
                         const lzx_ntt_plan *plan, size_t count, hipStream_t s)
{
    const uint32_t blocks = (uint32_t)std::min<size_t>((count + RQ2_TILE - 1) / RQ2_TILE, RQ2_MAX_BLOCKS);
    const uint64_t bar = plan->mod.barrett;
    if (mode == RQ2_MUL)
        hipLaunchKernelGGL((rq2_kernel<LOG_N, AR, RQ2_MUL>), dim3(blocks), dim3(64), 0, s, a, b, out, ok,
                           plan->dev_tab, bar, count);
    else if (mode == RQ2_NORM)
        hipLaunchKernelGGL((rq2_kernel<LOG_N, AR, RQ2_NORM>), dim3(blocks), dim3(64), 0, s, a, b, out, ok,
                           plan->dev_tab, bar, count);
    else
        hipLaunchKernelGGL((rq2_kernel<LOG_N, AR, RQ2_INV>), dim3(blocks), dim3(64), 0, s, a, b, out, ok,
                           plan->dev_tab, bar, count);
}

// resident blocks of one rq2_f64_kernel instance on a CU (queried once)
template <int LOG_N, int MODE>
static uint32_t rq2_f64_per_cu()
{
    static const uint32_t per_cu = [] {
        int v = 0;
        const hipError_t e =
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, reinterpret_cast<const void *>(&rq2_f64_kernel<LOG_N, MODE>), 64, 0);
        return (uint32_t)(e == hipSuccess && v > 0 ? v : 4);
    }();
    return per_cu;
}

template <int LOG_N, int MODE>
static void launch_f64_mode(const uint32_t *a, const uint32_t *b, uint32_t *out, uint8_t *ok, const lzx_ntt_plan *plan,
                            size_t count, hipStream_t s)
{
    // persistent: every resident block takes every gridDim.x-th tile
    const size_t tiles = (count + RQ2_TILE - 1) / RQ2_TILE;
    const size_t resident = (size_t)rq2_f64_per_cu<LOG_N, MODE>() * (size_t)std::max(1, plan->cus);
    const uint32_t blocks = (uint32_t)std::max<size_t>(1, std::min<size_t>(tiles, resident));
    hipLaunchKernelGGL((rq2_f64_kernel<LOG_N, MODE>), dim3(blocks), dim3(64), 0, s, a, b, out, ok, plan->dev_tab,
                       count);
}

template <int LOG_N>
static void launch_fused_f64(int mode, const uint32_t *a, const uint32_t *b, uint32_t *out, uint8_t *ok,
                             const lzx_ntt_plan *plan, size_t count, hipStream_t s)
{
    if (mode == RQ2_MUL)
        launch_f64_mode<LOG_N, RQ2_MUL>(a, b, out, ok, plan, count, s);
    else if (mode == RQ2_NORM)
        launch_f64_mode<LOG_N, RQ2_NORM>(a, b, out, ok, plan, count, s);
    else
        launch_f64_mode<LOG_N, RQ2_INV>(a, b, out, ok, plan, count, s);
}

static void dispatch_fused_f64(int mode, const uint32_t *a, const uint32_t *b, uint32_t *out, uint8_t *ok,
                               const lzx_ntt_plan *plan, size_t count, hipStream_t s)
{
    switch (plan->log_n) {
    case 2: launch_fused_f64<2>(mode, a, b, out, ok, plan, count, s); break;
    case 3: launch_fused_f64<3>(mode, a, b, out, ok, plan, count, s); break;
    case 4: launch_fused_f64<4>(mode, a, b, out, ok, plan, count, s); break;
    case 5: launch_fused_f64<5>(mode, a, b, out, ok, plan, count, s); break;
    default: launch_fused_f64<6>(mode, a, b, out, ok, plan, count, s); break;
    }
}

template <int AR>
static void dispatch_fused(int mode, const uint32_t *a, const uint32_t *b, uint32_t *out, uint8_t *ok,
                           const lzx_ntt_plan *plan, size_t count, hipStream_t s)
{
    switch (plan->log_n) {
    case 2: launch_fused<2, AR>(mode, a, b, out, ok, plan, count, s); break;
    case 3: launch_fused<3, AR>(mode, a, b, out, ok, plan, count, s); break;
    case 4: launch_fused<4, AR>(mode, a, b, out, ok, plan, count, s); break;
    case 5: launch_fused<5, AR>(mode, a, b, out, ok, plan, count, s); break;
    default: launch_fused<6, AR>(mode, a, b, out, ok, plan, count, s); break;
    }
}

// The composed path: scratch copies of the operands, forward launches, the
// slot-wise kernel into out, the inverse launch on out.
static lzx_status rq2_composed(int mode, const uint32_t *a, const uint32_t *b, uint32_t *out, uint8_t *ok,
                               const lzx_ntt_plan *plan, size_t count, hipStream_t s)
{
    const uint32_t n = plan->n, log_n = plan->log_n;
    const size_t elem = 2 * (size_t)n, nops = mode == RQ2_MUL ? 2 : 1;
    uint32_t *tmp = nullptr;
    LZX_HIP(scratch_alloc((void **)&tmp, nops * count * elem * sizeof(uint32_t), s));
    uint32_t *A = tmp, *B = mode == RQ2_MUL ? tmp + count * elem : nullptr;
    lzx_status st = LZX_OK;
    hipError_t e = hipMemcpyAsync(A, a, count * elem * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && B) e = hipMemcpyAsync(B, b, count * elem * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) st = device_fail(e, "extension-ring staging copy");
    if (st == LZX_OK) st = ntt_launch(&plan, 1, tmp, nops * count * 2, NTT_FWD, 0, s);
    const uint32_t g = rq2_grid(count * n, 256);
    const uint64_t bar = plan->mod.barrett;
    auto slot = [&](int m) {
        if (st != LZX_OK) return;
        hipLaunchKernelGGL(rq2_slot_kernel, dim3(g), dim3(256), 0, s, m, A, B, out, ok, log_n, count, plan->q, bar);
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) st = device_fail(le, "rq2_slot_kernel");
    };
    if (mode == RQ2_MUL) {
        slot(SLOT_MUL);
    } else if (mode == RQ2_NORM) {
        slot(SLOT_NORM);
    } else if (st == LZX_OK) {
        hipLaunchKernelGGL(rq2_fill_u8_kernel, dim3(rq2_grid(count, 256)), dim3(256), 0, s, ok, count, (uint8_t)1);
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) st = device_fail(le, "rq2_fill_u8_kernel");
        slot(SLOT_INV_FLAG);
        slot(SLOT_INV);
    }
    if (st == LZX_OK) st = ntt_launch(&plan, 1, out, count * (mode == RQ2_NORM ? 1 : 2), NTT_INV, 0, s);
    (void)scratch_free(tmp, s);
    return st;
}

static lzx_status rq2_run(int mode, const lzx_ntt_plan *plan, const uint32_t *a, const uint32_t *b, uint32_t *out,
                          uint8_t *ok, size_t count, void *stream)
{
    if (!plan) return LZX_INVALID_ARGUMENT;
    if (count == 0) return LZX_OK; // an empty batch may pass null pointers
    if (!a || !out || (mode == RQ2_MUL && !b) || (mode == RQ2_INV && !ok)) return LZX_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)stream;
    const bool aligned = (((uintptr_t)a | (uintptr_t)(b ? b : a) | (uintptr_t)out) & 15) == 0;
    // The fused pass writes each output tile after reading the same tile of
    // its inputs.  That is safe in place only when out is exactly an input of
    // the same [count][2][n] layout; a norm (out is [count][n]) or a shifted
    // overlap would overwrite other tiles' unread inputs, so those take the
    // composed path, which copies the inputs to scratch first.
    const size_t in_bytes = count * 2 * (size_t)plan->n * sizeof(uint32_t);
    const size_t out_bytes = mode == RQ2_NORM ? in_bytes / 2 : in_bytes;
    auto overlaps = [&](const uint32_t *x) {
        const uintptr_t xo = (uintptr_t)x, oo = (uintptr_t)out;
        return x && xo < oo + out_bytes && oo < xo + in_bytes;
    };
    auto unsafe = [&](const uint32_t *x) { return overlaps(x) && (mode == RQ2_NORM || x != out); };
    const bool alias_safe = !unsafe(a) && !unsafe(mode == RQ2_MUL ? b : nullptr);
    if (plan->log_n >= 2 && plan->log_n <= 6 && aligned && alias_safe) {
        if (plan->arith == ARITH_MONT) // q > 2^31: exact FP64 in place of Montgomery
            dispatch_fused_f64(mode, a, b, out, ok, plan, count, s);
        else
            dispatch_fused<ARITH_SHOUP>(mode, a, b, out, ok, plan, count, s);
        LZX_LAUNCH_CHECK("rq2_kernel");
        return LZX_OK;
    }
    return rq2_composed(mode, a, b, out, ok, plan, count, s);
}

static lzx_status rq2_linear(int op, const lzx_ntt_plan *plan, const uint32_t *a, const uint32_t *b, uint32_t *out,
                             size_t count, void *stream)
{
    if (!plan) return LZX_INVALID_ARGUMENT;
    if (count == 0) return LZX_OK; // an empty batch may pass null pointers
    if (!a || !out || (op <= RQ2_SUB && !b)) return LZX_INVALID_ARGUMENT;
    hipLaunchKernelGGL(rq2_linear_kernel, dim3(rq2_grid(count * 2 * plan->n, 256)), dim3(256), 0,
                       (hipStream_t)stream, op, a, b, out, plan->log_n, count, plan->q);
    LZX_LAUNCH_CHECK("rq2_linear_kernel");
    return LZX_OK;
}

} // namespace lzx

using namespace lzx;

extern "C" lzx_status lzx_rq2_mul_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, const uint32_t *b_dev,
                                        uint32_t *out_dev, size_t count, void *stream)
{
    return rq2_run(RQ2_MUL, plan, a_dev, b_dev, out_dev, nullptr, count, stream);
}

extern "C" lzx_status lzx_rq2_norm_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, uint32_t *out_dev,
                                         size_t count, void *stream)
{
    return rq2_run(RQ2_NORM, plan, a_dev, nullptr, out_dev, nullptr, count, stream);
}

extern "C" lzx_status lzx_rq2_inverse_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, uint32_t *out_dev,
                                            uint8_t *ok_dev, size_t count, void *stream)
{
    return rq2_run(RQ2_INV, plan, a_dev, nullptr, out_dev, ok_dev, count, stream);
}

extern "C" lzx_status lzx_rq2_add_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, const uint32_t *b_dev,
                                        uint32_t *out_dev, size_t count, void *stream)
{
    return rq2_linear(RQ2_ADD, plan, a_dev, b_dev, out_dev, count, stream);
}

extern "C" lzx_status lzx_rq2_sub_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, const uint32_t *b_dev,
                                        uint32_t *out_dev, size_t count, void *stream)
{
    return rq2_linear(RQ2_SUB, plan, a_dev, b_dev, out_dev, count, stream);
}

extern "C" lzx_status lzx_rq2_neg_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, uint32_t *out_dev,
                                        size_t count, void *stream)
{
    return rq2_linear(RQ2_NEG, plan, a_dev, nullptr, out_dev, count, stream);
}

extern "C" lzx_status lzx_rq2_conjugate_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, uint32_t *out_dev,
                                              size_t count, void *stream)
{
    return rq2_linear(RQ2_CONJ, plan, a_dev, nullptr, out_dev, count, stream);
}

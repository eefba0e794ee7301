// The ~2^50 quadratic-slot ring Z_q[X]/(X^n + 1), n = 2^k <= 128, over a
// prime q < 2^50 of two-adicity 7 (crates/lattice-ring/src/modulus50.rs:44-499).
//
// With only 128th roots of unity, X^n + 1 splits into n/2 quadratics
// X^2 - gamma_p; the forward transform is the incomplete DIF of
// QuadSlotTables::forward (:317-342): level d, block i: lo' = lo + w hi,
// hi' = lo - w hi with w = zeta^(2^(5-d) (2 rev_d(i) + 1)).  The inverse
// undoes each butterfly exactly (lo = (lo'+hi')/2, hi = (lo'-hi')/(2w),
// :346-375) and the product is slotwise Karatsuba mod (X^2 - gamma_p)
// (:380-402).  All values are canonical residues, so the device results equal
// the reference's.
//
// Device layout: a wave owns a tile of 128 consecutive u64 (128/n
// polynomials); lane l holds indices l and l + 64.  A butterfly partner is 64
// away (the same lane, n = 128, level 0) or `half` lanes away (a 64-bit lane
// exchange).  The hi lane multiplies by the twiddle and the pair swap one
// value, so each lane does one Shoup product per level.  Constant products
// (twiddles, gamma, 1/2) use 64-bit Shoup companions floor(w 2^64 / q); a
// general product reduces its 100-bit result as hi (2^64 mod q) + lo.
#include <algorithm>
#include <cstring>
#include <vector>

#include "modarith.hpp"
#include "runtime.hpp"

namespace lzx {

constexpr uint32_t QS_BLOCK = 256;
constexpr uint32_t QS_TILE = 128;

enum : int { QS_FWD = 0, QS_INV = 1, QS_SLOT = 2, QS_MUL = 3 };

// constants of one plan as the kernels see them
struct QuadDev {
    const uint64_t *tab; // [tw | twp | itw | itwp] 64 each, [gam | gamp] 64 each
    uint64_t q, inv2, inv2p, c64, c64p, onep;
};

// a w mod q for a < 2^64, w < q < 2^63, wp = floor(w 2^64 / q)
__device__ __forceinline__ uint64_t q50_shoup(uint64_t a, uint64_t w, uint64_t wp, uint64_t q)
{
    const uint64_t r = a * w - mulhi64(a, wp) * q; // in [0, 2q)
    return r >= q ? r - q : r;
}

__device__ __forceinline__ uint64_t q50_add(uint64_t a, uint64_t b, uint64_t q)
{
    const uint64_t s = a + b;
    return s >= q ? s - q : s;
}

__device__ __forceinline__ uint64_t q50_sub(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + (q - b); }

// a b mod q for a, b < q < 2^50: a b = hi 2^64 + lo = hi (2^64 mod q) + lo
__device__ __forceinline__ uint64_t q50_mul(uint64_t a, uint64_t b, const QuadDev &T)
{
    uint64_t hi, lo;
    mul128(a, b, hi, lo);
    return q50_add(q50_shoup(hi, T.c64, T.c64p, T.q), q50_shoup(lo, 1, T.onep, T.q), T.q);
}

__device__ __forceinline__ uint64_t lane_xor64(uint64_t v, uint32_t m) { return __shfl_xor(v, (int)m); }

template <int LOG_N>
__device__ __forceinline__ void quad_forward(uint64_t (&x)[2], const uint64_t *tw, const uint64_t *twp, uint32_t lane,
                                             uint64_t q)
{
    constexpr uint32_t N = 1u << LOG_N;
#pragma unroll
    for (int d = 0; d + 1 < LOG_N; d++) {
        const uint32_t half = N >> (d + 1);
        if (half == 64) { // n = 128, level 0: both halves in this lane
            const uint64_t t = q50_shoup(x[1], tw[0], twp[0], q);
            const uint64_t lo = x[0];
            x[0] = q50_add(lo, t, q);
            x[1] = q50_sub(lo, t, q);
            continue;
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t c = (lane + 64 * h) & (N - 1);
            const uint32_t k = (1u << d) - 1 + (c >> (LOG_N - d));
            const bool hi = c & half;
            const uint64_t t = q50_shoup(x[h], tw[k], twp[k], q);
            const uint64_t recv = lane_xor64(hi ? t : x[h], half);
            x[h] = hi ? q50_sub(recv, t, q) : q50_add(x[h], recv, q);
        }
    }
}

template <int LOG_N>
__device__ __forceinline__ void quad_inverse(uint64_t (&x)[2], const uint64_t *itw, const uint64_t *itwp,
                                             const QuadDev &T, uint32_t lane)
{
    constexpr uint32_t N = 1u << LOG_N;
#pragma unroll
    for (int d = LOG_N - 2; d >= 0; d--) {
        const uint32_t half = N >> (d + 1);
        if (half == 64) {
            const uint64_t s = q50_add(x[0], x[1], T.q), df = q50_sub(x[0], x[1], T.q);
            x[0] = q50_shoup(s, T.inv2, T.inv2p, T.q);
            x[1] = q50_shoup(df, itw[0], itwp[0], T.q);
            continue;
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t c = (lane + 64 * h) & (N - 1);
            const uint32_t k = (1u << d) - 1 + (c >> (LOG_N - d));
            const bool hi = c & half;
            const uint64_t recv = lane_xor64(x[h], half);
            // lo = (lo' + hi') / 2 on the lo lane, hi = (lo' - hi') / (2 w) on the hi lane
            const uint64_t s = hi ? q50_sub(recv, x[h], T.q) : q50_add(x[h], recv, T.q);
            x[h] = hi ? q50_shoup(s, itw[k], itwp[k], T.q) : q50_shoup(s, T.inv2, T.inv2p, T.q);
        }
    }
}

// slot p = (positions 2p, 2p+1): the even lane forms c0 = u0 v0 + gamma u1 v1,
// the odd lane c1 = (u0 + u1)(v0 + v1) - u0 v0 - u1 v1
template <int LOG_N>
__device__ __forceinline__ void quad_slotmul(uint64_t (&x)[2], const uint64_t (&y)[2], const uint64_t *gam,
                                             const uint64_t *gamp, const QuadDev &T, uint32_t lane)
{
    constexpr uint32_t N = 1u << LOG_N;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint32_t c = (lane + 64 * h) & (N - 1);
        const bool odd = c & 1;
        const uint64_t up = lane_xor64(x[h], 1), vp = lane_xor64(y[h], 1);
        const uint64_t own = q50_mul(x[h], y[h], T);
        const uint64_t par = lane_xor64(own, 1);
        const uint64_t p0 = odd ? par : own, p1 = odd ? own : par;
        const uint64_t cross = q50_mul(q50_add(x[h], up, T.q), q50_add(y[h], vp, T.q), T);
        const uint32_t p = c >> 1;
        x[h] = odd ? q50_sub(cross, q50_add(p0, p1, T.q), T.q)
                   : q50_add(p0, q50_shoup(p1, gam[p], gamp[p], T.q), T.q);
    }
}

template <int LOG_N, int MODE>
__global__ __launch_bounds__(QS_BLOCK) void quad_kernel(const uint64_t *a, const uint64_t *b, uint64_t *out,
                                                        size_t len, QuadDev T)
{
    __shared__ uint64_t tab[6 * 64];
    for (uint32_t i = threadIdx.x; i < 6 * 64; i += QS_BLOCK) tab[i] = T.tab[i];
    __syncthreads();
    const uint64_t *tw = tab, *twp = tab + 64, *itw = tab + 128, *itwp = tab + 192, *gam = tab + 256,
                   *gamp = tab + 320;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t tiles = (len + QS_TILE - 1) / QS_TILE;
    for (size_t tile = (size_t)blockIdx.x * (QS_BLOCK / 64) + wave; tile < tiles;
         tile += (size_t)gridDim.x * (QS_BLOCK / 64)) {
        const size_t i0 = tile * QS_TILE + lane, i1 = i0 + 64;
        // a partial last tile holds whole polynomials (n divides 128 and len), so
        // lanes past the end only ever pair with each other; out may alias a
        // (in-place transforms): each index is read, then written, by one lane
        uint64_t x[2] = {i0 < len ? a[i0] : 0, i1 < len ? a[i1] : 0};
        uint64_t y[2] = {0, 0};
        if (MODE == QS_SLOT || MODE == QS_MUL) y[0] = i0 < len ? b[i0] : 0, y[1] = i1 < len ? b[i1] : 0;
        if (MODE == QS_FWD || MODE == QS_MUL) quad_forward<LOG_N>(x, tw, twp, lane, T.q);
        if (MODE == QS_MUL) quad_forward<LOG_N>(y, tw, twp, lane, T.q);
        if (MODE == QS_SLOT || MODE == QS_MUL) quad_slotmul<LOG_N>(x, y, gam, gamp, T, lane);
        if (MODE == QS_INV || MODE == QS_MUL) quad_inverse<LOG_N>(x, itw, itwp, T, lane);
        if (i0 < len) out[i0] = x[0];
        if (i1 < len) out[i1] = x[1];
    }
}

} // namespace lzx

using namespace lzx;

struct lzx_quad_plan {
    uint64_t q, zeta, generator;
    uint32_t log_n, two_adicity;
    std::vector<uint64_t> host; // the device table image
    uint64_t *dev = nullptr;
    QuadDev kd{};
    std::vector<uint64_t> gamma;
};

namespace {

uint64_t pw(uint64_t b, uint64_t e, uint64_t q)
{
    unsigned __int128 acc = 1 % q, x = b % q;
    while (e) {
        if (e & 1) acc = acc * x % q;
        x = x * x % q;
        e >>= 1;
    }
    return (uint64_t)acc;
}

uint64_t companion(uint64_t w, uint64_t q) { return (uint64_t)(((unsigned __int128)w << 64) / q); }

uint32_t rev(uint32_t x, uint32_t bits)
{
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

template <int MODE>
void launch(const lzx_quad_plan *p, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t len, hipStream_t s)
{
    const size_t tiles = (len + QS_TILE - 1) / QS_TILE;
    const uint32_t grid = (uint32_t)std::max<size_t>(1, std::min<size_t>((tiles + 3) / 4, 8192));
#define QS_CASE(L)                                                                                                   \
    case L: hipLaunchKernelGGL((quad_kernel<L, MODE>), dim3(grid), dim3(QS_BLOCK), 0, s, a, b, out, len, p->kd); break;
    switch (p->log_n) {
        QS_CASE(1)
        QS_CASE(2)
        QS_CASE(3)
        QS_CASE(4)
        QS_CASE(5)
        QS_CASE(6)
    default:
        QS_CASE(7)
    }
#undef QS_CASE
}

template <int MODE>
lzx_status run(const lzx_quad_plan *p, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t len, void *stream)
{
    if (!p || !a || !out || ((MODE == QS_SLOT || MODE == QS_MUL) && !b)) return LZX_INVALID_ARGUMENT;
    if (len % ((size_t)1 << p->log_n)) return LZX_MOD50_LENGTH_MISMATCH; // Mod50Error::LengthMismatch
    if (len == 0) return LZX_OK;
    launch<MODE>(p, a, b, out, len, (hipStream_t)stream);
    LZX_LAUNCH_CHECK("quad_kernel");
    return LZX_OK;
}

} // namespace

extern "C" lzx_status lzx_quad_plan_create(uint64_t q, uint32_t two_adicity, uint64_t generator, uint32_t log_n,
                                           lzx_quad_plan **out)
{
    if (!out) return LZX_INVALID_ARGUMENT;
    *out = nullptr;
    // QuadSlotTables::new (modulus50.rs:241-300)
    if (log_n == 0 || log_n > 7) return LZX_MOD50_LOG_N_OUT_OF_RANGE;
    if (two_adicity < 7) return LZX_MOD50_ADICITY_TOO_SMALL;
    if (q < 3 || !(q & 1) || q >= (1ULL << 50)) return LZX_UNSUPPORTED; // the kernels assume an odd q < 2^50
    const uint64_t zeta = pw(generator, (q - 1) >> 7, q);
    if (pw(zeta, 64, q) != q - 1) return LZX_MOD50_GENERATOR_INVALID;
    auto *p = new lzx_quad_plan();
    p->q = q;
    p->zeta = zeta;
    p->generator = generator;
    p->log_n = log_n;
    p->two_adicity = two_adicity;
    const uint64_t inv2 = pw(2, q - 2, q);
    p->host.assign(6 * 64, 0);
    uint64_t *tw = p->host.data(), *twp = tw + 64, *itw = tw + 128, *itwp = tw + 192, *gam = tw + 256,
             *gamp = tw + 320;
    for (uint32_t d = 0; d + 1 < log_n; d++)
        for (uint32_t i = 0; i < (1u << d); i++) {
            const uint32_t k = (1u << d) - 1 + i;
            const uint64_t w = pw(zeta, (((uint64_t)(2 * rev(i, d) + 1)) << (5 - d)) % 128, q);
            tw[k] = w;
            twp[k] = companion(w, q);
            itw[k] = (uint64_t)((unsigned __int128)pw(w, q - 2, q) * inv2 % q); // 1 / (2 w)
            itwp[k] = companion(itw[k], q);
        }
    for (uint32_t s = 0; s < (1u << (log_n - 1)); s++) {
        gam[s] = pw(zeta, (((uint64_t)(2 * rev(s, log_n - 1) + 1)) << (7 - log_n)) % 128, q);
        gamp[s] = companion(gam[s], q);
        p->gamma.push_back(gam[s]);
    }
    hipError_t e = hipMalloc(&p->dev, p->host.size() * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemcpy(p->dev, p->host.data(), p->host.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (p->dev) (void)hipFree(p->dev);
        delete p;
        return device_fail(e, "quad plan tables");
    }
    const uint64_t c64 = (uint64_t)((((unsigned __int128)1) << 64) % q);
    p->kd = QuadDev{p->dev, q, inv2, companion(inv2, q), c64, companion(c64, q), companion(1, q)};
    *out = p;
    return LZX_OK;
}

extern "C" void lzx_quad_plan_destroy(lzx_quad_plan *plan)
{
    if (!plan) return;
    if (plan->dev) (void)hipFree(plan->dev);
    delete plan;
}

extern "C" lzx_status lzx_quad_slot_constants(const lzx_quad_plan *plan, uint64_t *gamma_out, uint64_t *zeta_out)
{
    if (!plan) return LZX_INVALID_ARGUMENT;
    if (gamma_out) memcpy(gamma_out, plan->gamma.data(), plan->gamma.size() * sizeof(uint64_t));
    if (zeta_out) *zeta_out = plan->zeta;
    return LZX_OK;
}

extern "C" lzx_status lzx_quad_forward_batch(const lzx_quad_plan *plan, uint64_t *data_dev, size_t len, void *stream)
{
    return run<QS_FWD>(plan, data_dev, nullptr, data_dev, len, stream);
}

extern "C" lzx_status lzx_quad_inverse_batch(const lzx_quad_plan *plan, uint64_t *data_dev, size_t len, void *stream)
{
    return run<QS_INV>(plan, data_dev, nullptr, data_dev, len, stream);
}

extern "C" lzx_status lzx_quad_slotwise_mul_batch(const lzx_quad_plan *plan, const uint64_t *a_dev,
                                                  const uint64_t *b_dev, uint64_t *out_dev, size_t len, void *stream)
{
    return run<QS_SLOT>(plan, a_dev, b_dev, out_dev, len, stream);
}

extern "C" lzx_status lzx_quad_mul_batch(const lzx_quad_plan *plan, const uint64_t *a_dev, const uint64_t *b_dev,
                                         uint64_t *out_dev, size_t len, void *stream)
{
    return run<QS_MUL>(plan, a_dev, b_dev, out_dev, len, stream);
}

// The ~2^50 quadratic-slot ring Z_q[X]/(X^n + 1), n = 2^k <= 128, over a
// prime q < 2^50 of two-adicity 7 (crates/lattice-ring/src/modulus50.rs:44-499).
//
// With only 128th roots of unity, X^n + 1 splits into n/2 quadratics
// X^2 - gamma_p; the forward transform is the incomplete DIF of
// QuadSlotTables::forward (:317-342): level d, block i: lo' = lo + w hi,
// hi' = lo - w hi with w = zeta^(2^(5-d) (2 rev_d(i) + 1)).  The inverse
// undoes each butterfly exactly (lo = (lo'+hi')/2, hi = (lo'-hi')/(2w),
// :346-375) and the product is slotwise Karatsuba mod (X^2 - gamma_p)
// (:380-402).  Inputs are canonical residues (< q, as the reference's ring
// elements hold them); intermediates are congruent to the reference's values
// and the outputs are reduced to canonical form, so they equal the reference's.
//
// Device layout: a wave owns a tile of 128 consecutive u64 (128/n
// polynomials), two values per lane, and each lane holds both operands of its
// butterflies; between levels lanes trade one value with a partner (DPP row
// moves for lane bits 0-3, v_permlane16/32_swap for bits 4-5).  Products run
// in FP64 on lazily reduced signed values (exact; the scheme and the ranges
// are stated beside d50_mulc), and the slot product is Karatsuba within a
// lane.
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
    const double *tab; // [tw | itw | gam], 64 (w, fl(w/q)) pairs each
    double q, qinv;
};

// Residues travel as doubles: every integer below 2^53 is exact, and the FP64
// FMA is full rate (the 64-bit integer products it replaces cost about ten
// quarter-rate 32-bit multiplies each).  Values are signed integer doubles in
// a lazy range, reduced only where a bound requires it.
//
// Product: for integers |a| < 2^52 and |b| < 2^52 with a b / q estimated to
// relative error e, a b = hi + lo exactly (hi = fl(a b), lo = fma(a, b, -hi)),
// t = rint(estimate) and r = fma(-t, q, hi) + lo = a b - t q exactly, with
// |r| <= q/2 + |a b| e (an integer below 2^53, so neither step rounds).  The
// estimate is fl(a fl(w/q)) for a constant 0 <= w < q (e <= 2^-52), else
// fl(fl(a b) fl(1/q)) (e <= 1.5 2^-52).  With q < 2^50 a constant product thus
// lands in q (1/2 + |a| / 4q).
//
// Ranges (u = max |x| / q; u < 4 keeps |x| < 2^52):
//   forward from canonical input: u 1 -> 1.75 -> 2.69 -> 3.86 over three
//     levels (u' = u + 1/2 + u/4); a centring (|x| <= q/2 (1 + 2^-50)) before
//     the fourth level restarts from 0.5 (-> 1.13 -> 1.91 -> 2.89);
//   slot product on centred inputs: own products <= 0.6q, cross <= 0.88q,
//     c0 <= 1.25q, c1 <= 2.1q;
//   inverse from u <= 1: s <= 2u, mulc <= q (1/2 + u/2) <= q;
//   the store reduces to [0, q) once.
#pragma clang fp contract(off)

// a w mod q up to a multiple of q (|result| <= q (1/2 + |a| 2^-52)); wq = fl(w / q)
__device__ __forceinline__ double d50_mulc(double a, double w, double wq, double q)
{
    const double hi = a * w, lo = __fma_rn(a, w, -hi);
    const double t = __builtin_rint(a * wq);
    return __fma_rn(-t, q, hi) + lo;
}

// a b mod q up to a multiple of q (|result| <= q/2 + |a b| 1.5 2^-52)
__device__ __forceinline__ double d50_mul(double a, double b, double q, double qinv)
{
    const double hi = a * b, lo = __fma_rn(a, b, -hi);
    const double t = __builtin_rint(hi * qinv);
    return __fma_rn(-t, q, hi) + lo;
}

// x - rint(x / q) q, |result| <= q/2 (1 + 2^-50) for integer |x| < 2^52
__device__ __forceinline__ double d50_centre(double x, double q, double qinv)
{
    return __fma_rn(-__builtin_rint(x * qinv), q, x);
}

// x mod q in [0, q) for integer |x| < 2^52: the quotient estimate is within
// |x / q| 2^-52 < 1/q of x/q, so floor() is exact unless x = k q, where it may
// return k - 1 and leave exactly q
__device__ __forceinline__ double d50_canonical(double x, double q, double qinv)
{
    const double r = __fma_rn(-__builtin_floor(x * qinv), q, x);
    return r >= q ? 0.0 : r;
}

// exact u64 <-> double for integers below 2^52: the integer sits in the
// mantissa of 2^52 + v
__device__ __forceinline__ double d50_from_u64(uint64_t v)
{
    return __longlong_as_double((long long)(v | 0x4330000000000000ull)) - 4503599627370496.0;
}

__device__ __forceinline__ uint64_t d50_to_u64(double v)
{
    return (uint64_t)__double_as_longlong(v + 4503599627370496.0) & 0x000FFFFFFFFFFFFFull;
}

// Paired layout: a lane holds both operands of its butterflies.  In layout s
// (0 <= s < log n) lane l holds tile elements e = ins0(l, s) (x0) and e + 2^s
// (x1), where ins0 inserts a zero bit at position s; the pair differs in
// element bit s and the other six element bits are the lane index.  Going
// from layout s to s - 1 (or back) exchanges one value with the lane that
// differs in lane bit s - 1: a lane with that bit clear keeps x0 and takes
// its partner's x0 as x1, a lane with it set keeps x1 and takes its
// partner's x1 as x0.
__device__ __forceinline__ uint32_t ins0(uint32_t l, int s) { return ((l >> s) << (s + 1)) | (l & ((1u << s) - 1)); }

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

template <int B>
__device__ __forceinline__ void xchg32(uint32_t &a, uint32_t &b, uint32_t lane)
{
    if constexpr (B == 5) { // rows 2-3 of a <-> rows 0-1 of b
        const u32x2 r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
        a = r[0], b = r[1];
    } else if constexpr (B == 4) { // odd rows of a <-> even rows of b
        const u32x2 r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
        a = r[0], b = r[1];
    } else if constexpr (B >= 2) {
        // banks of 4 lanes: the bit-clear banks read l + 2^B (row_shl), the
        // bit-set banks l - 2^B (row_shr); masked-off banks keep their value
        constexpr int K = 1 << B;
        constexpr uint32_t clear = B == 2 ? 0x5 : 0x3;
        const uint32_t a0 = a;
        a = __builtin_amdgcn_update_dpp(a, b, 0x110 + K, 0xF, 0xF & ~clear, false);
        b = __builtin_amdgcn_update_dpp(b, a0, 0x100 + K, 0xF, clear, false);
    } else { // quad_perm lane ^ 1 or lane ^ 2, then a per-lane choice
        constexpr int QP = B == 0 ? 0xB1 : 0x4E;
        const bool set = lane & (1u << B);
        const uint32_t pa = __builtin_amdgcn_update_dpp(0u, a, QP, 0xF, 0xF, false);
        const uint32_t pb = __builtin_amdgcn_update_dpp(0u, b, QP, 0xF, 0xF, false);
        a = set ? pb : a;
        b = set ? b : pa;
    }
}

// layout exchange across lane bit B (B < 0: none)
template <int B>
__device__ __forceinline__ void xchg(double &x0, double &x1, uint32_t lane)
{
    if constexpr (B >= 0) {
        const uint64_t u0 = (uint64_t)__double_as_longlong(x0), u1 = (uint64_t)__double_as_longlong(x1);
        uint32_t a0 = (uint32_t)u0, a1 = (uint32_t)(u0 >> 32), b0 = (uint32_t)u1, b1 = (uint32_t)(u1 >> 32);
        xchg32<B>(a0, b0, lane);
        xchg32<B>(a1, b1, lane);
        x0 = __longlong_as_double((long long)(((uint64_t)a1 << 32) | a0));
        x1 = __longlong_as_double((long long)(((uint64_t)b1 << 32) | b0));
    }
}

// x / 2 mod q up to a multiple of q: x/2 when x is even, (x + q)/2 when odd;
// |result| <= |x|/2 + q/2 (exact: h is a half-integer, fract(h) is 0 or 1/2)
__device__ __forceinline__ double d50_half(double x, double q)
{
    const double h = x * 0.5;
    return __fma_rn(__builtin_amdgcn_fract(h), q, h);
}

// canonical input in layout log n - 1 -> slot form in layout 0 (lazy, |x| < 4q);
// level D works in layout log n - 1 - D
template <int LOG_N, int D = 0>
__device__ __forceinline__ void quad_forward(double &x0, double &x1, const double *tw, const QuadDev &T, uint32_t lane)
{
    if constexpr (D + 1 < LOG_N) {
        constexpr uint32_t N = 1u << LOG_N;
        constexpr int s = LOG_N - 1 - D;
        if (D == 3) x0 = d50_centre(x0, T.q, T.qinv), x1 = d50_centre(x1, T.q, T.qinv);
        const uint32_t c = ins0(lane, s) & (N - 1);
        const uint32_t k = (1u << D) - 1 + (c >> (s + 1));
        const double t = d50_mulc(x1, tw[2 * k], tw[2 * k + 1], T.q);
        const double lo = x0;
        x0 = lo + t;
        x1 = lo - t;
        xchg<s - 1>(x0, x1, lane);
        quad_forward<LOG_N, D + 1>(x0, x1, tw, T, lane);
    }
}

// centred slot form in layout 0 -> coefficients in layout log n - 1 (|x| <= 2q);
// the level of block size 2^(s+1) works in layout s
template <int LOG_N, int S = 1>
__device__ __forceinline__ void quad_inverse(double &x0, double &x1, const double *itw, const QuadDev &T,
                                             uint32_t lane)
{
    if constexpr (S < LOG_N) {
        constexpr uint32_t N = 1u << LOG_N;
        constexpr int d = LOG_N - 1 - S;
        xchg<S - 1>(x0, x1, lane);
        if (S == 4) x0 = d50_centre(x0, T.q, T.qinv), x1 = d50_centre(x1, T.q, T.qinv);
        const uint32_t c = ins0(lane, S) & (N - 1);
        const uint32_t k = (1u << d) - 1 + (c >> (S + 1));
        // lo = (lo' + hi') / 2, hi = (lo' - hi') / (2 w)
        const double sum = x0 + x1, df = x0 - x1;
        x0 = d50_half(sum, T.q);
        x1 = d50_mulc(df, itw[2 * k], itw[2 * k + 1], T.q);
        quad_inverse<LOG_N, S + 1>(x0, x1, itw, T, lane);
    }
}

// slot p = (u0 + u1 X)(v0 + v1 X) mod (X^2 - gamma_p) on centred inputs:
// c0 = u0 v0 + gamma u1 v1, c1 = (u0 + u1)(v0 + v1) - u0 v0 - u1 v1
__device__ __forceinline__ void quad_slotmul(double &x0, double &x1, double y0, double y1, double g, double gq,
                                             const QuadDev &T)
{
    const double p0 = d50_mul(x0, y0, T.q, T.qinv), p1 = d50_mul(x1, y1, T.q, T.qinv);
    const double cross = d50_mul(x0 + x1, y0 + y1, T.q, T.qinv);
    x0 = p0 + d50_mulc(p1, g, gq, T.q);
    x1 = cross - (p0 + p1);
}

template <int LOG_N, int MODE>
__global__ __launch_bounds__(QS_BLOCK) void quad_kernel(const uint64_t *a, const uint64_t *b, uint64_t *out,
                                                        size_t len, QuadDev T)
{
    constexpr uint32_t N = 1u << LOG_N;
    constexpr int S0 = LOG_N - 1;
    __shared__ double tab[6 * 64];
    for (uint32_t i = threadIdx.x; i < 6 * 64; i += QS_BLOCK) tab[i] = T.tab[i];
    __syncthreads();
    const double *tw = tab, *itw = tab + 128, *gam = tab + 256; // (w, fl(w/q)) pairs
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t e0 = ins0(lane, S0), e1 = e0 + (1u << S0); // layout S0
    const uint32_t p = lane & (N / 2 - 1);                     // slot of layout 0
    const size_t tiles = (len + QS_TILE - 1) / QS_TILE;
    for (size_t tile = (size_t)blockIdx.x * (QS_BLOCK / 64) + wave; tile < tiles;
         tile += (size_t)gridDim.x * (QS_BLOCK / 64)) {
        const size_t base = tile * QS_TILE;
        // both values of a lane lie in one polynomial (they differ below bit
        // log n), and len is a whole number of polynomials, so a lane is
        // entirely in range or out; out-of-range lanes only exchange with each
        // other.  out may alias a: each index is read, then written, by one lane.
        double x0 = 0.0, x1 = 0.0, y0 = 0.0, y1 = 0.0;
        if (MODE == QS_FWD || MODE == QS_MUL) {
            if (base + e0 < len) {
                x0 = d50_from_u64(a[base + e0]), x1 = d50_from_u64(a[base + e1]);
                if (MODE == QS_MUL) y0 = d50_from_u64(b[base + e0]), y1 = d50_from_u64(b[base + e1]);
            }
        } else if (base + 2 * lane < len) { // layout 0: the lane's pair is adjacent
            const u64x2 va = *(const u64x2 *)(a + base + 2 * lane);
            x0 = d50_from_u64(va[0]), x1 = d50_from_u64(va[1]);
            if (MODE == QS_SLOT) {
                const u64x2 vb = *(const u64x2 *)(b + base + 2 * lane);
                y0 = d50_from_u64(vb[0]), y1 = d50_from_u64(vb[1]);
            }
        }
        if (MODE == QS_FWD || MODE == QS_MUL) quad_forward<LOG_N>(x0, x1, tw, T, lane);
        if (MODE == QS_MUL) quad_forward<LOG_N>(y0, y1, tw, T, lane);
        if (MODE != QS_FWD) x0 = d50_centre(x0, T.q, T.qinv), x1 = d50_centre(x1, T.q, T.qinv);
        if (MODE == QS_SLOT || MODE == QS_MUL) {
            y0 = d50_centre(y0, T.q, T.qinv), y1 = d50_centre(y1, T.q, T.qinv);
            quad_slotmul(x0, x1, y0, y1, gam[2 * p], gam[2 * p + 1], T);
        }
        if (MODE == QS_MUL) x0 = d50_centre(x0, T.q, T.qinv), x1 = d50_centre(x1, T.q, T.qinv);
        if (MODE == QS_INV || MODE == QS_MUL) quad_inverse<LOG_N>(x0, x1, itw, T, lane);
        const uint64_t r0 = d50_to_u64(d50_canonical(x0, T.q, T.qinv));
        const uint64_t r1 = d50_to_u64(d50_canonical(x1, T.q, T.qinv));
        // stores through a resource over the tile's in-range words, at system
        // scope (sc0 sc1: written through L2, so the launch ends with no dirty
        // lines to write back); lanes past len fall outside it and are dropped
        const uint64_t ob = (uint64_t)(out + base);
        const uint64_t obu = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)ob) |
                             (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(ob >> 32)) << 32;
        const uint32_t nw = (uint32_t)__builtin_amdgcn_readfirstlane(
            (uint32_t)(len - base < (size_t)QS_TILE ? len - base : (size_t)QS_TILE));
        const __amdgpu_buffer_rsrc_t ro = __builtin_amdgcn_make_buffer_rsrc((void *)obu, 0, (int)(8 * nw), 0x00020000);
        if (MODE == QS_INV || MODE == QS_MUL) {
            typedef unsigned qs_v2u __attribute__((ext_vector_type(2)));
            __builtin_amdgcn_raw_buffer_store_b64(qs_v2u{(unsigned)r0, (unsigned)(r0 >> 32)}, ro, 8 * e0, 0, 0x11);
            __builtin_amdgcn_raw_buffer_store_b64(qs_v2u{(unsigned)r1, (unsigned)(r1 >> 32)}, ro, 8 * e1, 0, 0x11);
        } else {
            typedef unsigned qs_v4u __attribute__((ext_vector_type(4)));
            const qs_v4u v = {(unsigned)r0, (unsigned)(r0 >> 32), (unsigned)r1, (unsigned)(r1 >> 32)};
            __builtin_amdgcn_raw_buffer_store_b128(v, ro, 16 * lane, 0, 0x11);
        }
    }
}

} // namespace lzx

using namespace lzx;

struct lzx_quad_plan {
    uint64_t q, zeta, generator;
    uint32_t log_n, two_adicity;
    std::vector<double> host; // the device table image
    double *dev = nullptr;
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
    if (!p) return LZX_INVALID_ARGUMENT;
    if (len % ((size_t)1 << p->log_n)) return LZX_MOD50_LENGTH_MISMATCH; // Mod50Error::LengthMismatch
    if (len == 0) return LZX_OK; // an empty batch may pass null pointers
    if (!a || !out || ((MODE == QS_SLOT || MODE == QS_MUL) && !b)) return LZX_INVALID_ARGUMENT;
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
    // each constant w (< 2^50, exact as a double) beside fl(w / q), the quotient estimate's factor
    p->host.assign(6 * 64, 0.0);
    double *tw = p->host.data(), *itw = tw + 128, *gam = tw + 256;
    const double qd = (double)q;
    for (uint32_t d = 0; d + 1 < log_n; d++)
        for (uint32_t i = 0; i < (1u << d); i++) {
            const uint32_t k = (1u << d) - 1 + i;
            const uint64_t w = pw(zeta, (((uint64_t)(2 * rev(i, d) + 1)) << (5 - d)) % 128, q);
            const uint64_t iw = (uint64_t)((unsigned __int128)pw(w, q - 2, q) * inv2 % q); // 1 / (2 w)
            tw[2 * k] = (double)w, tw[2 * k + 1] = (double)w / qd;
            itw[2 * k] = (double)iw, itw[2 * k + 1] = (double)iw / qd;
        }
    for (uint32_t s = 0; s < (1u << (log_n - 1)); s++) {
        const uint64_t g = pw(zeta, (((uint64_t)(2 * rev(s, log_n - 1) + 1)) << (7 - log_n)) % 128, q);
        gam[2 * s] = (double)g, gam[2 * s + 1] = (double)g / qd;
        p->gamma.push_back(g);
    }
    hipError_t e = hipMalloc(&p->dev, p->host.size() * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(p->dev, p->host.data(), p->host.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (p->dev) (void)hipFree(p->dev);
        delete p;
        return device_fail(e, "quad plan tables");
    }
    p->kd = QuadDev{p->dev, qd, 1.0 / qd};
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

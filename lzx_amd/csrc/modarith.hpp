// Modular arithmetic for the gfx950 kernels (u32 NTT primes and Goldilocks).
//
// Every routine returns the canonical residue, so results are bit-identical
// to the reference's Barrett arithmetic (lattice-ring/src/modulus.rs:65-112)
// and Goldilocks field (lattice-core/src/field.rs:28-105): canonical modular
// values do not depend on the reduction method.
//
// Two constant-multiplication schemes are used for the u32 primes:
//   - q < 2^31: Shoup.  w' = floor(w * 2^32 / q); r = a*w - mulhi(a, w')*q is
//     in [0, 2q) for every 32-bit a, and 2q < 2^32, so one conditional
//     subtraction gives the canonical value (3 multiplies).
//   - q > 2^31 (e.g. 3*2^30+1): Montgomery with R = 2^32 and the constant in
//     Montgomery form w*R mod q.  (a*wR - m*q)/R = hi(a*wR) - hi(m*q) lies in
//     (-q, q); one conditional add gives the canonical value (4 multiplies).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define LZX_HD __host__ __device__ __forceinline__

namespace lzx {

LZX_HD uint32_t mulhi32(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

LZX_HD uint64_t mulhi64(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// Constants of one u32 prime as the kernels see them.
struct Mod32 {
    uint32_t q;
    uint32_t qinv;   // q^{-1} mod 2^32 (Montgomery)
    uint64_t barrett; // floor((2^64-1)/q), modulus.rs:15-17
};

// ---- canonical add / sub ---------------------------------------------------

// q < 2^31: a + b < 2^32.  min(s, s-q) selects s-q exactly when s >= q.
LZX_HD uint32_t add_small(uint32_t a, uint32_t b, uint32_t q)
{
    uint32_t s = a + b;
    uint32_t t = s - q;
    return t < s ? t : s;
}

LZX_HD uint32_t sub_small(uint32_t a, uint32_t b, uint32_t q)
{
    uint32_t d = a - b;
    uint32_t e = d + q;
    return e < d ? e : d;
}

// any q < 2^32.
LZX_HD uint32_t add_big(uint32_t a, uint32_t b, uint32_t q)
{
    uint64_t s = (uint64_t)a + b;
    return s >= q ? (uint32_t)(s - q) : (uint32_t)s;
}

LZX_HD uint32_t sub_big(uint32_t a, uint32_t b, uint32_t q)
{
    return a >= b ? a - b : a + (q - b);
}

// ---- multiplication by a precomputed constant -----------------------------

LZX_HD uint32_t shoup_mul(uint32_t a, uint32_t w, uint32_t wp, uint32_t q)
{
    uint32_t qh = mulhi32(a, wp);
    uint32_t r = a * w - qh * q; // in [0, 2q)
    uint32_t t = r - q;
    return t < r ? t : r;
}

// a * b * 2^{-32} mod q, canonical.  Requires min(a, b) < q.
LZX_HD uint32_t mont_mul(uint32_t a, uint32_t b, uint32_t q, uint32_t qinv)
{
    uint64_t t = (uint64_t)a * b;
    uint32_t m = (uint32_t)t * qinv;
    uint32_t mh = mulhi32(m, q);
    uint32_t th = (uint32_t)(t >> 32);
    uint32_t r = th - mh;
    return th < mh ? r + q : r;
}

// Arithmetic classes of a u32 prime, chosen once per plan:
//   ARITH_SHOUP  q < 2^31: Shoup constant multiplication (3 multiplies) and
//                canonical add/sub in 32 bits;
//   ARITH_MONT   q > 2^31: Montgomery (4 multiplies), 33-bit add/sub;
//   ARITH_F64    any q < 2^32, register transforms only: exact FP64 products
//                on lazily reduced values (ntt_regs_f64); the register
//                kernels use it in place of ARITH_MONT.
enum Arith { ARITH_SHOUP = 0, ARITH_F64 = 1, ARITH_MONT = 2 };

struct QConst {
    uint32_t q, qinv, c, k;
};

// a * w mod q with (w, wp) from the plan's constant tables: the Shoup pair
// for q < 2^31, (w * 2^32 mod q, unused) for q > 2^31.
template <int AR>
LZX_HD uint32_t mulq(uint32_t a, uint32_t w, uint32_t wp, const QConst &Q)
{
    if (AR == ARITH_MONT) return mont_mul(a, w, Q.q, Q.qinv);
    return shoup_mul(a, w, wp, Q.q);
}

template <int AR>
LZX_HD uint32_t addq(uint32_t a, uint32_t b, const QConst &Q)
{
    return AR == ARITH_MONT ? add_big(a, b, Q.q) : add_small(a, b, Q.q);
}

template <int AR>
LZX_HD uint32_t subq(uint32_t a, uint32_t b, const QConst &Q)
{
    return AR == ARITH_MONT ? sub_big(a, b, Q.q) : sub_small(a, b, Q.q);
}

// Barrett reduction of any u64 (modulus.rs:65-76 restated with a 64-bit high
// product): qhat in {floor(x/q)-1, floor(x/q)}, so <= 2 corrections.
LZX_HD uint32_t reduce_u64(uint64_t x, uint32_t q, uint64_t barrett)
{
    uint64_t qhat = mulhi64(x, barrett);
    uint64_t r = x - qhat * q;
    if (r >= q) r -= q;
    if (r >= q) r -= q;
    return (uint32_t)r;
}

// ---- host-side constant helpers -------------------------------------------

inline uint32_t host_pow(uint32_t b, uint64_t e, uint32_t q)
{
    uint64_t acc = 1 % q, base = b % q;
    while (e) {
        if (e & 1) acc = acc * base % q;
        base = base * base % q;
        e >>= 1;
    }
    return (uint32_t)acc;
}

inline uint32_t host_inv32(uint32_t q) // q^{-1} mod 2^32 by Newton iteration
{
    uint32_t x = q; // correct to 3 bits for odd q
    for (int i = 0; i < 5; i++) x *= 2u - q * x;
    return x;
}

inline uint32_t shoup_companion(uint32_t w, uint32_t q) { return (uint32_t)(((uint64_t)w << 32) / q); }

inline uint32_t to_mont(uint32_t w, uint32_t q) { return (uint32_t)(((uint64_t)w << 32) % q); }

// ---- Goldilocks p = 2^64 - 2^32 + 1 ---------------------------------------

constexpr uint64_t GL_P = 0xFFFFFFFF00000001ULL;
constexpr uint64_t GL_EPS = 0xFFFFFFFFULL; // 2^64 mod p

LZX_HD uint64_t gl_canon(uint64_t x) { return x >= GL_P ? x - GL_P : x; }

LZX_HD uint64_t gl_add(uint64_t a, uint64_t b) // field.rs:70-77
{
    uint64_t s = a + b;
    if (s < a) s += GL_EPS; // a wrapped 2^64 is worth 2^32 - 1
    return gl_canon(s);
}

LZX_HD uint64_t gl_sub(uint64_t a, uint64_t b) // field.rs:82-89
{
    uint64_t d = a - b;
    if (a < b) d -= GL_EPS;
    return gl_canon(d);
}

// a - b for canonical a, b: a - b, or a - b + 2^64 - eps = a - b + p when a < b,
// both already in [0, p).
LZX_HD uint64_t gl_sub_c(uint64_t a, uint64_t b)
{
    uint64_t d = a - b;
    if (a < b) d -= GL_EPS;
    return d;
}

// A representative in [0, 2^64) of a + b for any a and canonical b (not
// reduced below p): a wrapped sum is below b < p, so adding eps cannot wrap.
LZX_HD uint64_t gl_add_nc(uint64_t a, uint64_t b)
{
    uint64_t s = a + b;
    if (s < b) s += GL_EPS;
    return s;
}

// Full 128-bit product: four 32x32->64 multiply-adds on the device.
LZX_HD void mul128(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t p00 = (uint64_t)a0 * b0;
    const uint64_t t = (uint64_t)a0 * b1 + (p00 >> 32);      // < 2^64
    const uint64_t u = (uint64_t)a1 * b0 + (uint32_t)t;      // < 2^64
    lo = (uint32_t)p00 | (u << 32);
    hi = (uint64_t)a1 * b1 + (t >> 32) + (u >> 32);          // <= 2^64 - 1
#else
    const unsigned __int128 pr = (unsigned __int128)a * b;
    hi = (uint64_t)(pr >> 64);
    lo = (uint64_t)pr;
#endif
}

// Reduce hi*2^64 + lo (any hi, lo) to [0, p) with 2^64 = 2^32 - 1 and
// 2^96 = -1 (mod p).
LZX_HD uint64_t gl_reduce128(uint64_t hi, uint64_t lo)
{
    const uint32_t hh = (uint32_t)(hi >> 32), hl = (uint32_t)hi;
    uint64_t t0 = lo - hh;
    if (lo < hh) t0 -= GL_EPS; // borrowed 2^64 = p + eps
    uint64_t sh = (uint64_t)hl << 32;
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(sh)); // keeps hl * (2^32 - 1) a subtraction, not a 64-bit multiply
#endif
    const uint64_t t1 = sh - hl; // hl * (2^32 - 1)
    uint64_t r = t0 + t1;
    if (r < t1) r += GL_EPS;
    return gl_canon(r);
}

LZX_HD uint64_t gl_mul(uint64_t a, uint64_t b)
{
    uint64_t hi, lo;
    mul128(a, b, hi, lo);
    return gl_reduce128(hi, lo);
}

// a b as any representative below 2^64 (gl_mul without its final reduction
// below p), for products that only feed lazy sums or further products
LZX_HD uint64_t gl_mul_lazy(uint64_t a, uint64_t b)
{
    uint64_t hi, lo;
    mul128(a, b, hi, lo);
    const uint32_t hh = (uint32_t)(hi >> 32), hl = (uint32_t)hi;
    uint64_t t0 = lo - hh;
    if (lo < hh) t0 -= GL_EPS;
    uint64_t sh = (uint64_t)hl << 32;
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(sh));
#endif
    const uint64_t t1 = sh - hl;
    uint64_t r = t0 + t1;
    if (r < t1) r += GL_EPS;
    return r;
}

// Lazy sum: true value = acc + carries * 2^64.
struct GlAcc {
    uint64_t acc;
    uint32_t carries;
};

LZX_HD void gl_acc_add(GlAcc &s, uint64_t v)
{
    uint64_t n = s.acc + v;
    s.carries += n < v;
    s.acc = n;
}

// A representative in [0, 2^64) of the lazy sum: carries * 2^64 + acc =
// carries * (2^32 - 1) + acc (mod p).  c <= (2^32 - 1)^2, so a wrapped r is
// below c and adding eps cannot wrap again.
LZX_HD uint64_t gl_acc_value(const GlAcc &s)
{
    uint64_t sh = (uint64_t)s.carries << 32;
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(sh)); // a subtraction, not a 64-bit multiply
#endif
    const uint64_t c = sh - s.carries;
    uint64_t r = s.acc + c;
    if (r < c) r += GL_EPS;
    return r;
}

LZX_HD uint64_t gl_acc_finish(const GlAcc &s) { return gl_canon(gl_acc_value(s)); }

// A representative below 2^64 of a + b for any a, b below 2^64: a wrapped
// sum is worth 2^64 = eps more; adding eps wraps again only when the wrapped
// sum is at least 2^64 - eps, and then the result is below eps.
LZX_HD uint64_t gl_add_lazy(uint64_t a, uint64_t b)
{
    uint64_t s = a + b;
    if (s < b) {
        s += GL_EPS;
        if (s < GL_EPS) s += GL_EPS;
    }
    return s;
}

// p - x for any representative x below 2^64, as a representative below 2^64.
LZX_HD uint64_t gl_neg_lazy(uint64_t x)
{
    x = gl_canon(x);
    return GL_P - x; // p for x = 0, a representative of 0
}

// Lazy sum of 64-bit values and full 128-bit products: the true value is
// top * 2^128 + hi * 2^64 + lo.  Products are added without reducing them; the
// sum is reduced once at the end, with 2^128 = (2^32 - 1)^2 = -2^32 (mod p).
// A product's high word is at most 2^64 - 2^33 + 1, so top counts the high
// word's wraps: at most one per added product.
struct GlAcc2 {
    uint64_t lo, hi;
    uint32_t top;
};

LZX_HD void gl_acc2_add(GlAcc2 &s, uint64_t v)
{
    s.lo += v;
    s.hi += s.lo < v; // below 2^64 for fewer than 2^64 additions
}

LZX_HD void gl_acc2_add_mul(GlAcc2 &s, uint64_t a, uint64_t b)
{
    uint64_t hi, lo;
    mul128(a, b, hi, lo);
    s.lo += lo;
    hi += s.lo < lo; // <= 2^64 - 2^33 + 2
    s.hi += hi;
    s.top += s.hi < hi;
}

LZX_HD uint64_t gl_acc2_finish(const GlAcc2 &s)
{
    // top * 2^32 <= 2^64 - 2^32 < p is canonical
    return gl_sub(gl_reduce128(s.hi, s.lo), (uint64_t)s.top << 32);
}

// a b + c as 128 bits for any a, b, c below 2^64.  On the device c's halves
// ride in the 32x32 multiply-adds: a0 b0 + c0, a0 b1 + carry and
// a1 b0 + t0 + c1 each stay below 2^64.
LZX_HD void mul128_add(uint64_t a, uint64_t b, uint64_t c, uint64_t &hi, uint64_t &lo)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t p00 = (uint64_t)a0 * b0 + (uint32_t)c;
    const uint64_t t = (uint64_t)a0 * b1 + (p00 >> 32);
    const uint64_t u = (uint64_t)a1 * b0 + ((uint64_t)(uint32_t)t + (c >> 32));
    lo = (uint32_t)p00 | (u << 32);
    hi = (uint64_t)a1 * b1 + (t >> 32) + (u >> 32);
#else
    const unsigned __int128 pr = (unsigned __int128)a * b + c;
    hi = (uint64_t)(pr >> 64);
    lo = (uint64_t)pr;
#endif
}

// a + (b - a) r for canonical a, b, r (bind_first_half_in_place's fold,
// field_simd.rs:297-313), canonical: the 128-bit a + (b - a) r is reduced once.
LZX_HD uint64_t gl_fold(uint64_t a, uint64_t b, uint64_t r)
{
    uint64_t hi, lo;
    mul128_add(gl_sub_c(b, a), r, a, hi, lo);
    return gl_reduce128(hi, lo);
}

} // namespace lzx

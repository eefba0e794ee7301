// Witness packing on the device, feeding the commitment:
//   pack_field_elements / unpack_field_elements  crates/lattice-ring/src/packing.rs:30-91
//   crt_pack / crt_unpack                        crates/lattice-ring/src/packing.rs:95-185
//   pack_bits / unpack_bits                      crates/lattice-commitment/src/sparse.rs:30-71
//   commit_onehot_columns                        crates/lattice-commitment/src/sparse.rs:118-144
//   GadgetDecomposition::decompose / recompose   crates/lattice-core/src/decomposition.rs:47-93
// Every kernel is a grid-stride map over independent outputs (HBM-bound
// integer work); range checks raise a device flag that the host turns into
// the reference's error after one synchronization.
#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "ajtai.hpp"
#include "modarith.hpp"
#include "ntt.hpp"
#include "runtime.hpp"

namespace lzx {

constexpr uint32_t PK_BLOCK = 256;
constexpr uint32_t PK_LIMB_BITS = 22;
constexpr uint64_t PK_LIMB_MASK = (1ULL << PK_LIMB_BITS) - 1;
constexpr uint32_t PK_BITS_PER_COEFF = 31;
constexpr uint32_t PK_MAX_CRT = 64;

static uint32_t pk_grid(size_t work)
{
    return (uint32_t)std::max<size_t>(1, std::min<size_t>((work + PK_BLOCK - 1) / PK_BLOCK, 65536));
}

#define PK_FOR(i, total) for (size_t i = (size_t)blockIdx.x * PK_BLOCK + threadIdx.x; i < (total); i += (size_t)gridDim.x * PK_BLOCK)

// coefficient g of the packed vector: limb g % 3 of value g / 3 (0 past the end), reduced mod q (from_signed)
__global__ __launch_bounds__(PK_BLOCK) void pack_field_kernel(const uint64_t *__restrict__ v, size_t count, uint32_t q,
                                                              uint32_t *__restrict__ out, size_t total)
{
    PK_FOR(g, total)
    {
        uint64_t l = 0;
        if (g < 3 * count) {
            const uint64_t raw = v[g / 3], j = g % 3;
            l = j == 0 ? raw & PK_LIMB_MASK : j == 1 ? (raw >> PK_LIMB_BITS) & PK_LIMB_MASK : raw >> (2 * PK_LIMB_BITS);
        }
        out[g] = (uint32_t)(l < q ? l : l % q); // limbs are < 2^22: reduced only for small q
    }
}

// Four values per thread: two 16-byte loads, three 16-byte stores of their 12
// limbs (output words 12 t .. 12 t + 11); words past the last full group of 12
// take the one-word form.  v and out are 16-byte aligned.
__global__ __launch_bounds__(PK_BLOCK) void pack_field_wide_kernel(const uint64_t *__restrict__ v, size_t count,
                                                                   uint32_t q, uint32_t *__restrict__ out,
                                                                   size_t total)
{
    const size_t groups = total / 12;
    PK_FOR(t, groups)
    {
        uint64_t x[4];
        if (4 * t + 3 < count) {
            const ulonglong2 a = ((const ulonglong2 *)v)[2 * t], b = ((const ulonglong2 *)v)[2 * t + 1];
            x[0] = a.x, x[1] = a.y, x[2] = b.x, x[3] = b.y;
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) x[k] = 4 * t + k < count ? v[4 * t + k] : 0;
        }
        uint32_t w[12];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            w[3 * k] = (uint32_t)(x[k] & PK_LIMB_MASK);
            w[3 * k + 1] = (uint32_t)((x[k] >> PK_LIMB_BITS) & PK_LIMB_MASK);
            w[3 * k + 2] = (uint32_t)(x[k] >> (2 * PK_LIMB_BITS));
        }
        if (q <= PK_LIMB_MASK) // limbs are < 2^22: reduced only for small q
#pragma unroll
            for (int k = 0; k < 12; k++) w[k] %= q;
        uint4 *o = (uint4 *)(out + 12 * t);
        o[0] = make_uint4(w[0], w[1], w[2], w[3]);
        o[1] = make_uint4(w[4], w[5], w[6], w[7]);
        o[2] = make_uint4(w[8], w[9], w[10], w[11]);
    }
    PK_FOR(r, total - 12 * groups)
    {
        const size_t g = 12 * groups + r;
        uint64_t l = 0;
        if (g < 3 * count) {
            const uint64_t raw = v[g / 3], j = g % 3;
            l = j == 0 ? raw & PK_LIMB_MASK : j == 1 ? (raw >> PK_LIMB_BITS) & PK_LIMB_MASK : raw >> (2 * PK_LIMB_BITS);
        }
        out[g] = (uint32_t)(l < q ? l : l % q);
    }
}

__device__ __forceinline__ int64_t balanced(uint32_t c, uint32_t q) { return c <= q / 2 ? (int64_t)c : (int64_t)c - q; }

__global__ __launch_bounds__(PK_BLOCK) void unpack_field_kernel(const uint32_t *__restrict__ pk, size_t total,
                                                                uint32_t q, uint64_t *__restrict__ out,
                                                                uint32_t *flag)
{
    const int64_t bound = 1LL << PK_LIMB_BITS;
    PK_FOR(g, total)
    {
        const int64_t b = balanced(pk[g], q);
        if (b <= -bound || b >= bound) atomicOr(flag, 1u);
    }
    PK_FOR(i, total / 3)
    {
        const int64_t b0 = balanced(pk[3 * i], q), b1 = balanced(pk[3 * i + 1], q), b2 = balanced(pk[3 * i + 2], q);
        const uint64_t raw = ((uint64_t)b0 & PK_LIMB_MASK) | (((uint64_t)b1 & PK_LIMB_MASK) << PK_LIMB_BITS) |
                             ((uint64_t)b2 << (2 * PK_LIMB_BITS));
        out[i] = raw >= GL_P ? raw - GL_P : raw; // Goldilocks::from_u64
    }
}

__global__ __launch_bounds__(PK_BLOCK) void pack_bits_kernel(const uint8_t *__restrict__ bits, size_t nbits,
                                                             uint32_t *__restrict__ out, size_t total)
{
    PK_FOR(g, total)
    {
        uint32_t v = 0;
        const size_t base = g * PK_BITS_PER_COEFF;
#pragma unroll
        for (uint32_t b = 0; b < PK_BITS_PER_COEFF; b++)
            if (base + b < nbits && bits[base + b]) v |= 1u << b;
        out[g] = v;
    }
}

// The same packing from 16-byte loads: a block owns chunks of PK_BLOCK
// coefficients (31 * 256 = 7936 input bytes = 496 aligned 16-byte pieces,
// since the chunk start 7936 c is 16-byte aligned), compresses each piece to a
// 16-bit mask in LDS, and each thread cuts its 31-bit window out of the masks.
constexpr uint32_t PB_BYTES = PK_BLOCK * PK_BITS_PER_COEFF; // 7936
constexpr uint32_t PB_PIECES = PB_BYTES / 16;               // 496

__device__ __forceinline__ uint32_t nz4(uint32_t w) // bit t = (byte t != 0)
{
    return ((w & 0xFFu) != 0) | (((w >> 8) & 0xFFu) != 0) << 1 | (((w >> 16) & 0xFFu) != 0) << 2 |
           ((w >> 24) != 0) << 3;
}

__global__ __launch_bounds__(PK_BLOCK) void pack_bits_wide_kernel(const uint8_t *__restrict__ bits, size_t nbits,
                                                                  uint32_t *__restrict__ out, size_t total)
{
    __shared__ uint16_t m16[PB_PIECES + 1]; // + 1 zero piece: the last window reads one past its bits
    const size_t chunks = (total + PK_BLOCK - 1) / PK_BLOCK;
    if (threadIdx.x == 0) m16[PB_PIECES] = 0;
    for (size_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        const size_t b0 = c * PB_BYTES;
        for (uint32_t k = threadIdx.x; k < PB_PIECES; k += PK_BLOCK) {
            const size_t off = b0 + 16 * (size_t)k;
            uint32_t m = 0;
            if (off + 16 <= nbits) {
                const uint4 v = *(const uint4 *)(bits + off);
                m = nz4(v.x) | nz4(v.y) << 4 | nz4(v.z) << 8 | nz4(v.w) << 12;
            } else {
                for (uint32_t t = 0; t < 16; t++)
                    if (off + t < nbits && bits[off + t]) m |= 1u << t;
            }
            m16[k] = (uint16_t)m;
        }
        __syncthreads();
        const size_t g = c * PK_BLOCK + threadIdx.x;
        if (g < total) {
            const uint32_t bit = threadIdx.x * PK_BITS_PER_COEFF, k = bit >> 4, sh = bit & 15;
            const uint64_t w = (uint64_t)m16[k] | (uint64_t)m16[k + 1] << 16 | (uint64_t)m16[k + 2] << 32;
            out[g] = (uint32_t)(w >> sh) & 0x7FFFFFFFu;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(PK_BLOCK) void unpack_bits_kernel(const uint32_t *__restrict__ el, size_t nelems,
                                                               uint32_t n, uint8_t *__restrict__ bits, size_t nbits)
{
    const size_t per = (size_t)PK_BITS_PER_COEFF * n;
    PK_FOR(idx, nbits)
    {
        const size_t e = idx / per, within = idx % per;
        bits[idx] = e < nelems ? (el[e * n + within / PK_BITS_PER_COEFF] >> (within % PK_BITS_PER_COEFF)) & 1 : 0;
    }
}

// x = q b + r by a Barrett estimate: with m = floor((2^64 - 1) / b), the
// high product floor(x m / 2^64) is floor(x / b) or at most 2 below it.
__device__ __forceinline__ uint64_t div_barrett(uint64_t x, uint64_t b, uint64_t m, uint64_t &r)
{
    uint64_t q = __umul64hi(x, m);
    r = x - q * b;
    if (r >= b) r -= b, q++;
    if (r >= b) r -= b, q++;
    return q;
}

// balanced digits: d = r if r <= b/2 else r - b; the next remainder is
// (rem - d) / b = rem / b, plus one when d borrowed b (exact in u64).
// STAGE: a block's PK_BLOCK values stage their digits in LDS (dynamic, one
// PK_BLOCK * nd * sizeof(DT) tile) and leave as coalesced 4-byte stores.
// POW2: base = 2^m, so the quotient and remainder are a shift and a mask.
//
// POW2 with m nd <= 120 (off = o sum_{k<nd} 2^(m k), o = base/2 - 1, as two
// words): the digits come from y = x + off without the borrow chain.  The
// recurrence's digits lie in [-o, base - 1 - o], a complete residue system mod
// base, so the representation of x with nd such digits plus a top part is
// unique; y's plain base-2^m digits less o are such a representation, so
// they are the recurrence's digits and y >> (m nd) is its final rem.
template <typename DT, bool STAGE, bool POW2>
__global__ __launch_bounds__(PK_BLOCK) void decompose_kernel(const uint64_t *__restrict__ x, size_t count, uint64_t base,
                                                             uint64_t m, uint32_t nd, DT *__restrict__ digits,
                                                             uint32_t *flag, uint64_t off_lo, uint64_t off_hi,
                                                             uint32_t closed)
{
    extern __shared__ uint8_t pk_stage[];
    const uint64_t half = base / 2;
    const size_t chunks = (count + PK_BLOCK - 1) / PK_BLOCK;
    for (size_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        const size_t i = c * PK_BLOCK + threadIdx.x;
        DT *dst = STAGE ? (DT *)pk_stage + (size_t)threadIdx.x * nd : digits + i * nd;
        if (i < count && POW2 && closed) {
            const uint64_t xi = x[i];
            uint64_t lo = xi + off_lo, hi = off_hi + (lo < xi ? 1 : 0);
            const int64_t o = (int64_t)(half - 1);
            for (uint32_t k = 0; k < nd; k++) {
                dst[k] = (DT)((int64_t)(lo & (base - 1)) - o);
                lo = (lo >> m) | (hi << (64 - m));
                hi >>= m;
            }
            if (lo | hi) atomicOr(flag, 1u); // DecompositionError::ValueTooLarge
        } else if (i < count) {
            uint64_t rem = x[i];
            for (uint32_t k = 0; k < nd; k++) {
                uint64_t r, qt;
                if (POW2)
                    r = rem & (base - 1), qt = rem >> m;
                else
                    qt = div_barrett(rem, base, m, r);
                const bool borrow = r > half;
                dst[k] = (DT)(borrow ? (int64_t)r - (int64_t)base : (int64_t)r);
                rem = qt + (borrow ? 1 : 0);
            }
            if (rem) atomicOr(flag, 1u); // DecompositionError::ValueTooLarge
        }
        if (STAGE) {
            __syncthreads();
            const size_t nv = count - c * PK_BLOCK < PK_BLOCK ? count - c * PK_BLOCK : PK_BLOCK;
            const size_t bytes = nv * nd * sizeof(DT);
            // 4-byte aligned: the base is, and PK_BLOCK * nd * sizeof(DT) is a multiple of 4
            uint8_t *g = (uint8_t *)(digits + c * PK_BLOCK * nd);
            const uint32_t *s4 = (const uint32_t *)pk_stage;
            for (size_t w = threadIdx.x; w < bytes / 4; w += PK_BLOCK) ((uint32_t *)g)[w] = s4[w];
            for (size_t b = (bytes & ~(size_t)3) + threadIdx.x; b < bytes; b += PK_BLOCK) g[b] = pk_stage[b];
            __syncthreads();
        }
    }
}

// flag bit 0: a digit outside [-b/2, b/2]; bit 1: the value leaves [0, 2^64)
template <typename DT>
__global__ __launch_bounds__(PK_BLOCK) void recompose_kernel(const DT *__restrict__ digits, size_t count, uint64_t base,
                                                             uint32_t nd, uint64_t *__restrict__ x, uint32_t *flag)
{
    const int64_t half = (int64_t)(base / 2);
    PK_FOR(i, count)
    {
        __int128 acc = 0, power = 1;
        bool bad = false;
        for (uint32_t k = 0; k < nd; k++) {
            const int64_t d = (int64_t)digits[i * nd + k];
            bad |= d < -half || d > half;
            acc += (__int128)d * power;
            power *= (__int128)base;
        }
        if (bad)
            atomicOr(flag, 1u);
        else if (acc < 0 || acc > (__int128)UINT64_MAX)
            atomicOr(flag, 2u);
        x[i] = (uint64_t)acc;
    }
}

struct CrtArgs {
    uint64_t m[PK_MAX_CRT];   // the moduli
    uint64_t inv[PK_MAX_CRT]; // (prod of the earlier moduli)^{-1} mod m
    uint32_t k, slots;        // group size, slots used per group (min(k, n))
};

__global__ __launch_bounds__(PK_BLOCK) void crt_pack_kernel(const uint64_t *__restrict__ v, size_t count, CrtArgs A,
                                                            uint32_t q, uint32_t n, uint32_t *__restrict__ out,
                                                            size_t elems, uint32_t *flag)
{
    PK_FOR(e, elems)
    {
        uint64_t combined = 0, prod = 1;
        for (uint32_t s = 0; s < A.slots; s++) {
            const size_t idx = e * A.k + s;
            if (idx >= count) break;
            const uint64_t m = A.m[s], val = v[idx];
            if (val >= m) {
                atomicOr(flag, 1u); // PackingError::ValueOutOfRange
                break;
            }
            // x = combined + prod t with t = (val - combined) prod^{-1} mod m (m < q < 2^32: no overflow)
            const uint64_t t = (val + m - combined % m) % m * A.inv[s] % m;
            combined += prod * t;
            prod *= m;
        }
        out[e * n] = (uint32_t)(combined % q);
        for (uint32_t c = 1; c < n; c++) out[e * n + c] = 0;
    }
}

__global__ __launch_bounds__(PK_BLOCK) void crt_unpack_kernel(const uint32_t *__restrict__ pk, uint32_t n, CrtArgs A,
                                                              uint64_t *__restrict__ out, size_t expected)
{
    PK_FOR(o, expected) out[o] = pk[(o / A.k) * n] % A.m[o % A.k];
}

__global__ __launch_bounds__(PK_BLOCK) void zero_kernel(uint32_t *__restrict__ s, size_t total)
{
    PK_FOR(g, total) s[g] = 0;
}

__global__ __launch_bounds__(PK_BLOCK) void scatter_set_kernel(uint32_t *__restrict__ s, const uint64_t *__restrict__ idx,
                                                               const uint32_t *__restrict__ val, size_t nnz)
{
    PK_FOR(i, nnz) s[idx[i]] = val[i];
}

// flag word: allocated, cleared, launched into, read back through a pinned
// staging word (a pageable copy would be a slower synchronous copy)
struct Flag {
    uint32_t *dev = nullptr;
    HostStage hs{};
    bool inflight = false;
    hipStream_t s;
    explicit Flag(hipStream_t st) : s(st) {}
    lzx_status init()
    {
        hipError_t e = scratch_alloc((void **)&dev, sizeof(uint32_t), s);
        if (e == hipSuccess) e = stage_acquire(hs, sizeof(uint32_t));
        if (e == hipSuccess) {
            hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(PK_BLOCK), 0, s, dev, (size_t)1);
            e = hipGetLastError();
        }
        return e == hipSuccess ? LZX_OK : device_fail(e, "packing flag");
    }
    lzx_status read(uint32_t *v)
    {
        hipError_t e = hipMemcpyAsync(hs.p, dev, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
        inflight = e == hipSuccess;
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess) inflight = false, *v = *(const uint32_t *)hs.p;
        return e == hipSuccess ? LZX_OK : device_fail(e, "packing flag readback");
    }
    ~Flag()
    {
        if (dev) (void)scratch_free(dev, s);
        if (inflight) (void)hipStreamSynchronize(s); // the pinned word may still be written
        stage_release(hs);
    }
};

} // namespace lzx

using namespace lzx;

extern "C" size_t lzx_pack_field_elements_len(size_t count, uint32_t n)
{
    return n ? std::max<size_t>(1, (3 * count + n - 1) / n) : 0;
}

extern "C" lzx_status lzx_pack_field_elements(const uint64_t *values_dev, size_t count, uint32_t q, uint32_t n,
                                              uint32_t *out_dev, void *stream)
{
    if (!out_dev || (count && !values_dev) || n == 0 || q < 2) return LZX_INVALID_ARGUMENT;
    const size_t total = lzx_pack_field_elements_len(count, n) * n;
    if ((uintptr_t)values_dev % 16 == 0 && (uintptr_t)out_dev % 16 == 0)
        hipLaunchKernelGGL(pack_field_wide_kernel, dim3(pk_grid(total / 12 + 1)), dim3(PK_BLOCK), 0,
                           (hipStream_t)stream, values_dev, count, q, out_dev, total);
    else
        hipLaunchKernelGGL(pack_field_kernel, dim3(pk_grid(total)), dim3(PK_BLOCK), 0, (hipStream_t)stream,
                           values_dev, count, q, out_dev, total);
    LZX_LAUNCH_CHECK("pack_field_kernel");
    return LZX_OK;
}

extern "C" lzx_status lzx_unpack_field_elements(const uint32_t *packed_dev, size_t elems, uint32_t q, uint32_t n,
                                                uint64_t *out_dev, void *stream)
{
    if (!packed_dev || n == 0 || q < 2 || (elems * n >= 3 && !out_dev)) return LZX_INVALID_ARGUMENT;
    const size_t total = elems * n;
    if (total == 0) return LZX_OK;
    hipStream_t s = (hipStream_t)stream;
    Flag f(s);
    lzx_status st = f.init();
    if (st != LZX_OK) return st;
    hipLaunchKernelGGL(unpack_field_kernel, dim3(pk_grid(total)), dim3(PK_BLOCK), 0, s, packed_dev, total, q, out_dev,
                       f.dev);
    LZX_LAUNCH_CHECK("unpack_field_kernel");
    uint32_t v = 0;
    st = f.read(&v);
    if (st != LZX_OK) return st;
    return v ? LZX_PACKING_LIMB_OUT_OF_RANGE : LZX_OK;
}

extern "C" size_t lzx_pack_bits_len(size_t nbits, uint32_t n)
{
    const size_t per = (size_t)PK_BITS_PER_COEFF * n;
    return n ? std::max<size_t>(1, (nbits + per - 1) / per) : 0;
}

extern "C" lzx_status lzx_pack_bits(const uint8_t *bits_dev, size_t nbits, uint32_t n, uint32_t *out_dev, void *stream)
{
    if (!out_dev || (nbits && !bits_dev) || n == 0) return LZX_INVALID_ARGUMENT;
    const size_t total = lzx_pack_bits_len(nbits, n) * n;
    if ((uintptr_t)bits_dev % 16 == 0)
        hipLaunchKernelGGL(pack_bits_wide_kernel, dim3(pk_grid(total)), dim3(PK_BLOCK), 0, (hipStream_t)stream,
                           bits_dev, nbits, out_dev, total);
    else
        hipLaunchKernelGGL(pack_bits_kernel, dim3(pk_grid(total)), dim3(PK_BLOCK), 0, (hipStream_t)stream, bits_dev,
                           nbits, out_dev, total);
    LZX_LAUNCH_CHECK("pack_bits_kernel");
    return LZX_OK;
}

extern "C" lzx_status lzx_unpack_bits(const uint32_t *elems_dev, size_t elems, uint32_t n, uint8_t *bits_dev,
                                      size_t nbits, void *stream)
{
    if ((nbits && !bits_dev) || (elems && !elems_dev) || n == 0) return LZX_INVALID_ARGUMENT;
    if (nbits == 0) return LZX_OK;
    hipLaunchKernelGGL(unpack_bits_kernel, dim3(pk_grid(nbits)), dim3(PK_BLOCK), 0, (hipStream_t)stream, elems_dev,
                       elems, n, bits_dev, nbits);
    LZX_LAUNCH_CHECK("unpack_bits_kernel");
    return LZX_OK;
}

extern "C" lzx_status lzx_gadget_decompose(const uint64_t *x_dev, size_t count, uint64_t base, uint32_t num_digits,
                                           void *digits_dev, uint32_t digit_bytes, void *stream)
{
    if ((count && (!x_dev || !digits_dev)) || base < 2) return LZX_INVALID_ARGUMENT;
    if (digit_bytes != 1 && digit_bytes != 8) return LZX_INVALID_ARGUMENT;
    if (digit_bytes == 1 && base > 255) return LZX_UNSUPPORTED; // balanced digits reach base/2
    if (count == 0) return LZX_OK;
    hipStream_t s = (hipStream_t)stream;
    Flag f(s);
    lzx_status st = f.init();
    if (st != LZX_OK) return st;
    const bool pow2 = (base & (base - 1)) == 0;
    uint64_t m = UINT64_MAX / base; // Barrett multiplier, or the shift for a power of two
    if (pow2)
        for (m = 0; (1ULL << m) != base; m++) {
        }
    // the closed form for power-of-two bases (see decompose_kernel)
    uint64_t off_lo = 0, off_hi = 0;
    const uint32_t closed = pow2 && m * num_digits <= 120 ? 1u : 0u;
    if (closed) {
        unsigned __int128 off = 0;
        for (uint32_t k = 0; k < num_digits; k++) off += (unsigned __int128)(base / 2 - 1) << (m * k);
        off_lo = (uint64_t)off, off_hi = (uint64_t)(off >> 64);
    }
    const size_t stage = (size_t)PK_BLOCK * num_digits * digit_bytes;
    const bool staged = stage <= 32768 && (uintptr_t)digits_dev % 4 == 0;
    const dim3 g(pk_grid(count));
    const size_t sh = staged ? stage : 0;
#define PK_DECOMPOSE(DT, ST, P2)                                                                                    \
    hipLaunchKernelGGL((decompose_kernel<DT, ST, P2>), g, dim3(PK_BLOCK), sh, s, x_dev, count, base, m, num_digits,    \
                       (DT *)digits_dev, f.dev, off_lo, off_hi, closed)
    if (digit_bytes == 1 && staged && pow2)
        PK_DECOMPOSE(int8_t, true, true);
    else if (digit_bytes == 1 && staged)
        PK_DECOMPOSE(int8_t, true, false);
    else if (digit_bytes == 1 && pow2)
        PK_DECOMPOSE(int8_t, false, true);
    else if (digit_bytes == 1)
        PK_DECOMPOSE(int8_t, false, false);
    else if (staged && pow2)
        PK_DECOMPOSE(int64_t, true, true);
    else if (staged)
        PK_DECOMPOSE(int64_t, true, false);
    else if (pow2)
        PK_DECOMPOSE(int64_t, false, true);
    else
        PK_DECOMPOSE(int64_t, false, false);
#undef PK_DECOMPOSE
    LZX_LAUNCH_CHECK("decompose_kernel");
    uint32_t v = 0;
    st = f.read(&v);
    if (st != LZX_OK) return st;
    return v ? LZX_DECOMP_VALUE_TOO_LARGE : LZX_OK;
}

extern "C" lzx_status lzx_gadget_recompose(const void *digits_dev, uint32_t digit_bytes, size_t count, uint64_t base,
                                           uint32_t num_digits, uint64_t *x_dev, void *stream)
{
    if ((count && (!x_dev || !digits_dev)) || base < 2) return LZX_INVALID_ARGUMENT;
    if (digit_bytes != 1 && digit_bytes != 8) return LZX_INVALID_ARGUMENT;
    // the running power base^num_digits <= 2^(ceil(log2 base) num_digits) must stay inside 128-bit integers
    uint64_t bits = 0;
    for (uint64_t b = base - 1; b; b >>= 1) bits += 1;
    if (bits * num_digits > 125) return LZX_UNSUPPORTED;
    if (count == 0) return LZX_OK;
    hipStream_t s = (hipStream_t)stream;
    Flag f(s);
    lzx_status st = f.init();
    if (st != LZX_OK) return st;
    if (digit_bytes == 1)
        hipLaunchKernelGGL(recompose_kernel<int8_t>, dim3(pk_grid(count)), dim3(PK_BLOCK), 0, s,
                           (const int8_t *)digits_dev, count, base, num_digits, x_dev, f.dev);
    else
        hipLaunchKernelGGL(recompose_kernel<int64_t>, dim3(pk_grid(count)), dim3(PK_BLOCK), 0, s,
                           (const int64_t *)digits_dev, count, base, num_digits, x_dev, f.dev);
    LZX_LAUNCH_CHECK("recompose_kernel");
    uint32_t v = 0;
    st = f.read(&v);
    if (st != LZX_OK) return st;
    if (v & 1) return LZX_DECOMP_DIGIT_OUT_OF_RANGE;
    return v & 2 ? LZX_DECOMP_VALUE_OUT_OF_RANGE : LZX_OK;
}

// crt_pack's modulus checks (packing.rs:100-118) and the per-slot inverses
static lzx_status crt_args(const uint64_t *moduli, size_t k, uint32_t q, uint32_t n, CrtArgs &A)
{
    if (k == 0) return LZX_PACKING_LENGTH_MISMATCH;
    if (k > PK_MAX_CRT) return LZX_UNSUPPORTED;
    uint64_t product = 1;
    for (size_t i = 0; i < k; i++) {
        const uint64_t m = moduli[i];
        if (m == 0 || m >= q) return LZX_PACKING_MODULUS_TOO_LARGE;
        if ((unsigned __int128)product * m >= q) return LZX_PACKING_MODULUS_TOO_LARGE;
        product *= m;
    }
    memset(&A, 0, sizeof(A));
    A.k = (uint32_t)k;
    A.slots = (uint32_t)std::min<size_t>(k, n);
    uint64_t prod = 1;
    for (uint32_t s = 0; s < A.slots; s++) {
        const uint64_t m = moduli[s];
        int64_t r0 = (int64_t)(prod % m), r1 = (int64_t)m, s0 = 1, s1 = 0; // extended Euclid
        while (r1) {
            const int64_t qq = r0 / r1, t = r0 - qq * r1, ts = s0 - qq * s1;
            r0 = r1, r1 = t, s0 = s1, s1 = ts;
        }
        if (m <= 1 || r0 != 1) return LZX_PACKING_NOT_COPRIME;
        A.inv[s] = (uint64_t)(((s0 % (int64_t)m) + (int64_t)m) % (int64_t)m);
        prod *= m;
    }
    for (size_t i = 0; i < k; i++) A.m[i] = moduli[i];
    return LZX_OK;
}

extern "C" lzx_status lzx_crt_pack(const uint64_t *values_dev, size_t count, const uint64_t *moduli, size_t k,
                                   uint32_t q, uint32_t n, uint32_t *out_dev, void *stream)
{
    if (!out_dev || (count && !values_dev) || (k && !moduli) || n == 0) return LZX_INVALID_ARGUMENT;
    CrtArgs A;
    lzx_status st = crt_args(moduli, k, q, n, A);
    if (st != LZX_OK) return st;
    const size_t elems = count ? (count + k - 1) / k : 1;
    hipStream_t s = (hipStream_t)stream;
    Flag f(s);
    st = f.init();
    if (st != LZX_OK) return st;
    hipLaunchKernelGGL(crt_pack_kernel, dim3(pk_grid(elems)), dim3(PK_BLOCK), 0, s, values_dev, count, A, q, n, out_dev,
                       elems, f.dev);
    LZX_LAUNCH_CHECK("crt_pack_kernel");
    uint32_t v = 0;
    st = f.read(&v);
    if (st != LZX_OK) return st;
    return v ? LZX_PACKING_VALUE_OUT_OF_RANGE : LZX_OK;
}

extern "C" lzx_status lzx_crt_unpack(const uint32_t *packed_dev, size_t elems, uint32_t n, const uint64_t *moduli,
                                     size_t k, size_t expected, uint64_t *out_dev, void *stream)
{
    if ((expected && !out_dev) || (elems && !packed_dev) || (k && !moduli) || n == 0) return LZX_INVALID_ARGUMENT;
    if (k > PK_MAX_CRT) return LZX_UNSUPPORTED;
    if (k == 0 || expected > elems * k) return LZX_PACKING_LENGTH_MISMATCH; // packing.rs:175-181
    if (expected == 0) return LZX_OK;
    CrtArgs A{};
    A.k = (uint32_t)k;
    for (size_t i = 0; i < k; i++) {
        if (moduli[i] == 0) return LZX_INVALID_ARGUMENT;
        A.m[i] = moduli[i];
    }
    hipLaunchKernelGGL(crt_unpack_kernel, dim3(pk_grid(expected)), dim3(PK_BLOCK), 0, (hipStream_t)stream, packed_dev,
                       n, A, out_dev, expected);
    LZX_LAUNCH_CHECK("crt_unpack_kernel");
    return LZX_OK;
}

// commit_onehot_columns (sparse.rs:118-144): column c packs its indicator
// (bit `position` of `len`) at 31 bits per coefficient into ring elements,
// padded to m; s[j] sums every column's element j mod q and t = commit(s).
extern "C" lzx_status lzx_ajtai_commit_onehot(const lzx_ajtai_key *key, const uint64_t *positions,
                                              const uint64_t *lens, size_t ncols, uint32_t *s_dev, uint32_t *t_dev,
                                              void *stream)
{
    if (!key || !s_dev || !t_dev || (ncols && (!positions || !lens))) return LZX_INVALID_ARGUMENT;
    const uint32_t q = key->plan->q, n = (uint32_t)key->n;
    const size_t m = key->m;
    const size_t per = (size_t)PK_BITS_PER_COEFF * n;
    std::unordered_map<uint64_t, uint64_t> acc; // coefficient index -> sum mod q
    for (size_t c = 0; c < ncols; c++) {
        if (lzx_pack_bits_len(lens[c], n) > m) return LZX_AJTAI_DIMENSION_MISMATCH; // OnehotColumn::pack
        if (positions[c] >= lens[c]) continue;                                  // an all-zero indicator
        const uint64_t p = positions[c], e = p / per, w = p % per;
        const uint64_t idx = e * n + w / PK_BITS_PER_COEFF;
        uint64_t &a = acc[idx];
        a = (a + (1ULL << (w % PK_BITS_PER_COEFF))) % q;
    }
    std::vector<uint64_t> idx;
    std::vector<uint32_t> val;
    for (const auto &kv : acc) {
        idx.push_back(kv.first);
        val.push_back((uint32_t)kv.second);
    }
    hipStream_t s = (hipStream_t)stream;
    const size_t total = m * n;
    hipLaunchKernelGGL(zero_kernel, dim3(pk_grid(total)), dim3(PK_BLOCK), 0, s, s_dev, total);
    LZX_LAUNCH_CHECK("zero_kernel");
    if (!idx.empty()) {
        uint64_t *d_idx = nullptr;
        LZX_HIP(scratch_alloc((void **)&d_idx, idx.size() * (sizeof(uint64_t) + sizeof(uint32_t)), s));
        uint32_t *d_val = (uint32_t *)(d_idx + idx.size());
        hipError_t e = hipMemcpyAsync(d_idx, idx.data(), idx.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_val, val.data(), val.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(scatter_set_kernel, dim3(pk_grid(idx.size())), dim3(PK_BLOCK), 0, s, s_dev, d_idx,
                               d_val, idx.size());
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s); // the host vectors feed the copies
        (void)scratch_free(d_idx, s);
        if (e != hipSuccess) return device_fail(e, "one-hot scatter");
    }
    return lzx_ajtai_commit_batch(key, s_dev, m, 1, t_dev, stream);
}

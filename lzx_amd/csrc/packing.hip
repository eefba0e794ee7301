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
        out[g] = (uint32_t)(l % q);
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

// balanced digits: d = r if r <= b/2 else r - b; the next remainder is
// (rem - d) / b = rem / b, plus one when d borrowed b (exact in u64)
template <typename DT>
__global__ __launch_bounds__(PK_BLOCK) void decompose_kernel(const uint64_t *__restrict__ x, size_t count, uint64_t base,
                                                             uint32_t nd, DT *__restrict__ digits, uint32_t *flag)
{
    const uint64_t half = base / 2;
    PK_FOR(i, count)
    {
        uint64_t rem = x[i];
        for (uint32_t k = 0; k < nd; k++) {
            const uint64_t r = rem % base;
            const bool borrow = r > half;
            digits[i * nd + k] = (DT)(borrow ? (int64_t)r - (int64_t)base : (int64_t)r);
            rem = rem / base + (borrow ? 1 : 0);
        }
        if (rem) atomicOr(flag, 1u); // DecompositionError::ValueTooLarge
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

// flag word: allocated, cleared, launched into, read back
struct Flag {
    uint32_t *dev = nullptr;
    hipStream_t s;
    explicit Flag(hipStream_t st) : s(st) {}
    lzx_status init()
    {
        hipError_t e = scratch_alloc((void **)&dev, sizeof(uint32_t), s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(PK_BLOCK), 0, s, dev, (size_t)1);
            e = hipGetLastError();
        }
        return e == hipSuccess ? LZX_OK : device_fail(e, "packing flag");
    }
    lzx_status read(uint32_t *v)
    {
        hipError_t e = hipMemcpyAsync(v, dev, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        return e == hipSuccess ? LZX_OK : device_fail(e, "packing flag readback");
    }
    ~Flag()
    {
        if (dev) (void)scratch_free(dev, s);
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
    hipLaunchKernelGGL(pack_field_kernel, dim3(pk_grid(total)), dim3(PK_BLOCK), 0, (hipStream_t)stream, values_dev,
                       count, q, out_dev, total);
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
    hipLaunchKernelGGL(pack_bits_kernel, dim3(pk_grid(total)), dim3(PK_BLOCK), 0, (hipStream_t)stream, bits_dev, nbits,
                       out_dev, total);
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
    if (digit_bytes == 1)
        hipLaunchKernelGGL(decompose_kernel<int8_t>, dim3(pk_grid(count)), dim3(PK_BLOCK), 0, s, x_dev, count, base,
                           num_digits, (int8_t *)digits_dev, f.dev);
    else
        hipLaunchKernelGGL(decompose_kernel<int64_t>, dim3(pk_grid(count)), dim3(PK_BLOCK), 0, s, x_dev, count, base,
                           num_digits, (int64_t *)digits_dev, f.dev);
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

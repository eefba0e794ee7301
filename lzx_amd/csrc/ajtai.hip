// Ajtai commitment t = A * s over R_q = Z_q[X]/(X^n+1), NTT-domain path.
//
// Reference: crates/lattice-commitment/src/ajtai.rs
//   from_seed (:87-103)  A[i*m+j] = uniform_from_seed("ajtai-A", seed, i*m+j),
//                        A^ cached as the forward NTT of every entry;
//   commit    (:118-153) forward NTT of each s_j, pointwise MAC of A^_ij * s^_j
//                        into k accumulators, one inverse NTT per row.
//
// Device pipeline of a batched commit (W witnesses) where the key holds no
// matrix-core fragments (ajtai_nttmm.hip computes the same sum on int8 MFMA
// where it does):
//   1. s^ = forward NTT of every witness element (scratch, [W][m][n]);
//   2. ajtai_mac_kernel: split-K over j.  For each coefficient t the product
//      is a (W x m) * (m x k) matrix product mod q; a thread owns one t and a
//      4 x 4 tile of (witness, row) pairs, accumulating the exact 64-bit
//      products lazily (64-bit sum + carry count), reduced once per chunk;
//   3. ajtai_reduce_kernel sums the chunk partials mod q;
//   4. inverse NTT of the [W][k][n] accumulators.
// All-zero witness elements contribute zero, which is exactly the
// reference's zero-skip (:130-137).
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "ajtai.hpp"
#include "keccak.hpp"
#include "seeded.hpp"

namespace lzx {

constexpr int MAC_TB = 64;  // coefficients per block
constexpr int MAC_TILE = 4; // witnesses x rows per thread

__global__ __launch_bounds__(1024) void ajtai_mac_kernel(const uint32_t *__restrict__ Ahat,
                                                         const uint32_t *__restrict__ shat,
                                                         uint32_t *__restrict__ partial, uint32_t W, uint32_t k,
                                                         uint32_t m, uint32_t n, uint32_t tb, uint32_t wg,
                                                         uint32_t ig, uint32_t jchunk, uint32_t q, uint64_t barrett,
                                                         uint32_t r64)
{
    // blockIdx.x = t-block + (n/tb) * w-block ; blockIdx.y = row block ; blockIdx.z = chunk
    const uint32_t tblocks = n / tb;
    const uint32_t tblk = blockIdx.x % tblocks, wblk = blockIdx.x / tblocks;
    const uint32_t tt = threadIdx.x % tb;
    const uint32_t rest = threadIdx.x / tb;
    const uint32_t wsub = rest % wg, isub = rest / wg;
    const uint32_t t = tblk * tb + tt;
    const uint32_t w0 = (wblk * wg + wsub) * MAC_TILE;
    const uint32_t i0 = (blockIdx.y * ig + isub) * MAC_TILE;
    const uint32_t j0 = blockIdx.z * jchunk;
    const uint32_t j1 = min(m, j0 + jchunk);

    uint64_t acc[MAC_TILE][MAC_TILE];
    uint32_t cnt[MAC_TILE][MAC_TILE];
#pragma unroll
    for (int a = 0; a < MAC_TILE; a++)
#pragma unroll
        for (int b = 0; b < MAC_TILE; b++) {
            acc[a][b] = 0;
            cnt[a][b] = 0;
        }
    const size_t mn = (size_t)m * n;
    for (uint32_t j = j0; j < j1; j++) {
        uint32_t sv[MAC_TILE], av[MAC_TILE];
#pragma unroll
        for (int a = 0; a < MAC_TILE; a++)
            sv[a] = (w0 + a < W) ? shat[(size_t)(w0 + a) * mn + (size_t)j * n + t] : 0u;
#pragma unroll
        for (int b = 0; b < MAC_TILE; b++)
            av[b] = (i0 + b < k) ? Ahat[((size_t)(i0 + b) * m + j) * n + t] : 0u;
#pragma unroll
        for (int a = 0; a < MAC_TILE; a++)
#pragma unroll
            for (int b = 0; b < MAC_TILE; b++) {
                const uint64_t p = (uint64_t)sv[a] * av[b];
                const uint64_t s = acc[a][b] + p;
                cnt[a][b] += s < p;
                acc[a][b] = s;
            }
    }
    const size_t plane = (size_t)W * k * n;
#pragma unroll
    for (int a = 0; a < MAC_TILE; a++)
#pragma unroll
        for (int b = 0; b < MAC_TILE; b++) {
            if (w0 + a < W && i0 + b < k) {
                // value = cnt * 2^64 + acc, 2^64 = r64 (mod q)
                const uint32_t lo = reduce_u64(acc[a][b], q, barrett);
                const uint32_t hi = reduce_u64((uint64_t)cnt[a][b] * r64, q, barrett);
                partial[(size_t)blockIdx.z * plane + ((size_t)(w0 + a) * k + (i0 + b)) * n + t] = add_big(lo, hi, q);
            }
        }
}

__global__ void ajtai_reduce_kernel(const uint32_t *__restrict__ partial, uint32_t *__restrict__ out, size_t plane,
                                    uint32_t chunks, uint32_t q, uint64_t barrett)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < plane; e += (size_t)gridDim.x * blockDim.x) {
        uint64_t s = 0;
        for (uint32_t c = 0; c < chunks; c++) s += partial[(size_t)c * plane + e];
        out[e] = reduce_u64(s, q, barrett);
    }
}

// int8 balanced digits -> canonical residues (ring.rs:232-243 from_signed).
__global__ void digits_to_residues_kernel(const int8_t *__restrict__ d, uint32_t *__restrict__ out, size_t len,
                                          uint32_t q)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < len; e += (size_t)gridDim.x * blockDim.x) {
        const int32_t v = d[e];
        out[e] = v < 0 ? q - (uint32_t)(-v) : (uint32_t)v;
    }
}

static uint32_t grid_for(size_t len, uint32_t block = 256)
{
    return (uint32_t)std::min<size_t>((len + block - 1) / block, 1u << 16);
}

lzx_status ajtai_commit_ntt(const lzx_ajtai_key *key, const uint32_t *s_dev, const int8_t *digits_dev, size_t W,
                            uint32_t *out_dev, hipStream_t stream)
{
    if (W == 0) return LZX_OK;
    if (ajtai_nttmm_enabled(key)) {
        // the same sum with the multiply-accumulate on the matrix cores
        const lzx_status st = ajtai_commit_nttmm(key, s_dev, digits_dev, W, out_dev, stream);
        if (st != LZX_UNSUPPORTED) return st;
    }
    const size_t n = key->n, m = key->m, k = key->k;
    const uint32_t q = key->plan->q;
    const uint64_t barrett = key->plan->mod.barrett;
    const size_t slen = W * m * n;
    if (W > 0xFFFFFFFFull / 4 || slen / n > 0x7FFFFFFFull) return LZX_UNSUPPORTED;

    const uint32_t tb = (uint32_t)std::min<size_t>(n, MAC_TB);
    const uint32_t wg = (uint32_t)std::min<size_t>(4, (W + MAC_TILE - 1) / MAC_TILE);
    const uint32_t ig = (uint32_t)std::min<size_t>(4, (k + MAC_TILE - 1) / MAC_TILE);
    const uint32_t wblocks = (uint32_t)((W + (size_t)wg * MAC_TILE - 1) / ((size_t)wg * MAC_TILE));
    const uint32_t iblocks = (uint32_t)((k + (size_t)ig * MAC_TILE - 1) / ((size_t)ig * MAC_TILE));
    const uint32_t xy = (uint32_t)(n / tb) * wblocks * iblocks;
    // enough chunks for ~1024 blocks in flight, at least 8 ring elements per chunk
    uint32_t chunks = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)((m + 7) / 8), (1024 + xy - 1) / xy));
    chunks = std::min<uint32_t>(chunks, 65535);
    const uint32_t jchunk = (uint32_t)((m + chunks - 1) / chunks);
    chunks = (uint32_t)((m + jchunk - 1) / jchunk);
    const size_t plane = W * k * n;

    uint32_t *shat = nullptr, *partial = nullptr;
    LZX_HIP(scratch_alloc((void **)&shat, slen * sizeof(uint32_t), stream));
    hipError_t e = scratch_alloc((void **)&partial, (size_t)chunks * plane * sizeof(uint32_t), stream);
    if (e != hipSuccess) {
        (void)scratch_free(shat, stream);
        return device_fail(e, "ajtai partial alloc");
    }
    lzx_status st = LZX_OK;
    // digit witnesses: conversion fused into the forward NTT where the shape allows
    bool transformed = false;
    if (digits_dev) {
        st = ntt_forward_digits(key->plan, digits_dev, shat, W * m, stream);
        transformed = st == LZX_OK;
        if (st == LZX_UNSUPPORTED) st = LZX_OK;
    }
    if (st == LZX_OK && !transformed) {
        if (digits_dev) {
            hipLaunchKernelGGL(digits_to_residues_kernel, dim3(grid_for(slen)), dim3(256), 0, stream, digits_dev,
                               shat, slen, q);
            e = hipGetLastError();
        } else {
            e = hipMemcpyAsync(shat, s_dev, slen * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream);
        }
        if (e != hipSuccess) st = device_fail(e, "ajtai witness staging");
        if (st == LZX_OK)
            st = ntt_launch((const lzx_ntt_plan *const *)&key->plan, 1, shat, W * m, NTT_FWD, 0, stream);
    }
    if (st == LZX_OK) {
        const uint32_t r64 = (uint32_t)((((unsigned __int128)1) << 64) % q);
        hipLaunchKernelGGL(ajtai_mac_kernel, dim3((uint32_t)(n / tb) * wblocks, iblocks, chunks), dim3(tb * wg * ig), 0, stream, key->Ahat_dev, shat,
                           partial, (uint32_t)W, (uint32_t)k, (uint32_t)m, (uint32_t)n, tb, wg, ig, jchunk, q,
                           barrett, r64);
        e = hipGetLastError();
        if (e != hipSuccess) st = device_fail(e, "ajtai_mac_kernel");
    }
    if (st == LZX_OK) {
        hipLaunchKernelGGL(ajtai_reduce_kernel, dim3(grid_for(plane)), dim3(256), 0, stream, partial, out_dev, plane,
                           chunks, q, barrett);
        e = hipGetLastError();
        if (e != hipSuccess) st = device_fail(e, "ajtai_reduce_kernel");
    }
    if (st == LZX_OK) st = ntt_launch((const lzx_ntt_plan *const *)&key->plan, 1, out_dev, W * k, NTT_INV, 0, stream);
    (void)scratch_free(partial, stream);
    (void)scratch_free(shat, stream);
    return st;
}

// Ahat = NTT(A) and the int8 limb planes, from A already in key->A_dev.
static lzx_status key_finish_device(lzx_ajtai_key *key)
{
    const size_t words = key->k * key->m * key->n;
    LZX_HIP(hipMalloc(&key->Ahat_dev, words * sizeof(uint32_t)));
    LZX_HIP(hipMemcpy(key->Ahat_dev, key->A_dev, words * sizeof(uint32_t), hipMemcpyDeviceToDevice));
    lzx_status st = ntt_launch((const lzx_ntt_plan *const *)&key->plan, 1, key->Ahat_dev, key->k * key->m, NTT_FWD,
                               0, nullptr);
    if (st != LZX_OK) return st;
    if (ajtai_int8_supported(key)) {
        st = ajtai_build_limbs(key);
        if (st != LZX_OK) return st;
    }
    if (ajtai_nttmm_supported(key)) {
        st = ajtai_build_nttmm(key);
        if (st != LZX_OK) return st;
    }
    LZX_HIP(hipDeviceSynchronize());
    return LZX_OK;
}

static lzx_status key_finish(lzx_ajtai_key *key, const uint32_t *A_host)
{
    const size_t words = key->k * key->m * key->n;
    LZX_HIP(hipMalloc(&key->A_dev, words * sizeof(uint32_t)));
    LZX_HIP(hipMemcpy(key->A_dev, A_host, words * sizeof(uint32_t), hipMemcpyHostToDevice));
    return key_finish_device(key);
}

// ---- seeded derivation on the device ----------------------------------------
//
// RingConfig::uniform_from_seed (ring.rs:102-180) for element idx: one SHAKE-256
// stream XOF("uniform", salt = domain || seed || u64le(idx)) (transcript.rs:
// 229-237 framing), read as 4-byte little-endian words from byte 8 on; a word
// v < 2^32 - (2^32 mod q) gives the next coefficient v mod q.  The reference
// first reads n + max(n, 16) words and then re-squeezes a fresh XOF for each
// further word, taking its last 4 bytes; SHAKE output is prefix-stable, so
// those are exactly the next words of the same stream, and the kernel simply
// keeps squeezing until n coefficients are accepted.
//
// The absorbed message ("LZXX" || u32le 7 || "uniform" || u32le |salt| ||
// salt) fits one 136-byte block; the host builds that padded block with
// idx = 0 and the kernel XORs each thread's idx in at byte offset `io`.
struct XofBlock {
    uint64_t lane[17];
};

__global__ __launch_bounds__(256) void derive_uniform_kernel(XofBlock base, uint32_t io, uint32_t q, uint32_t n,
                                                             uint64_t total, uint32_t *__restrict__ A)
{
    const uint32_t limit = (uint32_t)((1ULL << 32) - ((1ULL << 32) % q));
    const uint32_t L = io >> 3, sh = 8 * (io & 7);
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t st[25];
#pragma unroll
        for (int l = 0; l < 25; l++) st[l] = l < 17 ? base.lane[l] : 0;
#pragma unroll
        for (int l = 0; l < 16; l++) {
            if (l == (int)L) st[l] ^= idx << sh;
            if (l == (int)L + 1 && sh) st[l] ^= idx >> (64 - sh);
        }
        keccak_f1600(st);
        uint32_t *out = A + idx * n;
        uint32_t got = 0;
        for (uint32_t blk = 0; got < n; blk++) {
            if (blk) keccak_f1600(st);
#pragma unroll
            for (int w = 0; w < 34; w++) { // the block's 4-byte words; stream words 0 and 1 are skipped
                const uint32_t v = (uint32_t)(st[w >> 1] >> (32 * (w & 1)));
                if ((blk || w >= 2) && got < n && v < limit) out[got++] = v % q;
            }
        }
    }
}

// Host side of the block: the framed XOF message
//   "LZXX" || u32le |xdom| || xdom || u32le |msg| || msg,   msg = pre1 || pre2 || var
// (transcript.rs:229-237) with the var_len variable bytes zero, padded as one
// SHAKE-256 block; the kernel XORs its variable bytes in at offset io.
// Returns false if the message does not fit one block.
static bool build_xof_block(const char *xdom, const uint8_t *pre1, size_t len1, const uint8_t *pre2, size_t len2,
                            size_t var_len, XofBlock &blk, uint32_t &io)
{
    const size_t xdlen = strlen(xdom);
    if (12 + xdlen + len1 + len2 + var_len > 135) return false;
    uint8_t b[136] = {};
    size_t o = 0;
    auto put = [&](const void *p, size_t len) {
        if (len) memcpy(b + o, p, len);
        o += len;
    };
    auto put32 = [&](uint32_t v) {
        const uint8_t x[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
        put(x, 4);
    };
    put("LZXX", 4);
    put32((uint32_t)xdlen);
    put(xdom, xdlen);
    put32((uint32_t)(len1 + len2 + var_len));
    put(pre1, len1);
    put(pre2, len2);
    io = (uint32_t)o;
    o += var_len;
    b[o] ^= 0x1f;  // SHAKE-256 domain suffix
    b[135] ^= 0x80; // pad10*1
    for (int l = 0; l < 17; l++) {
        uint64_t v = 0;
        for (int i = 0; i < 8; i++) v |= (uint64_t)b[8 * l + i] << (8 * i);
        blk.lane[l] = v;
    }
    return true;
}

// XORs the 8-byte little-endian word v into the absorbed block at byte io.
__device__ __forceinline__ void xor_u64_at(uint64_t (&st)[25], uint32_t io, uint64_t v)
{
    const uint32_t L = io >> 3, sh = 8 * (io & 7);
#pragma unroll
    for (int l = 0; l < 17; l++) {
        if (l == (int)L) st[l] ^= v << sh;
        if (l == (int)L + 1 && sh) st[l] ^= v >> (64 - sh);
    }
}

// sample_small_secret (ajtai.rs:307-348), one thread per ring element i: for
// ctr = 0, 1, ...: stream = XOF("ajtai-secret", seed || u64le(i) || u64le(ctr),
// words_per_try * 4 bytes), 4-byte LE words from byte 0; a word v < limit =
// floor((2^32 - 1) / span) * span gives the next coefficient (v mod span) -
// bound, stored as its residue mod q (and as an int8 digit when asked).
__global__ __launch_bounds__(256) void small_secret_kernel(XofBlock base, uint32_t io, uint32_t words_per_try,
                                                           uint32_t bound, uint32_t span, uint32_t limit, uint32_t q,
                                                           uint32_t n, uint64_t m, uint32_t *__restrict__ out,
                                                           int8_t *__restrict__ digits)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t got = 0;
        for (uint64_t ctr = 0; got < n; ctr++) {
            uint64_t st[25];
#pragma unroll
            for (int l = 0; l < 25; l++) st[l] = l < 17 ? base.lane[l] : 0;
            xor_u64_at(st, io, i);
            xor_u64_at(st, io + 8, ctr);
            for (uint32_t w0 = 0; w0 < words_per_try && got < n; w0 += 34) {
                keccak_f1600(st);
#pragma unroll
                for (int w = 0; w < 34; w++) {
                    const uint32_t v = (uint32_t)(st[w >> 1] >> (32 * (w & 1)));
                    if (w0 + w < words_per_try && got < n && v < limit) {
                        const int64_t c = (int64_t)(v % span) - (int64_t)bound;
                        const int64_t r = c % (int64_t)q;
                        if (out) out[i * n + got] = (uint32_t)(r < 0 ? r + q : r);
                        if (digits) digits[i * n + got] = (int8_t)c;
                        got++;
                    }
                }
            }
        }
    }
}

// commit_naive_reference (ajtai.rs:158-180): the k*m ring products A_ij * s_j
// one by one (NTT of A_ij and of s_j, pointwise product, inverse NTT per
// product), then a coefficient-domain sum over j.  P[i][j][t] = X[i][j][t] *
// S[j][t] forms the products in the NTT domain; naive_sum_kernel sums the
// inverse-transformed products per row.
__global__ void naive_pointwise_kernel(uint32_t *__restrict__ P, const uint32_t *__restrict__ S, size_t mn,
                                       size_t total, uint32_t q, uint64_t barrett)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
        P[e] = reduce_u64((uint64_t)P[e] * S[e % mn], q, barrett);
}

__global__ void naive_sum_kernel(const uint32_t *__restrict__ P, uint32_t *__restrict__ out, uint32_t k, uint32_t m,
                                 uint32_t n, uint32_t q)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < (size_t)k * n;
         e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e / n, t = e % n;
        uint32_t acc = 0;
        for (uint32_t j = 0; j < m; j++) acc = add_big(acc, P[(i * m + j) * n + t], q);
        out[e] = acc;
    }
}

static lzx_status key_new(uint32_t q, uint32_t adic, uint32_t g, uint32_t log_n, size_t k, size_t m,
                          uint32_t norm_bound, lzx_ajtai_key **out, lzx_ajtai_key **key)
{
    if (!out) return LZX_INVALID_ARGUMENT;
    *out = nullptr;
    if (k == 0 || m == 0) return LZX_INVALID_ARGUMENT;
    lzx_ntt_plan *plan = nullptr;
    lzx_status st = lzx_ntt_plan_create(q, adic, g, log_n, &plan);
    if (st != LZX_OK) return st;
    lzx_ajtai_key *kk = new lzx_ajtai_key();
    kk->plan = plan;
    kk->k = k;
    kk->m = m;
    kk->n = plan->n;
    kk->norm_bound = norm_bound;
    kk->A_dev = kk->Ahat_dev = nullptr;
    kk->limbs_dev = nullptr;
    kk->limbs_bytes = 0;
    kk->nm_afrag = nullptr;
    kk->nm_corr = nullptr;
    kk->nm_psi = nullptr;
    kk->device = plan->device;
    *key = kk;
    return LZX_OK;
}

} // namespace lzx

using namespace lzx;

extern "C" void lzx_ajtai_key_destroy(lzx_ajtai_key *key)
{
    if (!key) return;
    if (key->A_dev) (void)hipFree(key->A_dev);
    if (key->Ahat_dev) (void)hipFree(key->Ahat_dev);
    if (key->limbs_dev) (void)hipFree(key->limbs_dev);
    if (key->nm_afrag) (void)hipFree(key->nm_afrag);
    if (key->nm_corr) (void)hipFree(key->nm_corr);
    if (key->nm_psi) (void)hipFree(key->nm_psi);
    lzx_ntt_plan_destroy(key->plan);
    delete key;
}

extern "C" lzx_status lzx_ajtai_key_from_matrix(uint32_t q, uint32_t two_adicity, uint32_t generator, uint32_t log_n,
                                                size_t k, size_t m, uint32_t norm_bound, const uint32_t *A_host,
                                                lzx_ajtai_key **out)
{
    if (!A_host) return LZX_INVALID_ARGUMENT;
    lzx_ajtai_key *key = nullptr;
    lzx_status st = key_new(q, two_adicity, generator, log_n, k, m, norm_bound, out, &key);
    if (st != LZX_OK) return st;
    st = key_finish(key, A_host);
    if (st != LZX_OK) {
        lzx_ajtai_key_destroy(key);
        return st;
    }
    *out = key;
    return LZX_OK;
}

extern "C" lzx_status lzx_ajtai_key_from_seed(uint32_t q, uint32_t two_adicity, uint32_t generator, uint32_t log_n,
                                              size_t k, size_t m, uint32_t norm_bound, const uint8_t seed[32],
                                              lzx_ajtai_key **out)
{
    if (!seed) return LZX_INVALID_ARGUMENT;
    lzx_ajtai_key *key = nullptr;
    lzx_status st = key_new(q, two_adicity, generator, log_n, k, m, norm_bound, out, &key);
    if (st != LZX_OK) return st;
    const size_t total = k * m, n = key->n;
    static const uint8_t DOM[] = {'a', 'j', 't', 'a', 'i', '-', 'A'};
    // ajtai.rs:91-96: element idx = i*m + j, derived on the device
    XofBlock blk;
    uint32_t io = 0;
    build_xof_block("uniform", DOM, 7, seed, 32, 8, blk, io);
    hipError_t e = hipMalloc(&key->A_dev, total * n * sizeof(uint32_t));
    if (e == hipSuccess) {
        const uint32_t blocks = (uint32_t)std::min<size_t>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(derive_uniform_kernel, dim3(blocks), dim3(256), 0, nullptr, blk, io, q, (uint32_t)n,
                           (uint64_t)total, key->A_dev);
        e = hipGetLastError();
    }
    st = e == hipSuccess ? key_finish_device(key) : device_fail(e, "derive_uniform_kernel");
    if (st != LZX_OK) {
        lzx_ajtai_key_destroy(key);
        return st;
    }
    *out = key;
    return LZX_OK;
}

extern "C" lzx_status lzx_ajtai_key_matrix(const lzx_ajtai_key *key, uint32_t *A_host)
{
    if (!key || !A_host) return LZX_INVALID_ARGUMENT;
    LZX_HIP(hipMemcpy(A_host, key->A_dev, key->k * key->m * key->n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return LZX_OK;
}

extern "C" lzx_status lzx_ajtai_key_matrix_ntt(const lzx_ajtai_key *key, uint32_t *Ahat_host)
{
    if (!key || !Ahat_host) return LZX_INVALID_ARGUMENT;
    LZX_HIP(hipMemcpy(Ahat_host, key->Ahat_dev, key->k * key->m * key->n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return LZX_OK;
}

extern "C" lzx_status lzx_ajtai_commit_batch(const lzx_ajtai_key *key, const uint32_t *s_dev, size_t s_len, size_t W,
                                             uint32_t *out_dev, void *stream)
{
    if (!key) return LZX_INVALID_ARGUMENT;
    if (s_len != key->m) return LZX_AJTAI_DIMENSION_MISMATCH; // ajtai.rs:119-124
    if (W && (!s_dev || !out_dev)) return LZX_INVALID_ARGUMENT;
    return ajtai_commit_ntt(key, s_dev, nullptr, W, out_dev, (hipStream_t)stream);
}

// ---- LZX_COMMIT_AUTO: the faster path per shape, by measurement ---------
//
// Both digit paths give the same commitment.  The first AUTO call for a
// (device, q, n, k, m, W) times each path on the caller's stream (after one
// untimed run of each) and keeps the faster in the tuning cache (persisted
// through LZX_AUTOTUNE_CACHE); later calls run the kept path.  A key without
// int8 support always takes the NTT path; under stream capture, with no
// kept choice, AUTO takes int8 without measuring.
namespace {
std::string path_key(const lzx_ajtai_key *key, size_t W)
{
    char buf[128];
    snprintf(buf, sizeof buf, "ajtai_commit_path dev=%d q=%u n=%zu k=%zu m=%zu W=%zu", key->device, key->plan->q,
             key->n, key->k, key->m, W);
    return buf;
}

lzx_status run_path(const lzx_ajtai_key *key, int32_t path, const int8_t *digits_dev, size_t W, uint32_t *out_dev,
                    hipStream_t s)
{
    return path == LZX_COMMIT_INT8 ? ajtai_commit_int8(key, digits_dev, W, out_dev, s)
                                   : ajtai_commit_ntt(key, nullptr, digits_dev, W, out_dev, s);
}

lzx_status commit_auto(const lzx_ajtai_key *key, const int8_t *digits_dev, size_t W, uint32_t *out_dev,
                       hipStream_t s)
{
    if (!ajtai_int8_supported(key)) return run_path(key, LZX_COMMIT_NTT, digits_dev, W, out_dev, s);
    const std::string pk = path_key(key, W);
    uint32_t kept = 0;
    if (autotune_get(pk, kept) && (kept == LZX_COMMIT_NTT || kept == LZX_COMMIT_INT8))
        return run_path(key, (int32_t)kept, digits_dev, W, out_dev, s);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    LZX_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) return run_path(key, LZX_COMMIT_INT8, digits_dev, W, out_dev, s);
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    lzx_status st = LZX_OK;
    for (int i = 0; i < 3 && st == LZX_OK; i++)
        if (hipEventCreate(&ev[i]) != hipSuccess) st = device_fail(hipGetLastError(), "commit path events");
    float t_i8 = 0, t_ntt = 0;
    if (st == LZX_OK) st = run_path(key, LZX_COMMIT_INT8, digits_dev, W, out_dev, s); // untimed: first-call costs
    if (st == LZX_OK) st = run_path(key, LZX_COMMIT_NTT, digits_dev, W, out_dev, s);
    if (st == LZX_OK && hipEventRecord(ev[0], s) != hipSuccess) st = device_fail(hipGetLastError(), "event");
    if (st == LZX_OK) st = run_path(key, LZX_COMMIT_INT8, digits_dev, W, out_dev, s);
    if (st == LZX_OK && hipEventRecord(ev[1], s) != hipSuccess) st = device_fail(hipGetLastError(), "event");
    if (st == LZX_OK) st = run_path(key, LZX_COMMIT_NTT, digits_dev, W, out_dev, s); // out: the commitment
    if (st == LZX_OK && hipEventRecord(ev[2], s) != hipSuccess) st = device_fail(hipGetLastError(), "event");
    if (st == LZX_OK) {
        hipError_t e = hipEventSynchronize(ev[2]);
        if (e == hipSuccess) e = hipEventElapsedTime(&t_i8, ev[0], ev[1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&t_ntt, ev[1], ev[2]);
        if (e != hipSuccess) st = device_fail(e, "commit path timing");
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (st == LZX_OK) autotune_put(pk, t_i8 <= t_ntt ? LZX_COMMIT_INT8 : LZX_COMMIT_NTT);
    return st;
}
} // namespace

extern "C" lzx_status lzx_ajtai_commit_digits_batch(const lzx_ajtai_key *key, const int8_t *digits_dev, size_t s_len,
                                                    size_t W, uint32_t *out_dev, int32_t path, void *stream)
{
    if (!key) return LZX_INVALID_ARGUMENT;
    if (s_len != key->m) return LZX_AJTAI_DIMENSION_MISMATCH;
    if (W && (!digits_dev || !out_dev)) return LZX_INVALID_ARGUMENT;
    if (path < LZX_COMMIT_AUTO || path > LZX_COMMIT_INT8) return LZX_INVALID_ARGUMENT;
    if (path == LZX_COMMIT_INT8 && !ajtai_int8_supported(key)) return LZX_UNSUPPORTED;
    if (W == 0) return LZX_OK;
    if (path == LZX_COMMIT_AUTO) return commit_auto(key, digits_dev, W, out_dev, (hipStream_t)stream);
    return run_path(key, path, digits_dev, W, out_dev, (hipStream_t)stream);
}

extern "C" lzx_status lzx_ajtai_commit_path_choice(const lzx_ajtai_key *key, size_t W, int32_t *path_out)
{
    if (!key || !path_out) return LZX_INVALID_ARGUMENT;
    uint32_t kept = 0;
    if (!ajtai_int8_supported(key)) *path_out = LZX_COMMIT_NTT;
    else if (autotune_get(path_key(key, W), kept) && (kept == LZX_COMMIT_NTT || kept == LZX_COMMIT_INT8))
        *path_out = (int32_t)kept;
    else *path_out = LZX_COMMIT_AUTO;
    return LZX_OK;
}

extern "C" lzx_status lzx_ajtai_commit_host(const lzx_ajtai_key *key, const uint32_t *s_host, size_t s_len,
                                            uint32_t *out_host)
{
    if (!key || !s_host || !out_host) return LZX_INVALID_ARGUMENT;
    if (s_len != key->m) return LZX_AJTAI_DIMENSION_MISMATCH;
    const size_t sb = key->m * key->n * sizeof(uint32_t), ob = key->k * key->n * sizeof(uint32_t);
    return host_slice_call(key->device, s_host, sb, out_host, sb, ob, sb + ob, [&](void *d, hipStream_t s) {
        return ajtai_commit_ntt(key, (const uint32_t *)d, nullptr, 1, (uint32_t *)((uint8_t *)d + sb), s);
    });
}

// commit_naive_reference for W witnesses in HBM (see naive_pointwise_kernel).
static lzx_status commit_naive(const lzx_ajtai_key *key, const uint32_t *s_dev, size_t W, uint32_t *out_dev,
                               hipStream_t s)
{
    if (W == 0) return LZX_OK;
    const size_t n = key->n, m = key->m, k = key->k, km = k * m, mn = m * n;
    uint32_t *P = nullptr;
    LZX_HIP(scratch_alloc((void **)&P, (km + m) * n * sizeof(uint32_t), s));
    uint32_t *S = P + km * n;
    const uint32_t q = key->plan->q;
    lzx_status st = LZX_OK;
    auto grid = [](size_t len) { return (uint32_t)std::min<size_t>((len + 255) / 256, 65536); };
    for (size_t w = 0; w < W && st == LZX_OK; w++) {
        hipError_t e = hipMemcpyAsync(P, key->A_dev, km * n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(S, s_dev + w * mn, mn * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) st = device_fail(e, "naive commit staging");
        if (st == LZX_OK) st = ntt_launch((const lzx_ntt_plan *const *)&key->plan, 1, P, km + m, NTT_FWD, 0, s);
        if (st == LZX_OK) {
            hipLaunchKernelGGL(naive_pointwise_kernel, dim3(grid(km * n)), dim3(256), 0, s, P, S, mn, km * n, q,
                               key->plan->mod.barrett);
            e = hipGetLastError();
            if (e != hipSuccess) st = device_fail(e, "naive_pointwise_kernel");
        }
        if (st == LZX_OK) st = ntt_launch((const lzx_ntt_plan *const *)&key->plan, 1, P, km, NTT_INV, 0, s);
        if (st == LZX_OK) {
            hipLaunchKernelGGL(naive_sum_kernel, dim3(grid(k * n)), dim3(256), 0, s, P, out_dev + w * k * n,
                               (uint32_t)k, (uint32_t)m, (uint32_t)n, q);
            e = hipGetLastError();
            if (e != hipSuccess) st = device_fail(e, "naive_sum_kernel");
        }
    }
    (void)scratch_free(P, s);
    return st;
}

extern "C" lzx_status lzx_ajtai_commit_naive_batch(const lzx_ajtai_key *key, const uint32_t *s_dev, size_t s_len,
                                                   size_t W, uint32_t *out_dev, void *stream)
{
    if (!key) return LZX_INVALID_ARGUMENT;
    if (s_len != key->m) return LZX_AJTAI_DIMENSION_MISMATCH; // ajtai.rs:159-164
    if (W && (!s_dev || !out_dev)) return LZX_INVALID_ARGUMENT;
    if (key->k * key->m > 0xFFFFFFFFull / key->n) return LZX_UNSUPPORTED;
    return commit_naive(key, s_dev, W, out_dev, (hipStream_t)stream);
}

extern "C" lzx_status lzx_ajtai_commit_naive_host(const lzx_ajtai_key *key, const uint32_t *s_host, size_t s_len,
                                                  uint32_t *out_host)
{
    if (!key || !s_host || !out_host) return LZX_INVALID_ARGUMENT;
    if (s_len != key->m) return LZX_AJTAI_DIMENSION_MISMATCH;
    if (key->k * key->m > 0xFFFFFFFFull / key->n) return LZX_UNSUPPORTED;
    const size_t sb = key->m * key->n * sizeof(uint32_t), ob = key->k * key->n * sizeof(uint32_t);
    return host_slice_call(key->device, s_host, sb, out_host, sb, ob, sb + ob, [&](void *d, hipStream_t s) {
        return commit_naive(key, (const uint32_t *)d, 1, (uint32_t *)((uint8_t *)d + sb), s);
    });
}

extern "C" lzx_status lzx_ajtai_sample_small_secret(uint32_t q, uint32_t log_n, size_t m, uint32_t bound,
                                                    const uint8_t *seed, size_t seed_len, uint32_t *out_dev,
                                                    int8_t *digits_dev, void *stream)
{
    if ((seed_len && !seed) || q < 2) return LZX_INVALID_ARGUMENT;
    if (m == 0) return LZX_OK;
    if (!out_dev && !digits_dev) return LZX_INVALID_ARGUMENT;
    // span = 2*bound + 1 must fit a u32 word (otherwise no word is ever accepted)
    if (bound >= (1u << 31) || (digits_dev && bound > 127)) return LZX_INVALID_ARGUMENT;
    if (log_n > NTT_MAX_LOG_N) return LZX_UNSUPPORTED;
    XofBlock blk;
    uint32_t io = 0;
    if (!build_xof_block("ajtai-secret", seed, seed_len, nullptr, 0, 16, blk, io)) return LZX_UNSUPPORTED;
    const uint32_t n = 1u << log_n;
    const uint64_t span = 2ull * bound + 1;
    const uint32_t limit = (uint32_t)((0xFFFFFFFFull / span) * span);
    const uint32_t words = std::max<uint32_t>(16 * n, 64) / 4;
    const uint32_t blocks = (uint32_t)std::min<size_t>((m + 255) / 256, 65536);
    hipLaunchKernelGGL(small_secret_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, blk, io, words, bound,
                       (uint32_t)span, limit, q, n, (uint64_t)m, out_dev, digits_dev);
    LZX_LAUNCH_CHECK("small_secret_kernel");
    return LZX_OK;
}

extern "C" lzx_status lzx_ajtai_verify_opening_host(const lzx_ajtai_key *key, const uint32_t *t_host,
                                                    const uint32_t *s_host, size_t s_len)
{
    if (!key || !t_host || !s_host) return LZX_INVALID_ARGUMENT;
    if (s_len != key->m) return LZX_AJTAI_DIMENSION_MISMATCH;
    // ajtai.rs:216-224: infinity norm of the balanced representative
    const uint32_t q = key->plan->q, half = q / 2;
    for (size_t i = 0; i < key->m * key->n; i++) {
        const uint32_t c = s_host[i];
        const uint32_t norm = c <= half ? c : q - c;
        if (norm > key->norm_bound) return LZX_AJTAI_NORM_EXCEEDED;
    }
    std::vector<uint32_t> t(key->k * key->n);
    lzx_status st = lzx_ajtai_commit_host(key, s_host, s_len, t.data());
    if (st != LZX_OK) return st;
    for (size_t i = 0; i < t.size(); i++)
        if (t[i] != t_host[i]) return LZX_AJTAI_COMMITMENT_MISMATCH;
    return LZX_OK;
}

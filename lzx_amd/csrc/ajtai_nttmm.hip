// NTT-domain Ajtai commit with the multiply-accumulate on int8 matrix cores.
//
// Reference: crates/lattice-commitment/src/ajtai.rs:118-153 (commit: forward
// NTT of every s_j, t^_i = sum_j A^_ij (.) s^_j, one inverse NTT per row).  The
// sum is the reference's; only its evaluation differs.  For each coefficient
// c it is the matrix product
//   T_c[i][w] = sum_j A^_c[i][j] * S^_c[j][w]   (mod q),
// computed exactly on v_mfma_i32_32x32x32_i8 with both operands split into
// bytes:
//   - key side: the centred A^ value a (|a| < 2^31 - 2^24 for q < 2^32 - 2^25)
//     as four balanced limbs a = sum_La 256^La a_La, a_La in [-128, 127]; the
//     rows of the product are (i, La), row 4i + La;
//   - witness side: the canonical s^ value as four unsigned bytes b_Lb, each
//     stored as b_Lb - 128 (the byte XOR 0x80), so that
//     s^ = sum_Lb 256^Lb (b_Lb - 128) + 0x80808080; the columns are (w, Lb),
//     column 4w + Lb;
//   - one K step of an MFMA is 32 ring elements.  Every product is at most
//     2^14 in magnitude, so a chunk of up to 2^16 ring elements sums exactly
//     in int32;
//   - per output, sum_{La,Lb} 256^(La+Lb) P[La][Lb] is folded mod q in the
//     epilogue (the four La of an output sit in four registers of one lane,
//     the four Lb in a lane quad); the remaining 0x80808080 * sum_j a_ij[c]
//     depends on the key only and is added by the reduction over chunks.
// That is 1024 int8 MACs per ring-MAC (64 coefficients x 16 limb pairs),
// against 16384 for the coefficient-domain Toeplitz product, and the kernel is
// bound by HBM: the witness's NTT-domain bytes and the key's limb fragments
// are each read once.
//
// Fragment layouts (v_mfma_i32_32x32x32_i8, lane l: A row / B column l mod 32,
// K bytes 16 (l / 32) .. + 16), 1 KiB per fragment, 16 contiguous bytes per
// lane:
//   key      Afrag[c][jb][rb][l][b]: row 32 rb + l % 32, ring element
//            j = 32 jb + 16 (l / 32) + b;
//   witness  Bfrag[c][jb][cb][l][b]: column 32 cb + l % 32, same j.
// Rows past 4k, columns past 4W and ring elements past m are zero on the key
// side (so the witness side may hold anything there).
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "ajtai.hpp"
#include "ntt_regs.hpp"

namespace lzx {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr uint64_t NM_Q_LIMIT = (1ULL << 32) - (1ULL << 25); // balanced 4-limb split of the centred value
constexpr uint32_t NM_MAX_K = 16;                           // rows (i, La) in at most two 32-row blocks
constexpr uint32_t NM_MAX_JBC = 2048;                       // 2048 * 32 ring elements * 2^14 = 2^30 < 2^31
constexpr uint32_t NM_CORR_J = 256;                         // ring elements per partial sum of the key constant

// ---- key side ---------------------------------------------------------------

// One thread per (c, jb, rb, lane): 16 limb bytes of 16 consecutive ring elements.
__global__ void nm_key_frag_kernel(const uint32_t *__restrict__ Ahat, int8_t *__restrict__ frag, uint32_t k,
                                   uint32_t m, uint32_t n, uint32_t njb, uint32_t nrb, uint32_t q)
{
    const size_t total = (size_t)n * njb * nrb * 64;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const uint32_t lane = (uint32_t)(e & 63);
        const size_t f = e >> 6; // fragment index (c, jb, rb)
        const uint32_t rb = (uint32_t)(f % nrb);
        const uint32_t jb = (uint32_t)((f / nrb) % njb);
        const uint32_t c = (uint32_t)(f / ((size_t)nrb * njb));
        const uint32_t row = 32 * rb + (lane & 31), i = row >> 2, La = row & 3;
        const uint32_t j0 = 32 * jb + 16 * (lane >> 5);
        uint32_t out[4] = {0, 0, 0, 0};
        if (i < k) {
            for (uint32_t b = 0; b < 16; b++) {
                const uint32_t j = j0 + b;
                if (j >= m) break;
                const uint32_t a = Ahat[((size_t)i * m + j) * n + c];
                int64_t x = a <= (q - 1) / 2 ? (int64_t)a : (int64_t)a - (int64_t)q;
                int64_t d = 0;
                for (uint32_t t = 0; t <= La; t++) {
                    d = ((x + 128) & 255) - 128;
                    x = (x - d) >> 8;
                }
                out[b >> 2] |= (uint32_t)(uint8_t)(int8_t)d << (8 * (b & 3));
            }
        }
        *reinterpret_cast<uint4 *>(frag + 16 * e) = make_uint4(out[0], out[1], out[2], out[3]);
    }
}

// part[i][jc][c] = sum of A^[i][j][c] over the jc-th run of NM_CORR_J ring elements (u64).
__global__ void nm_key_colsum_kernel(const uint32_t *__restrict__ Ahat, uint64_t *__restrict__ part, uint32_t m,
                                     uint32_t n, uint32_t njc)
{
    const uint32_t i = blockIdx.y, jc = blockIdx.x;
    const uint32_t j0 = jc * NM_CORR_J, j1 = min(m, j0 + NM_CORR_J);
    for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
        uint64_t s = 0;
        for (uint32_t j = j0; j < j1; j++) s += Ahat[((size_t)i * m + j) * n + c];
        part[((size_t)i * njc + jc) * n + c] = s;
    }
}

// corr[i][c] = 0x80808080 * sum_j A^[i][j][c] mod q.
__global__ void nm_key_corr_kernel(const uint64_t *__restrict__ part, uint32_t *__restrict__ corr, uint32_t k,
                                   uint32_t n, uint32_t njc, uint32_t q, uint64_t barrett, uint32_t c80)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= k * n) return;
    const uint32_t i = e / n, c = e % n;
    uint64_t s = 0;
    for (uint32_t jc = 0; jc < njc; jc++) s += reduce_u64(part[((size_t)i * njc + jc) * n + c], q, barrett);
    corr[e] = reduce_u64((uint64_t)reduce_u64(s, q, barrett) * c80, q, barrett);
}

// ---- witness side -----------------------------------------------------------

// The byte Lb of each of four words, packed low to high.
__device__ __forceinline__ uint32_t nm_bytes(uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3, uint32_t Lb)
{
    // v_perm_b32(hi, lo, sel): selector bytes 0-3 pick lo's bytes, 4-7 hi's
    const uint32_t sel = Lb | (Lb + 4) << 8 | Lb << 16 | (Lb + 4) << 24;
    const uint32_t lo = __builtin_amdgcn_perm(v1, v0, sel); // v0.Lb, v1.Lb, ..
    const uint32_t hi = __builtin_amdgcn_perm(v3, v2, sel); // v2.Lb, v3.Lb, ..
    return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

// Forward NTT of 8 witnesses x 32 ring elements (one polynomial per thread, as
// in the register kernels), then every coefficient's B fragment of that block
// of columns and ring elements, one contiguous 1 KiB each.  In is int8_t
// (balanced digits, ring.rs:232-243 from_signed) or uint32_t (canonical
// residues).  Polynomials past W or m transform as zero.
template <int LOG_N, int AR, typename In>
__global__ __launch_bounds__(256) void nm_witness_frag_kernel(const In *__restrict__ s, int8_t *__restrict__ frag,
                                                              const uint32_t *__restrict__ tab, uint32_t W,
                                                              uint32_t m, uint32_t njb, uint32_t ncb)
{
    constexpr uint32_t N = 1u << LOG_N, RS = N + 1;
    // polynomial t at row t; its coefficient c at column (c + 2 (t / 32)) mod N, so the fragment reads (16
    // ring elements of one column, for 16 column/half pairs at once) touch 32 distinct banks per 32 lanes
    __shared__ uint32_t T[256 * RS];
    const uint32_t t = threadIdx.x;
    const uint32_t jb = blockIdx.x % njb, cb = blockIdx.x / njb;
    const uint32_t w = 8 * cb + (t >> 5), j = 32 * jb + (t & 31);
    const uint32_t q = tab[0];
    ntt_val_t<AR> a[N];
    if (w < W && j < m) {
        const In *src = s + ((size_t)w * m + j) * N;
        if constexpr (sizeof(In) == 1) {
#pragma unroll
            for (uint32_t i = 0; i < N / 16; i++) {
                const uint4 v = reinterpret_cast<const uint4 *>(src)[i];
                const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (uint32_t e = 0; e < 16; e++) {
                    const int x = (int8_t)(wd[e >> 2] >> (8 * (e & 3)));
                    a[16 * i + e] = x < 0 ? q - (uint32_t)(-x) : (uint32_t)x;
                }
            }
        } else {
#pragma unroll
            for (uint32_t i = 0; i < N / 4; i++) {
                const uint4 v = reinterpret_cast<const uint4 *>(src)[i];
                a[4 * i] = v.x, a[4 * i + 1] = v.y, a[4 * i + 2] = v.z, a[4 * i + 3] = v.w;
            }
        }
    } else {
#pragma unroll
        for (uint32_t i = 0; i < N; i++) a[i] = 0;
    }
    ntt_regs_any<LOG_N, AR, NTT_FWD>(a, tab);
    const uint32_t sw = 2 * (t >> 5);
#pragma unroll
    for (uint32_t c = 0; c < N; c++) T[t * RS + ((c + sw) & (N - 1))] = ntt_out<AR>(a[brv_bits(c, LOG_N)], tab, N);
    __syncthreads();
    // item (c, r): r = 8 g + w' (witness 8 cb + w', ring elements 32 jb + 16 g ..) -> lane slots 4 r + Lb
    for (uint32_t it = t; it < 16 * N; it += 256) {
        const uint32_t c = it >> 4, r = it & 15, g = r >> 3, wp = r & 7;
        const uint32_t col = (c + 2 * wp) & (N - 1);
        uint32_t v[16];
#pragma unroll
        for (uint32_t b = 0; b < 16; b++) v[b] = T[(32 * wp + 16 * g + b) * RS + col] ^ 0x80808080u;
        int8_t *dst = frag + (((size_t)c * njb + jb) * ncb + cb) * 1024 + 64 * r;
#pragma unroll
        for (uint32_t Lb = 0; Lb < 4; Lb++)
            *reinterpret_cast<uint4 *>(dst + 16 * Lb) =
                make_uint4(nm_bytes(v[0], v[1], v[2], v[3], Lb), nm_bytes(v[4], v[5], v[6], v[7], Lb),
                           nm_bytes(v[8], v[9], v[10], v[11], Lb), nm_bytes(v[12], v[13], v[14], v[15], Lb));
    }
}

// ---- the product --------------------------------------------------------------

// Block (c + n * chunk, column group): NWV waves, each CBW column blocks of 32
// (8 witnesses x 4 limbs) over NRB row blocks, for the chunk's ring elements.
// Fragments are loaded PF K steps ahead, every load unconditional (clamped to
// the chunk), so the waits for them are exact.  Epilogue: one u32 partial per
// (chunk, c, i, w).
template <int NRB, int CBW>
__global__ __launch_bounds__(512) void nm_gemm_kernel(const int8_t *__restrict__ afrag,
                                                      const int8_t *__restrict__ bfrag, uint32_t *__restrict__ partial,
                                                      uint32_t n, uint32_t njb, uint32_t ncb, uint32_t jbc,
                                                      uint32_t nwv, uint32_t W, uint32_t k, uint32_t q,
                                                      uint64_t barrett, uint32_t pw1, uint32_t pw2, uint32_t pw3)
{
    constexpr int PF = 4;
    const uint32_t c = blockIdx.x % n, chunk = blockIdx.x / n;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const uint32_t cb0 = (blockIdx.y * nwv + wave) * CBW;
    const uint32_t jb0 = chunk * jbc, jb1 = min(njb, jb0 + jbc);
    if (jb0 >= jb1 || cb0 >= ncb) return; // uniform per wave
    const int8_t *A = afrag + (size_t)c * njb * NRB * 1024 + 16 * lane;
    const int8_t *B = bfrag + (size_t)c * njb * ncb * 1024 + 16 * lane;
    uint32_t cbs[CBW];
#pragma unroll
    for (int u = 0; u < CBW; u++) cbs[u] = min(cb0 + u, ncb - 1);

    v16i acc[NRB][CBW];
#pragma unroll
    for (int rb = 0; rb < NRB; rb++)
#pragma unroll
        for (int u = 0; u < CBW; u++)
#pragma unroll
            for (int x = 0; x < 16; x++) acc[rb][u][x] = 0;
    v4i fa[PF][NRB], fb[PF][CBW];
    auto load = [&](uint32_t jb, v4i (&a)[NRB], v4i (&b)[CBW]) {
        jb = min(jb, jb1 - 1);
#pragma unroll
        for (int rb = 0; rb < NRB; rb++) a[rb] = *reinterpret_cast<const v4i *>(A + ((size_t)jb * NRB + rb) * 1024);
#pragma unroll
        for (int u = 0; u < CBW; u++) b[u] = *reinterpret_cast<const v4i *>(B + ((size_t)jb * ncb + cbs[u]) * 1024);
    };
#pragma unroll
    for (int s = 0; s < PF; s++) load(jb0 + s, fa[s], fb[s]);
    for (uint32_t jb = jb0; jb < jb1; jb += PF) {
#pragma unroll
        for (int s = 0; s < PF; s++) {
            if (jb + s < jb1) {
#pragma unroll
                for (int rb = 0; rb < NRB; rb++)
#pragma unroll
                    for (int u = 0; u < CBW; u++)
                        acc[rb][u] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[s][rb], fb[s][u], acc[rb][u], 0, 0, 0);
            }
            load(jb + s + PF, fa[s], fb[s]);
        }
    }
    // D layout: lane l, register x -> row 32 rb + 8 (x / 4) + 4 (l / 32) + x % 4 = 4 i + La,
    // column 32 cb + l % 32 = 4 w + Lb
    const uint32_t g = lane >> 5, Lb = lane & 3;
    const uint32_t pwl = Lb == 0 ? 1u : Lb == 1 ? pw1 : Lb == 2 ? pw2 : pw3; // 256^Lb mod q
#pragma unroll
    for (int rb = 0; rb < NRB; rb++)
#pragma unroll
        for (int u = 0; u < CBW; u++)
#pragma unroll
            for (int r4 = 0; r4 < 4; r4++) {
                const int64_t x = (int64_t)acc[rb][u][4 * r4] + ((int64_t)acc[rb][u][4 * r4 + 1] << 8) +
                                  ((int64_t)acc[rb][u][4 * r4 + 2] << 16) + ((int64_t)acc[rb][u][4 * r4 + 3] << 24);
                const uint64_t mag = (uint64_t)(x < 0 ? -x : x);
                uint32_t xr = reduce_u64(mag, q, barrett);
                if (x < 0 && xr) xr = q - xr;
                uint64_t y = reduce_u64((uint64_t)xr * pwl, q, barrett);
                y += __shfl_xor(y, 1);
                y += __shfl_xor(y, 2); // four values < q
                const uint32_t z = reduce_u64(y, q, barrett);
                const uint32_t i = 8 * rb + 2 * r4 + g;
                const uint32_t w = 8 * (cb0 + u) + ((lane & 31) >> 2);
                if (Lb == 0 && i < k && w < W && cb0 + u < ncb)
                    partial[(((size_t)chunk * n + c) * k + i) * W + w] = z;
            }
}

// out[w][i][c] = corr[i][c] + sum over chunks of partial[chunk][c][i][w], mod q
__global__ void nm_reduce_kernel(const uint32_t *__restrict__ partial, const uint32_t *__restrict__ corr,
                                 uint32_t *__restrict__ out, uint32_t n, uint32_t k, uint32_t W, uint32_t nchunks,
                                 uint32_t q, uint64_t barrett)
{
    const size_t total = (size_t)W * k * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const uint32_t c = (uint32_t)(e % n), i = (uint32_t)((e / n) % k), w = (uint32_t)(e / ((size_t)n * k));
        uint64_t s = corr[(size_t)i * n + c];
        for (uint32_t ch = 0; ch < nchunks; ch++) s += partial[(((size_t)ch * n + c) * k + i) * W + w];
        out[e] = reduce_u64(s, q, barrett);
    }
}

// ---- host -------------------------------------------------------------------

static uint32_t nm_grid(size_t items, uint32_t block = 256)
{
    return (uint32_t)std::max<size_t>(1, std::min<size_t>((items + block - 1) / block, 1u << 16));
}

bool ajtai_nttmm_supported(const lzx_ajtai_key *key)
{
    return key->plan->q < NM_Q_LIMIT && key->k >= 1 && key->k <= NM_MAX_K && key->plan->log_n >= 4 &&
           key->plan->log_n <= 6 && key->m >= 1 && key->m <= 0x7FFFFFFFull;
}

bool ajtai_nttmm_enabled(const lzx_ajtai_key *key)
{
    const char *v = std::getenv("LZX_AJTAI_MAC");
    return key->nm_afrag && !(v && std::strcmp(v, "valu") == 0);
}

lzx_status ajtai_build_nttmm(lzx_ajtai_key *key)
{
    const uint32_t k = (uint32_t)key->k, m = (uint32_t)key->m, n = (uint32_t)key->n, q = key->plan->q;
    const uint64_t barrett = key->plan->mod.barrett;
    const uint32_t njb = (m + 31) / 32, nrb = (4 * k + 31) / 32;
    const size_t bytes = (size_t)n * njb * nrb * 1024;
    LZX_HIP(hipMalloc(&key->nm_afrag, bytes));
    LZX_HIP(hipMalloc(&key->nm_corr, (size_t)k * n * sizeof(uint32_t)));
    hipLaunchKernelGGL(nm_key_frag_kernel, dim3(nm_grid(bytes / 16)), dim3(256), 0, nullptr, key->Ahat_dev,
                       key->nm_afrag, k, m, n, njb, nrb, q);
    LZX_LAUNCH_CHECK("nm_key_frag_kernel");
    const uint32_t njc = (m + NM_CORR_J - 1) / NM_CORR_J;
    uint64_t *part = nullptr;
    LZX_HIP(hipMalloc(&part, (size_t)k * njc * n * sizeof(uint64_t)));
    hipLaunchKernelGGL(nm_key_colsum_kernel, dim3(njc, k), dim3(64), 0, nullptr, key->Ahat_dev, part, m, n, njc);
    lzx_status st = LZX_OK;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) st = device_fail(e, "nm_key_colsum_kernel");
    if (st == LZX_OK) {
        const uint32_t c80 = (uint32_t)(0x80808080ull % q);
        hipLaunchKernelGGL(nm_key_corr_kernel, dim3((k * n + 255) / 256), dim3(256), 0, nullptr, part, key->nm_corr,
                           k, n, njc, q, barrett, c80);
        e = hipGetLastError();
        if (e != hipSuccess) st = device_fail(e, "nm_key_corr_kernel");
    }
    if (st == LZX_OK) {
        e = hipDeviceSynchronize();
        if (e != hipSuccess) st = device_fail(e, "nttmm key build");
    }
    (void)hipFree(part);
    return st;
}

lzx_status ajtai_commit_nttmm(const lzx_ajtai_key *key, const uint32_t *s_dev, const int8_t *digits_dev, size_t W,
                              uint32_t *out_dev, hipStream_t stream)
{
    if (W == 0) return LZX_OK;
    const uint32_t k = (uint32_t)key->k, m = (uint32_t)key->m, n = (uint32_t)key->n, q = key->plan->q;
    const uint64_t barrett = key->plan->mod.barrett;
    if (W > 0x0FFFFFFFull || (digits_dev && ((uintptr_t)digits_dev % 16)) || (s_dev && ((uintptr_t)s_dev % 16)))
        return LZX_UNSUPPORTED;
    const uint32_t njb = (m + 31) / 32, nrb = (4 * k + 31) / 32;
    const uint32_t ncb = (uint32_t)((4 * W + 31) / 32);
    const uint32_t W32 = (uint32_t)W;
    // chunks: about two blocks per CU, each chunk within the int32 bound
    const uint32_t cus = (uint32_t)std::max(1, key->plan->cus);
    uint32_t nchunks = std::max<uint32_t>((njb + NM_MAX_JBC - 1) / NM_MAX_JBC, (2 * cus + n - 1) / n);
    nchunks = std::min<uint32_t>(nchunks, njb);
    const uint32_t jbc = (njb + nchunks - 1) / nchunks;
    nchunks = (njb + jbc - 1) / jbc;
    const int cbw = ncb >= 2 ? 2 : 1;
    const uint32_t nwv = std::min<uint32_t>(8, (ncb + cbw - 1) / cbw);
    const uint32_t groups = (ncb + nwv * cbw - 1) / (nwv * cbw);

    int8_t *bfrag = nullptr;
    uint32_t *partial = nullptr;
    const size_t bbytes = (size_t)n * njb * ncb * 1024;
    LZX_HIP(scratch_alloc((void **)&bfrag, bbytes, stream));
    hipError_t e = scratch_alloc((void **)&partial, (size_t)nchunks * n * k * W * sizeof(uint32_t), stream);
    if (e != hipSuccess) {
        (void)scratch_free(bfrag, stream);
        return device_fail(e, "nttmm partial alloc");
    }
    lzx_status st = LZX_OK;
    const bool big = key->plan->arith == ARITH_MONT; // q > 2^31: the FP64 register transform
    const dim3 fgrid(njb * ncb);
#define LZX_NM_FRAG(LN, AR)                                                                                        \
    do {                                                                                                           \
        if (digits_dev)                                                                                            \
            hipLaunchKernelGGL((nm_witness_frag_kernel<LN, AR, int8_t>), fgrid, dim3(256), 0, stream, digits_dev,  \
                               bfrag, key->plan->dev_tab, W32, m, njb, ncb);                                       \
        else                                                                                                       \
            hipLaunchKernelGGL((nm_witness_frag_kernel<LN, AR, uint32_t>), fgrid, dim3(256), 0, stream, s_dev,     \
                               bfrag, key->plan->dev_tab, W32, m, njb, ncb);                                       \
    } while (0)
    switch (key->plan->log_n) {
    case 4: if (big) LZX_NM_FRAG(4, ARITH_F64); else LZX_NM_FRAG(4, ARITH_SHOUP); break;
    case 5: if (big) LZX_NM_FRAG(5, ARITH_F64); else LZX_NM_FRAG(5, ARITH_SHOUP); break;
    default: if (big) LZX_NM_FRAG(6, ARITH_F64); else LZX_NM_FRAG(6, ARITH_SHOUP); break;
    }
#undef LZX_NM_FRAG
    e = hipGetLastError();
    if (e != hipSuccess) st = device_fail(e, "nm_witness_frag_kernel");
    if (st == LZX_OK) {
        const uint32_t pw1 = 256 % q, pw2 = (uint32_t)(65536ull % q), pw3 = (uint32_t)((1ull << 24) % q);
        const dim3 grid(n * nchunks, groups), block(64 * nwv);
#define LZX_NM_GEMM(NRB, CBW)                                                                                      \
    hipLaunchKernelGGL((nm_gemm_kernel<NRB, CBW>), grid, block, 0, stream, key->nm_afrag, bfrag, partial, n, njb,  \
                       ncb, jbc, nwv, W32, k, q, barrett, pw1, pw2, pw3)
        if (nrb == 1) {
            if (cbw == 1) LZX_NM_GEMM(1, 1); else LZX_NM_GEMM(1, 2);
        } else {
            if (cbw == 1) LZX_NM_GEMM(2, 1); else LZX_NM_GEMM(2, 2);
        }
#undef LZX_NM_GEMM
        e = hipGetLastError();
        if (e != hipSuccess) st = device_fail(e, "nm_gemm_kernel");
    }
    if (st == LZX_OK) {
        hipLaunchKernelGGL(nm_reduce_kernel, dim3(nm_grid((size_t)W * k * n)), dim3(256), 0, stream, partial,
                           key->nm_corr, out_dev, n, k, W32, nchunks, q, barrett);
        e = hipGetLastError();
        if (e != hipSuccess) st = device_fail(e, "nm_reduce_kernel");
    }
    if (st == LZX_OK) st = ntt_launch((const lzx_ntt_plan *const *)&key->plan, 1, out_dev, W * k, NTT_INV, 0, stream);
    (void)scratch_free(partial, stream);
    (void)scratch_free(bfrag, stream);
    return st;
}

} // namespace lzx

// Ajtai commitment on int8 MFMA in the coefficient domain (d = 64).
//
// No counterpart exists in the reference: it computes the same t = A*s
// (crates/lattice-commitment/src/ajtai.rs:118-153) exactly, as a dense integer
// contraction.  For witness w, row i, output coefficient r:
//   t_{w,i}[r] = sum_j sum_c S_{w,j}[r][c] * a_{i,j}[c]   (mod q)
// where S_{w,j} is the 64x64 negacyclic (signed Toeplitz) matrix of the
// witness element s_{w,j}: S[r][c] = s[r-c] for c <= r, -s[r-c+64] otherwise.
// a_{i,j} is the centred coefficient vector of A_{i,j} split into four signed
// 8-bit limbs, a = sum_L 256^L a^L with a^L in [-128, 127] (exact for
// q < 2^32 - 2^25).  So the commitment is one int8 GEMM
//   Out[(w, r)][(i, L)] = sum_(j, c) S_{w,j}[r][c] * a^L_{i,j}[c]
// with M = 64 W, N = 4 k, K = 64 m, followed by the exact recombination
// sum_L 256^L Out[.][(i, L)] mod q.
//
// Kernel mapping (v_mfma_i32_16x16x64_i8: one ring element = one K step):
//   - a block of 4 waves owns ITW * 4 witnesses and a chunk of j; each wave
//     owns ITW witnesses, all 64 output coefficients (4 row tiles) and 16*NT
//     limb columns (NT column tiles), i.e. 4*NT accumulators per witness;
//   - B fragments (limb planes, layout [j][col][c], 16 contiguous bytes per
//     lane) come straight from HBM/L2 as 16-byte loads, reused across the
//     wave's witnesses and row tiles;
//   - A fragments are never materialised in HBM: per (w, j) the wave writes
//     R = (s reversed, then -s reversed) -- 128 bytes -- to LDS, and lane
//     (rho, g) reads one 64-byte window of R; because S is Toeplitz, the four
//     row-tile fragments are consecutive 16-byte slices of that window, each
//     realigned with v_alignbyte;
//   - every int8 digit is exact, -128 included.  -(-128) does not fit int8,
//     so a digit vector is split as s = s1 + s2 with s1 = s where s != -128
//     and -127 where s = -128, and s2 = -1 where s = -128, else 0; both
//     halves negate safely.  A 16-element digit batch whose ballot finds a
//     -128 byte gets a second window pass over R2 (the Toeplitz of s2) into
//     the same accumulators; batches without one pay only the ballot;
//   - int32 accumulation is exact for chunks of <= 2032 ring elements: the
//     Toeplitz entries of s are at most 128 in magnitude, limbs too, so one
//     ring element adds at most 64 * 128 * 128 = 2^20 per output, and every
//     running sum (the s1 pass of an element included) stays below 2^31;
//     chunk partials are summed in 64-bit and folded mod q per limb by
//     ajtai_i8_reduce_kernel;
//   - blocks that share a j-chunk are numbered congruent mod 8 so they land on
//     one XCD (round-robin dispatch) and share that XCD's L2 copy of B.
#include <algorithm>
#include <mutex>

#include "ajtai.hpp"

namespace lzx {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int I8_WAVES = 4;
constexpr int I8_ITW = 2;          // witnesses per wave
constexpr uint32_t I8_MAX_JCHUNK = 2032; // 2032 * 2^20 < 2^31
constexpr uint64_t I8_Q_LIMIT = (1ULL << 32) - (1ULL << 25);

// Bytewise two's-complement negation (each byte of x must be != 0x80).
__device__ __forceinline__ uint32_t neg_bytes(uint32_t x)
{
    const uint32_t n = ~x;
    return ((n & 0x7f7f7f7fu) + 0x01010101u) ^ (n & 0x80808080u);
}

// 0x80 in each byte of x that equals 0x80 (-128), 0 elsewhere.
__device__ __forceinline__ uint32_t min_bytes(uint32_t x)
{
    return x & ~((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) & 0x80808080u;
}

// Combines the four limb accumulators of one output held by a lane quad
// (lane L of the quad has limb L) into sum_L 256^L acc_L on lane L = 0.
__device__ __forceinline__ int64_t combine_limbs(int32_t acc)
{
    const int64_t x = acc;
    const int64_t y = x + ((int64_t)__shfl_down(acc, 1) << 8);          // L = 0, 2
    const int64_t hi = __shfl_down(y, 2);
    return y + (hi << 16);                                                // L = 0
}

// Two waves per SIMD (accumulators live in VGPRs, <= 256 registers): the
// second wave issues its VALU and LDS work while the first one's MFMAs run.
template <int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void ajtai_i8_kernel(const int8_t *__restrict__ digits,
                                                       const int8_t *__restrict__ limbs,
                                                       uint32_t *__restrict__ partial, uint32_t W, uint32_t m,
                                                       uint32_t k, uint32_t npad, uint32_t col0, uint32_t jchunk,
                                                       uint32_t nchunks, uint32_t groups, uint32_t q,
                                                       uint64_t barrett)
{
    constexpr int JB = 16; // ring elements per digit batch (one 1 KiB load per witness)
    __shared__ __attribute__((aligned(16))) uint32_t Rs[I8_WAVES][I8_ITW][JB][36];
    __shared__ __attribute__((aligned(16))) uint32_t Rs2[I8_WAVES][I8_ITW][JB][36]; // s2 windows (digit -128)
    const uint32_t b = blockIdx.x;
    const uint32_t chunk = (b / (8 * groups)) * 8 + (b % 8);
    const uint32_t grp = (b / 8) % groups;
    if (chunk >= nchunks) return;
    // wave index as a scalar, so witness rows and their digit pointers stay in SGPRs
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const uint32_t rho = lane & 15, g = lane >> 4;
    const uint32_t wbase = (grp * I8_WAVES + wave) * I8_ITW;
    const uint32_t j0 = chunk * jchunk, j1 = min(m, j0 + jchunk);
    const uint32_t omin = 15 - rho + 16 * g; // window start in R
    const uint32_t d0 = omin >> 2, sh = omin & 3;
    const uint32_t e = lane >> 2, kq = lane & 3; // digit batch: ring element e, 16-byte chunk kq

    // padding dwords 32..35 of every R slot stay zero
    for (uint32_t idx = lane; idx < I8_ITW * JB; idx += 64) {
        *reinterpret_cast<uint4 *>(&Rs[wave][idx / JB][idx % JB][32]) = make_uint4(0, 0, 0, 0);
        *reinterpret_cast<uint4 *>(&Rs2[wave][idx / JB][idx % JB][32]) = make_uint4(0, 0, 0, 0);
    }

    v4i acc[I8_ITW][4][NT];
#pragma unroll
    for (int x = 0; x < I8_ITW; x++)
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int u = 0; u < NT; u++) acc[x][t][u] = v4i{0, 0, 0, 0};

    // uniform base pointers + 32-bit lane offsets (saddr-form global loads)
    const uint32_t boff = (col0 + rho) * 64 + 16 * g;
    const uint32_t doff = e * 64 + 16 * kq;
    auto load_digits = [&](uint32_t jb, uint4 (&dst)[I8_ITW]) {
#pragma unroll
        for (int x = 0; x < I8_ITW; x++) {
            const uint32_t w = min(wbase + x, W - 1);
            const int8_t *base = digits + ((size_t)w * m + jb) * 64;
            dst[x] = (wbase + x < W && jb + e < j1) ? *reinterpret_cast<const uint4 *>(base + doff)
                                                     : make_uint4(0, 0, 0, 0);
        }
    };
    auto load_b = [&](uint32_t j, v4i (&bf)[NT]) {
        const int8_t *base = limbs + (size_t)j * npad * 64;
#pragma unroll
        for (int u = 0; u < NT; u++) bf[u] = *reinterpret_cast<const v4i *>(base + boff + 16 * 64 * u);
    };

    // One window of R per (witness, ring element): the four row-tile fragments.
    auto compute = [&](uint32_t (&Rw)[I8_ITW][JB][36], uint32_t jj, const v4i (&bf)[NT]) {
#pragma unroll
        for (int x = 0; x < I8_ITW; x++) {
            const uint32_t *R = Rw[x][jj];
            uint32_t win[17];
#pragma unroll
            for (int q4 = 0; q4 < 17; q4++) win[q4] = R[d0 + q4];
            // all four row-tile fragments first: each has its own registers, so
            // realigning the next window never waits for an MFMA's operand reads
            v4i af[4];
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int q4 = 0; q4 < 4; q4++)
                    af[t][q4] = (int)__builtin_amdgcn_alignbyte(win[4 * (3 - t) + q4 + 1], win[4 * (3 - t) + q4], sh);
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int u = 0; u < NT; u++)
                    acc[x][t][u] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[t], bf[u], acc[x][t][u], 0, 0, 0);
        }
    };

    // B of ring element j lives in register set (j - j0) % 4 and is loaded four
    // ring elements ahead (an L2 round trip outlasts one ring element's
    // MFMAs); digits are loaded one batch ahead.
    uint4 dnext[I8_ITW];
    load_digits(j0, dnext);
    v4i bs[4][NT];
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++)
        if (j0 + s4 < j1) load_b(j0 + s4, bs[s4]);
    bool split = false; // the current batch holds a -128 digit (wave-uniform)
    for (uint32_t j = j0; j < j1; j += 4) {
        const uint32_t jj = (j - j0) % JB;
        if (jj == 0) {
            // R for the batch: R[15 - i] = bswap(s4[i]) (s reversed), R[31 - i] = -bswap(s4[i]),
            // built from s1; R2 the same from s2 when the batch has a -128 digit
            uint32_t mins = 0;
#pragma unroll
            for (int x = 0; x < I8_ITW; x++)
                mins |= min_bytes(dnext[x].x) | min_bytes(dnext[x].y) | min_bytes(dnext[x].z) |
                        min_bytes(dnext[x].w);
            split = __ballot(mins != 0) != 0;
#pragma unroll
            for (int x = 0; x < I8_ITW; x++) {
                uint4 d = dnext[x];
                uint32_t *R = Rs[wave][x][e];
                if (split) {
                    // s1 = s + 1 where s = -128 (0x80 -> 0x81, no carry); s2 = -1 there
                    const uint32_t o0 = min_bytes(d.x) >> 7, o1 = min_bytes(d.y) >> 7, o2 = min_bytes(d.z) >> 7,
                                   o3 = min_bytes(d.w) >> 7;
                    d = make_uint4(d.x + o0, d.y + o1, d.z + o2, d.w + o3);
                    const uint32_t t0 = __builtin_bswap32(o0), t1 = __builtin_bswap32(o1), t2 = __builtin_bswap32(o2),
                                   t3 = __builtin_bswap32(o3);
                    uint32_t *R2 = Rs2[wave][x][e];
                    // s2 bytes are -1 (= 0x01 * 255), -s2 bytes are +1
                    *reinterpret_cast<uint4 *>(R2 + 12 - 4 * kq) = make_uint4(t3 * 255u, t2 * 255u, t1 * 255u, t0 * 255u);
                    *reinterpret_cast<uint4 *>(R2 + 28 - 4 * kq) = make_uint4(t3, t2, t1, t0);
                }
                const uint32_t r0 = __builtin_bswap32(d.x), r1 = __builtin_bswap32(d.y), r2 = __builtin_bswap32(d.z),
                               r3 = __builtin_bswap32(d.w);
                *reinterpret_cast<uint4 *>(R + 12 - 4 * kq) = make_uint4(r3, r2, r1, r0);
                *reinterpret_cast<uint4 *>(R + 28 - 4 * kq) =
                    make_uint4(neg_bytes(r3), neg_bytes(r2), neg_bytes(r1), neg_bytes(r0));
            }
            if (j + JB < j1) load_digits(j + JB, dnext);
        }
#pragma unroll
        for (int s4 = 0; s4 < 4; s4++) {
            if (j + s4 < j1) {
                compute(Rs[wave], jj + s4, bs[s4]);
                if (split) compute(Rs2[wave], jj + s4, bs[s4]);
                if (j + s4 + 4 < j1) load_b(j + s4 + 4, bs[s4]);
            }
        }
    }
    // D layout: lane (rho, g), register i -> row 4g + i of tile t, column 16u + rho = 4 * row_i + limb.
    // Combine limbs across the lane quad, fold mod q, one u32 per (chunk, w, i, r).
#pragma unroll
    for (int x = 0; x < I8_ITW; x++) {
        const uint32_t w = wbase + x;
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int u = 0; u < NT; u++)
#pragma unroll
                for (int ireg = 0; ireg < 4; ireg++) {
                    const int64_t v = combine_limbs(acc[x][t][u][ireg]);
                    const uint32_t irow = (col0 + 16 * u + rho) >> 2;
                    if ((lane & 3) == 0 && w < W && irow < k) {
                        const uint64_t mag = (uint64_t)(v < 0 ? -v : v);
                        uint32_t res = reduce_u64(mag, q, barrett);
                        if (v < 0 && res) res = q - res;
                        const uint32_t r = 16 * t + 4 * g + ireg;
                        partial[(((size_t)chunk * W + w) * k + irow) * 64 + r] = res;
                    }
                }
    }
}

// out[w][i][r] = sum over chunks of the per-chunk residues, mod q
__global__ void ajtai_i8_reduce_kernel(const uint32_t *__restrict__ partial, uint32_t *__restrict__ out, size_t plane,
                                       uint32_t nchunks, uint32_t q, uint64_t barrett)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < plane; e += (size_t)gridDim.x * blockDim.x) {
        uint64_t s = 0;
        for (uint32_t c = 0; c < nchunks; c++) s += partial[(size_t)c * plane + e];
        out[e] = reduce_u64(s, q, barrett);
    }
}

// limbs[j][col][c], col = 4i + L (columns >= 4k are zero): balanced base-256
// digits of the centred coefficient of A[i][j][c].
__global__ void ajtai_limbs_kernel(const uint32_t *__restrict__ A, int8_t *__restrict__ limbs, uint32_t k,
                                   uint32_t m, uint32_t npad, uint32_t q)
{
    const size_t total = (size_t)m * npad * 64;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const uint32_t c = e & 63;
        const size_t jc = e >> 6;
        const uint32_t col = jc % npad;
        const size_t j = jc / npad;
        const uint32_t i = col >> 2, L = col & 3;
        int8_t out = 0;
        if (i < k) {
            const uint32_t a = A[((size_t)i * m + j) * 64 + c];
            int64_t x = a <= (q - 1) / 2 ? (int64_t)a : (int64_t)a - (int64_t)q;
            int64_t d = 0;
            for (uint32_t t = 0; t <= L; t++) {
                d = ((x + 128) & 255) - 128;
                x = (x - d) >> 8;
            }
            out = (int8_t)d;
        }
        limbs[e] = out;
    }
}

bool ajtai_int8_supported(const lzx_ajtai_key *key) { return key->n == 64 && key->plan->q < I8_Q_LIMIT; }

static uint32_t npad_of(size_t k) { return (uint32_t)(16 * ((4 * k + 15) / 16)); }

lzx_status ajtai_build_limbs(lzx_ajtai_key *key)
{
    const uint32_t npad = npad_of(key->k);
    key->limbs_bytes = key->m * npad * 64;
    LZX_HIP(hipMalloc(&key->limbs_dev, key->limbs_bytes));
    const uint32_t blocks = (uint32_t)std::min<size_t>((key->limbs_bytes + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(ajtai_limbs_kernel, dim3(blocks), dim3(256), 0, nullptr, key->A_dev, key->limbs_dev,
                       (uint32_t)key->k, (uint32_t)key->m, npad, key->plan->q);
    LZX_LAUNCH_CHECK("ajtai_limbs_kernel");
    return LZX_OK;
}

// Chunking of one int8 commit: nchunks j-chunks of jchunk ring elements
// (jchunk <= I8_MAX_JCHUNK), each writing one [W][k][64] u32 partial plane.
struct I8Plan {
    uint32_t m, k, npad, groups, nchunks, jchunk, padded;
    size_t plane;
};

// max_chunks caps the chunk count (the partial buffer holds that many planes).
static lzx_status i8_plan(const lzx_ajtai_key *key, size_t W, uint32_t max_chunks, I8Plan &p)
{
    p.m = (uint32_t)key->m;
    p.k = (uint32_t)key->k;
    p.npad = npad_of(key->k);
    const uint32_t per_block = I8_WAVES * I8_ITW;
    p.groups = (uint32_t)((W + per_block - 1) / per_block);
    const uint32_t min_chunks = (p.m + I8_MAX_JCHUNK - 1) / I8_MAX_JCHUNK;
    if (min_chunks > max_chunks) return LZX_UNSUPPORTED;
    // ~2 blocks per CU in flight; every chunk within the int32 fold-back bound
    uint32_t nchunks = std::max<uint32_t>(min_chunks, (1024 + p.groups - 1) / p.groups);
    nchunks = std::min<uint32_t>(nchunks, (p.m + 15) / 16);
    nchunks = std::min<uint32_t>(nchunks, max_chunks);
    nchunks = std::max<uint32_t>(nchunks, min_chunks);
    nchunks = std::max<uint32_t>(nchunks, 1);
    p.jchunk = (p.m + nchunks - 1) / nchunks;
    p.nchunks = (p.m + p.jchunk - 1) / p.jchunk;
    p.padded = (p.nchunks + 7) / 8 * 8;
    if ((uint64_t)p.padded * p.groups > 0x7FFFFFFFull) return LZX_UNSUPPORTED;
    p.plane = W * (size_t)p.k * 64;
    return LZX_OK;
}

// The MFMA kernels over every 64-column slice of the limb planes, then the
// chunk reduction into out.  `partial` holds nchunks planes; every word the
// reduction reads is written first by the same call.
static lzx_status i8_launch(const lzx_ajtai_key *key, const I8Plan &p, const int8_t *digits_dev, size_t W,
                            uint32_t *partial, uint32_t *out_dev, hipStream_t stream)
{
    const uint32_t q = key->plan->q;
    const uint64_t barrett = key->plan->mod.barrett;
    for (uint32_t col0 = 0; col0 < p.npad; col0 += 64) {
        const uint32_t nt = std::min<uint32_t>(4, (p.npad - col0) / 16);
        const dim3 grid(p.padded * p.groups);
#define LZX_I8_LAUNCH(NT)                                                                                        \
    hipLaunchKernelGGL(ajtai_i8_kernel<NT>, grid, dim3(256), 0, stream, digits_dev, key->limbs_dev, partial,      \
                       (uint32_t)W, p.m, p.k, p.npad, col0, p.jchunk, p.nchunks, p.groups, q, barrett)
        switch (nt) {
        case 1: LZX_I8_LAUNCH(1); break;
        case 2: LZX_I8_LAUNCH(2); break;
        case 3: LZX_I8_LAUNCH(3); break;
        default: LZX_I8_LAUNCH(4); break;
        }
#undef LZX_I8_LAUNCH
        LZX_LAUNCH_CHECK("ajtai_i8_kernel");
    }
    const uint32_t blocks = (uint32_t)std::min<size_t>((p.plane + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(ajtai_i8_reduce_kernel, dim3(blocks), dim3(256), 0, stream, partial, out_dev, p.plane,
                       p.nchunks, q, barrett);
    LZX_LAUNCH_CHECK("ajtai_i8_reduce_kernel");
    return LZX_OK;
}

lzx_status ajtai_commit_int8(const lzx_ajtai_key *key, const int8_t *digits_dev, size_t W, uint32_t *out_dev,
                             hipStream_t stream)
{
    if (W == 0) return LZX_OK;
    if (!ajtai_int8_supported(key) || !key->limbs_dev) return LZX_UNSUPPORTED;
    if ((uintptr_t)digits_dev % 16) return LZX_INVALID_ARGUMENT;
    I8Plan p;
    lzx_status st = i8_plan(key, W, UINT32_MAX, p);
    if (st != LZX_OK) return st;
    uint32_t *partial = nullptr;
    LZX_HIP(scratch_alloc((void **)&partial, (size_t)p.nchunks * p.plane * sizeof(uint32_t), stream));
    st = i8_launch(key, p, digits_dev, W, partial, out_dev, stream);
    (void)scratch_free(partial, stream);
    return st;
}

// ---- allocation-free variant --------------------------------------------
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): the chunk partials live in this fixed array
// instead of pool scratch.  Each device the library runs on holds its own
// copy at its own address until the process ends.  Host code never passes
// the array's name to a kernel (in host code it is a stand-in, not a device
// address); it uses the address hipGetSymbolAddress returns for the device.
constexpr size_t NOALLOC_WORDS = LZX_NOALLOC_SCRATCH_BYTES / sizeof(uint32_t);
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
uint32_t *g_noalloc_ptr[NOALLOC_MAX_DEVICES]; // per-device address of `scratch`
} // namespace

// Address of `scratch` on device dev, looked up once per device with that
// device current.
static lzx_status noalloc_scratch(int dev, uint32_t **out)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return LZX_UNSUPPORTED;
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        int cur = 0;
        LZX_HIP(hipGetDevice(&cur));
        if (cur != dev) LZX_HIP(hipSetDevice(dev));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) return device_fail(e, "hipGetSymbolAddress(scratch)");
        g_noalloc_ptr[dev] = (uint32_t *)ptr;
    }
    *out = g_noalloc_ptr[dev];
    return LZX_OK;
}

} // namespace lzx

using namespace lzx;

extern "C" lzx_status lzx_ajtai_commit_digits_int8_noalloc(const lzx_ajtai_key *key, const int8_t *digits_dev,
                                                          size_t s_len, size_t W, uint32_t *out_dev, void *stream)
{
    if (!key) return LZX_INVALID_ARGUMENT;
    if (s_len != key->m) return LZX_AJTAI_DIMENSION_MISMATCH;
    if (W == 0) return LZX_OK;
    if (!digits_dev || !out_dev || ((uintptr_t)digits_dev % 16)) return LZX_INVALID_ARGUMENT;
    if (!ajtai_int8_supported(key) || !key->limbs_dev) return LZX_UNSUPPORTED;
    // the partials of at least ceil(m / 2032) chunks must fit the fixed array
    const size_t plane = W * key->k * 64;
    if (plane > NOALLOC_WORDS) {
        note_error("lzx_ajtai_commit_digits_int8_noalloc: W*k*64*4 bytes exceed LZX_NOALLOC_SCRATCH_BYTES");
        return LZX_UNSUPPORTED;
    }
    I8Plan p;
    if (i8_plan(key, W, (uint32_t)std::min<size_t>(NOALLOC_WORDS / plane, UINT32_MAX), p) != LZX_OK) {
        note_error("lzx_ajtai_commit_digits_int8_noalloc: ceil(m/2032)*W*k*64*4 bytes exceed "
                   "LZX_NOALLOC_SCRATCH_BYTES");
        return LZX_UNSUPPORTED;
    }
    hipStream_t s = (hipStream_t)stream;
    hipDevice_t dev = 0;
    LZX_HIP(hipStreamGetDevice(s, &dev));
    uint32_t *partial = nullptr;
    lzx_status st = noalloc_scratch(dev, &partial);
    if (st != LZX_OK) return st;
    return i8_launch(key, p, digits_dev, W, partial, out_dev, s);
}

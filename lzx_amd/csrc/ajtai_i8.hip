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
// Kernel mapping (v_mfma_i32_32x32x32_i8: one ring element = two K steps):
//   - a block of 4 waves owns ITW * 4 witnesses and a chunk of j; each wave
//     owns ITW witnesses, all 64 output coefficients (2 row blocks of 32) and
//     32*NT limb columns (NT column blocks), i.e. 2*NT accumulators of 16
//     registers per witness, 4*NT MFMAs per witness and ring element;
//   - B fragments come straight from HBM/L2, reused across the wave's
//     witnesses and row blocks.  The limb planes are stored in fragment order
//     (see ajtai_limbs_kernel), so each fragment is one contiguous 1 KiB
//     load: with a column-major layout every 16-byte lane load would touch
//     its own cache line and the loads, not the MFMAs, would set the pace;
//   - A fragments are never materialised in HBM: per (w, j) the wave writes
//     R = (s reversed, then -s reversed) -- 128 bytes -- to LDS.  S is
//     Toeplitz (S[r][c] = R[63 - r + c]), so the fragment of row block rb and
//     K half kb depends on kb - rb only: lane (rho, g) takes the three
//     16-byte slices of R at 31 - rho + 16g + 32t, t = 0, 1, 2, each
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
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int I8_WAVES = 4;
constexpr int I8_ITW = 2;          // witnesses per wave
constexpr uint32_t I8_MAX_JCHUNK = 2032; // 2032 * 2^20 < 2^31, a multiple of 16
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
    const uint32_t rho = lane & 31, g = lane >> 5;
    const uint32_t wbase = (grp * I8_WAVES + wave) * I8_ITW;
    const uint32_t j0 = chunk * jchunk, j1 = min(m, j0 + jchunk);
    const uint32_t omin = 31 - rho + 16 * g; // first slice start in R (0..47); slices at omin + 32t
    const uint32_t d0 = omin >> 2, sh = omin & 3;
    const uint32_t e = lane >> 2, kq = lane & 3; // digit batch: ring element e, 16-byte chunk kq

    v16i acc[I8_ITW][2][NT];
#pragma unroll
    for (int x = 0; x < I8_ITW; x++)
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int u = 0; u < NT; u++)
#pragma unroll
                for (int i = 0; i < 16; i++) acc[x][rb][u][i] = 0;

    // uniform base pointers + 32-bit lane offsets (saddr-form global loads).
    // Every load is unconditional, from a clamped (valid) address: a load on
    // one path of a branch makes the compiler wait for all loads in flight
    // where the paths meet, which would serialise the B prefetch.
    const uint32_t boff = (col0 / 32) * 2048 + 16 * lane;
    // (witnesses past W read row W - 1: their results are never stored)
    auto load_digits = [&](uint32_t jb, uint4 (&dst)[I8_ITW]) {
        const uint32_t je = min(jb + e, m - 1);
#pragma unroll
        for (int x = 0; x < I8_ITW; x++) {
            const uint32_t w = min(wbase + x, W - 1);
            dst[x] = *reinterpret_cast<const uint4 *>(digits + ((size_t)w * m + je) * 64 + 16 * kq);
        }
    };
    // B fragment (u, kb): limb column col0 + 32u + rho, bytes 32kb + 16g .. +16, stored contiguously
    auto load_b = [&](uint32_t j, v4i (&bf)[NT][2]) {
        const int8_t *base = limbs + (size_t)min(j, m - 1) * npad * 64;
#pragma unroll
        for (int u = 0; u < NT; u++)
#pragma unroll
            for (int kb = 0; kb < 2; kb++)
                bf[u][kb] = *reinterpret_cast<const v4i *>(base + boff + 2048 * u + 1024 * kb);
    };

    // One pass over R per (witness, ring element): three slices, 4*NT MFMAs.
    auto compute = [&](uint32_t (&Rw)[I8_ITW][JB][36], uint32_t jj, const v4i (&bf)[NT][2]) {
#pragma unroll
        for (int x = 0; x < I8_ITW; x++) {
            const uint32_t *R = Rw[x][jj];
            uint32_t win[3][5];
#pragma unroll
            for (int t = 0; t < 3; t++)
#pragma unroll
                for (int q4 = 0; q4 < 5; q4++) win[t][q4] = R[d0 + 8 * t + q4];
            // af[t] is the fragment for kb - rb = t - 1 (rows 32rb + rho, K bytes 32kb + 16g ..)
            v4i af[3];
#pragma unroll
            for (int t = 0; t < 3; t++)
#pragma unroll
                for (int q4 = 0; q4 < 4; q4++)
                    af[t][q4] = (int)__builtin_amdgcn_alignbyte(win[t][q4 + 1], win[t][q4], sh);
            // the two K halves of one accumulator are 2*NT MFMAs apart
#pragma unroll
            for (int kb = 0; kb < 2; kb++)
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
#pragma unroll
                    for (int u = 0; u < NT; u++)
                        acc[x][rb][u] =
                            __builtin_amdgcn_mfma_i32_32x32x32_i8(af[kb - rb + 1], bf[u][kb], acc[x][rb][u], 0, 0, 0);
        }
    };

    // B of ring element j lives in register set (j - j0) % 4 and is loaded four
    // ring elements ahead (an L2 round trip outlasts one ring element's
    // MFMAs); digits are loaded one batch ahead.  The batch loop is unrolled;
    // ring elements past j1 (last chunk only) skip their MFMAs, while their
    // loads still run (from clamped rows) so that every path issues the same
    // loads.
    v4i bs[4][NT][2];
    auto batch = [&](bool split, uint32_t jb) {
#pragma unroll
        for (int s = 0; s < JB; s++) {
            if (jb + s < j1) {
                compute(Rs[wave], s, bs[s % 4]);
                if (split) compute(Rs2[wave], s, bs[s % 4]);
            }
            load_b(jb + s + 4, bs[s % 4]);
            __builtin_amdgcn_sched_barrier(0); // one ring element at a time: bounds register pressure
        }
    };
    uint4 dnext[I8_ITW];
    load_digits(j0, dnext);
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) load_b(j0 + s4, bs[s4]);
    for (uint32_t jb = j0; jb < j1; jb += JB) {
        // R for the batch: R[15 - i] = bswap(s4[i]) (s reversed), R[31 - i] = -bswap(s4[i]),
        // built from s1; R2 the same from s2 when the batch has a -128 digit
        uint32_t mins = 0;
#pragma unroll
        for (int x = 0; x < I8_ITW; x++)
            mins |= min_bytes(dnext[x].x) | min_bytes(dnext[x].y) | min_bytes(dnext[x].z) | min_bytes(dnext[x].w);
        const bool split = __ballot(mins != 0) != 0; // wave-uniform
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
        load_digits(jb + JB, dnext);
        batch(split, jb);
    }
    // D layout: lane (rho, g), register i -> row 32rb + 8(i/4) + 4g + i%4, column 32u + rho = 4 * row_i + limb.
    // Combine limbs across the lane quad, fold mod q, one u32 per (chunk, w, i, r).
#pragma unroll
    for (int x = 0; x < I8_ITW; x++) {
        const uint32_t w = wbase + x;
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int u = 0; u < NT; u++)
#pragma unroll
                for (int ireg = 0; ireg < 16; ireg++) {
                    const int64_t v = combine_limbs(acc[x][rb][u][ireg]);
                    const uint32_t irow = (col0 + 32 * u + rho) >> 2;
                    if ((lane & 3) == 0 && w < W && irow < k) {
                        const uint64_t mag = (uint64_t)(v < 0 ? -v : v);
                        uint32_t res = reduce_u64(mag, q, barrett);
                        if (v < 0 && res) res = q - res;
                        const uint32_t r = 32 * rb + 8 * (ireg / 4) + 4 * g + (ireg % 4);
                        partial[(((size_t)chunk * W + w) * k + irow) * 64 + r] = res;
                    }
                }
    }
}

// out[w][i][r] = sum over chunks of the per-chunk residues, mod q.  A block
// owns 64 consecutive outputs; its 4 waves take every 4th chunk with 8 loads
// in flight per lane, and the 4 sums (each below nchunks q, well inside u64) meet in
// LDS.  (One thread per output summing every chunk kept one load in flight
// and 4 waves per CU: 39 us for 128 chunks of a 64 x 16 x 64 plane.)
__global__ __launch_bounds__(256) void ajtai_i8_reduce_kernel(const uint32_t *__restrict__ partial,
                                                              uint32_t *__restrict__ out, size_t plane,
                                                              uint32_t nchunks, uint32_t q, uint64_t barrett)
{
    __shared__ uint64_t S[4][64];
    const uint32_t lane = threadIdx.x & 63, slice = threadIdx.x >> 6;
    for (size_t e0 = (size_t)blockIdx.x * 64; e0 < plane; e0 += (size_t)gridDim.x * 64) {
        const size_t e = e0 + lane;
        uint64_t s = 0;
        if (e < plane) {
            uint32_t c = slice;
            for (; c + 28 < nchunks; c += 32) {
                uint32_t v[8];
#pragma unroll
                for (uint32_t u = 0; u < 8; u++) v[u] = partial[(size_t)(c + 4 * u) * plane + e];
#pragma unroll
                for (uint32_t u = 0; u < 8; u++) s += v[u];
            }
            for (; c < nchunks; c += 4) s += partial[(size_t)c * plane + e];
        }
        S[slice][lane] = s;
        __syncthreads();
        if (slice == 0 && e < plane) out[e] = reduce_u64(S[0][lane] + S[1][lane] + S[2][lane] + S[3][lane], q, barrett);
        __syncthreads(); // S is free for the next 64 outputs
    }
}

// Balanced base-256 digits of the centred coefficient of A[i][j][c], for limb
// column col = 4i + L (columns >= 4k are zero), in MFMA fragment order: byte
// (j, col, c) sits at j*npad*64 + ((2*(col/32) + c/32)*64 + lane)*16 + c%16
// with lane = col%32 + 32*((c/16)%2), so the 64 lanes' 16-byte pieces of one
// fragment (32 columns x 32 bytes of c) are one contiguous 1 KiB.
__global__ void ajtai_limbs_kernel(const uint32_t *__restrict__ A, int8_t *__restrict__ limbs, uint32_t k,
                                   uint32_t m, uint32_t npad, uint32_t q)
{
    const size_t total = (size_t)m * npad * 64;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        // e = j*npad*64 + (fragment f, lane, byte b); fragment f = 2*(col/32) + c/32
        const uint32_t b = e & 15, lane = (e >> 4) & 63;
        const uint32_t f = (uint32_t)((e >> 10) % (npad / 16));
        const size_t j = (e >> 10) / (npad / 16);
        const uint32_t col = 32 * (f >> 1) + (lane & 31);
        const uint32_t c = 32 * (f & 1) + 16 * (lane >> 5) + b;
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

static uint32_t npad_of(size_t k) { return (uint32_t)(32 * ((4 * k + 31) / 32)); }

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
    // whole digit batches of 16 ring elements (I8_MAX_JCHUNK is one too)
    p.jchunk = std::min<uint32_t>(I8_MAX_JCHUNK, ((p.m + nchunks - 1) / nchunks + 15) / 16 * 16);
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
        const uint32_t nt = std::min<uint32_t>(2, (p.npad - col0) / 32);
        const dim3 grid(p.padded * p.groups);
#define LZX_I8_LAUNCH(NT)                                                                                        \
    hipLaunchKernelGGL(ajtai_i8_kernel<NT>, grid, dim3(256), 0, stream, digits_dev, key->limbs_dev, partial,      \
                       (uint32_t)W, p.m, p.k, p.npad, col0, p.jchunk, p.nchunks, p.groups, q, barrett)
        if (nt == 1)
            LZX_I8_LAUNCH(1);
        else
            LZX_I8_LAUNCH(2);
#undef LZX_I8_LAUNCH
        LZX_LAUNCH_CHECK("ajtai_i8_kernel");
    }
    const uint32_t blocks = (uint32_t)std::min<size_t>((p.plane + 63) / 64, 1 << 16);
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

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
#include <vector>

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

// The block's LDS tile of NTT-domain values, coefficient-major: value c of
// polynomial t (witness w' = t / 32, ring element jj = t % 32) at word
// (16 c + v) * 20 + jj % 16 with v = 2 w' + jj / 16, so the 16 ring elements a
// fragment lane needs are one 64-byte run (four 16-byte reads); rows are padded
// to 20 words so that the 16 runs of one coefficient start in different banks.
constexpr uint32_t NM_ROW = 20;
__device__ __forceinline__ uint32_t nm_tidx(uint32_t c, uint32_t t)
{
    return (16 * c + 2 * (t >> 5) + ((t & 31) >> 4)) * NM_ROW + (t & 15);
}

// Every coefficient's B fragment of one block of 8 witnesses x 32 ring
// elements, from the tile.  A wave writes one whole fragment per store
// instruction (1 KiB contiguous): lane s = 32 g + 4 w' + Lb takes byte Lb of
// the 16 values of witness 8 cb + w', ring elements 32 jb + 16 g .. + 15 (the
// four lanes of a (w', g) read the same run).
template <uint32_t N>
__device__ __forceinline__ void nm_write_frags(const uint32_t *T, int8_t *__restrict__ frag, uint32_t jb, uint32_t cb,
                                               uint32_t njb, uint32_t ncb, uint32_t t)
{
    const uint32_t lane = t & 63, wave = t >> 6;
    const uint32_t Lb = lane & 3, v = 2 * ((lane >> 2) & 7) + (lane >> 5);
    // v_perm_b32(hi, lo, sel): selector bytes 0-3 pick lo's bytes, 4-7 hi's
    const uint32_t sel = Lb | (Lb + 4) << 8 | Lb << 16 | (Lb + 4) << 24;
    for (uint32_t c = wave; c < N; c += 4) {
        const uint4 *run = reinterpret_cast<const uint4 *>(T + (16 * c + v) * NM_ROW);
        uint32_t d[4];
#pragma unroll
        for (uint32_t e = 0; e < 4; e++) {
            const uint4 x = run[e];
            const uint32_t lo = __builtin_amdgcn_perm(x.y, x.x, sel); // x.Lb, y.Lb, ..
            const uint32_t hi = __builtin_amdgcn_perm(x.w, x.z, sel);
            d[e] = __builtin_amdgcn_perm(hi, lo, 0x05040100u) ^ 0x80808080u;
        }
        *reinterpret_cast<uint4 *>(frag + (((size_t)c * njb + jb) * ncb + cb) * 1024 + 16 * lane) =
            make_uint4(d[0], d[1], d[2], d[3]);
    }
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
    constexpr uint32_t N = 1u << LOG_N;
    __shared__ __attribute__((aligned(16))) uint32_t T[16 * N * NM_ROW];
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
#pragma unroll
    for (uint32_t c = 0; c < N; c++) T[nm_tidx(c, t)] = ntt_out<AR>(a[brv_bits(c, LOG_N)], tab, N);
    __syncthreads();
    nm_write_frags<N>(T, frag, jb, cb, njb, ncb, t);
}

// The same fragments for digit witnesses with n = 64, the forward NTT on the
// matrix cores: s^[c] = sum_t psi^((2c+1) t) d_t (the forward transform of
// the register kernels, natural output order) is the 64 x 64 matrix Psi times
// the int8 digit vector.  Psi is split into four balanced byte limbs
// (psi_frag[La][rb][kh][lane][16]: row c = 32 rb + lane % 32, column
// t = 32 kh + 16 (lane / 32) + b), so each limb's product sums exactly in
// int32 (64 * 128 * 128 = 2^20), and V = L + 2^16 H with L = P0 + 2^8 P1,
// H = P2 + 2^8 P3 (|V| < 2^45) is exact in FP64: one rounded quotient and one
// fused multiply-add give V mod q.  A wave transforms its two witnesses'
// 32-element column blocks with 16 MFMAs each and writes the values to the
// block's LDS tile; the fragments then leave as in the register kernel.  The
// blocks are persistent (Psi, 16 KiB of fragments, stays in registers) and
// load the next tile's digits while transforming the current one.
__global__ __launch_bounds__(256) void nm_digits_frag_kernel(const int8_t *__restrict__ digits,
                                                             int8_t *__restrict__ frag,
                                                             const int8_t *__restrict__ psi_frag, uint32_t W,
                                                             uint32_t m, uint32_t njb, uint32_t ncb, double qd,
                                                             double qinvd)
{
    constexpr uint32_t N = 64;
    __shared__ __attribute__((aligned(16))) uint32_t T[16 * N * NM_ROW];
    const uint32_t t = threadIdx.x, lane = t & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const uint32_t ntiles = njb * ncb;
    // persistent: Psi stays in registers across the block's tiles (jb fastest)
    v4i P[4][2][2];
#pragma unroll
    for (int La = 0; La < 4; La++)
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int kh = 0; kh < 2; kh++)
                P[La][rb][kh] = *reinterpret_cast<const v4i *>(psi_frag + ((((La * 2 + rb) * 2 + kh) * 64) + lane) * 16);
    const uint32_t jl = lane & 31, g = lane >> 5;
    // the wave's two witnesses' digits of a tile, from clamped (valid) rows
    auto load_b = [&](uint32_t tile, v4i (&B)[2][2]) {
        tile = min(tile, ntiles - 1);
        const uint32_t jb = tile % njb, cb = tile / njb;
        const uint32_t j = min(32 * jb + jl, m - 1);
#pragma unroll
        for (uint32_t x = 0; x < 2; x++) {
            const uint32_t w = min(8 * cb + 2 * wave + x, W - 1);
            const int8_t *src = digits + ((size_t)w * m + j) * 64 + 16 * g;
#pragma unroll
            for (int kh = 0; kh < 2; kh++) B[x][kh] = *reinterpret_cast<const v4i *>(src + 32 * kh);
        }
    };
    v4i Bn[2][2];
    load_b(blockIdx.x, Bn);
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint32_t jb = tile % njb, cb = tile / njb;
        v4i B[2][2];
#pragma unroll
        for (uint32_t x = 0; x < 2; x++) {
            // polynomials past W or m transform as zero
            const bool live = 8 * cb + 2 * wave + x < W && 32 * jb + jl < m;
#pragma unroll
            for (int kh = 0; kh < 2; kh++) B[x][kh] = live ? Bn[x][kh] : v4i{0, 0, 0, 0};
        }
        load_b(tile + gridDim.x, Bn);
#pragma unroll
        for (uint32_t x = 0; x < 2; x++) {
            const uint32_t wp = 2 * wave + x;
            const uint32_t tp = 32 * wp + jl; // this lane's polynomial in the tile
#pragma unroll
            for (int rb = 0; rb < 2; rb++) {
                v16i acc[4];
#pragma unroll
                for (int La = 0; La < 4; La++) {
#pragma unroll
                    for (int r = 0; r < 16; r++) acc[La][r] = 0;
#pragma unroll
                    for (int kh = 0; kh < 2; kh++)
                        acc[La] = __builtin_amdgcn_mfma_i32_32x32x32_i8(P[La][rb][kh], B[x][kh], acc[La], 0, 0, 0);
                }
                // D: register r of lane l holds row c = 32 rb + 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int32_t L = acc[0][r] + acc[1][r] * 256, H = acc[2][r] + acc[3][r] * 256;
                    const double V = __builtin_fma((double)H, 65536.0, (double)L);
                    const double qq = __builtin_rint(V * qinvd);
                    double rr = __builtin_fma(-qq, qd, V); // exact, in (-q, q)
                    rr = rr < 0.0 ? rr + qd : rr;
                    const uint32_t c = 32 * rb + 8 * (r >> 2) + 4 * g + (r & 3);
                    T[nm_tidx(c, tp)] = (uint32_t)rr;
                }
            }
        }
        __syncthreads();
        nm_write_frags<N>(T, frag, jb, cb, njb, ncb, t);
        __syncthreads(); // the tile is free for the next one
    }
}

// ---- the product --------------------------------------------------------------

// Block (c + n * chunk, column group): NWV waves, each CBW column blocks of 32
// (8 witnesses x 4 limbs) over NRB row blocks, for the chunk's ring elements.
// Fragments are loaded PF K steps ahead through buffer resources; the loop
// has no branch, so the waits for them are exact.  Epilogue: one u32 partial
// per (chunk, c, i, w).
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
    // Buffer loads: one resource per operand over this chunk's K steps (SGPRs) and a 32-bit lane offset.
    // Steps past the chunk's end are out of the resource's range and read zeros, so every step runs the
    // same loads and MFMAs with no branch, and the waits for the prefetched steps stay exact.  (With
    // per-lane 64-bit addresses the compiler recycled prefetch registers for address arithmetic, and a
    // branch around the tail's MFMAs made every wait a full drain.)
    const uint32_t nk = jb1 - jb0;
    const __amdgpu_buffer_rsrc_t Ar = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(afrag + ((size_t)c * njb + jb0) * NRB * 1024), 0, (int)(nk * NRB * 1024), 0x00020000);
    const __amdgpu_buffer_rsrc_t Br = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(bfrag + ((size_t)c * njb + jb0) * ncb * 1024), 0, (int)(nk * ncb * 1024), 0x00020000);
    const uint32_t lo = 16 * lane;
    uint32_t bo[CBW];
#pragma unroll
    for (int u = 0; u < CBW; u++) bo[u] = lo + 1024 * min(cb0 + u, ncb - 1);

    v16i acc[NRB][CBW];
#pragma unroll
    for (int rb = 0; rb < NRB; rb++)
#pragma unroll
        for (int u = 0; u < CBW; u++)
#pragma unroll
            for (int x = 0; x < 16; x++) acc[rb][u][x] = 0;
    v4i fa[PF][NRB], fb[PF][CBW];
    auto load = [&](uint32_t st, v4i (&a)[NRB], v4i (&b)[CBW]) {
#pragma unroll
        for (int rb = 0; rb < NRB; rb++)
            a[rb] = __builtin_bit_cast(v4i,
                                       __builtin_amdgcn_raw_buffer_load_b128(Ar, lo + 1024 * (st * NRB + rb), 0, 0));
#pragma unroll
        for (int u = 0; u < CBW; u++)
            b[u] = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(Br, bo[u] + 1024 * st * ncb, 0, 0));
    };
#pragma unroll
    for (int s = 0; s < PF; s++) {
        load(s, fa[s], fb[s]);
        __builtin_amdgcn_sched_barrier(0); // stages in order, as the loop reloads them
    }
    for (uint32_t st = 0; st < nk; st += PF) {
#pragma unroll
        for (int s = 0; s < PF; s++) {
#pragma unroll
            for (int rb = 0; rb < NRB; rb++)
#pragma unroll
                for (int u = 0; u < CBW; u++)
                    acc[rb][u] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[s][rb], fb[s][u], acc[rb][u], 0, 0, 0);
            load(st + s + PF, fa[s], fb[s]);
            __builtin_amdgcn_sched_barrier(0); // step by step: each wait covers only the oldest stage
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

// Psi[c][t] = psi^((2c+1) t) mod q for n = 64, as balanced byte limbs in the
// A-fragment order of nm_digits_frag_kernel (16 KiB).
static std::vector<int8_t> nm_psi_table(const lzx_ntt_plan *p)
{
    const uint32_t n = p->n, q = p->q;
    std::vector<uint64_t> pw(2 * n);
    pw[0] = 1;
    for (uint32_t e = 1; e < 2 * n; e++) pw[e] = pw[e - 1] * p->psi % q;
    std::vector<int8_t> out(16 * 1024);
    for (uint32_t La = 0; La < 4; La++)
        for (uint32_t rb = 0; rb < 2; rb++)
            for (uint32_t kh = 0; kh < 2; kh++)
                for (uint32_t lane = 0; lane < 64; lane++)
                    for (uint32_t b = 0; b < 16; b++) {
                        const uint32_t c = 32 * rb + lane % 32, t = 32 * kh + 16 * (lane / 32) + b;
                        const uint64_t v = pw[((2 * c + 1) * t) % (2 * n)];
                        int64_t x = v <= (q - 1) / 2 ? (int64_t)v : (int64_t)v - (int64_t)q;
                        int64_t d = 0;
                        for (uint32_t L = 0; L <= La; L++) {
                            d = ((x + 128) & 255) - 128;
                            x = (x - d) >> 8;
                        }
                        out[(((La * 2 + rb) * 2 + kh) * 64 + lane) * 16 + b] = (int8_t)d;
                    }
    return out;
}

lzx_status ajtai_build_nttmm(lzx_ajtai_key *key)
{
    const uint32_t k = (uint32_t)key->k, m = (uint32_t)key->m, n = (uint32_t)key->n, q = key->plan->q;
    const uint64_t barrett = key->plan->mod.barrett;
    const uint32_t njb = (m + 31) / 32, nrb = (4 * k + 31) / 32;
    const size_t bytes = (size_t)n * njb * nrb * 1024;
    LZX_HIP(hipMalloc(&key->nm_afrag, bytes));
    LZX_HIP(hipMalloc(&key->nm_corr, (size_t)k * n * sizeof(uint32_t)));
    if (n == 64) {
        const std::vector<int8_t> psi = nm_psi_table(key->plan);
        LZX_HIP(hipMalloc(&key->nm_psi, psi.size()));
        LZX_HIP(hipMemcpy(key->nm_psi, psi.data(), psi.size(), hipMemcpyHostToDevice));
    }
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
    // each coefficient's fragments are one buffer resource (< 2^31 bytes)
    if ((uint64_t)njb * ncb * 1024 > 0x7FFFFFFFull || (uint64_t)njb * nrb * 1024 > 0x7FFFFFFFull)
        return LZX_UNSUPPORTED;
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
    if (digits_dev && n == 64 && key->nm_psi) {
        // digit witnesses: the forward transform on the matrix cores
        const dim3 pgrid(std::min<uint32_t>(njb * ncb, 2 * cus));
        hipLaunchKernelGGL(nm_digits_frag_kernel, pgrid, dim3(256), 0, stream, digits_dev, bfrag, key->nm_psi, W32,
                           m, njb, ncb, (double)q, 1.0 / (double)q);
    } else {
        switch (key->plan->log_n) {
        case 4: if (big) LZX_NM_FRAG(4, ARITH_F64); else LZX_NM_FRAG(4, ARITH_SHOUP); break;
        case 5: if (big) LZX_NM_FRAG(5, ARITH_F64); else LZX_NM_FRAG(5, ARITH_SHOUP); break;
        default: if (big) LZX_NM_FRAG(6, ARITH_F64); else LZX_NM_FRAG(6, ARITH_SHOUP); break;
        }
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

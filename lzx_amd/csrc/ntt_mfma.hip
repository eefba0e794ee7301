// Batched negacyclic NTT of degree 64 on int8 MFMA (2^20 < q < 2^31 - 2^16).
//
// The transforms are the reference's (ntt.rs:108-216, outputs in natural
// order):
//   forward  X[j] = sum_c x[c] psi^((2j+1)c)
//   inverse  x[j] = n^-1 sum_c X[c] psi^(-(2c+1)j)
// Each is one fixed 64x64 matrix per modulus, applied to 32 polynomials at a
// time as an exact int8 product on the matrix cores, with one radix-2 step
// in VALU that halves the matrix work:
//   forward: X[j0 + 32 h] = E[j0] + (-1)^h O[j0], j0 < 32, with E / O the
//     32x32 products over the even / odd coefficients (an output butterfly);
//   inverse: x[2k + h] = n^-1 sum_{c<32} psi^(-(2c+1)(2k+h)) (X[c] + (-1)^h
//     X[c+32]) (an input butterfly, then 32x32 products for the even and the
//     odd outputs).
// Exact integer products:
//   - the data side needs no conversion: any 32-bit word x is its four
//     little-endian bytes, x = sum_i b_i 2^(8i).  Flipping the top bit of
//     every byte (x ^ 0x80808080) makes them signed, s_i = b_i - 128, so
//     x = sum_i s_i 2^(8i) + 0x80808080; the constant's contribution
//     0x80808080 * sum_c M[j][c] is added back per row.  This holds for any
//     word, so the inverse's input butterflies stay unreduced (X + Y and
//     X + q - Y, both below 2^32);
//   - the matrix side folds the byte weights and the Montgomery factor into
//     M'[j][(c, i)] = M[j][c] 2^(8i) 2^32 mod q, centred (|M'| < q/2 < 2^30)
//     and split into four signed base-256 limbs M' = sum_L 256^L M'_L;
//   - so a row's value is REDC(sum_L 256^L a_L[j]) + corr[j], with the limb
//     products a_L = M'_L . s on v_mfma_i32_16x16x64_i8.  K = 128 bytes (32
//     coefficients), so |a_L| <= 128 * 128 * 128 = 2^21: int32 accumulation
//     is exact.
// The reduction: the L = 0 and L = 2 accumulators start at a bias
// (2^30 + 2^23 + delta, one delta per plan and direction, and 2^30 + 2^23)
// that makes lo = a_0 + 256 a_1 and hi = a_2 + 256 a_3 unsigned 32-bit
// values, V = lo + 2^16 hi < 2^48; one Montgomery step m = V q^-1 mod 2^32,
// t = V_hi - mulhi(m, q) gives t = V 2^-32 mod q in (-q, 2^16); the per-row
// constant K_j in [q, 2q - 2^16) (the byte offset and the biases; delta is
// chosen so that every row has such a representative) lifts t + K_j into
// [0, 2q), and one conditional subtraction makes it canonical.
//
// Kernel mapping (one persistent block of 8 waves per CU):
//   - the block holds one plan's limb matrices (32 KiB) and row constants in
//     LDS; each wave transforms whole tiles of 32 polynomials (8 KiB);
//   - every global access is a whole 1 KiB per wave instruction (4
//     polynomials), and the wave's next tile is loaded into VGPRs while the
//     current one is multiplied.  Loads and stores go through one buffer
//     resource per tile sized to its polynomials: lanes of a ragged last tile
//     read zeros and their stores are dropped by the hardware range check, so
//     every tile issues the same 8 loads and 8 stores with no branch.  (With
//     branch-skipped stores the compiler cannot count the stores that follow
//     the prefetch, and its wait for the prefetched tile then also waits for
//     the previous tile's first store.)  The transposes between that layout and the
//     MFMA operand layout go through a per-wave 8 KiB LDS tile in a swizzled
//     order, conflict-free on both sides (contiguous 16-byte operand loads of
//     every polynomial at 64-byte granularity measured 10% slower in the
//     vector memory path than whole 1 KiB lines);
//   - operand layouts of v_mfma_i32_16x16x64_i8 (lane = rho + 16 g): A (a
//     limb matrix, rows) row rho, K bytes [16g, 16g + 16) -- one 16-byte LDS
//     read per lane, stored in exactly that order; B (the data, columns =
//     polynomials) column rho, K bytes [16g, 16g + 16), the K order of the
//     bytes chosen so that each lane's operand is two 16-byte chunks of its
//     polynomial; D: lane holds rows 4g .. 4g + 3 of column rho, four
//     outputs of one polynomial, written back as one 16-byte chunk;
//   - per tile and row block: 2 groups (E/O or even/odd) x 2 K blocks x 4
//     limbs A fragments, each reused by the tile's 2 column blocks: 64 MFMAs
//     per tile.
#include <string.h>

#include <algorithm>
#include <vector>

#include "ntt.hpp"
#include "ntt_mfma.hpp"

namespace lzx {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) uint32_t *c_u32p;

constexpr int C0 = (int)0x80808080;
// cache-policy operand of a buffer store on gfx950: bit 0 sc0, bit 1 nt,
// bit 4 sc1; sc0 | sc1 is system scope
constexpr int MF_STORE_POLICY = 0x11;

// x in [0, 2q) -> x mod q
__device__ __forceinline__ uint32_t mf_red(uint32_t x, uint32_t q)
{
    const uint32_t t = x - q;
    return t < x ? t : x;
}

// canonical residue of row value i of the four limb accumulators acc[0..3]
__device__ __forceinline__ uint32_t mf_row(const v4i (&acc)[4], int i, uint32_t k, uint32_t q, uint32_t qinv)
{
    const uint32_t lo = (uint32_t)acc[0][i] + ((uint32_t)acc[1][i] << 8);
    const uint32_t hi = (uint32_t)acc[2][i] + ((uint32_t)acc[3][i] << 8);
    const uint64_t V = (uint64_t)hi * 65536u + lo;
    const uint32_t m = (uint32_t)V * qinv;
    return mf_red((uint32_t)(V >> 32) + k - __umulhi(m, q), q);
}

// The per-wave LDS tile: tile block i (polynomials 4i .. 4i + 3) keeps chunk
// k (16 bytes) of polynomial 4i + a at lane position 16 a + ((k + 4 a + i)
// mod 16) of its 1 KiB.  Both the contiguous side (lane 16 a + s holds chunk
// s) and the operand side (16 consecutive lanes hold 16 polynomials' chunk
// k) then touch 16 distinct 16-byte bank groups per 16 lanes.
__device__ __forceinline__ uint32_t mfc_pos(uint32_t i, uint32_t a, uint32_t k)
{
    return 1024 * i + 16 * (16 * a + ((k + 4 * a + i) & 15));
}

__device__ __forceinline__ void mfc_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int MODE>
__global__ __launch_bounds__(64 * MF_WAVES) __attribute__((amdgpu_waves_per_eu(2, 2))) void ntt_mfma_kernel(
    uint32_t *__restrict__ data, MfSet ms, uint32_t polys, uint32_t tiles_per_plan, uint32_t total)
{
    constexpr bool FWD = MODE == NTT_FWD;
    constexpr uint32_t TP = MF_TP, WAVES = MF_WAVES;
    __shared__ __attribute__((aligned(16))) uint8_t S[MF_DIR + WAVES * 8192];
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t rho = lane & 15, g = lane >> 4;
    const uint32_t la = lane >> 4, ls = lane & 15; // contiguous side: polynomial 4i + la, chunk ls
    const uint32_t per = (total + gridDim.x - 1) / gridDim.x;
    const uint32_t t0 = blockIdx.x * per, t1 = min(total, t0 + per);
    uint8_t *const T = S + MF_DIR + wave * 8192;
    // operand chunk (kb, h) of lane (rho, g): forward coefficients
    // 32kb + 8g + 4h .. + 3, inverse X[c] and X[c + 32] for c = 16kb + 4g ..
    auto in_chunk = [&](int kb, int h) -> uint32_t { return FWD ? 2 * g + 8 * kb + h : g + 4 * kb + 8 * h; };

    auto locate = [&](uint32_t tt, uint32_t plan, uint32_t &np) -> uint8_t * {
        const uint32_t p0 = (tt - plan * tiles_per_plan) * TP;
        np = min(TP, polys - p0);
        return (uint8_t *)(data + ((size_t)plan * polys + p0) * 64);
    };
    // np polynomials from base: offsets past np * 256 bytes read zeros / drop stores
    auto rsrc_of = [&](uint8_t *base, uint32_t np) {
        return __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)(np * 256), 0x00020000);
    };
    auto fetch = [&](__amdgpu_buffer_rsrc_t r, v4i (&X)[8]) {
#pragma unroll
        for (int i = 0; i < 8; i++)
            X[i] = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(r, 1024 * i + 16 * lane, 0, 0));
    };

    for (uint32_t s = t0; s < t1;) {
        const uint32_t plan = s / tiles_per_plan;
        const uint32_t s_end = min(t1, (plan + 1) * tiles_per_plan);
        const uint8_t *mat = ms.p[plan] + MODE * MF_DIR;
        const c_u32p aux = (c_u32p)(mat + MF_MAT);
        const uint32_t q = aux[0], qinv = aux[1], b2 = aux[2];
        v4i X[8];
        uint32_t tt = s + wave;
        __syncthreads(); // every wave is done with the previous segment's matrices
        for (uint32_t c = wave; c < MF_DIR / 1024; c += WAVES)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(mat + 1024 * c + 16 * lane),
                                             (__attribute__((address_space(3))) void *)(S + 1024 * c), 16, 0, 0);
        // the resource of tile t (an empty one past the segment: no traffic)
        auto tile_rsrc = [&](uint32_t t) {
            uint32_t n = 0;
            uint8_t *b = (uint8_t *)data;
            if (t < s_end) b = locate(t, plan, n);
            return rsrc_of(b, n);
        };
        // Pipeline per wave: the tile in T is multiplied while the next one is
        // in flight into X; its results are stored only after X has been
        // written to T.  vmcnt does not order loads against stores, so a wait
        // for X also waits for every store in flight: issuing the stores after
        // that wait leaves a whole tile's compute between each store and the
        // next wait.
        __amdgpu_buffer_rsrc_t rs = tile_rsrc(tt);
        fetch(rs, X);
        __syncthreads();
        // contiguous words -> swizzled tile
#pragma unroll
        for (int i = 0; i < 8; i++) *reinterpret_cast<v4i *>(T + mfc_pos(i, la, ls)) = X[i];
        __amdgpu_buffer_rsrc_t rn = tile_rsrc(tt + WAVES);
        fetch(rn, X);
        mfc_wave_sync();

        for (; tt < s_end; tt += WAVES) {
            // swizzled tile -> operands
            v4i B[2][2][2];
#pragma unroll
            for (int pb = 0; pb < 2; pb++)
#pragma unroll
                for (int kb = 0; kb < 2; kb++) {
                    const uint32_t i = 4 * pb + rho / 4, a = rho % 4;
                    const v4i x = *reinterpret_cast<const v4i *>(T + mfc_pos(i, a, in_chunk(kb, 0)));
                    const v4i y = *reinterpret_cast<const v4i *>(T + mfc_pos(i, a, in_chunk(kb, 1)));
                    if (FWD) {
                        B[pb][0][kb] = v4i{x[0], x[2], y[0], y[2]} ^ C0;
                        B[pb][1][kb] = v4i{x[1], x[3], y[1], y[3]} ^ C0;
                    } else {
                        const int qq = (int)q;
                        B[pb][0][kb] = (x + y) ^ C0;
                        B[pb][1][kb] = (x + qq - y) ^ C0;
                    }
                }
            const v4i i2 = {(int)b2, (int)b2, (int)b2, (int)b2};
            const v4i z = {0, 0, 0, 0};
            mfc_wave_sync(); // the operand reads precede the result writes
#pragma unroll 1
            for (int jb = 0; jb < 2; jb++) {
                uint32_t r0[2][4];
#pragma unroll
                for (int gr = 0; gr < 2; gr++) {
                    __builtin_amdgcn_sched_barrier(0);
                    const v4i i0 =
                        *reinterpret_cast<const v4i *>(S + MF_MAT + 4 * (MF_B0 + 32 * gr + 16 * jb + 4 * g));
                    v4i acc[2][4];
#pragma unroll
                    for (int kb = 0; kb < 2; kb++)
#pragma unroll
                        for (int L = 0; L < 4; L++) {
                            const v4i A = *reinterpret_cast<const v4i *>(
                                S + 16 * (((((gr * 4 + L) * 2 + jb) * 2 + kb) * 64) + lane));
#pragma unroll
                            for (int pp = 0; pp < 2; pp++)
                                acc[pp][L] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                    A, B[pp][gr][kb], kb == 0 ? (L == 0 ? i0 : L == 2 ? i2 : z) : acc[pp][L], 0, 0,
                                    0);
                        }
                    const v4i kj =
                        *reinterpret_cast<const v4i *>(S + MF_MAT + 4 * (MF_K + 32 * gr + 16 * jb + 4 * g));
#pragma unroll
                    for (int pp = 0; pp < 2; pp++) {
                        if (gr == 0) {
#pragma unroll
                            for (int r = 0; r < 4; r++) r0[pp][r] = mf_row(acc[pp], r, (uint32_t)kj[r], q, qinv);
                            continue;
                        }
                        uint32_t r1[4];
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const uint32_t e = r0[pp][r];
                            const uint32_t o = mf_row(acc[pp], r, (uint32_t)kj[r], q, qinv);
                            if (FWD) {
                                r0[pp][r] = mf_red(e + o, q); // X[j0]
                                const uint32_t d = e - o;
                                r1[r] = min(d, d + q);        // X[j0 + 32]
                            } else {
                                r1[r] = o;                    // x[2k + 1] beside x[2k]
                            }
                        }
                        // results to the tile: polynomial 16 pp + rho
                        const uint32_t i = 4 * pp + rho / 4, a = rho % 4;
                        const uint32_t *rr = r0[pp];
                        if (FWD) {
                            *reinterpret_cast<v4i *>(T + mfc_pos(i, a, 4 * jb + g)) =
                                v4i{(int)rr[0], (int)rr[1], (int)rr[2], (int)rr[3]};
                            *reinterpret_cast<v4i *>(T + mfc_pos(i, a, 8 + 4 * jb + g)) =
                                v4i{(int)r1[0], (int)r1[1], (int)r1[2], (int)r1[3]};
                        } else {
                            *reinterpret_cast<v4i *>(T + mfc_pos(i, a, 8 * jb + 2 * g)) =
                                v4i{(int)rr[0], (int)r1[0], (int)rr[1], (int)r1[1]};
                            *reinterpret_cast<v4i *>(T + mfc_pos(i, a, 8 * jb + 2 * g + 1)) =
                                v4i{(int)rr[2], (int)r1[2], (int)rr[3], (int)r1[3]};
                        }
                    }
                }
            }
            mfc_wave_sync();
            // swizzled tile -> registers; the next tile -> T; then the stores
            v4i V[8];
#pragma unroll
            for (int i = 0; i < 8; i++) V[i] = *reinterpret_cast<const v4i *>(T + mfc_pos(i, la, ls));
            mfc_wave_sync();
#pragma unroll
            for (int i = 0; i < 8; i++) *reinterpret_cast<v4i *>(T + mfc_pos(i, la, ls)) = X[i];
            // system-scope stores (sc0 sc1) write through L2: the launch ends
            // with no dirty lines for the release to write back (1-2 us of a
            // 27 us pass at 4 x 2^16 polynomials, measured)
#pragma unroll
            for (int i = 0; i < 8; i++)
                __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, V[i]), rs, 1024 * i + 16 * lane, 0,
                                                       MF_STORE_POLICY);
            rs = rn;
            rn = tile_rsrc(tt + 2 * WAVES);
            fetch(rn, X);
            mfc_wave_sync();
        }
        s = s_end;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static uint32_t mulmod(uint64_t a, uint64_t b, uint32_t q) { return (uint32_t)(a * b % q); }

bool ntt_mfma_supported(const lzx_ntt_plan *p) { return p->n == 64 && p->q > MF_Q_MIN && p->q < MF_Q_LIMIT; }

// The limb matrices and row constants of both directions, MF_DIR bytes each
// (empty if some row has no offset in range: the plan then keeps the
// butterfly kernels).
std::vector<uint8_t> ntt_mfma_tables(const lzx_ntt_plan *p)
{
    const uint32_t q = p->q;
    std::vector<uint8_t> buf(2 * (size_t)MF_DIR, 0);
    const uint32_t r32 = (uint32_t)((1ull << 32) % q);
    const uint32_t inv32 = host_pow(r32, q - 2, q); // 2^-32 mod q
    const uint32_t c0 = 0x80808080u % q;
    const uint32_t bias = (1u << 30) + (1u << 23);
    const uint32_t b2 = mulmod(bias, 1u << 16, q); // 2^16 * bias mod q
    for (uint32_t dir = 0; dir < 2; dir++) {
        // entry (group, row, K-slot coefficient c) of the 32x32 products
        auto coef = [&](uint32_t gr, uint32_t row, uint32_t c) -> uint32_t {
            if (dir == 0) return host_pow(p->psi, (uint64_t)(2 * row + 1) * c, q); // c has parity gr
            return mulmod(p->n_inv, host_pow(p->psi_inv, (uint64_t)(2 * c + 1) * (2 * row + gr), q), q);
        };
        // K slot (kb, g, byte b) -> coefficient of the group's input
        auto slot = [&](uint32_t gr, uint32_t kb, uint32_t g, uint32_t b) -> uint32_t {
            return dir == 0 ? 32 * kb + 8 * g + 2 * (b / 4) + gr : 16 * kb + 4 * g + b / 4;
        };
        uint8_t *mat = buf.data() + (size_t)dir * MF_DIR;
        for (uint32_t gr = 0; gr < 2; gr++)
            for (uint32_t L = 0; L < 4; L++)
                for (uint32_t jb = 0; jb < 2; jb++)
                    for (uint32_t kb = 0; kb < 2; kb++)
                        for (uint32_t lane = 0; lane < 64; lane++)
                            for (uint32_t b = 0; b < 16; b++) {
                                const uint32_t row = 16 * jb + (lane & 15), i = b % 4;
                                const uint32_t c = slot(gr, kb, lane >> 4, b);
                                const uint32_t w = mulmod(mulmod(coef(gr, row, c), 1u << (8 * i), q), r32, q);
                                int64_t x = w <= q / 2 ? (int64_t)w : (int64_t)w - q;
                                int64_t d = 0;
                                for (uint32_t t = 0; t <= L; t++) {
                                    d = ((x + 128) & 255) - 128;
                                    x = (x - d) >> 8;
                                }
                                mat[16 * (((((gr * 4 + L) * 2 + jb) * 2 + kb) * 64) + lane) + b] = (uint8_t)(int8_t)d;
                            }
        uint32_t *aux = reinterpret_cast<uint32_t *>(mat + MF_MAT);
        aux[0] = q;
        aux[1] = host_inv32(q);
        aux[2] = bias;
        // one L = 0 bias for every row (2^30 + 2^23 + delta): q > 2^20, so a
        // row misses the window [0, q - 2^16) with probability below 2^-4,
        // and each delta moves every row's residue by 2^-32
        uint32_t corr[2][32];
        for (uint32_t gr = 0; gr < 2; gr++)
            for (uint32_t row = 0; row < 32; row++) {
                uint64_t rowsum = 0;
                for (uint32_t c = 0; c < 32; c++) rowsum += coef(gr, row, dir == 0 ? 2 * c + gr : c);
                corr[gr][row] = mulmod(c0, rowsum % q, q);
            }
        bool found = false;
        for (uint32_t delta = 0; delta < (1u << 20) && !found; delta++) {
            const uint32_t b0 = bias + delta;
            const uint32_t off = mulmod((b0 % q + (uint64_t)b2) % q, inv32, q); // (B0 + 2^16 B2) 2^-32
            found = true;
            for (uint32_t r = 0; r < 64 && found; r++) {
                const uint32_t k = (uint32_t)(((uint64_t)corr[r / 32][r % 32] + q - off) % q);
                if (k < q - (1u << 16)) {
                    aux[MF_B0 + r] = b0;
                    aux[MF_K + r] = k + q;
                } else {
                    found = false;
                }
            }
            aux[3] = b0;
        }
        if (!found) return {};
    }
    return buf;
}

void ntt_mfma_launch(const lzx_ntt_plan *const *plans, size_t nplans, uint32_t *data, uint32_t polys, NttMode mode,
                     hipStream_t stream)
{
    MfSet ms{};
    for (size_t i = 0; i < nplans; i++) ms.p[i] = plans[i]->mfma_dev;
    const uint32_t tpp = (polys + MF_TP - 1) / MF_TP;
    const uint32_t total = tpp * (uint32_t)nplans;
    // persistent: one block per CU (the block's LDS is 97 KiB of 160), each
    // wave two tiles or more
    const uint32_t grid =
        std::max<uint32_t>(1, std::min<uint32_t>((total + 2 * MF_WAVES - 1) / (2 * MF_WAVES), (uint32_t)plans[0]->cus));
    if (mode == NTT_FWD)
        hipLaunchKernelGGL((ntt_mfma_kernel<NTT_FWD>), dim3(grid), dim3(64 * MF_WAVES), 0, stream, data, ms, polys, tpp,
                           total);
    else
        hipLaunchKernelGGL((ntt_mfma_kernel<NTT_INV>), dim3(grid), dim3(64 * MF_WAVES), 0, stream, data, ms, polys, tpp,
                           total);
}

} // namespace lzx

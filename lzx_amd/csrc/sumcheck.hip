// Goldilocks sum-check prover (crates/lattice-sumcheck/src/sumcheck.rs:62-371).
//
// Round j of the reference (half = 2^(num_vars-j-1), variable 0 = MSB):
//   g_j(t) = sum_{p<half} sum_terms c * prod_f (f[p] + t (f[p+half] - f[p]))
//   for t = 0..d (round_evals_single_bind, :172-226), then the transcript
//   yields r_j, and every factor is folded f[p] <- f[p] + (f[p+half]-f[p]) r_j
//   (bind_first_half_in_place, field_simd.rs:297-313).
// Every value is a canonical field element, so the device results equal the
// reference's for any evaluation order.
//
// Device schedule: round 0 is evaluated by sc_eval_kernel; afterwards one
// sc_fold_eval_kernel per round folds with r_j AND evaluates round j+1: a
// thread folds the pairs (p, p+half) and (p+half/2, p+half+half/2), whose
// results are exactly its pair for the next round.  So each round streams the
// tables once (read 2*half, write half per factor).  Round 0 reads the
// caller's tables (left untouched) and writes the folded tables to scratch;
// later rounds fold in place in scratch.  Each large kernel publishes
// per-block totals and a one-block sc_close_kernel sums them and runs the
// round's transcript step; once a fold's half drops to the kernel set's tail threshold, a single
// sc_tail_kernel block runs every remaining round.  All rounds are enqueued
// at once and the host synchronizes once.
//
// Arithmetic: each point's last product (by the last common factor or by
// E(p)) is added to a 128-bit lazy sum unreduced, and a fold is one 128-bit
// multiply-add reduced once.  The virtual polynomial's term structure is
// interpreted at run time (ScGenericShape), or, for the shapes the reference's
// provers build most (ScShape), fixed at compile time so each point is
// straight-line code.  Consecutive large kernels visit their pairs in
// opposite orders, so each starts on the entries its predecessor touched last.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <type_traits>
#include <vector>

#include "modarith.hpp"
#include "runtime.hpp"
#include "keccak.hpp"
#include "seeded.hpp"

namespace lzx {

constexpr int SC_MAXF = 16; // factors of a virtual polynomial (the widest kernel set)
constexpr int SC_MAXT = 16;
constexpr int SC_MAXD = 15; // degree: a round message of up to 16 values
constexpr int SC_BLOCK = 256;
constexpr uint32_t SC_MAX_BLOCKS = 2048;

// Terms with the factor powers shared by every term pulled out:
// sum_t c_t prod_f v_f^e_tf = prod_f v_f^common_f * sum_t c_t prod_f v_f^(e_tf - common_f).
// For eq * (a*b - c) this is eq * (a*b - c): two products per point instead of three.
struct ScTerms {
    uint32_t nterms, nf, d, has_common;
    uint32_t npts;                  // points evaluated per pair: d + 1, or d with the split eq factor
    int32_t eq_f;                   // the split eq factor (never read), or -1
    uint64_t coeff[SC_MAXT];
    uint8_t exps[SC_MAXT][SC_MAXF]; // multiplicity of factor f in term t beyond the common part
    uint8_t common[SC_MAXF];        // multiplicity of factor f in every term
};

// eq(w, .) over a round's pair index p without a table: E(p) =
// hi[p >> lb] * lo[p & (2^lb - 1)], the prefix/suffix factorisation of eq
// over the variables after the one being bound.
struct ScEqArgs {
    const uint64_t *hi, *lo;
    uint32_t lb, pad;
};

__device__ __forceinline__ uint64_t eq_split_at(const ScEqArgs &E, uint64_t p)
{
    return gl_mul(E.hi[p >> E.lb], E.lo[p & ((1ULL << E.lb) - 1)]);
}

struct ScTables {
    const uint64_t *src[SC_MAXF];
    uint64_t *dst[SC_MAXF];
};

// sum_terms c * prod_f v[f]^e over the point values v (any representatives
// below 2^64); the result is a representative below 2^64.
template <int NF>
__device__ __forceinline__ uint64_t eval_terms(const uint64_t (&v)[NF], const ScTerms &T)
{
    GlAcc acc{0, 0};
    for (uint32_t t = 0; t < T.nterms; t++) {
        uint64_t prod = 1;
        bool seeded = false;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const uint32_t e = T.exps[t][f];
            for (uint32_t x = 0; x < e; x++) {
                prod = seeded ? gl_mul_lazy(prod, v[f]) : v[f];
                seeded = true;
            }
        }
        const uint64_t c = T.coeff[t];
        if (c == GL_P - 1) {
            prod = gl_canon(prod); // a lone factor or a lazy product may be unreduced
            prod = prod == 0 ? 0 : GL_P - prod;
        } else if (c != 1) {
            prod = gl_mul_lazy(prod, c);
        }
        gl_acc_add(acc, prod);
    }
    uint64_t sum = gl_acc_value(acc); // exact mod p, a representative below 2^64
    if (T.has_common) {
#pragma unroll
        for (int f = 0; f < NF; f++)
            for (uint32_t x = 0; x < T.common[f]; x++) sum = gl_mul_lazy(sum, v[f]);
    }
    return sum; // the caller's lazy sums reduce it
}

// A term structure fixed at compile time, for the shapes the reference's
// provers build most: at most 4 factors and 2 terms, each factor at most
// once per term beyond the common part.  Term t multiplies the factors of bit
// mask M[t] by its coefficient, of class C[t]: 1 for +1, 2 for -1, 0 for
// T.coeff[t]; CM is the set of factors every term shares (T.common).  With
// the structure known the compiler lays out each point's products as
// straight-line code; ScGenericShape interprets T at run time instead.
struct ScGenericShape {
};

template <uint32_t NT_, uint32_t CM_, uint32_t M0_, uint32_t C0_, uint32_t M1_, uint32_t C1_, bool SPLIT_>
struct ScShape {
    static constexpr uint32_t NT = NT_, CM = CM_;
    static constexpr bool SPLIT = SPLIT_; // the split eq factor (T.eq_f >= 0, never in a mask)
    static constexpr uint32_t mask(int t) { return t == 0 ? M0_ : M1_; }
    static constexpr uint32_t cls(int t) { return t == 0 ? C0_ : C1_; }
    static constexpr uint32_t used() { return CM | M0_ | (NT > 1 ? M1_ : 0); }
    // A one-factor term with coefficient -1 whose factor appears nowhere
    // else is negated in the pair itself: -(lo + t (hi - lo)) = (p - lo) + t (lo - hi).
    static constexpr uint32_t neg_mask()
    {
        uint32_t n = 0;
        for (int t = 0; t < (int)NT; t++) {
            const uint32_t m = mask(t), others = CM | (t == 0 ? (NT > 1 ? M1_ : 0) : M0_);
            if (cls(t) == 2 && m && (m & (m - 1)) == 0 && !(m & others)) n |= m;
        }
        return n;
    }
};

template <class SH>
constexpr bool sc_generic() { return std::is_same<SH, ScGenericShape>::value; }

// Factor f's table is read: it is in the shape, or with a generic shape one
// of the T.nf factors other than the split eq factor.
template <class SH>
__device__ __forceinline__ bool factor_live(int f, const ScTerms &T)
{
    if constexpr (sc_generic<SH>()) return (uint32_t)f < T.nf && f != T.eq_f;
    else return (SH::used() >> f) & 1;
}

// One point of a compile-time shape: the term sum, times the common factors,
// times E(p) with the split eq factor, added to acc without reducing the last
// product.
template <class SH, int NF>
__device__ __forceinline__ void shape_point(const uint64_t (&v)[NF], const ScTerms &T, uint64_t eq_w, GlAcc2 &acc)
{
    constexpr bool split = SH::SPLIT;
    constexpr uint32_t NEG = SH::neg_mask();
    // s_one: the term sum is the constant 1 (a pure product of common factors)
    constexpr bool s_one = SH::NT == 1 && SH::mask(0) == 0 && SH::cls(0) == 1;
    uint64_t s = 0;
    if constexpr (!s_one) {
#pragma unroll
        for (int t = 0; t < (int)SH::NT; t++) {
            const uint32_t m = SH::mask(t), c = SH::cls(t);
            uint64_t x = 0;
            bool seeded = false;
#pragma unroll
            for (int f = 0; f < NF; f++)
                if ((m >> f) & 1) {
                    x = seeded ? gl_mul_lazy(x, v[f]) : v[f];
                    seeded = true;
                }
            if (!seeded) x = c == 1 ? 1 : (c == 2 ? GL_P - 1 : T.coeff[t]);
            else if (c == 2 && !(m & NEG)) x = gl_neg_lazy(x);
            else if (c == 0) x = gl_mul_lazy(x, T.coeff[t]);
            s = t == 0 ? x : gl_add_lazy(s, x);
        }
    }
    // the common factors; the last product goes to the accumulator unreduced
    constexpr uint32_t CM = SH::CM;
    constexpr int LASTC = CM ? 31 - __builtin_clz(CM) : -1;
    bool one = s_one;
#pragma unroll
    for (int f = 0; f < NF; f++)
        if (((CM >> f) & 1) && (f != LASTC || split)) {
            s = one ? v[f] : gl_mul_lazy(s, v[f]);
            one = false;
        }
    if constexpr (split) {
        if (one) gl_acc2_add(acc, eq_w);
        else gl_acc2_add_mul(acc, s, eq_w);
    } else if constexpr (LASTC >= 0) {
        if (one) gl_acc2_add(acc, v[LASTC]);
        else gl_acc2_add_mul(acc, s, v[LASTC]);
    } else {
        gl_acc2_add(acc, s);
    }
}

// Accumulate g(t) (t < T.npts) of one pair (lo, hi) into acc[]; with the
// split eq factor every point's term sum is weighted by eq_w = E(p).
template <int NF, int D, class SH>
__device__ __forceinline__ void eval_pair(const uint64_t (&lo)[NF], const uint64_t (&hi)[NF], const ScTerms &T,
                                          uint64_t eq_w, GlAcc2 (&acc)[D + 1])
{
    const bool split = T.eq_f >= 0;
    uint64_t diff[NF], cur[NF];
    if constexpr (sc_generic<SH>()) {
        auto put = [&](GlAcc2 &a, uint64_t v) {
            if (split) gl_acc2_add_mul(a, v, eq_w);
            else gl_acc2_add(a, v);
        };
#pragma unroll
        for (int f = 0; f < NF; f++) {
            diff[f] = gl_sub_c(hi[f], lo[f]);
            cur[f] = lo[f];
        }
        put(acc[0], eval_terms<NF>(cur, T));
        if (T.npts > 1) put(acc[1], eval_terms<NF>(hi, T));
#pragma unroll
        for (int f = 0; f < NF; f++) cur[f] = hi[f];
#pragma unroll
        for (int t = 2; t <= D; t++) {
            if ((uint32_t)t >= T.npts) break;
#pragma unroll
            for (int f = 0; f < NF; f++) cur[f] = gl_add_nc(cur[f], diff[f]); // lo + t*(hi-lo), unreduced
            put(acc[t], eval_terms<NF>(cur, T));
        }
    } else {
        // points 0 .. D: degree D = |CM| + the largest term, D + 1 points, or D
        // with the split eq factor (its point D is extrapolated)
        constexpr uint32_t USED = SH::used(), NEG = SH::neg_mask();
        constexpr int NPTS = SH::SPLIT ? D : D + 1;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            cur[f] = diff[f] = 0;
            if ((USED >> f) & 1) {
                cur[f] = (NEG >> f) & 1 ? GL_P - lo[f] : lo[f];
                diff[f] = (NEG >> f) & 1 ? gl_sub_c(lo[f], hi[f]) : gl_sub_c(hi[f], lo[f]);
            }
        }
        shape_point<SH, NF>(cur, T, eq_w, acc[0]);
        if constexpr (NPTS > 1) {
#pragma unroll
            for (int f = 0; f < NF; f++)
                if ((USED >> f) & 1) cur[f] = (NEG >> f) & 1 ? GL_P - hi[f] : hi[f];
            shape_point<SH, NF>(cur, T, eq_w, acc[1]);
        }
#pragma unroll
        for (int t = 2; t < NPTS; t++) {
#pragma unroll
            for (int f = 0; f < NF; f++)
                if ((USED >> f) & 1) cur[f] = gl_add_nc(cur[f], diff[f]);
            shape_point<SH, NF>(cur, T, eq_w, acc[t]);
        }
    }
}

// ---- cross-lane sums on DPP (no LDS round trips) --------------------------
// update_dpp with old = 0: a lane whose source is outside its row, or whose
// row is masked off, reads 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, false);
}

// One step of a lazy 64-bit sum with a carry count (true value x + c 2^64).
template <int CTRL, int ROWS>
__device__ __forceinline__ void dpp_lazy_step(uint64_t &x, uint32_t &c)
{
    const uint64_t o = (uint64_t)dpp_u32<CTRL, ROWS>((uint32_t)(x >> 32)) << 32 | dpp_u32<CTRL, ROWS>((uint32_t)x);
    const uint32_t oc = dpp_u32<CTRL, ROWS>(c);
    x += o;
    c += oc + (x < o);
}

// Sum of v over the wave's 64 lanes, canonical in lane 63: an inclusive scan
// within each 16-lane row (row_shr 1, 2, 4, 8), then row 0's total into row
// 1 and row 2's into row 3 (row_bcast:15), then rows 0-1 into rows 2-3
// (row_bcast:31).
__device__ __forceinline__ uint64_t wave_sum_last(uint64_t v)
{
    uint32_t c = 0;
    dpp_lazy_step<0x111, 0xF>(v, c);
    dpp_lazy_step<0x112, 0xF>(v, c);
    dpp_lazy_step<0x114, 0xF>(v, c);
    dpp_lazy_step<0x118, 0xF>(v, c);
    dpp_lazy_step<0x142, 0xA>(v, c);
    dpp_lazy_step<0x143, 0xC>(v, c);
    return gl_acc_finish(GlAcc{v, c});
}

// Canonical sum of the canonical values on each aligned group of G lanes
// (G = 4, 8, 16), in every lane of the group: quad_perm [1,0,3,2] and
// [2,3,0,1], then half-row and row mirrors (each lane then meets the other
// half's sum).
template <int G>
__device__ __forceinline__ uint64_t group_sum(uint64_t v)
{
    auto step = [&](auto ctrl) {
        constexpr int C = decltype(ctrl)::value;
        const uint64_t o = (uint64_t)dpp_u32<C, 0xF>((uint32_t)(v >> 32)) << 32 | dpp_u32<C, 0xF>((uint32_t)v);
        v = gl_add(v, o);
    };
    step(std::integral_constant<int, 0xB1>{});
    step(std::integral_constant<int, 0x4E>{});
    if constexpr (G >= 8) step(std::integral_constant<int, 0x141>{});
    if constexpr (G >= 16) step(std::integral_constant<int, 0x140>{});
    return v;
}

constexpr int SC_TAIL_BLOCK = 512;
constexpr int SC_MAX_WAVES = SC_TAIL_BLOCK / 64;
// The rounds whose fold has half <= sc_tail_half<NF>() run in one
// single-block kernel, its tables in LDS (NF * half * 8 <= 128 KiB).
constexpr uint64_t SC_TAIL_HALF = 2048;
template <int NF>
constexpr uint64_t sc_tail_half()
{
    return NF <= 8 ? SC_TAIL_HALF : SC_TAIL_HALF * 8 / NF;
}
// Folds with at most this many blocks close the previous round themselves.
constexpr uint32_t SC_FUSE_BLOCKS = 256;

struct ScShared {
    uint64_t wsum[SC_MAX_WAVES][SC_MAXD + 1];
    uint64_t ev[SC_MAXD + 1]; // the round polynomial's values at 0..d
    uint64_t chal;            // the latest challenge (tail kernel)
    uint64_t given[64];       // caller-supplied challenges (tail kernel)
    uint64_t eqc[64][2];      // split eq constants per round (tail kernel)
    ScEqArgs eqr[64];         // split eq tables per round (tail kernel)
    uint8_t msg[192]; // 22 + 8 (d+1) + 30 <= 180 transcript bytes per round
    uint32_t status;
};

__device__ __forceinline__ uint64_t acc_finish(const GlAcc &a) { return gl_acc_finish(a); }
__device__ __forceinline__ uint64_t acc_finish(const GlAcc2 &a) { return gl_acc2_finish(a); }

// Canonical block totals of the per-thread lazy sums acc[0..dd] into out[].
// The loop runs to the compile-time D so acc[] stays in registers.
template <int D, int WAVES, class ACC>
__device__ __forceinline__ void block_totals(ACC (&acc)[D + 1], uint32_t dd, ScShared &sh, uint64_t *out)
{
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int t = 0; t <= D; t++)
        if ((uint32_t)t <= dd) {
            const uint64_t v = wave_sum_last(acc_finish(acc[t]));
            if (lane == 63) sh.wsum[wave][t] = v;
        }
    __syncthreads();
    if (threadIdx.x <= dd) {
        uint64_t s = sh.wsum[0][threadIdx.x];
#pragma unroll
        for (int w = 1; w < WAVES; w++) s = gl_add(s, sh.wsum[w][threadIdx.x]);
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// Thread 0 publishes the block's totals for the close kernel.
template <int D>
__device__ __forceinline__ void publish_totals(GlAcc2 (&acc)[D + 1], uint32_t dd, ScShared &sh,
                                               uint64_t *__restrict__ partial)
{
    block_totals<D, SC_BLOCK / 64>(acc, dd, sh, sh.ev);
    if (threadIdx.x <= dd) partial[(size_t)blockIdx.x * (dd + 1) + threadIdx.x] = sh.ev[threadIdx.x];
}

// ---- device-resident rounds -------------------------------------------------
//
// The prover's Fiat-Shamir transcript (transcript.rs:36-238) lives in HBM
// during a prove, so all rounds are enqueued at once with no host round trip.
// Closing round j sums the blocks' totals into g_j, absorbs it
// (append_field_slice "sumcheck-round", sumcheck.rs:110-112), draws the
// challenge (challenge_fields "sumcheck-challenge": "CHAL" framing, squeeze
// from a copy of the raw state, 16-byte chunks whose first 8 bytes LE are
// accepted below p), checks g_j(0) + g_j(1) against the running claim
// (sumcheck.rs:120-123) and interpolates the next claim.  The first failure is
// kept in `status`; later rounds then leave the transcript alone.  Large rounds
// are closed by sc_close_kernel between fold kernels; the small rounds at the
// end run in sc_tail_kernel, one block that folds, sums and runs the
// transcript round after round with the sponge held in registers.
struct ScDevTx {
    uint64_t st[25];
    uint32_t pos, pad;
    uint64_t query_count, query_budget;
};

struct ScCtl {
    ScDevTx tx;
    uint64_t current;  // running claim
    uint64_t eq_scale; // split eq factor: prod_{i<j} eq(w_i, r_i) over the bound variables
    uint32_t status;   // lzx_status of the first failing round (0: none)
    uint32_t pad;
};

struct ScLagrange {
    uint64_t w[SC_MAXD + 1]; // w_i = prod_{j != i} (i - j)^{-1}, nodes 0..d
    uint64_t x[SC_MAXD];     // S(d) = sum_{i<d} x_i S(i) for S of degree < d (split eq)
};

struct ScRoundArgs {
    ScCtl *ctl;
    uint64_t *rounds;      // [num_vars][width]
    uint64_t *chal;        // [num_vars]
    const uint64_t *given; // caller-supplied challenges, or nullptr for the transcript
    uint64_t *claims;      // [nf] the fully folded tables
    const uint64_t *eq_c;  // split eq: [num_vars][2] = (1 - w_j, 2 w_j - 1), else nullptr
    const ScEqArgs *eq_rounds; // split eq: round j's E(p) tables
    uint32_t width;        // d + 1 values per round message
    uint32_t pw;           // values per block total: d + 1, or d with the split eq factor
    ScLagrange L;
    uint8_t frame[52];     // the round message's framing: 22 bytes before the values, 30 after
};

// The transcript and running claim as wave 0 holds them: lane l < 50 keeps
// half l % 2 of sponge word l / 2 in the bit-interleaved form.
struct TxRegs {
    uint32_t a;
    uint32_t pos, status;
    uint64_t qc, budget, current, eq_scale;
    uint64_t lw, lx; // Lagrange weights of node `lane` (A.L.w, A.L.x), loaded up front
};

__device__ __forceinline__ TxRegs tx_load(const ScRoundArgs &A, const ScCtl *ctl, uint32_t lane)
{
    TxRegs s;
    s.lw = lane < A.width ? A.L.w[lane] : 0;
    s.lx = lane < A.width - 1 ? A.L.x[lane] : 0;
    const uint64_t word = lane < 50 ? ctl->tx.st[lane >> 1] : 0;
    s.a = (uint32_t)(bit_unshuffle64(word) >> (32 * (lane & 1)));
    s.pos = ctl->tx.pos;
    s.status = ctl->status;
    s.qc = ctl->tx.query_count;
    s.budget = ctl->tx.query_budget;
    s.current = ctl->current;
    s.eq_scale = ctl->eq_scale;
    return s;
}

__device__ __forceinline__ void tx_store(ScCtl *ctl, const TxRegs &s, uint32_t lane)
{
    const uint32_t other = lane_get(s.a, lane ^ 1);
    if (lane < 50 && (lane & 1) == 0) ctl->tx.st[lane >> 1] = bit_shuffle64((uint64_t)s.a | (uint64_t)other << 32);
    if (lane == 0) {
        ctl->tx.pos = s.pos;
        ctl->tx.query_count = s.qc;
        ctl->tx.query_budget = s.budget;
        ctl->current = s.current;
        ctl->eq_scale = s.eq_scale;
        ctl->status = s.status;
    }
}

// The round's transcript bytes in one buffer: append_field_slice framing
// u32le|label| "sumcheck-round" u32le(8 (d+1)), the d+1 values (u64 LE, written
// per round), then challenge_fields' framing "CHAL" u32le|label|
// "sumcheck-challenge" u32le(1) (transcript.rs:70-120).  The host lays out the
// 22 + 30 framing bytes in A.frame (sc_frame_bytes).
static void sc_frame_bytes(uint32_t width, uint8_t (&fr)[52])
{
    const char *l1 = "sumcheck-round", *l2 = "sumcheck-challenge";
    uint32_t k = 0;
    auto u32 = [&](uint32_t v) {
        for (int i = 0; i < 4; i++) fr[k++] = (uint8_t)(v >> (8 * i));
    };
    u32(14);
    for (int i = 0; i < 14; i++) fr[k++] = (uint8_t)l1[i];
    u32(8 * width);
    fr[k++] = 'C', fr[k++] = 'H', fr[k++] = 'A', fr[k++] = 'L';
    u32(18);
    for (int i = 0; i < 18; i++) fr[k++] = (uint8_t)l2[i];
    u32(1);
}

// Wave 0: the constant framing bytes of sh.msg.
__device__ __forceinline__ void tx_frame_init(ScShared &sh, const ScRoundArgs &A, uint32_t lane)
{
    if (lane < 52) sh.msg[lane < 22 ? lane : 8 * A.width + lane] = A.frame[lane];
}

// XOR len bytes of msg into the sponge at its position (rate 136 = 17 words),
// permuting whenever the rate fills.
__device__ __forceinline__ void tx_absorb(TxRegs &s, const uint8_t *msg, uint32_t len, uint32_t lane)
{
    for (uint32_t off = 0; off < len;) {
        const uint32_t take = min(136u - s.pos, len - off);
        if (lane < 34) { // sponge word lane / 2 takes bytes [p0, p1) of the rate
            const int32_t w0 = 8 * (int32_t)(lane >> 1), p0 = (int32_t)s.pos, p1 = p0 + (int32_t)take;
            const int32_t k = (int32_t)off + w0 - p0; // message index of the word's byte 0
            uint64_t xw = 0;
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const bool in = w0 + b >= p0 && w0 + b < p1;
                const uint32_t byte = msg[in ? k + b : 0];
                xw |= (uint64_t)(in ? byte : 0) << (8 * b);
            }
            s.a ^= (uint32_t)(bit_unshuffle64(xw) >> (32 * (lane & 1)));
        }
        s.pos += take;
        off += take;
        if (s.pos == 136) {
            s.a = keccak_f1600_lanes(s.a, lane);
            s.pos = 0;
        }
    }
}

// Wave 0: close round rnd from sh.ev (transcript step, claim check, next
// claim).  sh.msg holds the framing bytes (tx_frame_init).
template <int D>
__device__ void tx_round(TxRegs &s, ScShared &sh, const ScRoundArgs &A, uint32_t rnd, uint64_t given, uint64_t omw,
                         uint64_t tw, bool write)
{
    const uint32_t lane = threadIdx.x & 63, width = A.width;
    if (s.status) return; // an earlier round failed
    constexpr int LANES = D + 1 <= 4 ? 4 : (D + 1 <= 8 ? 8 : 16);
    if (A.eq_c) {
        // sh.ev[t] = S(t) = sum_p E(p) Q_t(p) for t < d, where eq's value at
        // point t is scale (1 - w_j + t (2 w_j - 1)) E(p): extrapolate S(d)
        // (degree < d) and scale, g(t) = c_t S(t)
        const uint32_t d = width - 1;
        uint64_t sv = lane < d ? sh.ev[lane] : 0;
        uint64_t part = lane < d ? gl_mul(sv, s.lx) : 0;
        part = group_sum<LANES>(part);
        const uint64_t sd = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(part >> 32)) << 32) |
                            (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)part);
        if (lane == d) sv = sd;
        const uint64_t g = gl_mul(gl_mul(s.eq_scale, gl_add(omw, gl_mul(tw, lane))), sv);
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
        if (lane < width) sh.ev[lane] = g;
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
    }
    if (write && lane < width) A.rounds[(size_t)rnd * width + lane] = sh.ev[lane];
    // g(ch) = sum_i ev_i w_i prod_{j != i} (ch - j) (sumcheck.rs:286-306): node i on lane i;
    // ev_i w_i does not depend on the challenge
    const uint64_t evw = lane < width ? gl_mul(sh.ev[lane], s.lw) : 0;
    uint64_t ch = given;
    if (!A.given) {
        for (uint32_t i = lane; i < 8 * width; i += 64) sh.msg[22 + i] = (uint8_t)(sh.ev[i >> 3] >> (8 * (i & 7)));
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
        // append_field_slice, then challenge_fields' budget check (the count moves
        // and the append stays absorbed, transcript.rs:109-120) and framing
        s.qc++;
        const bool over = s.qc > s.budget;
        tx_absorb(s, sh.msg, 22 + 8 * width + (over ? 0 : 30), lane);
        if (over) {
            s.status = LZX_TRANSCRIPT_REJECTION_BUDGET;
            return;
        }
        // squeeze from a copy: candidate k is stream word 2k (chunk 16k), block (2k) / 17
        uint32_t cp = s.a;
        int32_t blk = -1;
        bool found = false;
        for (uint32_t k = 0; k < 4096 && !found; k++) { // each candidate is rejected with probability 2^-32
            const uint32_t w = 2 * k;
            while ((int32_t)(w / 17) > blk) {
                cp = keccak_f1600_lanes(cp, lane);
                blk++;
            }
            const uint32_t lo = lane_get(cp, 2 * (w % 17)), hi = lane_get(cp, 2 * (w % 17) + 1);
            const uint64_t cand = bit_shuffle64((uint64_t)lo | (uint64_t)hi << 32);
            if (cand < GL_P) {
                ch = cand;
                found = true;
            }
        }
        if (!found) { // unreachable in practice (2^-131072); reported rather than looped on
            s.status = LZX_DEVICE_ERROR;
            return;
        }
    }
    if (write && lane == 0) A.chal[rnd] = ch;
    sh.chal = ch;
    if (A.eq_c) s.eq_scale = gl_mul(s.eq_scale, gl_add(omw, gl_mul(tw, ch))); // eq(w_j, r_j)
    if (gl_add(sh.ev[0], sh.ev[1]) != s.current) {
        s.status = LZX_SUMCHECK_CLAIM_MISMATCH;
        return;
    }
    // the D+1 leaves ev_i w_i and (ch - j), j != i (1 beyond the width), as a product tree
    constexpr int LEAVES = LANES;
    uint64_t x[LEAVES];
#pragma unroll
    for (int k = 0; k < LEAVES; k++) {
        const uint32_t j = (uint32_t)k < lane ? (uint32_t)k : (uint32_t)k + 1; // skip node i
        x[k] = (uint32_t)k + 1 < width ? gl_sub(ch, j) : ((uint32_t)k + 1 == width ? evw : 1);
    }
#pragma unroll
    for (int n = LEAVES / 2; n > 0; n >>= 1)
#pragma unroll
        for (int k = 0; k < n; k++) x[k] = gl_mul(x[k], x[k + n]);
    uint64_t term = lane < width ? x[0] : 0;
    term = group_sum<LEAVES>(term); // nodes sit on lanes < LEAVES
    s.current = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(term >> 32)) << 32) |
                (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)term); // lane 0's group holds the sum
}

// The transcript state after round j lives in A.ctl[j & 1] (the initial state
// in A.ctl[1]), and round j's block totals in the partial buffer j & 1, so a
// kernel that closes one round and publishes the next never overwrites what
// its own blocks still read.
//
// Close round rnd in this block from the nblocks block totals at part: sum
// them, then wave 0 runs the transcript step.  With write set the block
// stores the round's values, challenge and state; every block that calls it
// ends with the same challenge (returned) and status (sh.status).
template <int D, int BLOCK>
__device__ __forceinline__ uint64_t close_in_block(const uint64_t *__restrict__ part, uint32_t nblocks,
                                                   const ScRoundArgs &A, uint32_t rnd, ScShared &sh, TxRegs &s,
                                                   bool write)
{
    const uint32_t pw = A.pw;
    uint64_t given = 0, omw = 0, tw = 0;
    if (threadIdx.x < 64) { // issued before the partial loads: one round trip for all
        s = tx_load(A, A.ctl + ((rnd + 1) & 1), threadIdx.x);
        if (A.given) given = A.given[rnd];
        if (A.eq_c) omw = A.eq_c[2 * rnd], tw = A.eq_c[2 * rnd + 1];
        tx_frame_init(sh, A, threadIdx.x);
    }
    GlAcc acc[D + 1];
#pragma unroll
    for (int t = 0; t <= D; t++) acc[t] = GlAcc{0, 0};
    constexpr int U = 4; // loads in flight per thread and pass
    for (uint32_t b0 = threadIdx.x; b0 < nblocks; b0 += U * BLOCK) {
        uint64_t v[U][D + 1];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t b = b0 + u * BLOCK;
#pragma unroll
            for (int t = 0; t <= D; t++) v[u][t] = b < nblocks && (uint32_t)t < pw ? part[(size_t)b * pw + t] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int t = 0; t <= D; t++) gl_acc_add(acc[t], v[u][t]);
    }
    block_totals<D, BLOCK / 64>(acc, pw - 1, sh, sh.ev);
    if (threadIdx.x < 64) {
        tx_round<D>(s, sh, A, rnd, given, omw, tw, write);
        if (write) tx_store(A.ctl + (rnd & 1), s, threadIdx.x);
        if (threadIdx.x == 0) sh.status = s.status;
    }
    __syncthreads();
    return sh.chal;
}

// One block: sum the eval / fold kernel's block totals and close round rnd.
template <int D>
__global__ __launch_bounds__(SC_BLOCK) void sc_close_kernel(const uint64_t *__restrict__ partial, uint32_t nblocks,
                                                            ScRoundArgs A, uint32_t rnd)
{
    __shared__ ScShared sh;
    TxRegs s{};
    close_in_block<D, SC_BLOCK>(partial, nblocks, A, rnd, sh, s, true);
}

// Round 0: evaluate g_0 (or S_0 with the split eq factor) over the caller's tables.
// Index of the i-th pair a grid visits: ascending, or descending (rev) so a
// kernel starts on the entries its predecessor touched last, which the
// device's last-level cache may still hold.
__device__ __forceinline__ uint64_t sc_pair_index(uint64_t i, uint64_t count, uint32_t rev)
{
    return rev ? count - 1 - i : i;
}

template <int NF, int D, class SH>
__global__ __launch_bounds__(SC_BLOCK) void sc_eval_kernel(ScTables tabs, ScTerms T, uint64_t half, ScEqArgs E,
                                                           uint64_t *__restrict__ partial, uint32_t rev)
{
    __shared__ ScShared sh;
    GlAcc2 acc[D + 1];
#pragma unroll
    for (int t = 0; t <= D; t++) acc[t] = GlAcc2{0, 0, 0};
    for (uint64_t i = (uint64_t)blockIdx.x * SC_BLOCK + threadIdx.x; i < half; i += (uint64_t)gridDim.x * SC_BLOCK) {
        const uint64_t p = sc_pair_index(i, half, rev);
        uint64_t lo[NF], hi[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const bool live = factor_live<SH>(f, T);
            lo[f] = live ? tabs.src[f][p] : 0;
            hi[f] = live ? tabs.src[f][p + half] : 0;
        }
        eval_pair<NF, D, SH>(lo, hi, T, T.eq_f >= 0 ? eq_split_at(E, p) : 1, acc);
    }
    publish_totals<D>(acc, T.npts - 1, sh, partial);
}

// Fold pairs p < quarter of tables of length 2*half with challenge r (in place
// from the second round on) and accumulate the next round's values (E: the
// next round's split eq tables).  A thread reads and writes only its own
// pairs' entries, so the visiting order is free.
template <int NF, class SH>
__device__ __forceinline__ void fold_load(const ScTables &tabs, const ScTerms &T, uint64_t half, uint64_t p,
                                          uint64_t (&v)[NF][4])
{
    const uint64_t quarter = half >> 1;
#pragma unroll
    for (int f = 0; f < NF; f++) {
        if (factor_live<SH>(f, T)) {
            const uint64_t *s = tabs.src[f];
            v[f][0] = s[p];
            v[f][1] = s[p + half];
            v[f][2] = s[p + quarter];
            v[f][3] = s[p + half + quarter];
        } else {
            v[f][0] = v[f][1] = v[f][2] = v[f][3] = 0;
        }
    }
}

template <int NF, int D, class SH>
__device__ __forceinline__ void fold_eval_pair(const ScTables &tabs, const ScTerms &T, uint64_t half, uint64_t r,
                                               const ScEqArgs &E, uint64_t p, const uint64_t (&v)[NF][4],
                                               GlAcc2 (&acc)[D + 1])
{
    const uint64_t quarter = half >> 1;
    uint64_t n0[NF], n1[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        if (factor_live<SH>(f, T)) {
            n0[f] = gl_fold(v[f][0], v[f][1], r);
            n1[f] = gl_fold(v[f][2], v[f][3], r);
            tabs.dst[f][p] = n0[f];
            tabs.dst[f][p + quarter] = n1[f];
        } else {
            n0[f] = n1[f] = 0;
        }
    }
    eval_pair<NF, D, SH>(n0, n1, T, T.eq_f >= 0 ? eq_split_at(E, p) : 1, acc);
}

template <int NF, int D, class SH>
__device__ __forceinline__ void fold_eval_range(const ScTables &tabs, const ScTerms &T, uint64_t half, uint64_t r,
                                                const ScEqArgs &E, uint64_t i0, uint64_t stride, uint32_t rev,
                                                GlAcc2 (&acc)[D + 1])
{
    const uint64_t quarter = half >> 1;
    for (uint64_t i = i0; i < quarter; i += stride) {
        const uint64_t p = sc_pair_index(i, quarter, rev);
        uint64_t v[NF][4];
        fold_load<NF, SH>(tabs, T, half, p, v);
        fold_eval_pair<NF, D, SH>(tabs, T, half, r, E, p, v, acc);
    }
}

// Fold with challenge rnd (half >= 2) and evaluate round rnd+1.
template <int NF, int D, class SH>
__global__ __launch_bounds__(SC_BLOCK) void sc_fold_eval_kernel(ScTables tabs, ScTerms T, uint64_t half,
                                                                const uint64_t *__restrict__ rp, ScEqArgs E,
                                                                uint64_t *__restrict__ partial, uint32_t rev)
{
    __shared__ ScShared sh;
    const uint64_t r = *rp; // written on the device by the close kernel
    GlAcc2 acc[D + 1];
#pragma unroll
    for (int t = 0; t <= D; t++) acc[t] = GlAcc2{0, 0, 0};
    fold_eval_range<NF, D, SH>(tabs, T, half, r, E, (uint64_t)blockIdx.x * SC_BLOCK + threadIdx.x,
                               (uint64_t)gridDim.x * SC_BLOCK, rev, acc);
    publish_totals<D>(acc, T.npts - 1, sh, partial);
}

// Close round rnd from the previous kernel's nprev block totals (every block
// redundantly, block 0 writing), then fold with its challenge and evaluate
// round rnd+1.  For small grids: each thread's first pairs are loaded before
// the close, so their latency hides behind the transcript step.
template <int NF, int D, class SH>
__global__ __launch_bounds__(SC_BLOCK) void sc_fold_close_kernel(ScTables tabs, ScTerms T, uint64_t half, ScEqArgs E,
                                                                 const uint64_t *__restrict__ part_in, uint32_t nprev,
                                                                 uint64_t *__restrict__ part_out, ScRoundArgs A,
                                                                 uint32_t rnd, uint32_t rev)
{
    __shared__ ScShared sh;
    const uint64_t quarter = half >> 1, stride = (uint64_t)gridDim.x * SC_BLOCK;
    uint64_t i = (uint64_t)blockIdx.x * SC_BLOCK + threadIdx.x;
    uint64_t v[NF][4];
    if (i < quarter) fold_load<NF, SH>(tabs, T, half, sc_pair_index(i, quarter, rev), v);
    TxRegs s{};
    const uint64_t r = close_in_block<D, SC_BLOCK>(part_in, nprev, A, rnd, sh, s, blockIdx.x == 0);
    GlAcc2 acc[D + 1];
#pragma unroll
    for (int t = 0; t <= D; t++) acc[t] = GlAcc2{0, 0, 0};
    for (bool first = true; i < quarter; i += stride, first = false) {
        const uint64_t p = sc_pair_index(i, quarter, rev);
        if (!first) fold_load<NF, SH>(tabs, T, half, p, v);
        fold_eval_pair<NF, D, SH>(tabs, T, half, r, E, p, v, acc);
    }
    publish_totals<D>(acc, T.npts - 1, sh, part_out);
}

// Rounds r0 .. num_vars-1 in one block: close round r0 from the previous
// kernel's nprev block totals, then round after round fold with the latest
// challenge, sum, close the next round; the last fold leaves one entry per
// table (the factor claims, sumcheck.rs:139-143).  tabs.src is the input of
// round r0's fold; its output, half0 entries per factor, and every later
// fold stay in the block's LDS (the dynamic allocation of sc_tail_lds bytes).
// The state after the last round goes to A.ctl[(num_vars - 1) & 1].
template <int NF, int D, class SH>
__global__ __launch_bounds__(SC_TAIL_BLOCK) void sc_tail_kernel(ScTables tabs, ScTerms T, uint64_t half0, uint32_t r0,
                                                                uint32_t num_vars, ScRoundArgs A,
                                                                const uint64_t *__restrict__ part, uint32_t nprev)
{
    __shared__ ScShared sh;
    extern __shared__ uint64_t tail_tabs[];
#pragma unroll
    for (int f = 0; f < NF; f++) tabs.dst[f] = tail_tabs + (size_t)f * half0;
    const uint32_t lane = threadIdx.x & 63;
    if (threadIdx.x < 64) {
        for (uint32_t i = lane; r0 + 1 + i < num_vars && i < 64; i += 64) { // slot i: round r0 + 1 + i
            if (A.given) sh.given[i] = A.given[r0 + 1 + i];
            if (A.eq_c) {
                sh.eqc[i][0] = A.eq_c[2 * (r0 + 1 + i)];
                sh.eqc[i][1] = A.eq_c[2 * (r0 + 1 + i) + 1];
                sh.eqr[i] = A.eq_rounds[r0 + 1 + i];
            }
        }
    }
    TxRegs s{};
    close_in_block<D, SC_TAIL_BLOCK>(part, nprev, A, r0, sh, s, true); // its barriers order sh's slots
    uint64_t half = half0;
    for (uint32_t r = r0; r < num_vars && !sh.status; r++, half >>= 1) {
        const uint64_t ch = sh.chal;
        if (half == 1) { // the last fold
            if (threadIdx.x == 0) {
#pragma unroll
                for (int f = 0; f < NF; f++)
                    if ((uint32_t)f < T.nf && f != T.eq_f) {
                        const uint64_t a = tabs.src[f][0], b = tabs.src[f][1];
                        A.claims[f] = gl_add(a, gl_mul(gl_sub(b, a), ch));
                    }
            }
            break;
        }
        GlAcc2 acc[D + 1];
#pragma unroll
        for (int t = 0; t <= D; t++) acc[t] = GlAcc2{0, 0, 0};
        const uint32_t k = r - r0; // round r+1's slot in the preloaded tables
        ScEqArgs E{};
        if (A.eq_c) E = sh.eqr[k];
        fold_eval_range<NF, D, SH>(tabs, T, half, ch, E, threadIdx.x, SC_TAIL_BLOCK, 0, acc);
#pragma unroll
        for (int f = 0; f < NF; f++) tabs.src[f] = tabs.dst[f]; // later rounds fold in place
        block_totals<D, SC_MAX_WAVES>(acc, T.npts - 1, sh, sh.ev); // its barriers also order the folded tables
        if (threadIdx.x < 64) {
            tx_round<D>(s, sh, A, r + 1, A.given ? sh.given[k] : 0, A.eq_c ? sh.eqc[k][0] : 0,
                        A.eq_c ? sh.eqc[k][1] : 0, true);
            if (lane == 0) sh.status = s.status;
        }
        __syncthreads();
    }
    if (threadIdx.x < 64 && num_vars > r0 + 1) tx_store(A.ctl + ((num_vars - 1) & 1), s, lane);
}

// Sums the block partials into the d+1 values of a round (sum_over_hypercube).
__global__ void sc_reduce_kernel(const uint64_t *__restrict__ partial, uint32_t nblocks, uint32_t width,
                                 uint64_t *__restrict__ out)
{
    __shared__ uint64_t red[SC_BLOCK];
    for (uint32_t t = 0; t < width; t++) {
        uint64_t s = 0;
        for (uint32_t b = threadIdx.x; b < nblocks; b += SC_BLOCK) s = gl_add(s, partial[(size_t)b * width + t]);
        red[threadIdx.x] = s;
        __syncthreads();
        for (uint32_t k = SC_BLOCK / 2; k > 0; k >>= 1) {
            if (threadIdx.x < k) red[threadIdx.x] = gl_add(red[threadIdx.x], red[threadIdx.x + k]);
            __syncthreads();
        }
        if (threadIdx.x == 0) out[t] = red[0];
        __syncthreads();
    }
}

// eq[b] = E_hi[b >> lo_bits] * E_lo[b & mask] (exact: field products commute).
__global__ void eq_combine_kernel(const uint64_t *__restrict__ ehi, const uint64_t *__restrict__ elo, uint32_t lo_bits,
                                  uint64_t total, uint64_t *__restrict__ out)
{
    const uint64_t mask = (1ULL << lo_bits) - 1;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < total; b += (uint64_t)gridDim.x * blockDim.x)
        out[b] = gl_mul(ehi[b >> lo_bits], elo[b & mask]);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static uint64_t h_interpolate(const uint64_t *ev, size_t n, uint64_t r) // sumcheck.rs:286-306
{
    auto inv = [](uint64_t a) {
        uint64_t acc = 1, e = GL_P - 2;
        while (e) {
            if (e & 1) acc = gl_mul(acc, a);
            a = gl_mul(a, a);
            e >>= 1;
        }
        return acc;
    };
    uint64_t acc = 0;
    for (size_t i = 0; i < n; i++) {
        uint64_t w = 1;
        for (size_t j = 0; j < n; j++) {
            if (i == j) continue;
            w = gl_mul(w, gl_mul(gl_sub(r, j), inv(gl_sub(i, j))));
        }
        acc = gl_add(acc, gl_mul(ev[i], w));
    }
    return acc;
}

struct ScProblem {
    size_t num_vars, nf, nterms, d;
    uint8_t exps[SC_MAXT][SC_MAXF]; // full multiplicities (host-side evaluation)
    ScTerms T;                      // device form, common factor powers pulled out
};

// sum_t c_t prod_f v[f]^e_tf on the host (sumcheck.rs:144-154 final claim).
static uint64_t host_eval_terms(const ScProblem &P, const uint64_t *v)
{
    uint64_t tot = 0;
    for (size_t t = 0; t < P.nterms; t++) {
        uint64_t prod = P.T.coeff[t];
        for (size_t f = 0; f < P.nf; f++)
            for (uint32_t x = 0; x < P.exps[t][f]; x++) prod = gl_mul(prod, v[f]);
        tot = gl_add(tot, prod);
    }
    return tot;
}

static lzx_status build_problem(size_t num_vars, size_t nfactors, const uint64_t *const *factors, size_t nterms,
                                const uint64_t *coeffs, const uint32_t *term_off, const uint32_t *term_ids,
                                ScProblem &P, int no_table = -1)
{
    if (nterms && (!coeffs || !term_off || !term_ids)) return LZX_INVALID_ARGUMENT;
    if (nfactors && !factors) return LZX_INVALID_ARGUMENT;
    if (nfactors > (size_t)SC_MAXF || nterms > (size_t)SC_MAXT) return LZX_UNSUPPORTED;
    if (num_vars > 40) return LZX_UNSUPPORTED;
    P.num_vars = num_vars;
    P.nf = nfactors;
    P.nterms = nterms;
    memset(P.exps, 0, sizeof(P.exps));
    memset(&P.T, 0, sizeof(P.T));
    P.T.nterms = (uint32_t)nterms;
    P.T.nf = (uint32_t)nfactors;
    size_t d = 0;
    for (size_t t = 0; t < nterms; t++) {
        if (term_off[t + 1] < term_off[t]) return LZX_INVALID_ARGUMENT;
        const size_t len = term_off[t + 1] - term_off[t];
        if (len == 0) return LZX_SUMCHECK_VIRTUAL_POLY; // VirtualPolyError::EmptyProduct
        d = std::max(d, len);
        if (coeffs[t] >= GL_P) return LZX_INVALID_ARGUMENT;
        P.T.coeff[t] = coeffs[t];
        for (uint32_t k = term_off[t]; k < term_off[t + 1]; k++) {
            if (term_ids[k] >= nfactors) return LZX_SUMCHECK_VIRTUAL_POLY; // add_term index check
            P.exps[t][term_ids[k]]++;
        }
    }
    if (d > (size_t)SC_MAXD) return LZX_UNSUPPORTED;
    for (size_t f = 0; f < nfactors; f++)
        if (!factors[f] && (int)f != no_table) return LZX_INVALID_ARGUMENT;
    P.d = d;
    P.T.d = (uint32_t)d;
    P.T.npts = (uint32_t)d + 1;
    P.T.eq_f = -1;
    for (size_t f = 0; f < nfactors; f++) {
        uint8_t c = nterms ? 255 : 0;
        for (size_t t = 0; t < nterms; t++) c = std::min(c, P.exps[t][f]);
        P.T.common[f] = c;
        P.T.has_common |= c != 0;
        for (size_t t = 0; t < nterms; t++) P.T.exps[t][f] = (uint8_t)(P.exps[t][f] - c);
    }
    return LZX_OK;
}

// eval / fold launchers; rev = 1 visits the pairs in descending order
struct ScKernels {
    void (*eval)(dim3, hipStream_t, const ScTables &, const ScTerms &, uint64_t, const ScEqArgs &, uint64_t *,
                 uint32_t);
    void (*fold)(dim3, hipStream_t, const ScTables &, const ScTerms &, uint64_t, const uint64_t *, const ScEqArgs &,
                 uint64_t *, uint32_t);
    void (*close)(hipStream_t, const uint64_t *, uint32_t, const ScRoundArgs &, uint32_t);
    void (*fold_close)(dim3, hipStream_t, const ScTables &, const ScTerms &, uint64_t, const ScEqArgs &,
                       const uint64_t *, uint32_t, uint64_t *, const ScRoundArgs &, uint32_t, uint32_t);
    void (*tail)(hipStream_t, const ScTables &, const ScTerms &, uint64_t, uint32_t, uint32_t, const ScRoundArgs &,
                 const uint64_t *, uint32_t);
    uint64_t tail_half; // rounds whose fold has half <= tail_half run in the tail kernel
};

template <int NF, int D, class SH>
static void launch_eval(dim3 g, hipStream_t s, const ScTables &tb, const ScTerms &T, uint64_t half,
                        const ScEqArgs &E, uint64_t *part, uint32_t rev)
{
    hipLaunchKernelGGL((sc_eval_kernel<NF, D, SH>), g, dim3(SC_BLOCK), 0, s, tb, T, half, E, part, rev);
}

template <int NF, int D, class SH>
static void launch_fold(dim3 g, hipStream_t s, const ScTables &tb, const ScTerms &T, uint64_t half,
                        const uint64_t *rp, const ScEqArgs &E, uint64_t *part, uint32_t rev)
{
    hipLaunchKernelGGL((sc_fold_eval_kernel<NF, D, SH>), g, dim3(SC_BLOCK), 0, s, tb, T, half, rp, E, part, rev);
}

template <int D>
static void launch_close(hipStream_t s, const uint64_t *part, uint32_t nblocks, const ScRoundArgs &A, uint32_t rnd)
{
    hipLaunchKernelGGL((sc_close_kernel<D>), dim3(1), dim3(SC_BLOCK), 0, s, part, nblocks, A, rnd);
}

template <int NF, int D, class SH>
static void launch_fold_close(dim3 g, hipStream_t s, const ScTables &tb, const ScTerms &T, uint64_t half,
                              const ScEqArgs &E, const uint64_t *part_in, uint32_t nprev, uint64_t *part_out,
                              const ScRoundArgs &A, uint32_t rnd, uint32_t rev)
{
    hipLaunchKernelGGL((sc_fold_close_kernel<NF, D, SH>), g, dim3(SC_BLOCK), 0, s, tb, T, half, E, part_in, nprev,
                       part_out, A, rnd, rev);
}

// LDS for the tail's tables: half0 entries of each of the NF factor slots
// (half0 <= sc_tail_half<NF>(), so at most 128 KiB beside ScShared).
template <int NF>
static size_t sc_tail_lds(uint64_t half0)
{
    return (size_t)NF * half0 * sizeof(uint64_t);
}

template <int NF, int D, class SH>
static void launch_tail(hipStream_t s, const ScTables &tb, const ScTerms &T, uint64_t half0, uint32_t r0,
                        uint32_t num_vars, const ScRoundArgs &A, const uint64_t *part, uint32_t nprev)
{
    static const hipError_t attr = hipFuncSetAttribute((const void *)sc_tail_kernel<NF, D, SH>,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       (int)sc_tail_lds<NF>(sc_tail_half<NF>()));
    (void)attr; // a failure shows as the launch's error
    hipLaunchKernelGGL((sc_tail_kernel<NF, D, SH>), dim3(1), dim3(SC_TAIL_BLOCK), sc_tail_lds<NF>(half0), s, tb, T,
                       half0, r0, num_vars, A, part, nprev);
}

// fold_close is null for the widest set: its registers are full without the
// close's, so its rounds close in a kernel of their own.
template <int NF, int D, class SH = ScGenericShape>
static ScKernels kernels_for()
{
    if constexpr (NF <= 8)
        return ScKernels{launch_eval<NF, D, SH>, launch_fold<NF, D, SH>, launch_close<D>,
                         launch_fold_close<NF, D, SH>, launch_tail<NF, D, SH>, sc_tail_half<NF>()};
    else
        return ScKernels{launch_eval<NF, D, SH>, launch_fold<NF, D, SH>, launch_close<D>, nullptr,
                         launch_tail<NF, D, SH>, sc_tail_half<NF>()};
}

// The compile-time shape P.T has, when it has one of the registered ones.
struct ScShapeKey {
    uint32_t nt, cm, m0, c0, m1, c1, split;
    // a registered coefficient class 0 reads T.coeff, so it takes any coefficient
    bool runs(const ScShapeKey &reg) const
    {
        auto cls_ok = [](uint32_t have, uint32_t want) { return want == 0 || have == want; };
        return nt == reg.nt && cm == reg.cm && m0 == reg.m0 && cls_ok(c0, reg.c0) &&
               (nt < 2 || (m1 == reg.m1 && cls_ok(c1, reg.c1))) && split == reg.split;
    }
};

static bool shape_key(const ScProblem &P, ScShapeKey &k)
{
    if (P.nf > 4 || P.nterms < 1 || P.nterms > 2) return false;
    uint32_t m[2] = {0, 0}, c[2] = {0, 0}, cm = 0;
    for (size_t f = 0; f < P.nf; f++) {
        if (P.T.common[f] > 1) return false;
        if (P.T.common[f]) cm |= 1u << f;
    }
    for (size_t t = 0; t < P.nterms; t++) {
        for (size_t f = 0; f < P.nf; f++) {
            if (P.T.exps[t][f] > 1) return false;
            if (P.T.exps[t][f]) m[t] |= 1u << f;
        }
        c[t] = P.T.coeff[t] == 1 ? 1 : (P.T.coeff[t] == GL_P - 1 ? 2 : 0);
    }
    k = ScShapeKey{(uint32_t)P.nterms, cm, m[0], c[0], m[1], c[1], P.T.eq_f >= 0};
    return true;
}

template <uint32_t NT, uint32_t CM, uint32_t M0, uint32_t C0, uint32_t M1, uint32_t C1, bool SPLIT>
static constexpr ScShapeKey key_of(ScShape<NT, CM, M0, C0, M1, C1, SPLIT>)
{
    return ScShapeKey{NT, CM, M0, C0, M1, C1, SPLIT};
}

// Registered shapes (factor indices in the caller's order):
//   eq * (a*b - c) with eq a table (eq, a, b, c)              sumcheck over R1CS rows
//   (a*b - c) with the split eq factor at index 0             prove_fast_with_eq of the same
//   f * e                                                    pcs.rs:95-98, salsa_response.rs:106-109
//   rho_0 f_0 e_0 + rho_1 f_1 e_1                             pcs.rs:225-231, salsa_binding.rs:182-188
using ShEqABC = ScShape<2, 0x1, 0x6, 1, 0x8, 2, false>;
using ShSplitABC = ScShape<2, 0x0, 0x6, 1, 0x8, 2, true>;
using ShProd2 = ScShape<1, 0x3, 0x0, 1, 0, 0, false>;
using ShBatch2 = ScShape<2, 0x0, 0x3, 0, 0xC, 0, false>;

static ScKernels pick_kernels(const ScProblem &P)
{
    const size_t nf = P.nf, d = P.d;
    ScShapeKey k;
    if (!getenv("LZX_SC_GENERIC") && shape_key(P, k)) {
        if (k.runs(key_of(ShEqABC{})) && nf == 4) return kernels_for<4, 3, ShEqABC>();
        if (k.runs(key_of(ShSplitABC{})) && nf == 4) return kernels_for<4, 3, ShSplitABC>();
        if (k.runs(key_of(ShProd2{})) && nf == 2) return kernels_for<2, 2, ShProd2>();
        if (k.runs(key_of(ShBatch2{})) && nf == 4) return kernels_for<4, 2, ShBatch2>();
    }
    if (nf <= 2 && d <= 2) return kernels_for<2, 2>();
    if (nf <= 4 && d <= 3) return kernels_for<4, 3>();
    if (nf <= 8 && d <= 8) return kernels_for<8, 8>();
    return kernels_for<SC_MAXF, SC_MAXD>();
}

static uint32_t sc_grid(uint64_t work)
{
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((work + SC_BLOCK - 1) / SC_BLOCK, SC_MAX_BLOCKS));
}

// The prover loop shared by the transcript and explicit-challenge forms: every
// round is enqueued up front (eval, then one fold per round, each closing its
// round on the device), and the host synchronizes once at the end.
// eq(w_a .. w_{b-1}, x) for x in {0,1}^(b-a), the first variable the most
// significant bit (field_simd.rs:391-407 doubling order).
static void host_eq_append(const uint64_t *w, size_t a, size_t b, std::vector<uint64_t> &out)
{
    std::vector<uint64_t> t(1, 1);
    for (size_t i = a; i < b; i++) {
        std::vector<uint64_t> n(t.size() * 2);
        const uint64_t om = gl_sub(1, w[i]);
        for (size_t k = 0; k < t.size(); k++) {
            n[2 * k] = gl_mul(t[k], om);
            n[2 * k + 1] = gl_mul(t[k], w[i]);
        }
        t.swap(n);
    }
    out.insert(out.end(), t.begin(), t.end());
}

// ---- whole-prove graphs -------------------------------------------------------
// A prove's launches depend only on its sizes, term structure, factor tables
// and workspace.  From the second call with the same ones, the whole
// sequence (upload, round kernels, readback) is captured once into a hipGraph
// on a workspace and pinned stage owned by a cache entry; later calls fill the
// stage and launch the graph.  At most SC_GRAPH_MAX keys are remembered and
// SC_GRAPH_BYTES of workspace held, least recently used entries going first.
// A call holds its entry (pinned against eviction, its lock taken) from the
// stage fill to the readback, so callers that share an entry take turns.
// LZX_SC_GRAPH=0 turns the cache off.
struct ScGraphKey {
    std::vector<uint8_t> b;
    template <class T> ScGraphKey &add(const T &v)
    {
        const uint8_t *p = (const uint8_t *)&v;
        b.insert(b.end(), p, p + sizeof(T));
        return *this;
    }
};

struct ScGraph {
    std::vector<uint8_t> key;
    uint32_t calls = 0, pins = 0;
    uint64_t last = 0;
    uint8_t *dev = nullptr;
    void *stage = nullptr;
    size_t bytes = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::mutex mu;
    void release_memory()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        if (dev) (void)hipFree(dev);
        if (stage) (void)hipHostFree(stage);
        exec = nullptr, graph = nullptr, cap = nullptr, dev = nullptr, stage = nullptr, bytes = 0;
    }
    ~ScGraph() { release_memory(); }
};

constexpr size_t SC_GRAPH_MAX = 8;
constexpr size_t SC_GRAPH_BYTES = 8ull << 30;

static bool sc_graph_enabled()
{
    const char *e = getenv("LZX_SC_GRAPH");
    return !(e && e[0] == '0');
}

static std::mutex g_sc_graph_mu;
static std::vector<std::unique_ptr<ScGraph>> g_sc_graphs;
static uint64_t g_sc_graph_clock = 0;

// Drops the least recently used entry that no call holds (its memory too);
// false when every entry is held.  g_sc_graph_mu is held.
static bool sc_graph_evict_one()
{
    size_t best = SIZE_MAX;
    for (size_t i = 0; i < g_sc_graphs.size(); i++)
        if (g_sc_graphs[i]->pins == 0 && (best == SIZE_MAX || g_sc_graphs[i]->last < g_sc_graphs[best]->last))
            best = i;
    if (best == SIZE_MAX) return false;
    g_sc_graphs.erase(g_sc_graphs.begin() + best);
    return true;
}

// Releases a held entry at the end of a call.
struct ScGraphHold {
    ScGraph *g = nullptr;
    ~ScGraphHold()
    {
        if (!g) return;
        g->mu.unlock();
        std::lock_guard<std::mutex> lk(g_sc_graph_mu);
        g->pins--;
    }
};

// The entry for key, held, with its workspace of dev_bytes and stage of
// area bytes allocated; none (hold.g == nullptr, run directly) for a key's
// first call or when the cache cannot make room.
static lzx_status sc_graph_get(const ScGraphKey &key, size_t dev_bytes, size_t area, ScGraphHold &hold)
{
    std::unique_lock<std::mutex> lk(g_sc_graph_mu);
    ScGraph *g = nullptr;
    for (auto &e : g_sc_graphs)
        if (e->key == key.b) g = e.get();
    if (!g) {
        if (g_sc_graphs.size() >= SC_GRAPH_MAX && !sc_graph_evict_one()) return LZX_OK;
        auto e = std::make_unique<ScGraph>();
        e->key = key.b;
        e->calls = 1;
        e->last = ++g_sc_graph_clock;
        g_sc_graphs.push_back(std::move(e));
        return LZX_OK;
    }
    g->last = ++g_sc_graph_clock;
    g->calls++;
    if (!g->dev) {
        size_t held = 0;
        for (auto &e : g_sc_graphs) held += e->bytes;
        while (held + dev_bytes > SC_GRAPH_BYTES) {
            size_t victim = SIZE_MAX; // drop the memory of the least recently used idle entry
            for (size_t i = 0; i < g_sc_graphs.size(); i++) {
                ScGraph *v = g_sc_graphs[i].get();
                if (v != g && v->pins == 0 && v->dev && (victim == SIZE_MAX || v->last < g_sc_graphs[victim]->last))
                    victim = i;
            }
            if (victim == SIZE_MAX) return LZX_OK;
            held -= g_sc_graphs[victim]->bytes;
            g_sc_graphs[victim]->release_memory();
        }
        if (hipMalloc((void **)&g->dev, dev_bytes) != hipSuccess ||
            hipHostMalloc(&g->stage, area, hipHostMallocDefault) != hipSuccess ||
            hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError(); // out of memory here: the call runs directly
            g->release_memory();
            return LZX_OK;
        }
        g->bytes = dev_bytes;
    }
    g->pins++;
    lk.unlock();
    g->mu.lock();
    hold.g = g;
    return LZX_OK;
}

// eq_w: the split eq factor's point (P.T.eq_f >= 0), else nullptr.
static lzx_status prove_impl(const ScProblem &P, const uint64_t *const *factors, uint64_t claim, lzx_transcript *tr,
                             const uint64_t *given, const uint64_t *eq_w, uint64_t *rounds_out,
                             uint64_t *challenges_out, uint64_t *final_claim, uint64_t *factor_claims_out,
                             hipStream_t stream)
{
    const size_t m = P.num_vars, d = P.d, width = d + 1, nf = P.nf;
    const uint64_t N = 1ULL << m;
    const ScKernels K = pick_kernels(P);

    ScRoundArgs A0{}; // the pointer fields are set per workspace (enqueue)
    A0.width = (uint32_t)width;
    sc_frame_bytes(A0.width, A0.frame);
    for (size_t i = 0; i < width; i++) { // Lagrange weights for nodes 0..d
        uint64_t den = 1;
        for (size_t j = 0; j < width; j++)
            if (j != i) den = gl_mul(den, gl_sub(i, j));
        uint64_t inv = 1, base = den, e = GL_P - 2;
        while (e) {
            if (e & 1) inv = gl_mul(inv, base);
            base = gl_mul(base, base);
            e >>= 1;
        }
        A0.L.w[i] = inv;
    }
    const bool split = P.T.eq_f >= 0;
    A0.pw = split ? (uint32_t)d : (uint32_t)width;
    // split eq: per-round (1 - w_j, 2 w_j - 1), round j's E tables (pairs over
    // variables j+1..m-1: hi over all but the last lb, lo over the last lb)
    // and S(d) = sum_{i<d} x_i S(i) for the degree-(d-1) sum S
    std::vector<uint64_t> eqc, arena;
    std::vector<uint64_t> hi_off(m), lo_off(m), lbits(m);
    if (split) {
        for (size_t i = 0; i < d; i++) {
            uint64_t num = 1, den = 1;
            for (size_t k = 0; k < d; k++)
                if (k != i) {
                    num = gl_mul(num, gl_sub(d, k));
                    den = gl_mul(den, gl_sub(i, k));
                }
            uint64_t inv = 1, base = den, e = GL_P - 2;
            while (e) {
                if (e & 1) inv = gl_mul(inv, base);
                base = gl_mul(base, base);
                e >>= 1;
            }
            A0.L.x[i] = gl_mul(num, inv);
        }
        const size_t LB = m / 2;
        arena.push_back(1); // the empty hi table
        const size_t o_lfix = arena.size();
        host_eq_append(eq_w, m - LB, m, arena);
        for (size_t j = 0; j < m; j++) {
            eqc.push_back(gl_sub(1, eq_w[j]));
            eqc.push_back(gl_sub(gl_add(eq_w[j], eq_w[j]), 1));
            const size_t R = m - j - 1;
            if (R > LB) {
                hi_off[j] = arena.size();
                host_eq_append(eq_w, j + 1, m - LB, arena);
                lo_off[j] = o_lfix;
                lbits[j] = LB;
            } else {
                hi_off[j] = 0;
                lo_off[j] = arena.size();
                host_eq_append(eq_w, j + 1, m, arena);
                lbits[j] = R;
            }
        }
    }
    // one device allocation: folded tables | block totals | control area
    // control area: ScCtl | rounds[m][width] | chal[m] | given[m] | claims[nf] (read back), then
    // eq_c[m][2] | eq_rounds[m] | eq arena (setup only)
    const size_t mm = std::max<size_t>(1, m), nff = std::max<size_t>(1, nf);
    const uint64_t halfN = m ? N / 2 : 1;
    const size_t tab_bytes = nff * halfN * sizeof(uint64_t), part_bytes = 2 * SC_MAX_BLOCKS * width * sizeof(uint64_t);
    const size_t o_rounds = 2 * sizeof(ScCtl), o_chal = o_rounds + mm * width * 8, o_given = o_chal + mm * 8,
                 o_claims = o_given + mm * 8, area_rb = o_claims + nff * 8;
    const size_t o_eqc = (area_rb + 15) & ~(size_t)15, o_eqr = o_eqc + eqc.size() * 8,
                 o_arena = o_eqr + (split ? m * sizeof(ScEqArgs) : 0), area = o_arena + arena.size() * 8;
    const size_t dev_bytes = tab_bytes + part_bytes + area;

    // The staged control area for a workspace at dev: initial state (in
    // A.ctl[1]), given challenges, split eq data.
    auto fill_stage = [&](uint8_t *h, uint8_t *dev) {
        uint8_t *dctl = dev + tab_bytes + part_bytes;
        memset(h, 0, area);
        ScCtl &c = ((ScCtl *)h)[1]; // the state before round 0
        if (tr) {
            for (int i = 0; i < 25; i++) c.tx.st[i] = tr->sponge.st[i];
            c.tx.pos = tr->sponge.pos;
            c.tx.query_count = tr->query_count;
            c.tx.query_budget = tr->query_budget;
        }
        c.current = claim;
        c.eq_scale = 1;
        if (given) memcpy(h + o_given, given, m * sizeof(uint64_t));
        if (split) {
            memcpy(h + o_eqc, eqc.data(), eqc.size() * 8);
            const uint64_t *ar = (const uint64_t *)(dctl + o_arena);
            ScEqArgs *er = (ScEqArgs *)(h + o_eqr);
            for (size_t j = 0; j < m; j++) er[j] = ScEqArgs{ar + hi_off[j], ar + lo_off[j], (uint32_t)lbits[j], 0};
            memcpy(h + o_arena, arena.data(), arena.size() * 8);
        }
    };
    // Every step of a prove on the workspace at dev with stage h: the upload,
    // the round kernels, the readback.  Nothing in it depends on a call's
    // values (only on the sizes, tables and pointers), so it can be captured.
    auto enqueue = [&](uint8_t *dev, uint8_t *h, hipStream_t s) -> lzx_status {
        uint8_t *dctl = dev + tab_bytes + part_bytes;
        LZX_HIP(hipMemcpyAsync(dctl, h, area, hipMemcpyHostToDevice, s));
        uint64_t *partial = (uint64_t *)(dev + tab_bytes);
        ScRoundArgs A = A0;
        A.ctl = (ScCtl *)dctl;
        A.rounds = (uint64_t *)(dctl + o_rounds);
        A.chal = (uint64_t *)(dctl + o_chal);
        A.given = given ? (const uint64_t *)(dctl + o_given) : nullptr;
        A.claims = (uint64_t *)(dctl + o_claims);
        A.eq_c = split ? (const uint64_t *)(dctl + o_eqc) : nullptr;
        A.eq_rounds = split ? (const ScEqArgs *)(dctl + o_eqr) : nullptr;
        auto eq_args = [&](size_t j) {
            if (!split) return ScEqArgs{};
            const uint64_t *ar = (const uint64_t *)(dctl + o_arena);
            return ScEqArgs{ar + hi_off[j], ar + lo_off[j], (uint32_t)lbits[j], 0};
        };
        ScTables tabs{};
        for (size_t f = 0; f < nf; f++) {
            tabs.src[f] = factors[f];
            tabs.dst[f] = (uint64_t *)dev + f * halfN;
        }
        if (m > 0) {
            // round j's block totals go to part[j & 1]; consecutive large
            // kernels visit their pairs in opposite orders
            uint64_t *part[2] = {partial, partial + (size_t)SC_MAX_BLOCKS * width};
            uint32_t gprev = sc_grid(N / 2);
            K.eval(dim3(gprev), s, tabs, P.T, N / 2, eq_args(0), part[0], 0);
            LZX_LAUNCH_CHECK("sc_eval_kernel");
            size_t r = 0;
            for (; (N >> (r + 1)) > K.tail_half; r++) { // a grid per fold
                const uint64_t half = N >> (r + 1);
                const uint32_t g = sc_grid(half / 2);
                if (g > SC_FUSE_BLOCKS || !K.fold_close) { // a one-block kernel closes round r, then the fold
                    K.close(s, part[r & 1], gprev, A, (uint32_t)r);
                    LZX_LAUNCH_CHECK("sc_close_kernel");
                    K.fold(dim3(g), s, tabs, P.T, half, A.chal + r, eq_args(r + 1), part[(r + 1) & 1], r % 2 == 0);
                    LZX_LAUNCH_CHECK("sc_fold_eval_kernel");
                } else { // every block closes round r itself
                    K.fold_close(dim3(g), s, tabs, P.T, half, eq_args(r + 1), part[r & 1], gprev,
                                 part[(r + 1) & 1], A, (uint32_t)r, r % 2 == 0);
                    LZX_LAUNCH_CHECK("sc_fold_close_kernel");
                }
                gprev = g;
                for (size_t f = 0; f < nf; f++) tabs.src[f] = tabs.dst[f]; // later rounds fold in place
            }
            K.tail(s, tabs, P.T, N >> (r + 1), (uint32_t)r, (uint32_t)m, A, part[r & 1], gprev);
            LZX_LAUNCH_CHECK("sc_tail_kernel");
        }
        // zero rounds: the claims are the tables' single entries (sumcheck.rs:139-143)
        for (size_t f = 0; m == 0 && f < nf; f++) {
            if ((int)f == P.T.eq_f) continue; // eq over no variables is 1 (the scale)
            LZX_HIP(hipMemcpyAsync(dctl + o_claims + 8 * f, factors[f], sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        }
        LZX_HIP(hipMemcpyAsync(h, dctl, area_rb, hipMemcpyDeviceToHost, s)); // one readback of the control area
        return LZX_OK;
    };
    // The proof from the read-back control area at h.
    auto finish = [&](const uint8_t *h) -> lzx_status {
        const ScCtl &c = ((const ScCtl *)h)[m ? (m - 1) & 1 : 1]; // the state after the last round
        if (tr) { // the caller's transcript advances exactly as the reference's would
            for (int i = 0; i < 25; i++) tr->sponge.st[i] = c.tx.st[i];
            tr->sponge.pos = c.tx.pos;
            tr->query_count = c.tx.query_count;
        }
        memcpy(rounds_out, h + o_rounds, m * width * sizeof(uint64_t));
        memcpy(challenges_out, h + o_chal, m * sizeof(uint64_t));
        if (c.status) return (lzx_status)c.status;
        memcpy(factor_claims_out, h + o_claims, nf * sizeof(uint64_t));
        if (split) factor_claims_out[P.T.eq_f] = c.eq_scale; // eq(w, r): the bound eq factor
        const uint64_t fin = host_eval_terms(P, factor_claims_out); // sumcheck.rs:144-154
        *final_claim = fin;
        return fin != c.current ? LZX_SUMCHECK_FINAL_CHECK_FAILED : LZX_OK;
    };

    if (m > 0 && sc_graph_enabled()) {
        // the key: everything the captured launches bake in
        ScGraphKey key;
        int dev_id = 0;
        LZX_HIP(hipGetDevice(&dev_id));
        key.add(dev_id).add(m).add(nf).add(d).add(given != nullptr).add(tr != nullptr).add(P.T);
        for (size_t f = 0; f < nf; f++) key.add(factors[f]);
        ScGraphHold hold;
        lzx_status gs = sc_graph_get(key, dev_bytes, area, hold);
        if (gs != LZX_OK) return gs;
        if (ScGraph *g = hold.g) { // held until the proof is read
            if (!g->exec) {
                if (g->graph) (void)hipGraphDestroy(g->graph); // an earlier instantiation failed
                g->graph = nullptr;
                LZX_HIP(hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal));
                lzx_status cs = enqueue(g->dev, (uint8_t *)g->stage, g->cap);
                hipGraph_t graph = nullptr;
                const hipError_t ce = hipStreamEndCapture(g->cap, &graph);
                if (cs != LZX_OK || ce != hipSuccess) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return cs != LZX_OK ? cs : device_fail(ce, "sumcheck graph capture");
                }
                g->graph = graph;
                LZX_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
            }
            fill_stage((uint8_t *)g->stage, g->dev);
            LZX_HIP(hipGraphLaunch(g->exec, stream));
            LZX_HIP(hipStreamSynchronize(stream));
            return finish((const uint8_t *)g->stage);
        }
    }

    uint8_t *dev = nullptr;
    HostStage hs;
    lzx_status st = LZX_OK;
    hipError_t e = scratch_alloc((void **)&dev, dev_bytes, stream);
    if (e == hipSuccess) e = stage_acquire(hs, area);
    if (e != hipSuccess) st = device_fail(e, "sumcheck scratch");
    bool enqueued = false;
    if (st == LZX_OK) {
        fill_stage((uint8_t *)hs.p, dev);
        st = enqueue(dev, (uint8_t *)hs.p, stream);
        enqueued = true;
    }
    if (st == LZX_OK) {
        e = hipStreamSynchronize(stream);
        enqueued = false;
        st = e == hipSuccess ? finish((const uint8_t *)hs.p) : device_fail(e, "sumcheck readback");
    }
    if (enqueued) (void)hipStreamSynchronize(stream); // the staging buffer may still be read
    if (dev) (void)scratch_free(dev, stream);
    stage_release(hs);
    return st;
}

// eq_index >= 0: that factor is eq(eq_w, .) and its table is never read (may
// be null); the split applies when it appears exactly once in every term.
static lzx_status prove_entry(size_t num_vars, size_t nfactors, const uint64_t *const *factors_dev, size_t nterms,
                              const uint64_t *coeffs, const uint32_t *term_off, const uint32_t *term_ids,
                              uint64_t claim, lzx_transcript *tr, const uint64_t *given, uint64_t *rounds_out,
                              uint64_t *challenges_out, uint64_t *final_claim, uint64_t *factor_claims_out,
                              size_t *degree_out, hipStream_t stream, int eq_index = -1,
                              const uint64_t *eq_w = nullptr)
{
    if (!final_claim || !degree_out || (num_vars && (!rounds_out || !challenges_out))) return LZX_INVALID_ARGUMENT;
    if (!tr && !given && num_vars) return LZX_INVALID_ARGUMENT;
    if (nterms == 0) {
        // sumcheck.rs:67-92: zero polynomial, degree-1 zero rounds
        if (claim != 0) return LZX_SUMCHECK_CLAIM_MISMATCH;
        *degree_out = 1;
        for (size_t r = 0; r < num_vars; r++) {
            rounds_out[2 * r] = rounds_out[2 * r + 1] = 0;
            if (tr) {
                lzx_status st = transcript_append_fields(tr, "sumcheck-round", rounds_out + 2 * r, 2);
                if (st == LZX_OK) st = transcript_challenge(tr, "sumcheck-challenge", &challenges_out[r]);
                if (st != LZX_OK) return st;
            } else {
                challenges_out[r] = given[r];
            }
        }
        *final_claim = 0;
        return LZX_OK;
    }
    ScProblem P;
    lzx_status st = build_problem(num_vars, nfactors, factors_dev, nterms, coeffs, term_off, term_ids, P, eq_index);
    if (st != LZX_OK) return st;
    if (!factor_claims_out) return LZX_INVALID_ARGUMENT;
    *degree_out = P.d;
    if (eq_index >= 0) {
        P.T.eq_f = eq_index;
        P.T.common[eq_index] = 0;
        P.T.has_common = 0;
        for (size_t f = 0; f < nfactors; f++) P.T.has_common |= P.T.common[f] != 0;
        P.T.npts = (uint32_t)P.d;
    }
    return prove_impl(P, factors_dev, claim, tr, given, eq_index >= 0 ? eq_w : nullptr, rounds_out, challenges_out,
                      final_claim, factor_claims_out, stream);
}

} // namespace lzx

using namespace lzx;

extern "C" lzx_status lzx_sumcheck_prove(size_t num_vars, size_t nfactors, const uint64_t *const *factors_dev,
                                         size_t nterms, const uint64_t *coeffs, const uint32_t *term_off,
                                         const uint32_t *term_ids, uint64_t claim, lzx_transcript *transcript,
                                         uint64_t *rounds_out, uint64_t *challenges_out, uint64_t *final_claim,
                                         uint64_t *factor_claims_out, size_t *degree_out, void *stream)
{
    if (!transcript) return LZX_INVALID_ARGUMENT;
    return prove_entry(num_vars, nfactors, factors_dev, nterms, coeffs, term_off, term_ids, claim, transcript,
                       nullptr, rounds_out, challenges_out, final_claim, factor_claims_out, degree_out,
                       (hipStream_t)stream);
}

extern "C" lzx_status lzx_sumcheck_prove_with_challenges(size_t num_vars, size_t nfactors,
                                                         const uint64_t *const *factors_dev, size_t nterms,
                                                         const uint64_t *coeffs, const uint32_t *term_off,
                                                         const uint32_t *term_ids, uint64_t claim,
                                                         const uint64_t *challenges, uint64_t *rounds_out,
                                                         uint64_t *final_claim, uint64_t *factor_claims_out,
                                                         size_t *degree_out, void *stream)
{
    if (num_vars && !challenges) return LZX_INVALID_ARGUMENT;
    std::vector<uint64_t> chal_copy(num_vars ? num_vars : 1);
    return prove_entry(num_vars, nfactors, factors_dev, nterms, coeffs, term_off, term_ids, claim, nullptr,
                       challenges, rounds_out, chal_copy.data(), final_claim, factor_claims_out, degree_out,
                       (hipStream_t)stream);
}

extern "C" lzx_status lzx_sumcheck_prove_eq(size_t num_vars, size_t nfactors, const uint64_t *const *factors_dev,
                                            size_t nterms, const uint64_t *coeffs, const uint32_t *term_off,
                                            const uint32_t *term_ids, uint64_t claim, lzx_transcript *transcript,
                                            const size_t *eq_indices, size_t n_eq, const uint64_t *w,
                                            uint64_t *rounds_out, uint64_t *challenges_out, uint64_t *final_claim,
                                            uint64_t *factor_claims_out, size_t *degree_out, void *stream)
{
    if (!transcript || (n_eq && !eq_indices) || (n_eq && num_vars && !w)) return LZX_INVALID_ARGUMENT;
    if (nfactors > (size_t)SC_MAXF || nterms > (size_t)SC_MAXT) return LZX_UNSUPPORTED;
    if (num_vars > 40) return LZX_UNSUPPORTED;
    if (nfactors && !factors_dev) return LZX_INVALID_ARGUMENT;
    if (nterms && (!coeffs || !term_off || !term_ids)) return LZX_INVALID_ARGUMENT;
    // fastprover.rs:245-282: an index names a factor to treat as eq(w, .); indices
    // matching no factor select nothing
    std::vector<int> eqf;
    for (size_t i = 0; i < n_eq; i++)
        if (eq_indices[i] < nfactors && std::find(eqf.begin(), eqf.end(), (int)eq_indices[i]) == eqf.end())
            eqf.push_back((int)eq_indices[i]);
    if (!eqf.empty())
        for (size_t j = 0; j < num_vars; j++)
            if (w[j] >= GL_P) return LZX_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)stream;
    if (nterms && eqf.size() == 1) { // split when eq is in every term exactly once
        bool once = true;
        for (size_t t = 0; t < nterms && once; t++) {
            if (term_off[t + 1] < term_off[t]) return LZX_INVALID_ARGUMENT;
            uint32_t c = 0;
            for (uint32_t k = term_off[t]; k < term_off[t + 1]; k++) c += term_ids[k] == (uint32_t)eqf[0];
            once = c == 1;
        }
        if (once)
            return prove_entry(num_vars, nfactors, factors_dev, nterms, coeffs, term_off, term_ids, claim, transcript,
                               nullptr, rounds_out, challenges_out, final_claim, factor_claims_out, degree_out, s,
                               eqf[0], w);
    }
    // otherwise the eq factors are materialised from w: the same proof
    std::vector<const uint64_t *> fac(factors_dev, factors_dev + nfactors);
    std::vector<uint64_t *> owned;
    lzx_status st = LZX_OK;
    for (int f : eqf) {
        uint64_t *tab = nullptr;
        const hipError_t e = scratch_alloc((void **)&tab, ((size_t)1 << num_vars) * sizeof(uint64_t), s);
        if (e != hipSuccess) {
            st = device_fail(e, "eq table");
            break;
        }
        owned.push_back(tab);
        st = lzx_eq_table(w, num_vars, tab, stream);
        if (st != LZX_OK) break;
        fac[f] = tab;
    }
    if (st == LZX_OK)
        st = prove_entry(num_vars, nfactors, fac.data(), nterms, coeffs, term_off, term_ids, claim, transcript,
                         nullptr, rounds_out, challenges_out, final_claim, factor_claims_out, degree_out, s);
    for (uint64_t *tab : owned) (void)scratch_free(tab, s);
    if (!owned.empty()) (void)hipStreamSynchronize(s);
    return st;
}

extern "C" lzx_status lzx_sumcheck_verify(const uint64_t *rounds, size_t num_rounds, size_t round_len,
                                          size_t num_vars, size_t max_degree, uint64_t claim,
                                          lzx_transcript *transcript, int32_t has_expected, uint64_t expected_final,
                                          uint64_t *point_out, uint64_t *final_claim)
{
    if (!transcript || !final_claim || (num_rounds && (!rounds || !point_out))) return LZX_INVALID_ARGUMENT;
    if (num_rounds != num_vars) return LZX_SUMCHECK_BAD_ROUND_SHAPE; // sumcheck.rs:332-337
    // Goldilocks values are canonical in the reference (field.rs:9-40); a
    // non-canonical word would give one proof two encodings with different
    // challenges, so it is refused before anything is absorbed.
    if (claim >= GL_P || (has_expected && expected_final >= GL_P)) return LZX_INVALID_ARGUMENT;
    for (size_t i = 0; i < num_rounds * round_len; i++)
        if (rounds[i] >= GL_P) return LZX_INVALID_ARGUMENT;
    uint64_t current = claim;
    for (size_t r = 0; r < num_rounds; r++) {
        const uint64_t *ev = rounds + r * round_len;
        // sumcheck.rs:341-346 (a one-entry round cannot satisfy the g(0)+g(1) check)
        if (round_len == 0 || round_len > max_degree + 1 || round_len < 2) return LZX_SUMCHECK_BAD_ROUND_SHAPE;
        lzx_status st = transcript_append_fields(transcript, "sumcheck-round", ev, round_len);
        uint64_t c = 0;
        if (st == LZX_OK) st = transcript_challenge(transcript, "sumcheck-challenge", &c);
        if (st != LZX_OK) return st;
        if (gl_add(ev[0], ev[1]) != current) return LZX_SUMCHECK_ROUND_CHECK_FAILED;
        current = h_interpolate(ev, round_len, c);
        point_out[r] = c;
    }
    if (has_expected && current != expected_final) return LZX_SUMCHECK_FINAL_CHECK_FAILED;
    *final_claim = current;
    return LZX_OK;
}

extern "C" lzx_status lzx_eq_table(const uint64_t *point_host, size_t num_vars, uint64_t *out_dev, void *stream)
{
    if (!out_dev || (num_vars && !point_host)) return LZX_INVALID_ARGUMENT;
    if (num_vars > 40) return LZX_UNSUPPORTED;
    for (size_t i = 0; i < num_vars; i++)
        if (point_host[i] >= GL_P) return LZX_INVALID_ARGUMENT;
    // split the variables: hi = variables 0..h-1 (most significant bits)
    const uint32_t hbits = (uint32_t)(num_vars / 2), lbits = (uint32_t)(num_vars - hbits);
    auto table = [&](size_t first, uint32_t bits) {
        std::vector<uint64_t> t((size_t)1 << bits);
        for (size_t b = 0; b < t.size(); b++) {
            uint64_t v = 1;
            for (uint32_t i = 0; i < bits; i++) {
                const uint64_t p = point_host[first + i];
                const bool bit = (b >> (bits - 1 - i)) & 1;
                v = gl_mul(v, bit ? p : gl_sub(1, p));
            }
            t[b] = v;
        }
        return t;
    };
    std::vector<uint64_t> ehi = table(0, hbits), elo = table(hbits, lbits);
    hipStream_t s = (hipStream_t)stream;
    uint64_t *dh = nullptr, *dl = nullptr;
    LZX_HIP(scratch_alloc((void **)&dh, ehi.size() * sizeof(uint64_t), s));
    LZX_HIP(scratch_alloc((void **)&dl, elo.size() * sizeof(uint64_t), s));
    LZX_HIP(hipMemcpyAsync(dh, ehi.data(), ehi.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    LZX_HIP(hipMemcpyAsync(dl, elo.data(), elo.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    const uint64_t total = 1ULL << num_vars;
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((total + 255) / 256, 1 << 16));
    hipLaunchKernelGGL(eq_combine_kernel, dim3(blocks), dim3(256), 0, s, dh, dl, lbits, total, out_dev);
    LZX_LAUNCH_CHECK("eq_combine_kernel");
    LZX_HIP(scratch_free(dh, s));
    LZX_HIP(scratch_free(dl, s));
    // the host vectors must outlive the async copies
    LZX_HIP(hipStreamSynchronize(s));
    return LZX_OK;
}

extern "C" lzx_status lzx_vp_sum_over_hypercube(size_t num_vars, size_t nfactors, const uint64_t *const *factors_dev,
                                                size_t nterms, const uint64_t *coeffs, const uint32_t *term_off,
                                                const uint32_t *term_ids, uint64_t *sum_out, void *stream)
{
    if (!sum_out) return LZX_INVALID_ARGUMENT;
    if (nterms == 0) {
        *sum_out = 0;
        return LZX_OK;
    }
    ScProblem P;
    lzx_status st = build_problem(num_vars, nfactors, factors_dev, nterms, coeffs, term_off, term_ids, P);
    if (st != LZX_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    const ScKernels K = pick_kernels(P);
    ScTables tabs{};
    for (size_t f = 0; f < P.nf; f++) tabs.src[f] = factors_dev[f];
    if (num_vars == 0) {
        std::vector<uint64_t> v(P.nf);
        for (size_t f = 0; f < P.nf; f++)
            LZX_HIP(hipMemcpy(&v[f], factors_dev[f], sizeof(uint64_t), hipMemcpyDeviceToHost));
        *sum_out = host_eval_terms(P, v.data());
        return LZX_OK;
    }
    const uint64_t half = 1ULL << (num_vars - 1);
    const uint32_t g = sc_grid(half);
    const size_t width = P.d + 1;
    uint64_t *partial = nullptr, *dev_out = nullptr;
    LZX_HIP(scratch_alloc((void **)&partial, ((size_t)g + 1) * width * sizeof(uint64_t), s));
    dev_out = partial + (size_t)g * width;
    K.eval(dim3(g), s, tabs, P.T, half, ScEqArgs{}, partial, 0);
    LZX_LAUNCH_CHECK("sc_eval_kernel");
    hipLaunchKernelGGL(sc_reduce_kernel, dim3(1), dim3(SC_BLOCK), 0, s, partial, g, (uint32_t)width, dev_out);
    LZX_LAUNCH_CHECK("sc_reduce_kernel");
    std::vector<uint64_t> ev(width);
    LZX_HIP(hipMemcpyAsync(ev.data(), dev_out, width * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    LZX_HIP(hipStreamSynchronize(s));
    LZX_HIP(scratch_free(partial, s));
    *sum_out = gl_add(ev[0], ev[1]); // sum over x0 = 0 and x0 = 1
    return LZX_OK;
}

// ---- zerocheck (zerocheck.rs:26-84) -------------------------------------------

// sum_x P(x) eq(w, x) without an eq table: round 0 of the split prover gives
// S(t) = sum_p E_0(p) P(t, p) (t = 0, 1) and the sum is (1 - w_0) S(0) + w_0 S(1).
static lzx_status split_eq_sum(size_t num_vars, const uint64_t *poly_dev, const uint64_t *w, uint64_t *sum_out,
                               hipStream_t s)
{
    const uint64_t *fac[2] = {poly_dev, nullptr};
    const uint64_t coeff[1] = {1};
    const uint32_t off[2] = {0, 2}, ids[2] = {0, 1};
    ScProblem P;
    lzx_status st = build_problem(num_vars, 2, fac, 1, coeff, off, ids, P, 1);
    if (st != LZX_OK) return st;
    P.T.eq_f = 1;
    P.T.common[1] = 0;
    P.T.has_common = P.T.common[0] != 0;
    P.T.npts = 2;
    if (num_vars == 0) { // P(w) is P's single entry
        LZX_HIP(hipMemcpy(sum_out, poly_dev, sizeof(uint64_t), hipMemcpyDeviceToHost));
        return LZX_OK;
    }
    const size_t m = num_vars, LB = m / 2, R = m - 1;
    std::vector<uint64_t> arena(1, 1);
    size_t hi_off = 0, lo_off, lb;
    if (R > LB) {
        hi_off = arena.size();
        host_eq_append(w, 1, m - LB, arena);
        lo_off = arena.size();
        host_eq_append(w, m - LB, m, arena);
        lb = LB;
    } else {
        lo_off = arena.size();
        host_eq_append(w, 1, m, arena);
        lb = R;
    }
    const uint64_t half = 1ULL << R;
    const uint32_t g = sc_grid(half);
    uint64_t *buf = nullptr;
    LZX_HIP(scratch_alloc((void **)&buf, ((size_t)g * 2 + 2 + arena.size()) * sizeof(uint64_t), s));
    uint64_t *partial = buf, *dev_out = buf + (size_t)g * 2, *dar = dev_out + 2;
    hipError_t e = hipMemcpyAsync(dar, arena.data(), arena.size() * 8, hipMemcpyHostToDevice, s);
    ScTables tabs{};
    tabs.src[0] = poly_dev;
    const ScKernels K = pick_kernels(P);
    if (e == hipSuccess) {
        K.eval(dim3(g), s, tabs, P.T, half, ScEqArgs{dar + hi_off, dar + lo_off, (uint32_t)lb, 0}, partial, 0);
        hipLaunchKernelGGL(sc_reduce_kernel, dim3(1), dim3(SC_BLOCK), 0, s, partial, g, 2u, dev_out);
        e = hipGetLastError();
    }
    uint64_t S2[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(S2, dev_out, sizeof(S2), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s); // the arena and S2 are host memory
    (void)scratch_free(buf, s);
    if (e != hipSuccess) return device_fail(e, "zerocheck evaluation");
    *sum_out = gl_add(gl_mul(gl_sub(1, w[0]), S2[0]), gl_mul(w[0], S2[1]));
    return LZX_OK;
}

extern "C" lzx_status lzx_zerocheck_prove(size_t num_vars, const uint64_t *poly_dev, lzx_transcript *transcript,
                                          uint64_t *point_out, uint64_t *rounds_out, uint64_t *challenges_out,
                                          uint64_t *final_claim, uint64_t *factor_claims_out, void *stream)
{
    if (!transcript || !poly_dev || !final_claim || !factor_claims_out ||
        (num_vars && (!point_out || !rounds_out || !challenges_out)))
        return LZX_INVALID_ARGUMENT;
    if (num_vars > 40) return LZX_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    // commit to the random point r (zerocheck.rs:31-34)
    std::vector<uint64_t> r(num_vars ? num_vars : 1);
    static const char LBL[] = "zerocheck-point";
    lzx_status st =
        lzx_transcript_challenge_fields(transcript, (const uint8_t *)LBL, sizeof(LBL) - 1, num_vars, r.data());
    if (st != LZX_OK) return st;
    if (num_vars) memcpy(point_out, r.data(), num_vars * sizeof(uint64_t));
    // the claim P(r) = sum_x P(x) eq(r, x) must be zero (zerocheck.rs:37-45)
    uint64_t pr = 0;
    st = split_eq_sum(num_vars, poly_dev, r.data(), &pr, s);
    if (st != LZX_OK) return st;
    if (pr != 0) return LZX_ZEROCHECK_FINAL_CLAIM_NONZERO;
    // sum-check of P(x) eq(r, x), claim 0 (zerocheck.rs:46-54); eq is never materialised
    const uint64_t *fac[2] = {poly_dev, nullptr};
    const uint64_t coeff[1] = {1};
    const uint32_t off[2] = {0, 2}, ids[2] = {0, 1};
    const size_t eqi[1] = {1};
    size_t deg = 0;
    return lzx_sumcheck_prove_eq(num_vars, 2, fac, 1, coeff, off, ids, 0, transcript, eqi, 1, r.data(), rounds_out,
                                 challenges_out, final_claim, factor_claims_out, &deg, stream);
}

extern "C" lzx_status lzx_zerocheck_verify(const uint64_t *rounds, size_t num_rounds, size_t round_len,
                                           size_t num_vars, lzx_transcript *transcript, int32_t has_expected,
                                           uint64_t expected_poly_eval, uint64_t *point_out)
{
    if (!transcript || (num_vars && !point_out)) return LZX_INVALID_ARGUMENT;
    if (num_vars > 40) return LZX_UNSUPPORTED;
    std::vector<uint64_t> r(num_vars ? num_vars : 1);
    static const char LBL[] = "zerocheck-point";
    lzx_status st =
        lzx_transcript_challenge_fields(transcript, (const uint8_t *)LBL, sizeof(LBL) - 1, num_vars, r.data());
    if (st != LZX_OK) return st;
    // degree-2 rounds (a product of two multilinears), claim 0 (zerocheck.rs:68-77)
    std::vector<uint64_t> point(num_rounds ? num_rounds : 1);
    uint64_t fin = 0;
    st = lzx_sumcheck_verify(rounds, num_rounds, round_len, num_vars, 2, 0, transcript, has_expected,
                             expected_poly_eval, point.data(), &fin);
    if (st != LZX_OK) return st;
    if (num_vars) memcpy(point_out, r.data(), num_vars * sizeof(uint64_t));
    return LZX_OK;
}

// ---- claim batching (batch.rs:33-296) -----------------------------------------

namespace lzx {

// Lift a table over r variables to r + k by ignoring its last k variables
// (the low index bits): out[i] = f[i >> k] (batch.rs:130-148).
__global__ void pad_table_kernel(const uint64_t *__restrict__ f, uint32_t k, uint64_t total, uint64_t *__restrict__ out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = f[i >> k];
}

// 2^k in the field (batch.rs:150-157 doubles k times; square-and-multiply gives the same element)
static uint64_t gl_pow2(uint64_t k)
{
    uint64_t acc = 1, b = 2;
    while (k) {
        if (k & 1) acc = gl_mul(acc, b);
        b = gl_mul(b, b);
        k >>= 1;
    }
    return acc;
}

// BatchError::ChallengeCount stands for every transcript failure while drawing rho
static lzx_status draw_rhos(lzx_transcript *tr, const char *label, size_t k, uint64_t *rhos)
{
    const lzx_status st = lzx_transcript_challenge_fields(tr, (const uint8_t *)label, strlen(label), k, rhos);
    return st == LZX_TRANSCRIPT_REJECTION_BUDGET || st == LZX_TRANSCRIPT_MESSAGE_TOO_LARGE ? LZX_BATCH_CHALLENGE_COUNT
                                                                                           : st;
}

// prove_batch (star = false) and prove_batch_star (star = true) share this: rho_i
// from the transcript, the combined virtual polynomial over the union of the
// claims' factors (rho_i c_t per term, a table passed twice folded once, lifted
// tables built on the device), one device prove, the factor claims mapped back.
static lzx_status prove_batch_impl(const lzx_batch_claim *claims, size_t k, bool star, lzx_transcript *tr,
                                   uint64_t *rhos_out, uint64_t *rounds_out, uint64_t *challenges_out,
                                   uint64_t *final_claim, uint64_t *factor_claims_out, size_t *num_vars_out,
                                   size_t *degree_out, hipStream_t s)
{
    if (!tr || !final_claim || !degree_out) return LZX_INVALID_ARGUMENT;
    if (k == 0) return LZX_BATCH_EMPTY; // batch.rs:37-39, 165-167
    if (!claims || !rhos_out) return LZX_INVALID_ARGUMENT;
    size_t N = 0, total_f = 0;
    for (size_t i = 0; i < k; i++) {
        const lzx_batch_claim &c = claims[i];
        if ((c.nfactors && !c.factors_dev) || (c.nterms && (!c.coeffs || !c.term_off || !c.term_ids)) ||
            c.claimed_sum >= GL_P)
            return LZX_INVALID_ARGUMENT;
        if (c.num_vars > 40) return LZX_UNSUPPORTED;
        N = std::max(N, c.num_vars);
        total_f += c.nfactors;
    }
    if (!star)
        for (size_t i = 0; i < k; i++)
            if (claims[i].num_vars != claims[0].num_vars) return LZX_BATCH_VARIABLE_COUNT_MISMATCH; // batch.rs:40-46
    if ((N && (!rounds_out || !challenges_out)) || (total_f && !factor_claims_out)) return LZX_INVALID_ARGUMENT;

    // the combined problem's shape first, so a refused batch leaves the transcript untouched
    std::vector<std::pair<const uint64_t *, uint32_t>> tabs; // distinct (table, lift)
    std::vector<uint32_t> gmap(total_f);                      // global factor -> distinct table
    std::vector<uint32_t> off(1, 0), ids;
    size_t g = 0, nterms_all = 0, degree = 0;
    for (size_t i = 0; i < k; i++) {
        const lzx_batch_claim &c = claims[i];
        const uint32_t lift = (uint32_t)(N - c.num_vars);
        const size_t base = g;
        for (size_t f = 0; f < c.nfactors; f++) {
            if (!c.factors_dev[f]) return LZX_INVALID_ARGUMENT;
            const std::pair<const uint64_t *, uint32_t> key(c.factors_dev[f], lift);
            size_t j = std::find(tabs.begin(), tabs.end(), key) - tabs.begin();
            if (j == tabs.size()) tabs.push_back(key);
            gmap[g++] = (uint32_t)j;
        }
        for (size_t t = 0; t < c.nterms; t++) {
            if (c.term_off[t + 1] < c.term_off[t] || c.coeffs[t] >= GL_P) return LZX_INVALID_ARGUMENT;
            if (c.term_off[t + 1] == c.term_off[t]) return LZX_SUMCHECK_VIRTUAL_POLY; // add_term: EmptyProduct
            for (uint32_t x = c.term_off[t]; x < c.term_off[t + 1]; x++) {
                if (c.term_ids[x] >= c.nfactors) return LZX_SUMCHECK_VIRTUAL_POLY;
                ids.push_back(gmap[base + c.term_ids[x]]);
            }
            off.push_back((uint32_t)ids.size());
            degree = std::max<size_t>(degree, c.term_off[t + 1] - c.term_off[t]);
            nterms_all++;
        }
    }
    if (tabs.size() > (size_t)SC_MAXF || nterms_all > (size_t)SC_MAXT || degree > (size_t)SC_MAXD)
        return LZX_UNSUPPORTED;
    lzx_status st = draw_rhos(tr, star ? "batch-star-rho" : "batch-rho", k, rhos_out);
    if (st != LZX_OK) return st;
    std::vector<uint64_t> coeffs;
    uint64_t combined = 0;
    for (size_t i = 0; i < k; i++) {
        const lzx_batch_claim &c = claims[i];
        const uint64_t rho = rhos_out[i];
        combined = gl_add(combined, gl_mul(gl_mul(rho, gl_pow2((uint32_t)(N - c.num_vars))), c.claimed_sum));
        for (size_t t = 0; t < c.nterms; t++) coeffs.push_back(gl_mul(rho, c.coeffs[t]));
    }
    if (num_vars_out) *num_vars_out = N;

    // lifted tables: 2^N entries each in scratch
    std::vector<const uint64_t *> fac(tabs.size());
    std::vector<uint64_t *> owned;
    for (size_t j = 0; j < tabs.size() && st == LZX_OK; j++) {
        fac[j] = tabs[j].first;
        if (tabs[j].second == 0) continue;
        uint64_t *dst = nullptr;
        const uint64_t total = 1ULL << N;
        hipError_t e = scratch_alloc((void **)&dst, total * sizeof(uint64_t), s);
        if (e == hipSuccess) {
            owned.push_back(dst);
            const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((total + 255) / 256, 1 << 16));
            hipLaunchKernelGGL(pad_table_kernel, dim3(blocks), dim3(256), 0, s, tabs[j].first, tabs[j].second, total,
                               dst);
            e = hipGetLastError();
        }
        if (e != hipSuccess) st = device_fail(e, "batch lift");
        fac[j] = dst;
    }
    std::vector<uint64_t> fclaims(std::max<size_t>(1, tabs.size()), 0);
    if (st == LZX_OK)
        st = prove_entry(N, tabs.size(), fac.data(), coeffs.size(), coeffs.data(), off.data(), ids.data(), combined,
                         tr, nullptr, rounds_out, challenges_out, final_claim, fclaims.data(), degree_out, s);
    for (uint64_t *p : owned) (void)scratch_free(p, s);
    if (st == LZX_OK && !coeffs.empty())
        for (size_t x = 0; x < total_f; x++) factor_claims_out[x] = fclaims[gmap[x]];
    return st;
}

} // namespace lzx

extern "C" lzx_status lzx_sumcheck_prove_batch(const lzx_batch_claim *claims, size_t nclaims,
                                               lzx_transcript *transcript, uint64_t *rhos_out, uint64_t *rounds_out,
                                               uint64_t *challenges_out, uint64_t *final_claim,
                                               uint64_t *factor_claims_out, size_t *degree_out, void *stream)
{
    return prove_batch_impl(claims, nclaims, false, transcript, rhos_out, rounds_out, challenges_out, final_claim,
                            factor_claims_out, nullptr, degree_out, (hipStream_t)stream);
}

extern "C" lzx_status lzx_sumcheck_prove_batch_star(const lzx_batch_claim *claims, size_t nclaims,
                                                    lzx_transcript *transcript, uint64_t *rhos_out,
                                                    uint64_t *rounds_out, uint64_t *challenges_out,
                                                    uint64_t *final_claim, uint64_t *factor_claims_out,
                                                    size_t *num_vars_out, size_t *degree_out, void *stream)
{
    return prove_batch_impl(claims, nclaims, true, transcript, rhos_out, rounds_out, challenges_out, final_claim,
                            factor_claims_out, num_vars_out, degree_out, (hipStream_t)stream);
}

extern "C" lzx_status lzx_sumcheck_verify_batch(const uint64_t *rounds, size_t num_rounds, size_t round_len,
                                                size_t num_vars, size_t max_degree, const uint64_t *claimed_sums,
                                                size_t nclaims, lzx_transcript *transcript, int32_t has_expected,
                                                uint64_t expected_final, uint64_t *point_out)
{
    if (!transcript || (nclaims && !claimed_sums)) return LZX_INVALID_ARGUMENT;
    std::vector<uint64_t> rhos(std::max<size_t>(1, nclaims));
    lzx_status st = draw_rhos(transcript, "batch-rho", nclaims, rhos.data());
    if (st != LZX_OK) return st;
    uint64_t combined = 0;
    for (size_t i = 0; i < nclaims; i++) {
        if (claimed_sums[i] >= GL_P) return LZX_INVALID_ARGUMENT;
        combined = gl_add(combined, gl_mul(rhos[i], claimed_sums[i]));
    }
    uint64_t fin = 0;
    return lzx_sumcheck_verify(rounds, num_rounds, round_len, num_vars, max_degree, combined, transcript,
                               has_expected, expected_final, point_out, &fin);
}

extern "C" lzx_status lzx_sumcheck_verify_batch_star(const uint64_t *rounds, size_t num_rounds, size_t round_len,
                                                     const size_t *num_vars, size_t n_num_vars, size_t max_degree,
                                                     const uint64_t *claimed_sums, size_t nclaims,
                                                     const uint64_t *rhos, size_t nrhos, lzx_transcript *transcript,
                                                     int32_t has_expected, uint64_t expected_final,
                                                     uint64_t *point_out)
{
    if (!transcript) return LZX_INVALID_ARGUMENT;
    if (n_num_vars == 0 || n_num_vars != nclaims) return LZX_BATCH_VARIABLE_COUNT_MISMATCH; // batch.rs:243-248
    if (!num_vars || !claimed_sums) return LZX_INVALID_ARGUMENT;
    std::vector<uint64_t> drawn;
    if (!rhos) { // verify_batch_star draws them; verify_batch_star_with_rhos is handed them
        drawn.resize(nclaims);
        const lzx_status st = draw_rhos(transcript, "batch-star-rho", nclaims, drawn.data());
        if (st != LZX_OK) return st;
        rhos = drawn.data();
        nrhos = nclaims;
    }
    if (nrhos != nclaims) return LZX_BATCH_CHALLENGE_COUNT; // batch.rs:284-286
    const size_t max_vars = *std::max_element(num_vars, num_vars + n_num_vars);
    uint64_t combined = 0;
    for (size_t i = 0; i < nclaims; i++) {
        if (claimed_sums[i] >= GL_P || rhos[i] >= GL_P) return LZX_INVALID_ARGUMENT;
        combined = gl_add(combined, gl_mul(gl_mul(rhos[i], gl_pow2(max_vars - num_vars[i])), claimed_sums[i]));
    }
    uint64_t fin = 0;
    return lzx_sumcheck_verify(rounds, num_rounds, round_len, max_vars, max_degree, combined, transcript,
                               has_expected, expected_final, point_out, &fin);
}

// ---- Goldilocks product rate (measurement aid) -------------------------------------
namespace lzx {

constexpr int GLR_CHAINS = 8;

__global__ __launch_bounds__(1024) void gl_mul_rate_kernel(uint64_t *__restrict__ out, uint32_t iters, uint64_t seed)
{
    uint64_t x[GLR_CHAINS];
    const uint64_t y = seed * 0x9E3779B97F4A7C15ULL + threadIdx.x;
#pragma unroll
    for (int i = 0; i < GLR_CHAINS; i++) x[i] = y + (uint64_t)i * 0x12345ULL;
    for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < GLR_CHAINS; i++) x[i] = gl_mul_lazy(x[i], y);
    }
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < GLR_CHAINS; i++) acc ^= x[i];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

} // namespace lzx

extern "C" lzx_status lzx_gl_mul_rate(uint32_t waves_per_simd, double *products_per_s)
{
    if (!products_per_s || (waves_per_simd != 4 && waves_per_simd != 8)) return LZX_INVALID_ARGUMENT;
    int dev = 0, cus = 0;
    LZX_HIP(hipGetDevice(&dev));
    LZX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t threads = 1024, blocks = (uint32_t)cus * (256 * waves_per_simd / threads), iters = 4096;
    uint64_t *out = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipMalloc((void **)&out, (size_t)blocks * threads * sizeof(uint64_t));
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    float best = 0;
    for (int rep = 0; rep < 3 && e == hipSuccess; rep++) { // the first run warms the clocks
        hipLaunchKernelGGL(gl_mul_rate_kernel, dim3(blocks), dim3(threads), 0, 0, out, iters, 7);
        e = hipEventRecord(a, 0);
        if (e == hipSuccess) hipLaunchKernelGGL(gl_mul_rate_kernel, dim3(blocks), dim3(threads), 0, 0, out, iters, 7);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(b, 0);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (e == hipSuccess && rep > 0 && (best == 0 || ms < best)) best = ms;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (out) (void)hipFree(out);
    if (e != hipSuccess) return device_fail(e, "gl_mul_rate");
    *products_per_s = (double)blocks * threads * iters * GLR_CHAINS / (best * 1e-3);
    return LZX_OK;
}

// ---- round-level session ---------------------------------------------------------
// The prover's rounds one call at a time, for callers that drive their own
// transcript between rounds (sumcheck.rs:62-162 with the absorb/challenge
// steps left to the caller): the session evaluates the current round's
// g_j(0..d) on the device and folds with the caller's challenge.

namespace lzx {

// The last fold: each factor's two entries become its claim (sumcheck.rs:139-143).
__global__ void sc_last_fold_kernel(ScTables tabs, uint32_t nf, const uint64_t *__restrict__ rp,
                                    uint64_t *__restrict__ claims)
{
    const uint32_t f = threadIdx.x;
    if (f < nf) {
        const uint64_t a = tabs.src[f][0], b = tabs.src[f][1], r = *rp;
        claims[f] = gl_add(a, gl_mul(gl_sub(b, a), r));
    }
}

} // namespace lzx

struct lzx_sumcheck_session {
    lzx::ScProblem P;
    lzx::ScKernels K;
    lzx::ScTables tabs;
    hipStream_t stream;
    size_t width, round;
    uint64_t half;             // pairs of the current round (0 once every variable is bound)
    uint8_t *dev = nullptr;    // folded tables | block partials | evals[width] | r | claims[nf]
    uint64_t *partial, *evals, *rdev, *claims;
    bool zero;                 // no terms: the zero polynomial's degree-1 zero rounds (sumcheck.rs:67-92)
};

namespace lzx {

static lzx_status session_reduce(lzx_sumcheck_session *S, uint32_t blocks)
{
    hipLaunchKernelGGL(sc_reduce_kernel, dim3(1), dim3(SC_BLOCK), 0, S->stream, S->partial, blocks,
                       (uint32_t)S->width, S->evals);
    LZX_LAUNCH_CHECK("sc_reduce_kernel");
    return LZX_OK;
}

} // namespace lzx

extern "C" lzx_status lzx_sumcheck_session_create(size_t num_vars, size_t nfactors, const uint64_t *const *factors_dev,
                                                  size_t nterms, const uint64_t *coeffs, const uint32_t *term_off,
                                                  const uint32_t *term_ids, void *stream,
                                                  lzx_sumcheck_session **out)
{
    if (!out) return LZX_INVALID_ARGUMENT;
    *out = nullptr;
    lzx_sumcheck_session *S = new lzx_sumcheck_session();
    S->stream = (hipStream_t)stream;
    S->round = 0;
    S->zero = nterms == 0;
    if (S->zero) {
        memset(&S->P, 0, sizeof(S->P));
        S->P.num_vars = num_vars;
        S->width = 2;
        S->half = num_vars ? 1ULL << (num_vars - 1) : 0;
        *out = S;
        return LZX_OK;
    }
    lzx_status st = build_problem(num_vars, nfactors, factors_dev, nterms, coeffs, term_off, term_ids, S->P);
    if (st != LZX_OK) {
        delete S;
        return st;
    }
    S->K = pick_kernels(S->P);
    S->width = S->P.d + 1;
    const size_t nf = S->P.nf, N = (size_t)1 << num_vars, halfN = num_vars ? N / 2 : 1;
    const size_t tab_bytes = std::max<size_t>(1, nf) * halfN * 8, part_bytes = (size_t)SC_MAX_BLOCKS * S->width * 8;
    const size_t bytes = tab_bytes + part_bytes + (S->width + 1 + std::max<size_t>(1, nf)) * 8;
    const hipError_t e = scratch_alloc((void **)&S->dev, bytes, S->stream);
    if (e != hipSuccess) {
        delete S;
        return device_fail(e, "sumcheck session scratch");
    }
    S->partial = (uint64_t *)(S->dev + tab_bytes);
    S->evals = S->partial + (size_t)SC_MAX_BLOCKS * S->width;
    S->rdev = S->evals + S->width;
    S->claims = S->rdev + 1;
    S->tabs = ScTables{};
    for (size_t f = 0; f < nf; f++) {
        S->tabs.src[f] = factors_dev[f];
        S->tabs.dst[f] = (uint64_t *)S->dev + f * halfN;
    }
    S->half = num_vars ? N / 2 : 0;
    if (num_vars) { // round 0's evaluations over the caller's tables
        const uint32_t g = sc_grid(S->half);
        S->K.eval(dim3(g), S->stream, S->tabs, S->P.T, S->half, ScEqArgs{}, S->partial, 0);
        hipError_t le = hipGetLastError();
        if (le == hipSuccess) st = session_reduce(S, g);
        else st = device_fail(le, "sc_eval_kernel");
    } else { // no variables: the claims are the tables' single entries
        for (size_t f = 0; f < nf && st == LZX_OK; f++) {
            const hipError_t ce =
                hipMemcpyAsync(S->claims + f, factors_dev[f], 8, hipMemcpyDeviceToDevice, S->stream);
            if (ce != hipSuccess) st = device_fail(ce, "session claims");
        }
    }
    if (st != LZX_OK) {
        (void)scratch_free(S->dev, S->stream);
        delete S;
        return st;
    }
    *out = S;
    return LZX_OK;
}

extern "C" size_t lzx_sumcheck_session_round_len(const lzx_sumcheck_session *S) { return S ? S->width : 0; }

extern "C" lzx_status lzx_sumcheck_session_round_eval(lzx_sumcheck_session *S, uint64_t *evals_out)
{
    if (!S || !evals_out) return LZX_INVALID_ARGUMENT;
    if (S->half == 0) return LZX_SUMCHECK_BAD_ROUND_SHAPE; // every variable is bound
    if (S->zero) {
        evals_out[0] = evals_out[1] = 0;
        return LZX_OK;
    }
    LZX_HIP(hipMemcpyAsync(evals_out, S->evals, S->width * 8, hipMemcpyDeviceToHost, S->stream));
    LZX_HIP(hipStreamSynchronize(S->stream));
    return LZX_OK;
}

extern "C" lzx_status lzx_sumcheck_session_fold(lzx_sumcheck_session *S, uint64_t r)
{
    if (!S || r >= GL_P) return LZX_INVALID_ARGUMENT;
    if (S->half == 0) return LZX_SUMCHECK_BAD_ROUND_SHAPE;
    S->round++;
    if (S->zero) {
        S->half >>= 1;
        return LZX_OK;
    }
    LZX_HIP(hipMemcpyAsync(S->rdev, &r, 8, hipMemcpyHostToDevice, S->stream));
    if (S->half == 1) {
        hipLaunchKernelGGL(sc_last_fold_kernel, dim3(1), dim3(64), 0, S->stream, S->tabs, (uint32_t)S->P.nf, S->rdev,
                           S->claims);
        LZX_LAUNCH_CHECK("sc_last_fold_kernel");
        S->half = 0;
    } else {
        const uint32_t g = sc_grid(S->half / 2);
        S->K.fold(dim3(g), S->stream, S->tabs, S->P.T, S->half, S->rdev, ScEqArgs{}, S->partial, 0);
        LZX_LAUNCH_CHECK("sc_fold_eval_kernel");
        for (size_t f = 0; f < S->P.nf; f++) S->tabs.src[f] = S->tabs.dst[f]; // later rounds fold in place
        S->half >>= 1;
        const lzx_status st = session_reduce(S, g);
        if (st != LZX_OK) return st;
    }
    // r is pageable host memory: the copy must land before the call returns
    LZX_HIP(hipStreamSynchronize(S->stream));
    return LZX_OK;
}

extern "C" lzx_status lzx_sumcheck_session_final(lzx_sumcheck_session *S, uint64_t *factor_claims_out,
                                                 uint64_t *final_claim)
{
    if (!S || !final_claim) return LZX_INVALID_ARGUMENT;
    if (S->half != 0) return LZX_SUMCHECK_BAD_ROUND_SHAPE; // variables left to bind
    if (S->zero) {
        *final_claim = 0;
        return LZX_OK;
    }
    if (S->P.nf && !factor_claims_out) return LZX_INVALID_ARGUMENT;
    LZX_HIP(hipMemcpyAsync(factor_claims_out, S->claims, S->P.nf * 8, hipMemcpyDeviceToHost, S->stream));
    LZX_HIP(hipStreamSynchronize(S->stream));
    *final_claim = host_eval_terms(S->P, factor_claims_out); // sumcheck.rs:144-154
    return LZX_OK;
}

extern "C" void lzx_sumcheck_session_destroy(lzx_sumcheck_session *S)
{
    if (!S) return;
    if (S->dev) {
        (void)scratch_free(S->dev, S->stream);
        (void)hipStreamSynchronize(S->stream);
    }
    delete S;
}

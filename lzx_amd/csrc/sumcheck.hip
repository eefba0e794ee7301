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
// later rounds fold in place in scratch.  Per-block partial sums (canonical)
// are reduced by sc_reduce_kernel; the d+1 values go to the host for the
// transcript.
#include <algorithm>
#include <cstring>
#include <vector>

#include "modarith.hpp"
#include "runtime.hpp"
#include "seeded.hpp"

namespace lzx {

constexpr int SC_MAXF = 8;
constexpr int SC_MAXT = 16;
constexpr int SC_MAXD = 8;
constexpr int SC_BLOCK = 256;
constexpr uint32_t SC_MAX_BLOCKS = 2048;

struct ScTerms {
    uint32_t nterms, nf, d;
    uint64_t coeff[SC_MAXT];
    uint8_t exps[SC_MAXT][SC_MAXF]; // multiplicity of factor f in term t
    uint8_t first[SC_MAXT];         // a factor with exps > 0 (seeds the product)
};

struct ScTables {
    const uint64_t *src[SC_MAXF];
    uint64_t *dst[SC_MAXF];
};

// sum_terms c * prod_f v[f]^e over the current point values v.
template <int NF>
__device__ __forceinline__ uint64_t eval_terms(const uint64_t (&v)[NF], const ScTerms &T)
{
    uint64_t total = 0;
    GlAcc acc{0, 0};
    for (uint32_t t = 0; t < T.nterms; t++) {
        uint64_t prod = 1;
        bool seeded = false;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const uint32_t e = T.exps[t][f];
            for (uint32_t x = 0; x < e; x++) {
                prod = seeded ? gl_mul(prod, v[f]) : v[f];
                seeded = true;
            }
        }
        const uint64_t c = T.coeff[t];
        if (c == GL_P - 1)
            prod = prod == 0 ? 0 : GL_P - prod;
        else if (c != 1)
            prod = gl_mul(prod, c);
        gl_acc_add(acc, prod);
    }
    (void)total;
    return gl_acc_finish(acc);
}

// Accumulate g(t), t = 0..d, of one pair (lo, hi) into acc[].
template <int NF, int D>
__device__ __forceinline__ void eval_pair(const uint64_t (&lo)[NF], const uint64_t (&hi)[NF], const ScTerms &T,
                                          GlAcc (&acc)[D + 1])
{
    uint64_t diff[NF], cur[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        diff[f] = gl_sub(hi[f], lo[f]);
        cur[f] = lo[f];
    }
    gl_acc_add(acc[0], eval_terms<NF>(cur, T));
    if (D >= 1) gl_acc_add(acc[1], eval_terms<NF>(hi, T));
#pragma unroll
    for (int f = 0; f < NF; f++) cur[f] = hi[f];
#pragma unroll
    for (int t = 2; t <= D; t++) {
        if ((uint32_t)t > T.d) break;
#pragma unroll
        for (int f = 0; f < NF; f++) cur[f] = gl_add(cur[f], diff[f]); // lo + t*(hi-lo)
        gl_acc_add(acc[t], eval_terms<NF>(cur, T));
    }
}

// Block-reduce the per-thread lazy sums and write canonical partials.
template <int D>
__device__ __forceinline__ void block_store(GlAcc (&acc)[D + 1], uint64_t *__restrict__ partial, uint32_t dd)
{
    __shared__ uint64_t red[SC_BLOCK];
    for (uint32_t t = 0; t <= dd; t++) {
        red[threadIdx.x] = gl_acc_finish(acc[t]);
        __syncthreads();
        for (uint32_t s = SC_BLOCK / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] = gl_add(red[threadIdx.x], red[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(size_t)blockIdx.x * (dd + 1) + t] = red[0];
        __syncthreads();
    }
}

template <int NF, int D>
__global__ __launch_bounds__(SC_BLOCK) void sc_eval_kernel(ScTables tabs, ScTerms T, uint64_t half,
                                                           uint64_t *__restrict__ partial)
{
    GlAcc acc[D + 1];
#pragma unroll
    for (int t = 0; t <= D; t++) acc[t] = GlAcc{0, 0};
    for (uint64_t p = (uint64_t)blockIdx.x * SC_BLOCK + threadIdx.x; p < half; p += (uint64_t)gridDim.x * SC_BLOCK) {
        uint64_t lo[NF], hi[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
            lo[f] = (uint32_t)f < T.nf ? tabs.src[f][p] : 0;
            hi[f] = (uint32_t)f < T.nf ? tabs.src[f][p + half] : 0;
        }
        eval_pair<NF, D>(lo, hi, T, acc);
    }
    block_store<D>(acc, partial, T.d);
}

// Fold with r (tables of length 2*half -> half) and evaluate the next round
// (pairs (p, p + half/2) of the folded tables) when half >= 2.
template <int NF, int D>
__global__ __launch_bounds__(SC_BLOCK) void sc_fold_eval_kernel(ScTables tabs, ScTerms T, uint64_t half, uint64_t r,
                                                                uint64_t *__restrict__ partial)
{
    GlAcc acc[D + 1];
#pragma unroll
    for (int t = 0; t <= D; t++) acc[t] = GlAcc{0, 0};
    const uint64_t quarter = half >> 1;
    if (quarter == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
#pragma unroll
            for (int f = 0; f < NF; f++)
                if ((uint32_t)f < T.nf) {
                    const uint64_t a = tabs.src[f][0], b = tabs.src[f][1];
                    tabs.dst[f][0] = gl_add(a, gl_mul(gl_sub(b, a), r));
                }
        }
        return;
    }
    for (uint64_t p = (uint64_t)blockIdx.x * SC_BLOCK + threadIdx.x; p < quarter;
         p += (uint64_t)gridDim.x * SC_BLOCK) {
        uint64_t n0[NF], n1[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
            if ((uint32_t)f < T.nf) {
                const uint64_t *s = tabs.src[f];
                const uint64_t a0 = s[p], b0 = s[p + half], a1 = s[p + quarter], b1 = s[p + half + quarter];
                n0[f] = gl_add(a0, gl_mul(gl_sub(b0, a0), r));
                n1[f] = gl_add(a1, gl_mul(gl_sub(b1, a1), r));
                tabs.dst[f][p] = n0[f];
                tabs.dst[f][p + quarter] = n1[f];
            } else {
                n0[f] = n1[f] = 0;
            }
        }
        eval_pair<NF, D>(n0, n1, T, acc);
    }
    block_store<D>(acc, partial, T.d);
}

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
    ScTerms T;
};

static lzx_status build_problem(size_t num_vars, size_t nfactors, const uint64_t *const *factors, size_t nterms,
                                const uint64_t *coeffs, const uint32_t *term_off, const uint32_t *term_ids,
                                ScProblem &P)
{
    if (nterms && (!coeffs || !term_off || !term_ids)) return LZX_INVALID_ARGUMENT;
    if (nfactors && !factors) return LZX_INVALID_ARGUMENT;
    if (nfactors > (size_t)SC_MAXF || nterms > (size_t)SC_MAXT) return LZX_UNSUPPORTED;
    if (num_vars > 40) return LZX_UNSUPPORTED;
    P.num_vars = num_vars;
    P.nf = nfactors;
    P.nterms = nterms;
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
            P.T.exps[t][term_ids[k]]++;
        }
        P.T.first[t] = (uint8_t)term_ids[term_off[t]];
    }
    if (d > (size_t)SC_MAXD) return LZX_UNSUPPORTED;
    for (size_t f = 0; f < nfactors; f++)
        if (!factors[f]) return LZX_INVALID_ARGUMENT;
    P.d = d;
    P.T.d = (uint32_t)d;
    return LZX_OK;
}

using EvalFn = void (*)(dim3, hipStream_t, const ScTables &, const ScTerms &, uint64_t, uint64_t *);
using FoldFn = void (*)(dim3, hipStream_t, const ScTables &, const ScTerms &, uint64_t, uint64_t, uint64_t *);

template <int NF, int D>
static void launch_eval(dim3 g, hipStream_t s, const ScTables &tb, const ScTerms &T, uint64_t half, uint64_t *part)
{
    hipLaunchKernelGGL((sc_eval_kernel<NF, D>), g, dim3(SC_BLOCK), 0, s, tb, T, half, part);
}

template <int NF, int D>
static void launch_fold(dim3 g, hipStream_t s, const ScTables &tb, const ScTerms &T, uint64_t half, uint64_t r,
                        uint64_t *part)
{
    hipLaunchKernelGGL((sc_fold_eval_kernel<NF, D>), g, dim3(SC_BLOCK), 0, s, tb, T, half, r, part);
}

static void pick_kernels(size_t nf, size_t d, EvalFn &ev, FoldFn &fo)
{
    if (nf <= 2 && d <= 2) {
        ev = launch_eval<2, 2>;
        fo = launch_fold<2, 2>;
    } else if (nf <= 4 && d <= 3) {
        ev = launch_eval<4, 3>;
        fo = launch_fold<4, 3>;
    } else {
        ev = launch_eval<SC_MAXF, SC_MAXD>;
        fo = launch_fold<SC_MAXF, SC_MAXD>;
    }
}

// The prover loop shared by the transcript and explicit-challenge forms.
static lzx_status prove_impl(const ScProblem &P, const uint64_t *const *factors, uint64_t claim, lzx_transcript *tr,
                             const uint64_t *given, uint64_t *rounds_out, uint64_t *challenges_out,
                             uint64_t *final_claim, uint64_t *factor_claims_out, hipStream_t stream)
{
    const size_t m = P.num_vars, d = P.d, width = d + 1, nf = P.nf;
    const uint64_t N = 1ULL << m;
    EvalFn evf;
    FoldFn fof;
    pick_kernels(nf, d, evf, fof);

    uint64_t *scratch = nullptr, *partial = nullptr, *dev_out = nullptr, *host_out = nullptr;
    const uint64_t halfN = m ? N / 2 : 1;
    lzx_status st = LZX_OK;
    hipError_t e = hipMallocAsync((void **)&scratch, std::max<size_t>(1, nf) * halfN * sizeof(uint64_t), stream);
    if (e == hipSuccess)
        e = hipMallocAsync((void **)&partial, (size_t)SC_MAX_BLOCKS * width * sizeof(uint64_t), stream);
    if (e == hipSuccess) e = hipMallocAsync((void **)&dev_out, std::max<size_t>(width, nf) * sizeof(uint64_t), stream);
    if (e == hipSuccess) e = hipHostMalloc((void **)&host_out, std::max<size_t>(width, nf) * sizeof(uint64_t), 0);
    if (e != hipSuccess) st = device_fail(e, "sumcheck scratch");

    auto grid_for = [](uint64_t work) {
        return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((work + SC_BLOCK - 1) / SC_BLOCK, SC_MAX_BLOCKS));
    };
    auto fetch_evals = [&](uint32_t nblocks, uint64_t *evals) -> lzx_status {
        hipLaunchKernelGGL(sc_reduce_kernel, dim3(1), dim3(SC_BLOCK), 0, stream, partial, nblocks, (uint32_t)width,
                           dev_out);
        hipError_t le = hipGetLastError();
        if (le != hipSuccess) return device_fail(le, "sc_reduce_kernel");
        le = hipMemcpyAsync(host_out, dev_out, width * sizeof(uint64_t), hipMemcpyDeviceToHost, stream);
        if (le == hipSuccess) le = hipStreamSynchronize(stream);
        if (le != hipSuccess) return device_fail(le, "sumcheck round readback");
        memcpy(evals, host_out, width * sizeof(uint64_t));
        return LZX_OK;
    };

    ScTables tabs{};
    for (size_t f = 0; f < nf; f++) {
        tabs.src[f] = factors[f];
        tabs.dst[f] = scratch + f * halfN;
    }
    uint64_t current = claim;
    if (st == LZX_OK && m > 0) {
        const uint32_t g0 = grid_for(N / 2);
        evf(dim3(g0), stream, tabs, P.T, N / 2, partial);
        e = hipGetLastError();
        if (e != hipSuccess) st = device_fail(e, "sc_eval_kernel");
        if (st == LZX_OK) st = fetch_evals(g0, rounds_out);
    }
    for (size_t r = 0; st == LZX_OK && r < m; r++) {
        uint64_t *ev = rounds_out + r * width;
        uint64_t ch;
        if (tr) {
            st = transcript_append_fields(tr, "sumcheck-round", ev, width); // sumcheck.rs:110-112
            if (st == LZX_OK) st = transcript_challenge(tr, "sumcheck-challenge", &ch);
            if (st != LZX_OK) break;
        } else {
            ch = given[r];
        }
        challenges_out[r] = ch;
        if (gl_add(ev[0], ev[1]) != current) { // sumcheck.rs:120-123
            st = LZX_SUMCHECK_CLAIM_MISMATCH;
            break;
        }
        current = h_interpolate(ev, width, ch);
        const uint64_t half = N >> (r + 1);
        const uint32_t g = grid_for(std::max<uint64_t>(1, half / 2));
        fof(dim3(g), stream, tabs, P.T, half, ch, partial);
        e = hipGetLastError();
        if (e != hipSuccess) {
            st = device_fail(e, "sc_fold_eval_kernel");
            break;
        }
        for (size_t f = 0; f < nf; f++) tabs.src[f] = tabs.dst[f]; // later rounds fold in place
        if (r + 1 < m) st = fetch_evals(g, rounds_out + (r + 1) * width);
    }
    if (st == LZX_OK) {
        // factor claims: the single remaining entry of every table (sumcheck.rs:139-143)
        for (size_t f = 0; f < nf && st == LZX_OK; f++) {
            e = hipMemcpyAsync(host_out + f, m ? tabs.dst[f] : factors[f], sizeof(uint64_t), hipMemcpyDeviceToHost,
                               stream);
            if (e != hipSuccess) st = device_fail(e, "factor claim readback");
        }
        if (st == LZX_OK) {
            e = hipStreamSynchronize(stream);
            if (e != hipSuccess) st = device_fail(e, "factor claim sync");
        }
        if (st == LZX_OK) {
            for (size_t f = 0; f < nf; f++) factor_claims_out[f] = host_out[f];
            uint64_t fin = 0; // sumcheck.rs:144-154
            for (size_t t = 0; t < P.nterms; t++) {
                uint64_t prod = P.T.coeff[t];
                for (size_t f = 0; f < nf; f++)
                    for (uint32_t x = 0; x < P.T.exps[t][f]; x++) prod = gl_mul(prod, factor_claims_out[f]);
                fin = gl_add(fin, prod);
            }
            *final_claim = fin;
            if (fin != current) st = LZX_SUMCHECK_FINAL_CHECK_FAILED;
        }
    }
    if (scratch) (void)hipFreeAsync(scratch, stream);
    if (partial) (void)hipFreeAsync(partial, stream);
    if (dev_out) (void)hipFreeAsync(dev_out, stream);
    if (host_out) {
        (void)hipStreamSynchronize(stream);
        (void)hipHostFree(host_out);
    }
    return st;
}

static lzx_status prove_entry(size_t num_vars, size_t nfactors, const uint64_t *const *factors_dev, size_t nterms,
                              const uint64_t *coeffs, const uint32_t *term_off, const uint32_t *term_ids,
                              uint64_t claim, lzx_transcript *tr, const uint64_t *given, uint64_t *rounds_out,
                              uint64_t *challenges_out, uint64_t *final_claim, uint64_t *factor_claims_out,
                              size_t *degree_out, hipStream_t stream)
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
    lzx_status st = build_problem(num_vars, nfactors, factors_dev, nterms, coeffs, term_off, term_ids, P);
    if (st != LZX_OK) return st;
    if (!factor_claims_out) return LZX_INVALID_ARGUMENT;
    *degree_out = P.d;
    return prove_impl(P, factors_dev, claim, tr, given, rounds_out, challenges_out, final_claim, factor_claims_out,
                      stream);
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

extern "C" lzx_status lzx_sumcheck_verify(const uint64_t *rounds, size_t num_rounds, size_t round_len,
                                          size_t num_vars, size_t max_degree, uint64_t claim,
                                          lzx_transcript *transcript, int32_t has_expected, uint64_t expected_final,
                                          uint64_t *point_out, uint64_t *final_claim)
{
    if (!transcript || !final_claim || (num_rounds && (!rounds || !point_out))) return LZX_INVALID_ARGUMENT;
    if (num_rounds != num_vars) return LZX_SUMCHECK_BAD_ROUND_SHAPE; // sumcheck.rs:332-337
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
    LZX_HIP(hipMallocAsync((void **)&dh, ehi.size() * sizeof(uint64_t), s));
    LZX_HIP(hipMallocAsync((void **)&dl, elo.size() * sizeof(uint64_t), s));
    LZX_HIP(hipMemcpyAsync(dh, ehi.data(), ehi.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    LZX_HIP(hipMemcpyAsync(dl, elo.data(), elo.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    const uint64_t total = 1ULL << num_vars;
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((total + 255) / 256, 1 << 16));
    hipLaunchKernelGGL(eq_combine_kernel, dim3(blocks), dim3(256), 0, s, dh, dl, lbits, total, out_dev);
    LZX_LAUNCH_CHECK("eq_combine_kernel");
    LZX_HIP(hipFreeAsync(dh, s));
    LZX_HIP(hipFreeAsync(dl, s));
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
    EvalFn evf;
    FoldFn fof;
    pick_kernels(P.nf, P.d, evf, fof);
    ScTables tabs{};
    for (size_t f = 0; f < P.nf; f++) tabs.src[f] = factors_dev[f];
    if (num_vars == 0) {
        std::vector<uint64_t> v(P.nf);
        for (size_t f = 0; f < P.nf; f++)
            LZX_HIP(hipMemcpy(&v[f], factors_dev[f], sizeof(uint64_t), hipMemcpyDeviceToHost));
        uint64_t tot = 0;
        for (size_t t = 0; t < P.nterms; t++) {
            uint64_t prod = P.T.coeff[t];
            for (size_t f = 0; f < P.nf; f++)
                for (uint32_t x = 0; x < P.T.exps[t][f]; x++) prod = gl_mul(prod, v[f]);
            tot = gl_add(tot, prod);
        }
        *sum_out = tot;
        return LZX_OK;
    }
    const uint64_t half = 1ULL << (num_vars - 1);
    const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((half + SC_BLOCK - 1) / SC_BLOCK,
                                                                           SC_MAX_BLOCKS));
    const size_t width = P.d + 1;
    uint64_t *partial = nullptr, *dev_out = nullptr;
    LZX_HIP(hipMallocAsync((void **)&partial, (size_t)g * width * sizeof(uint64_t), s));
    LZX_HIP(hipMallocAsync((void **)&dev_out, width * sizeof(uint64_t), s));
    evf(dim3(g), s, tabs, P.T, half, partial);
    LZX_LAUNCH_CHECK("sc_eval_kernel");
    hipLaunchKernelGGL(sc_reduce_kernel, dim3(1), dim3(SC_BLOCK), 0, s, partial, g, (uint32_t)width, dev_out);
    LZX_LAUNCH_CHECK("sc_reduce_kernel");
    std::vector<uint64_t> ev(width);
    LZX_HIP(hipMemcpyAsync(ev.data(), dev_out, width * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    LZX_HIP(hipStreamSynchronize(s));
    LZX_HIP(hipFreeAsync(partial, s));
    LZX_HIP(hipFreeAsync(dev_out, s));
    *sum_out = gl_add(ev[0], ev[1]); // sum over x0 = 0 and x0 = 1
    return LZX_OK;
}

/* CPU baseline restatements of the reference's CPU paths, timed by bench.py.
 *
 * TEST INFRASTRUCTURE ONLY (part of the oracle library): bench.py's
 * cpu_baseline leg and tests/ call these; the product (liblzx_gpu.so) never
 * does.
 *
 *   or_sc_prove_ref  sumcheck::prove (crates/lattice-sumcheck/src/sumcheck.rs:
 *                    62-162) with round_evals_single_bind (:172-226): per t,
 *                    bind every factor into a buffer (t >= 2), then one lazy
 *                    8-lane term-product pass per term (Sum8,
 *                    crates/lattice-core/src/field_simd.rs:461-535), then the
 *                    in-place first-variable fold of every factor.
 *                    simd = 0 is the reference's LZX_NO_SIMD=1 mode; simd = 1
 *                    its AVX-512F kernels (field_simd.rs:545-660 mul/add/sub,
 *                    :789-826 bind_half / bind_first_half, :892-922
 *                    accumulate_term), selected only when the CPU reports
 *                    avx512f (the reference's is_x86_feature_detected gate,
 *                    field_simd.rs:107-123).  threads > 1 splits each pass
 *                    over index ranges with OpenMP (each range its own Sum8;
 *                    the canonical total is order-independent), a labelled
 *                    all-cores column: the reference itself is
 *                    single-threaded.
 *   or_ntt_roundtrip_par, or_ajtai_commit_par
 *                    the scalar NTT forward+inverse (ntt.rs:118-216) and the
 *                    cached-NTT commit (ajtai.rs:118-153) over OpenMP threads
 *                    (polynomials; witness elements, then rows).
 * Every variant returns the same values as the scalar restatement.
 */
#include <immintrin.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lzx_oracle.h"

#define GLP 0xFFFFFFFF00000001ULL
#define GLC 0xFFFFFFFFULL /* 2^64 mod p */

/* ---- Sum8: 8 lazy u64 lanes with per-lane wrap counts (field_simd.rs:440-535).
 * Invariant: the lane's true sum = acc + car * (2^32 - 1)  (mod p). */
typedef struct {
    uint64_t acc[8];
    uint64_t car[8];
} sum8;

static uint64_t sum8_finish(const sum8 *s)
{
    uint64_t total = 0;
    for (int l = 0; l < 8; l++) {
        unsigned __int128 v = (unsigned __int128)s->acc[l] + (unsigned __int128)s->car[l] * GLC;
        total = or_gl_add(total, or_gl_from_u128((uint64_t)(v >> 64), (uint64_t)v));
    }
    return total;
}

/* scalar continuation on lane 0 (field_simd.rs:508-518) */
static inline void sum8_add0(sum8 *s, uint64_t prod)
{
    const uint64_t v = s->acc[0] + prod;
    s->car[0] += v < prod;
    s->acc[0] = v;
}

/* ---- AVX-512F kernels ---------------------------------------------------- */
#define AVX __attribute__((target("avx512f")))

AVX static inline __m512i v_q(void) { return _mm512_set1_epi64((long long)GLP); }
AVX static inline __m512i v_c(void) { return _mm512_set1_epi64((long long)GLC); }
AVX static inline __m512i v_one(void) { return _mm512_set1_epi64(1); }

/* one fold (hi, lo) -> lo + hi * (2^32 - 1) split into (hi', lo') in u64 lanes */
AVX static inline void v_fold(__m512i *hi, __m512i *lo)
{
    const __m512i sum = _mm512_add_epi64(*lo, _mm512_slli_epi64(*hi, 32));
    const __mmask8 cy = _mm512_cmplt_epu64_mask(sum, *lo);
    const __m512i top = _mm512_add_epi64(_mm512_srli_epi64(*hi, 32), _mm512_maskz_set1_epi64(cy, 1));
    const __mmask8 bw = _mm512_cmplt_epu64_mask(sum, *hi);
    *lo = _mm512_sub_epi64(sum, *hi);
    *hi = _mm512_mask_sub_epi64(top, bw, top, v_one());
}

/* canonical a*b mod p per lane, from four 32x32 products and three folds */
AVX static inline __m512i v_mul(__m512i a, __m512i b)
{
    const __m512i ah = _mm512_srli_epi64(a, 32), bh = _mm512_srli_epi64(b, 32);
    const __m512i ll = _mm512_mul_epu32(a, b), m1 = _mm512_mul_epu32(a, bh), m2 = _mm512_mul_epu32(ah, b);
    const __m512i hh = _mm512_mul_epu32(ah, bh);
    const __m512i mid = _mm512_add_epi64(m1, m2);
    const __mmask8 c1 = _mm512_cmplt_epu64_mask(mid, m1);
    __m512i lo = _mm512_add_epi64(_mm512_slli_epi64(mid, 32), ll);
    const __mmask8 c2 = _mm512_cmplt_epu64_mask(lo, ll);
    __m512i hi = _mm512_add_epi64(_mm512_add_epi64(hh, _mm512_srli_epi64(mid, 32)),
                                  _mm512_add_epi64(_mm512_maskz_set1_epi64(c2, 1),
                                                   _mm512_slli_epi64(_mm512_maskz_set1_epi64(c1, 1), 32)));
    v_fold(&hi, &lo);
    v_fold(&hi, &lo);
    v_fold(&hi, &lo);
    lo = _mm512_mask_add_epi64(lo, _mm512_cmpeq_epu64_mask(hi, v_one()), lo, v_c());
    return _mm512_mask_sub_epi64(lo, _mm512_cmpge_epu64_mask(lo, v_q()), lo, v_q());
}

AVX static inline __m512i v_add(__m512i a, __m512i b)
{
    __m512i s = _mm512_add_epi64(a, b);
    s = _mm512_mask_add_epi64(s, _mm512_cmplt_epu64_mask(s, a), s, v_c());
    return _mm512_mask_sub_epi64(s, _mm512_cmpge_epu64_mask(s, v_q()), s, v_q());
}

AVX static inline __m512i v_sub(__m512i a, __m512i b)
{
    __m512i d = _mm512_sub_epi64(a, b);
    d = _mm512_mask_sub_epi64(d, _mm512_cmplt_epu64_mask(a, b), d, v_c());
    return _mm512_mask_sub_epi64(d, _mm512_cmpge_epu64_mask(d, v_q()), d, v_q());
}

/* out[i] = lo[i] + (hi[i] - lo[i]) * r over [i0, i1) (8-aligned span) */
AVX static void v_bind_half(const uint64_t *lo, const uint64_t *hi, uint64_t r, uint64_t *out, size_t i0, size_t i1)
{
    const __m512i vr = _mm512_set1_epi64((long long)r);
    for (size_t i = i0; i + 8 <= i1; i += 8) {
        const __m512i a = _mm512_loadu_si512(lo + i), b = _mm512_loadu_si512(hi + i);
        _mm512_storeu_si512(out + i, v_add(a, v_mul(v_sub(b, a), vr)));
    }
}

/* acc += coeff * prod_f factors[f][p] for p in [p0, p1), 8 lanes */
AVX static void v_accumulate(sum8 *s, uint64_t coeff, const uint64_t *const *fs, size_t nf, size_t p0, size_t p1)
{
    __m512i acc = _mm512_loadu_si512(s->acc), car = _mm512_loadu_si512(s->car);
    const __m512i c = _mm512_set1_epi64((long long)coeff);
    for (size_t p = p0; p + 8 <= p1; p += 8) {
        __m512i prod = c;
        for (size_t f = 0; f < nf; f++) prod = v_mul(prod, _mm512_loadu_si512(fs[f] + p));
        const __m512i sum = _mm512_add_epi64(acc, prod);
        car = _mm512_add_epi64(car, _mm512_maskz_set1_epi64(_mm512_cmplt_epu64_mask(sum, acc), 1));
        acc = sum;
    }
    _mm512_storeu_si512(s->acc, acc);
    _mm512_storeu_si512(s->car, car);
}

int or_cpu_has_avx512f(void) { return __builtin_cpu_supports("avx512f") ? 1 : 0; }

/* ---- the passes, scalar or AVX-512, over [lo, hi) index ranges ------------ */

static void bind_half_range(int simd, const uint64_t *lo, const uint64_t *hi, uint64_t r, uint64_t *out, size_t i0,
                            size_t i1)
{
    size_t i = i0;
    if (simd) {
        const size_t n8 = i0 + (i1 - i0) / 8 * 8;
        v_bind_half(lo, hi, r, out, i0, n8);
        i = n8;
    }
    for (; i < i1; i++) out[i] = or_gl_add(lo[i], or_gl_mul(or_gl_sub(hi[i], lo[i]), r));
}

static void accumulate_range(int simd, sum8 *s, uint64_t coeff, const uint64_t *const *fs, size_t nf, size_t p0,
                             size_t p1)
{
    size_t p = p0;
    if (simd) {
        const size_t n8 = p0 + (p1 - p0) / 8 * 8;
        v_accumulate(s, coeff, fs, nf, p0, n8);
        p = n8;
    }
    for (; p < p1; p++) {
        uint64_t prod = coeff;
        for (size_t f = 0; f < nf; f++) prod = or_gl_mul(prod, fs[f][p]);
        sum8_add0(s, prod);
    }
}

/* [i0, i1) of thread `id` among `nt` over n items, 8-aligned boundaries */
static void split_range(size_t n, int id, int nt, size_t *i0, size_t *i1)
{
    const size_t chunks = (n + 7) / 8;
    const size_t a = chunks * (size_t)id / (size_t)nt, b = chunks * (size_t)(id + 1) / (size_t)nt;
    *i0 = a * 8 < n ? a * 8 : n;
    *i1 = b * 8 < n ? b * 8 : n;
}

/* sumcheck.rs:62-162 with round_evals_single_bind (:172-226); same contract
 * as or_sumcheck_prove, plus simd (0/1) and threads (>= 1).  Returns
 * OR_BAD_ARGUMENT if simd is asked for on a CPU without avx512f. */
int or_sc_prove_ref(size_t num_vars, size_t nfactors, const uint64_t *const *factors, size_t nterms,
                    const uint64_t *coeffs, const uint32_t *term_off, const uint32_t *term_ids, uint64_t claim,
                    or_transcript *tr, uint64_t *rounds_out, uint64_t *challenges_out, uint64_t *final_claim_out,
                    uint64_t *factor_claims_out, size_t *d_out, int simd, int threads)
{
    static const uint8_t LROUND[] = "sumcheck-round";
    static const uint8_t LCHAL[] = "sumcheck-challenge";
    if (simd && !or_cpu_has_avx512f()) return OR_BAD_ARGUMENT;
    if (threads < 1) threads = 1;
    if (nterms == 0 || nfactors == 0)
        return or_sumcheck_prove(num_vars, nfactors, factors, nterms, coeffs, term_off, term_ids, claim, tr,
                                 rounds_out, challenges_out, final_claim_out, factor_claims_out, d_out);
    size_t d = 0;
    for (size_t t = 0; t < nterms; t++)
        if (term_off[t + 1] - term_off[t] > d) d = term_off[t + 1] - term_off[t];
    *d_out = d;
    size_t len = (size_t)1 << num_vars;
    uint64_t **bound = (uint64_t **)malloc(sizeof(uint64_t *) * nfactors);
    uint64_t **buf = (uint64_t **)malloc(sizeof(uint64_t *) * nfactors);
    const uint64_t **vals = (const uint64_t **)malloc(sizeof(uint64_t *) * nfactors);
    for (size_t f = 0; f < nfactors; f++) {
        bound[f] = (uint64_t *)malloc(sizeof(uint64_t) * len);
        memcpy(bound[f], factors[f], sizeof(uint64_t) * len);
        buf[f] = (uint64_t *)malloc(sizeof(uint64_t) * (len / 2 ? len / 2 : 1));
    }
    size_t maxdeg = d ? d : 1;
    const uint64_t **fs_all = (const uint64_t **)malloc(sizeof(uint64_t *) * maxdeg * (size_t)threads);
    sum8 *parts = (sum8 *)malloc(sizeof(sum8) * (size_t)threads);
    uint64_t current = claim;
    int status = OR_OK;
    for (size_t r = 0; r < num_vars; r++) {
        const size_t points = len / 2;
        uint64_t *ev = rounds_out + r * (d + 1);
        for (size_t t = 0; t <= d; t++) {
            for (size_t f = 0; f < nfactors; f++) vals[f] = t == 0 ? bound[f] : t == 1 ? bound[f] + points : buf[f];
#pragma omp parallel num_threads(threads)
            {
                const int id = omp_get_thread_num(), nt = omp_get_num_threads();
                size_t i0, i1;
                split_range(points, id, nt, &i0, &i1);
                if (t >= 2)
                    for (size_t f = 0; f < nfactors; f++)
                        bind_half_range(simd, bound[f], bound[f] + points, (uint64_t)t, buf[f], i0, i1);
                sum8 *s = &parts[id];
                memset(s, 0, sizeof(*s));
                const uint64_t **fs = fs_all + (size_t)id * maxdeg;
                for (size_t tt = 0; tt < nterms; tt++) {
                    const size_t nf = term_off[tt + 1] - term_off[tt];
                    for (size_t k = 0; k < nf; k++) fs[k] = vals[term_ids[term_off[tt] + k]];
                    accumulate_range(simd, s, coeffs[tt], fs, nf, i0, i1);
                }
            }
            uint64_t total = 0;
            for (int i = 0; i < threads; i++) total = or_gl_add(total, sum8_finish(&parts[i]));
            ev[t] = total;
        }
        if ((status = or_transcript_append_field_slice(tr, LROUND, 14, ev, d + 1))) goto done;
        if ((status = or_transcript_challenge_field(tr, LCHAL, 18, &challenges_out[r]))) goto done;
        if (or_gl_add(ev[0], ev[1]) != current) {
            status = OR_SUMCHECK_CLAIM_MISMATCH;
            goto done;
        }
        current = or_interpolate_at(ev, d + 1, challenges_out[r]);
        const uint64_t ch = challenges_out[r];
#pragma omp parallel num_threads(threads)
        {
            size_t i0, i1;
            split_range(points, omp_get_thread_num(), omp_get_num_threads(), &i0, &i1);
            for (size_t f = 0; f < nfactors; f++) bind_half_range(simd, bound[f], bound[f] + points, ch, bound[f], i0, i1);
        }
        len /= 2;
    }
    {
        uint64_t fin = 0;
        for (size_t f = 0; f < nfactors; f++) factor_claims_out[f] = bound[f][0];
        for (size_t t = 0; t < nterms; t++) {
            uint64_t prod = coeffs[t];
            for (uint32_t k = term_off[t]; k < term_off[t + 1]; k++) prod = or_gl_mul(prod, factor_claims_out[term_ids[k]]);
            fin = or_gl_add(fin, prod);
        }
        *final_claim_out = fin;
        if (fin != current) status = OR_SUMCHECK_FINAL_CHECK_FAILED;
    }
done:
    for (size_t f = 0; f < nfactors; f++) {
        free(bound[f]);
        free(buf[f]);
    }
    free(bound);
    free(buf);
    free(vals);
    free(fs_all);
    free(parts);
    return status;
}

/* ntt.rs:118-216 forward then inverse of `polys` polynomials, threads-way. */
void or_ntt_roundtrip_par(const or_ntt *t, uint32_t *data, size_t polys, int threads)
{
    const size_t n = t->n;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (size_t p = 0; p < polys; p++) {
        or_ntt_forward(t, data + p * n, n);
        or_ntt_inverse(t, data + p * n, n);
    }
}

/* ajtai.rs:118-153 over `threads` threads: the witness transforms split over
 * j, the MAC and inverse over rows i.  Same values as or_ajtai_commit. */
int or_ajtai_commit_par(const or_ntt *t, size_t k, size_t m, const uint32_t *Ahat, const uint32_t *s, size_t s_len,
                        uint32_t *out, int threads)
{
    if (s_len != m) return OR_AJTAI_DIMENSION_MISMATCH;
    const size_t n = t->n;
    const or_mod32 *q = &t->mod;
    uint32_t *s_ntt = (uint32_t *)malloc(sizeof(uint32_t) * n * (m ? m : 1));
    uint8_t *nz = (uint8_t *)malloc(m ? m : 1);
#pragma omp parallel for num_threads(threads) schedule(static)
    for (size_t j = 0; j < m; j++) {
        uint8_t z = 0;
        for (size_t c = 0; c < n; c++) z |= s[j * n + c] != 0;
        nz[j] = z;
        memcpy(s_ntt + j * n, s + j * n, sizeof(uint32_t) * n);
        if (z) or_ntt_forward(t, s_ntt + j * n, n);
    }
#pragma omp parallel for num_threads(threads) schedule(static)
    for (size_t i = 0; i < k; i++) {
        uint32_t *acc = out + i * n;
        memset(acc, 0, sizeof(uint32_t) * n);
        for (size_t j = 0; j < m; j++) {
            if (!nz[j]) continue;
            const uint32_t *a = Ahat + (i * m + j) * n, *b = s_ntt + j * n;
            for (size_t c = 0; c < n; c++) acc[c] = or_add(q, acc[c], or_mul(q, a[c], b[c]));
        }
        or_ntt_inverse(t, acc, n);
    }
    free(nz);
    free(s_ntt);
    return OR_OK;
}

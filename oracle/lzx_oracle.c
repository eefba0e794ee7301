/*
 * lzx_oracle.c -- CPU restatement of the LZX lattice-commitment hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: the parity checker for
 * the HIP product path and the timed scalar CPU baseline.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing measured or shipped.
 *
 * Every function restates the algorithm of the Rust reference (idrees2516/lzx)
 * at the cited file:line (paths relative to the reference's crates/).  The
 * reference cannot be compiled here (no Rust toolchain), so it is pinned by:
 *   - the reference's own known answers (Goldilocks KATs field.rs:268-313,
 *     FIPS 202 KATs keccak.rs:416-446, reduce_i64 cases modulus.rs:254-265,
 *     Lagrange KAT sumcheck.rs:503-509), held in tests/golden/;
 *   - Python's hashlib SHA3/SHAKE (FIPS 202) for the Keccak permutation;
 *   - closed forms computed independently in Python big integers: the direct
 *     odd-power DFT for the NTT, schoolbook negacyclic convolution for ring
 *     products and Ajtai commitments, hypercube enumeration for sum-check
 *     round polynomials and the multilinear evaluation identity.
 * Parity unpinned: the byte-exact Fiat-Shamir challenge VALUES of
 * Transcript::challenge_fields (its squeeze is non-standard and the reference
 * holds no vector for it); only its structure is pinned, by code reading and
 * by self-consistency tests.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lzx_oracle.h"


/* ------------------------------------------------------------------------ */
/* Modulus32: lattice-ring/src/modulus.rs:22-162                             */
/* ------------------------------------------------------------------------ */


or_mod32 or_mod32_new(uint32_t q, uint32_t two_adicity, uint32_t generator)
{
    or_mod32 m;
    m.q = q;
    m.two_adicity = two_adicity;
    m.generator = generator;
    m.barrett = UINT64_MAX / (uint64_t)q;
    return m;
}

/* modulus.rs:65-76: Barrett with u128 intermediate, <=2 corrections. */
uint32_t or_reduce_u64(const or_mod32 *m, uint64_t x)
{
    u128 q = m->q;
    u128 qhat = ((u128)x * (u128)m->barrett) >> 64;
    u128 r = (u128)x - qhat * q;
    if (r >= q) r -= q;
    if (r >= q) r -= q;
    return (uint32_t)r;
}

/* modulus.rs:80-83: rem_euclid. */
uint32_t or_reduce_i64(const or_mod32 *m, int64_t x)
{
    int64_t q = (int64_t)m->q;
    int64_t r = x % q;
    if (r < 0) r += q;
    return (uint32_t)r;
}

uint32_t or_mul(const or_mod32 *m, uint32_t a, uint32_t b) /* modulus.rs:87 */
{
    return or_reduce_u64(m, (uint64_t)a * (uint64_t)b);
}

uint32_t or_add(const or_mod32 *m, uint32_t a, uint32_t b) /* modulus.rs:93-100 */
{
    uint64_t s = (uint64_t)a + (uint64_t)b;
    return s >= m->q ? (uint32_t)(s - m->q) : (uint32_t)s;
}

uint32_t or_sub(const or_mod32 *m, uint32_t a, uint32_t b) /* modulus.rs:104-112 */
{
    return a >= b ? a - b : a + (m->q - b);
}

uint32_t or_neg(const or_mod32 *m, uint32_t a) /* modulus.rs:116-122 */
{
    return a == 0 ? 0 : m->q - a;
}

uint32_t or_pow(const or_mod32 *m, uint32_t base, uint64_t e) /* modulus.rs:125-137 */
{
    uint32_t acc = 1;
    uint32_t b = or_reduce_u64(m, base);
    while (e > 0) {
        if (e & 1) acc = or_mul(m, acc, b);
        b = or_mul(m, b, b);
        e >>= 1;
    }
    return acc;
}

/* modulus.rs:140-146: Fermat inverse; returns 0 for a == 0 (None). */
uint32_t or_inv(const or_mod32 *m, uint32_t a)
{
    if (a == 0) return 0;
    return or_pow(m, a, (uint64_t)m->q - 2);
}

/* modulus.rs:149-156: g^((q-1) >> k); returns 0 when k > two_adicity (None). */
uint32_t or_root_of_unity(const or_mod32 *m, uint32_t k)
{
    if (k > m->two_adicity) return 0;
    return or_pow(m, m->generator, ((uint64_t)m->q - 1) >> k);
}

/* ------------------------------------------------------------------------ */
/* NttTables: lattice-ring/src/ntt.rs:16-342                                 */
/* ------------------------------------------------------------------------ */


void or_ntt_free(or_ntt *t)
{
    if (!t) return;
    free(t->pow_psi);
    free(t->pow_psi_inv);
    free(t->pow_omega);
    free(t->pow_omega_inv);
    free(t);
}

/* ntt.rs:46-106.  The per-level tables of ntt.rs:80-94 are views of
 * pow_omega with stride n/len, so forward/forward_fast read pow_omega
 * directly at index stride*t (identical values). */
int or_ntt_new(uint32_t q, uint32_t two_adicity, uint32_t generator, uint32_t log_n, or_ntt **out)
{
    *out = NULL;
    or_mod32 m = or_mod32_new(q, two_adicity, generator);
    if (log_n + 1 > m.two_adicity) return OR_NTT_ADICITY_TOO_SMALL;
    if (log_n > 24) return OR_BAD_ARGUMENT;
    uint32_t n = 1u << log_n;
    uint32_t psi = or_root_of_unity(&m, log_n + 1);
    uint32_t psi_inv = or_inv(&m, psi);
    uint32_t omega = or_mul(&m, psi, psi);
    uint32_t omega_inv = or_inv(&m, omega);
    uint32_t n_inv = or_pow(&m, n, (uint64_t)m.q - 2);
    if (psi == 0 || psi_inv == 0 || omega_inv == 0) return OR_NTT_ADICITY_TOO_SMALL;
    or_ntt *t = (or_ntt *)calloc(1, sizeof(or_ntt));
    uint32_t half = n / 2 > 0 ? n / 2 : 1;
    t->mod = m;
    t->log_n = log_n;
    t->n = n;
    t->psi = psi;
    t->n_inv = n_inv;
    t->pow_psi = (uint32_t *)malloc(sizeof(uint32_t) * n);
    t->pow_psi_inv = (uint32_t *)malloc(sizeof(uint32_t) * n);
    t->pow_omega = (uint32_t *)malloc(sizeof(uint32_t) * half);
    t->pow_omega_inv = (uint32_t *)malloc(sizeof(uint32_t) * half);
    uint32_t p = 1, pi = 1;
    for (uint32_t i = 0; i < n; i++) {
        t->pow_psi[i] = p;
        t->pow_psi_inv[i] = pi;
        p = or_mul(&m, p, psi);
        pi = or_mul(&m, pi, psi_inv);
    }
    uint32_t w = 1, wi = 1;
    for (uint32_t i = 0; i < half; i++) {
        t->pow_omega[i] = w;
        t->pow_omega_inv[i] = wi;
        w = or_mul(&m, w, omega);
        wi = or_mul(&m, wi, omega_inv);
    }
    *out = t;
    return OR_OK;
}

uint32_t or_ntt_psi(const or_ntt *t) { return t->psi; }
uint32_t or_ntt_n_inv(const or_ntt *t) { return t->n_inv; }

/* ntt.rs:334-342 */
static uint32_t reverse_bits(uint32_t x, uint32_t bits)
{
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

/* ntt.rs:320-332 */
static void bit_reverse(uint32_t *a, uint32_t n)
{
    if (n <= 1) return;
    uint32_t bits = (uint32_t)__builtin_ctz(n);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t j = reverse_bits(i, bits);
        if (i < j) {
            uint32_t tmp = a[i];
            a[i] = a[j];
            a[j] = tmp;
        }
    }
}

/* Cooley-Tukey DIT levels len = 2^(first+1) .. 2^(last) (ntt.rs:165-180). */
static void ct_levels(const or_ntt *t, uint32_t *a, uint32_t first_len, uint32_t max_levels)
{
    const or_mod32 *q = &t->mod;
    uint32_t n = t->n, applied = 0;
    for (uint32_t len = first_len; len <= n && applied < max_levels; len <<= 1, applied++) {
        uint32_t stride = n / len;
        for (uint32_t start = 0; start < n; start += len) {
            for (uint32_t j = 0; j < len / 2; j++) {
                uint32_t w = t->pow_omega[stride * j];
                uint32_t u = a[start + j];
                uint32_t v = or_mul(q, a[start + j + len / 2], w);
                a[start + j] = or_add(q, u, v);
                a[start + j + len / 2] = or_sub(q, u, v);
            }
        }
    }
}

/* ntt.rs:118-146 (forward_fast) == ntt.rs:153-182 (forward).  Output is
 * natural order: out[k] = sum_i a_i psi^(i(2k+1)). */
int or_ntt_forward(const or_ntt *t, uint32_t *a, size_t len)
{
    if (len != t->n) return OR_NTT_LENGTH_MISMATCH;
    for (uint32_t i = 0; i < t->n; i++) a[i] = or_mul(&t->mod, a[i], t->pow_psi[i]);
    bit_reverse(a, t->n);
    ct_levels(t, a, 2, 64);
    return OR_OK;
}

/* ntt.rs:187-216 (inverse_fast) == ntt.rs:220-250 (inverse). */
int or_ntt_inverse(const or_ntt *t, uint32_t *a, size_t len)
{
    if (len != t->n) return OR_NTT_LENGTH_MISMATCH;
    const or_mod32 *q = &t->mod;
    uint32_t n = t->n;
    for (uint32_t l = n; l >= 2; l >>= 1) {
        uint32_t stride = n / l;
        for (uint32_t start = 0; start < n; start += l) {
            for (uint32_t j = 0; j < l / 2; j++) {
                uint32_t w = t->pow_omega_inv[stride * j];
                uint32_t u = a[start + j];
                uint32_t v = a[start + j + l / 2];
                a[start + j] = or_add(q, u, v);
                a[start + j + l / 2] = or_mul(q, or_sub(q, u, v), w);
            }
        }
    }
    bit_reverse(a, n);
    for (uint32_t i = 0; i < n; i++) a[i] = or_mul(q, or_mul(q, a[i], t->pow_psi_inv[i]), t->n_inv);
    return OR_OK;
}

/* ntt.rs:257-286 */
int or_ntt_forward_partial(const or_ntt *t, uint32_t *a, size_t len, uint32_t levels)
{
    if (len != t->n || levels > t->log_n) return OR_NTT_LENGTH_MISMATCH;
    for (uint32_t i = 0; i < t->n; i++) a[i] = or_mul(&t->mod, a[i], t->pow_psi[i]);
    bit_reverse(a, t->n);
    ct_levels(t, a, 2, levels);
    return OR_OK;
}

/* ntt.rs:291-316 */
int or_ntt_complete_partial(const or_ntt *t, uint32_t *a, size_t len, uint32_t already)
{
    if (len != t->n || already > t->log_n) return OR_NTT_LENGTH_MISMATCH;
    if (already >= 31) return OR_OK;
    ct_levels(t, a, 1u << (already + 1), 64);
    return OR_OK;
}

/* Batched helpers (loops over the scalar transform; layout [poly][n]). */
int or_ntt_forward_batch(const or_ntt *t, uint32_t *a, size_t count)
{
    for (size_t p = 0; p < count; p++) or_ntt_forward(t, a + p * t->n, t->n);
    return OR_OK;
}

int or_ntt_inverse_batch(const or_ntt *t, uint32_t *a, size_t count)
{
    for (size_t p = 0; p < count; p++) or_ntt_inverse(t, a + p * t->n, t->n);
    return OR_OK;
}

/* ring.rs:309-340: negacyclic product via two forwards, pointwise, inverse. */
int or_ring_mul(const or_ntt *t, const uint32_t *a, const uint32_t *b, uint32_t *out)
{
    uint32_t n = t->n;
    uint32_t *tb = (uint32_t *)malloc(sizeof(uint32_t) * n);
    memcpy(out, a, sizeof(uint32_t) * n);
    memcpy(tb, b, sizeof(uint32_t) * n);
    or_ntt_forward(t, out, n);
    or_ntt_forward(t, tb, n);
    for (uint32_t i = 0; i < n; i++) out[i] = or_mul(&t->mod, out[i], tb[i]);
    or_ntt_inverse(t, out, n);
    free(tb);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* R_q[Y]/(Y^2+1): lattice-ring/src/extension.rs:27-176; an element is      */
/* [c0 | c1], n coefficients each                                             */
/* ------------------------------------------------------------------------ */

static void or_poly_add(const or_mod32 *m, const uint32_t *a, const uint32_t *b, uint32_t *o, uint32_t n)
{
    for (uint32_t i = 0; i < n; i++) o[i] = or_add(m, a[i], b[i]);
}

static void or_poly_sub(const or_mod32 *m, const uint32_t *a, const uint32_t *b, uint32_t *o, uint32_t n)
{
    for (uint32_t i = 0; i < n; i++) o[i] = or_sub(m, a[i], b[i]);
}

/* Rq2::mul (extension.rs:115-124): p0 = a0 b0, p1 = a1 b1, pc = (a0+a1)(b0+b1);
 * c0 = p0 - p1, c1 = pc - p0 - p1. */
int or_rq2_mul(const or_ntt *t, const uint32_t *a, const uint32_t *b, uint32_t *out)
{
    const uint32_t n = t->n;
    uint32_t *w = (uint32_t *)malloc(sizeof(uint32_t) * 5 * n);
    uint32_t *p0 = w, *p1 = w + n, *pc = w + 2 * n, *sa = w + 3 * n, *sb = w + 4 * n;
    or_ring_mul(t, a, b, p0);
    or_ring_mul(t, a + n, b + n, p1);
    or_poly_add(&t->mod, a, a + n, sa, n);
    or_poly_add(&t->mod, b, b + n, sb, n);
    or_ring_mul(t, sa, sb, pc);
    or_poly_sub(&t->mod, p0, p1, out, n);
    or_poly_sub(&t->mod, pc, p0, out + n, n);
    or_poly_sub(&t->mod, out + n, p1, out + n, n);
    free(w);
    return OR_OK;
}

/* Rq2::norm (extension.rs:142-147): c0^2 + c1^2, one R_q element. */
int or_rq2_norm(const or_ntt *t, const uint32_t *a, uint32_t *out)
{
    const uint32_t n = t->n;
    uint32_t *sq = (uint32_t *)malloc(sizeof(uint32_t) * n);
    or_ring_mul(t, a, a, out);
    or_ring_mul(t, a + n, a + n, sq);
    or_poly_add(&t->mod, out, sq, out, n);
    free(sq);
    return OR_OK;
}

/* Rq2::inverse (extension.rs:151-167): conj(z) N(z)^{-1} with N(z)^{-1} taken
 * slot-wise in the NTT domain; returns 1 and writes out when every slot of
 * N(z) is non-zero, else 0 (None) and leaves out untouched. */
int or_rq2_inverse(const or_ntt *t, const uint32_t *a, uint32_t *out)
{
    const uint32_t n = t->n;
    uint32_t *w = (uint32_t *)malloc(sizeof(uint32_t) * 3 * n);
    uint32_t *nrm = w, *c1n = w + n, *res = w + 2 * n;
    or_rq2_norm(t, a, nrm);
    or_ntt_forward(t, nrm, n);
    for (uint32_t i = 0; i < n; i++) {
        if (nrm[i] == 0) {
            free(w);
            return 0;
        }
        nrm[i] = or_inv(&t->mod, nrm[i]);
    }
    or_ntt_inverse(t, nrm, n);
    for (uint32_t i = 0; i < n; i++) c1n[i] = or_neg(&t->mod, a[n + i]);
    or_ring_mul(t, a, nrm, res);
    memcpy(out, res, sizeof(uint32_t) * n);
    or_ring_mul(t, c1n, nrm, res);
    memcpy(out + n, res, sizeof(uint32_t) * n);
    free(w);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Keccak-f[1600] / SHA3 / SHAKE256: lattice-core/src/keccak.rs:9-373       */
/* ------------------------------------------------------------------------ */

static const uint64_t KRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
    0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

/* rho offsets by lane index x + 5y (FIPS 202 table). */
static const unsigned KRHO[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                                  25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};

static inline uint64_t rotl64(uint64_t x, unsigned r) { return r ? (x << r) | (x >> (64 - r)) : x; }

/* keccak.rs:52 (unrolled) == keccak.rs reference loop: theta, rho, pi, chi, iota. */
void or_keccak_f1600(uint64_t st[25])
{
    for (int round = 0; round < 24; round++) {
        uint64_t c[5], b[25];
        for (int x = 0; x < 5; x++) c[x] = st[x] ^ st[x + 5] ^ st[x + 10] ^ st[x + 15] ^ st[x + 20];
        for (int x = 0; x < 5; x++) {
            uint64_t d = c[(x + 4) % 5] ^ rotl64(c[(x + 1) % 5], 1);
            for (int y = 0; y < 5; y++) st[x + 5 * y] ^= d;
        }
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++) {
                int nx = y, ny = (2 * x + 3 * y) % 5;
                b[nx + 5 * ny] = rotl64(st[x + 5 * y], KRHO[x + 5 * y]);
            }
        for (int y = 0; y < 5; y++)
            for (int x = 0; x < 5; x++)
                st[x + 5 * y] = b[x + 5 * y] ^ ((~b[(x + 1) % 5 + 5 * y]) & b[(x + 2) % 5 + 5 * y]);
        st[0] ^= KRC[round];
    }
}


static void sponge_init(or_sponge *s, size_t rate, uint8_t suffix)
{
    memset(s, 0, sizeof(*s));
    s->rate = rate;
    s->suffix = suffix;
}

/* keccak.rs:322-331 */
static void sponge_update(or_sponge *s, const uint8_t *data, size_t len)
{
    for (size_t i = 0; i < len; i++) {
        s->st[s->pos / 8] ^= (uint64_t)data[i] << (8 * (s->pos % 8));
        s->pos++;
        if (s->pos == s->rate) {
            or_keccak_f1600(s->st);
            s->pos = 0;
        }
    }
}

/* keccak.rs:334-355 */
static void sponge_finalize(or_sponge *s, uint8_t *out, size_t out_len)
{
    s->st[s->pos / 8] ^= (uint64_t)s->suffix << (8 * (s->pos % 8));
    size_t last = s->rate - 1;
    s->st[last / 8] ^= 0x80ULL << (8 * (last % 8));
    or_keccak_f1600(s->st);
    size_t done = 0;
    while (done < out_len) {
        size_t take = s->rate < out_len - done ? s->rate : out_len - done;
        for (size_t i = 0; i < take; i++) out[done + i] = (uint8_t)(s->st[i / 8] >> (8 * (i % 8)));
        done += take;
        if (done < out_len) or_keccak_f1600(s->st);
    }
}

void or_shake256(const uint8_t *in, size_t len, uint8_t *out, size_t out_len) /* keccak.rs:369 */
{
    or_sponge s;
    sponge_init(&s, 136, 0x1f);
    sponge_update(&s, in, len);
    sponge_finalize(&s, out, out_len);
}

void or_sha3_256(const uint8_t *in, size_t len, uint8_t out[32]) /* keccak.rs:359 */
{
    or_sponge s;
    sponge_init(&s, 136, 0x06);
    sponge_update(&s, in, len);
    sponge_finalize(&s, out, 32);
}

static void put_u32le(uint8_t *p, uint32_t v)
{
    for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i));
}

static void put_u64le(uint8_t *p, uint64_t v)
{
    for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i));
}

static uint32_t get_u32le(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

static uint64_t get_u64le(const uint8_t *p)
{
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = (v << 8) | p[i];
    return v;
}

/* transcript.rs:229-237: SHAKE256("LZXX" || u32le|d| || d || u32le|m| || m). */
void or_xof(const uint8_t *domain, size_t dlen, const uint8_t *msg, size_t mlen, uint8_t *out, size_t out_len)
{
    size_t blen = 12 + dlen + mlen;
    uint8_t *buf = (uint8_t *)malloc(blen);
    memcpy(buf, "LZXX", 4);
    put_u32le(buf + 4, (uint32_t)dlen);
    memcpy(buf + 8, domain, dlen);
    put_u32le(buf + 8 + dlen, (uint32_t)mlen);
    memcpy(buf + 12 + dlen, msg, mlen);
    or_shake256(buf, blen, out, out_len);
    free(buf);
}

/* ring.rs:102-180: uniform ring element by rejection over 4-byte LE windows
 * of the XOF stream starting at offset 8. */
void or_uniform_from_seed(uint32_t q, uint32_t log_n, const uint8_t *domain, size_t dlen, const uint8_t *seed,
                          size_t slen, uint64_t index, uint32_t *coeffs)
{
    size_t n = (size_t)1 << log_n;
    uint64_t limit = ((uint64_t)UINT32_MAX + 1) - (((uint64_t)UINT32_MAX + 1) % q);
    size_t salt_len = dlen + slen + 8;
    uint8_t *salt = (uint8_t *)malloc(salt_len);
    memcpy(salt, domain, dlen);
    memcpy(salt + dlen, seed, slen);
    put_u64le(salt + dlen + slen, index);
    size_t slack = n > 16 ? n : 16;
    size_t stream_len = 8 + 4 * (n + slack);
    uint8_t *stream = (uint8_t *)malloc(stream_len);
    or_xof((const uint8_t *)"uniform", 7, salt, salt_len, stream, stream_len);
    size_t got = 0;
    for (size_t w = 0; w < n + slack && got < n; w++) {
        uint64_t cand = get_u32le(stream + 8 + 4 * w);
        if (cand < limit) coeffs[got++] = (uint32_t)(cand % q);
    }
    if (got < n) {
        /* ring.rs:155-175: counter-indexed continuation (window `counter`
         * is the last 4 bytes of an XOF of length 8 + 4*counter + 4). */
        size_t counter = (stream_len - 8) / 4;
        while (got < n) {
            size_t blen = 8 + counter * 4 + 4;
            uint8_t *bytes = (uint8_t *)malloc(blen);
            or_xof((const uint8_t *)"uniform", 7, salt, salt_len, bytes, blen);
            uint64_t cand = get_u32le(bytes + blen - 4);
            if (cand < limit) coeffs[got++] = (uint32_t)(cand % q);
            free(bytes);
            counter++;
        }
    }
    free(stream);
    free(salt);
}

/* ajtai.rs:307-348: small secret with coefficients uniform in [-bound, bound]. */
void or_sample_small_secret(uint32_t q, uint32_t log_n, size_t m, uint32_t bound, const uint8_t *seed, size_t slen,
                            uint32_t *out /* m*n */)
{
    or_mod32 mod = or_mod32_new(q, 1, 1);
    size_t n = (size_t)1 << log_n;
    uint64_t span = 2ULL * bound + 1;
    uint64_t limit = ((uint64_t)UINT32_MAX / span) * span;
    size_t blen = n * 16 > 64 ? n * 16 : 64;
    uint8_t *bytes = (uint8_t *)malloc(blen);
    uint8_t *input = (uint8_t *)malloc(slen + 16);
    memcpy(input, seed, slen);
    for (size_t i = 0; i < m; i++) {
        size_t got = 0;
        uint64_t ctr = 0;
        while (got < n) {
            put_u64le(input + slen, (uint64_t)i);
            put_u64le(input + slen + 8, ctr);
            or_xof((const uint8_t *)"ajtai-secret", 12, input, slen + 16, bytes, blen);
            for (size_t c = 0; c + 4 <= blen && got < n; c += 4) {
                uint64_t v = get_u32le(bytes + c);
                if (v < limit) out[i * n + got++] = or_reduce_i64(&mod, (int64_t)(v % span) - (int64_t)bound);
            }
            ctr++;
        }
    }
    free(bytes);
    free(input);
}

/* ------------------------------------------------------------------------ */
/* Ajtai commitment: lattice-commitment/src/ajtai.rs:87-180                  */
/* ------------------------------------------------------------------------ */

/* ajtai.rs:87-103: A[idx] = uniform_from_seed("ajtai-A", seed, idx), idx = i*m + j;
 * Ahat[idx] = forward NTT of A[idx].  Either output may be NULL. */
int or_ajtai_from_seed(uint32_t q, uint32_t two_adicity, uint32_t generator, uint32_t log_n, size_t k, size_t m,
                       const uint8_t seed[32], uint32_t *A, uint32_t *Ahat)
{
    or_ntt *t = NULL;
    int st = or_ntt_new(q, two_adicity, generator, log_n, &t);
    if (st) return st;
    size_t n = t->n;
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (size_t idx = 0; idx < k * m; idx++) {
        or_uniform_from_seed(q, log_n, (const uint8_t *)"ajtai-A", 7, seed, 32, (uint64_t)idx, tmp);
        if (A) memcpy(A + idx * n, tmp, sizeof(uint32_t) * n);
        if (Ahat) {
            or_ntt_forward(t, tmp, n);
            memcpy(Ahat + idx * n, tmp, sizeof(uint32_t) * n);
        }
    }
    free(tmp);
    or_ntt_free(t);
    return OR_OK;
}

/* ajtai.rs:118-153: cached-NTT commit.  Ahat is k*m*n (row-major), s is m*n
 * canonical coefficients, out is k*n.  All-zero s_j are skipped. */
int or_ajtai_commit(const or_ntt *t, size_t k, size_t m, const uint32_t *Ahat, const uint32_t *s, size_t s_len,
                    uint32_t *out)
{
    if (s_len != m) return OR_AJTAI_DIMENSION_MISMATCH;
    size_t n = t->n;
    const or_mod32 *q = &t->mod;
    uint32_t *s_ntt = (uint32_t *)malloc(sizeof(uint32_t) * n * (m ? m : 1));
    uint8_t *nz = (uint8_t *)malloc(m ? m : 1);
    for (size_t j = 0; j < m; j++) {
        nz[j] = 0;
        for (size_t c = 0; c < n; c++) nz[j] |= s[j * n + c] != 0;
        memcpy(s_ntt + j * n, s + j * n, sizeof(uint32_t) * n);
        if (nz[j]) or_ntt_forward(t, s_ntt + j * n, n);
    }
    for (size_t i = 0; i < k; i++) {
        uint32_t *acc = out + i * n;
        memset(acc, 0, sizeof(uint32_t) * n);
        for (size_t j = 0; j < m; j++) {
            if (!nz[j]) continue;
            const uint32_t *a = Ahat + (i * m + j) * n;
            const uint32_t *b = s_ntt + j * n;
            for (size_t c = 0; c < n; c++) acc[c] = or_add(q, acc[c], or_mul(q, a[c], b[c]));
        }
        or_ntt_inverse(t, acc, n);
    }
    free(nz);
    free(s_ntt);
    return OR_OK;
}

/* ajtai.rs:158-180: per-product ring multiplication (differential path). */
int or_ajtai_commit_naive(const or_ntt *t, size_t k, size_t m, const uint32_t *A, const uint32_t *s, size_t s_len,
                          uint32_t *out)
{
    if (s_len != m) return OR_AJTAI_DIMENSION_MISMATCH;
    size_t n = t->n;
    uint32_t *prod = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (size_t i = 0; i < k; i++) {
        uint32_t *acc = out + i * n;
        memset(acc, 0, sizeof(uint32_t) * n);
        for (size_t j = 0; j < m; j++) {
            int nzj = 0;
            for (size_t c = 0; c < n; c++) nzj |= s[j * n + c] != 0;
            if (!nzj) continue;
            or_ring_mul(t, A + (i * m + j) * n, s + j * n, prod);
            for (size_t c = 0; c < n; c++) acc[c] = or_add(&t->mod, acc[c], prod[c]);
        }
    }
    free(prod);
    return OR_OK;
}

/* Schoolbook negacyclic product (ntt.rs:390-403 shape), exact in u128. */
void or_negacyclic_schoolbook(uint32_t q, uint32_t n, const uint32_t *a, const uint32_t *b, uint32_t *out)
{
    or_mod32 m = or_mod32_new(q, 1, 1);
    for (uint32_t i = 0; i < n; i++) out[i] = 0;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < n; j++) {
            uint32_t v = or_mul(&m, a[i], b[j]);
            if (i + j < n)
                out[i + j] = or_add(&m, out[i + j], v);
            else
                out[i + j - n] = or_sub(&m, out[i + j - n], v);
        }
}

/* ------------------------------------------------------------------------ */
/* Goldilocks: lattice-core/src/field.rs:9-209                               */
/* ------------------------------------------------------------------------ */

#define GL_P 0xFFFFFFFF00000001ULL

uint64_t or_gl_from_u64(uint64_t x) { return x >= GL_P ? x - GL_P : x; } /* field.rs:28-36 */

uint64_t or_gl_from_u128(uint64_t hi, uint64_t lo) /* field.rs:42-55 */
{
    u128 acc = ((u128)hi << 64) | lo;
    for (int i = 0; i < 6; i++) {
        if ((acc >> 64) == 0) break;
        uint64_t h = (uint64_t)(acc >> 64), l = (uint64_t)acc;
        acc = (u128)h * 0xFFFFFFFFULL + (u128)l;
    }
    return or_gl_from_u64((uint64_t)acc);
}

uint64_t or_gl_add(uint64_t a, uint64_t b) /* field.rs:70-77 */
{
    uint64_t s = a + b;
    if (s < a) s += 0xFFFFFFFFULL;
    return or_gl_from_u64(s);
}

uint64_t or_gl_sub(uint64_t a, uint64_t b) /* field.rs:82-89 */
{
    uint64_t d = a - b;
    if (a < b) d -= 0xFFFFFFFFULL;
    return or_gl_from_u64(d);
}

uint64_t or_gl_neg(uint64_t a) { return a == 0 ? 0 : or_gl_from_u64(GL_P - a); } /* field.rs:93-99 */

uint64_t or_gl_mul(uint64_t a, uint64_t b) /* field.rs:103-105 */
{
    u128 p = (u128)a * (u128)b;
    return or_gl_from_u128((uint64_t)(p >> 64), (uint64_t)p);
}

uint64_t or_gl_pow(uint64_t a, uint64_t e) /* field.rs:119-130 (single word) */
{
    uint64_t acc = 1;
    for (int i = 63; i >= 0; i--) {
        acc = or_gl_mul(acc, acc);
        if ((e >> i) & 1) acc = or_gl_mul(acc, a);
    }
    return acc;
}

uint64_t or_gl_inverse(uint64_t a) { return a == 0 ? 0 : or_gl_pow(a, GL_P - 2); } /* field.rs:137-143 */

/* ------------------------------------------------------------------------ */
/* Transcript: lattice-core/src/transcript.rs:36-238                        */
/* ------------------------------------------------------------------------ */


/* transcript.rs:46-59 */
or_transcript *or_transcript_new(const uint8_t *protocol, size_t plen, uint64_t budget)
{
    or_transcript *t = (or_transcript *)calloc(1, sizeof(or_transcript));
    sponge_init(&t->sponge, 136, 0x1f);
    uint8_t head[8];
    memcpy(head, "LZX1", 4);
    put_u32le(head + 4, (uint32_t)plen);
    sponge_update(&t->sponge, head, 8);
    sponge_update(&t->sponge, protocol, plen);
    t->query_budget = budget;
    return t;
}

void or_transcript_free(or_transcript *t) { free(t); }

or_transcript *or_transcript_clone(const or_transcript *t)
{
    or_transcript *c = (or_transcript *)malloc(sizeof(or_transcript));
    *c = *t;
    return c;
}

/* transcript.rs:67-78 */
int or_transcript_append_message(or_transcript *t, const uint8_t *label, size_t llen, const uint8_t *msg, size_t mlen)
{
    if (mlen > ((size_t)1 << 26)) return OR_TRANSCRIPT_MESSAGE_TOO_LARGE;
    uint8_t b4[4];
    put_u32le(b4, (uint32_t)llen);
    sponge_update(&t->sponge, b4, 4);
    sponge_update(&t->sponge, label, llen);
    put_u32le(b4, (uint32_t)mlen);
    sponge_update(&t->sponge, b4, 4);
    sponge_update(&t->sponge, msg, mlen);
    return OR_OK;
}

/* transcript.rs:90-100 */
int or_transcript_append_field_slice(or_transcript *t, const uint8_t *label, size_t llen, const uint64_t *v, size_t n)
{
    uint8_t *bytes = (uint8_t *)malloc(n * 8 + 1);
    for (size_t i = 0; i < n; i++) put_u64le(bytes + 8 * i, v[i]);
    int st = or_transcript_append_message(t, label, llen, bytes, n * 8);
    free(bytes);
    return st;
}

/* transcript.rs:180-205: squeeze from a COPY of the raw state; stream block i
 * is available after i+1 permutations (no padding is applied). */
static void transcript_squeeze(const or_transcript *t, size_t *offset, uint8_t *out, size_t n)
{
    const size_t RATE = 136;
    uint64_t st[25];
    memcpy(st, t->sponge.st, sizeof(st));
    size_t block_idx = *offset / RATE;
    for (size_t i = 0; i <= block_idx; i++) or_keccak_f1600(st);
    size_t pos = *offset % RATE, done = 0;
    while (done < n) {
        size_t take = RATE - pos < n - done ? RATE - pos : n - done;
        for (size_t i = 0; i < take; i++) out[done + i] = (uint8_t)(st[(pos + i) / 8] >> (8 * ((pos + i) % 8)));
        done += take;
        pos += take;
        if (pos == RATE) {
            or_keccak_f1600(st);
            pos = 0;
        }
    }
    *offset += n;
}

/* transcript.rs:109-148 */
int or_transcript_challenge_fields(or_transcript *t, const uint8_t *label, size_t llen, size_t n, uint64_t *out)
{
    t->query_count++;
    if (t->query_count > t->query_budget) return OR_TRANSCRIPT_REJECTION_BUDGET;
    uint8_t b4[4];
    sponge_update(&t->sponge, (const uint8_t *)"CHAL", 4);
    put_u32le(b4, (uint32_t)llen);
    sponge_update(&t->sponge, b4, 4);
    sponge_update(&t->sponge, label, llen);
    put_u32le(b4, (uint32_t)n);
    sponge_update(&t->sponge, b4, 4);
    size_t got = 0, offset = 0;
    while (got < n) {
        size_t need = (n - got) * 16;
        if (need < 32) need = 32;
        uint8_t *bytes = (uint8_t *)malloc(need);
        transcript_squeeze(t, &offset, bytes, need);
        for (size_t c = 0; c < need && got < n; c += 16) {
            uint8_t chunk[16] = {0};
            size_t cl = need - c < 16 ? need - c : 16;
            memcpy(chunk, bytes + c, cl);
            uint64_t cand = get_u64le(chunk);
            if (cand < GL_P) out[got++] = cand;
        }
        free(bytes);
    }
    return OR_OK;
}

int or_transcript_challenge_field(or_transcript *t, const uint8_t *label, size_t llen, uint64_t *out)
{
    return or_transcript_challenge_fields(t, label, llen, 1, out);
}

/* ------------------------------------------------------------------------ */
/* Multilinear pieces: lattice-core/src/mle.rs, field_simd.rs                */
/* ------------------------------------------------------------------------ */

/* field_simd.rs:391-407: eq table, variable 0 = most significant index bit. */
void or_eq_table(const uint64_t *point, size_t m, uint64_t *out /* 2^m */)
{
    out[0] = 1;
    size_t cur = 1;
    for (size_t v = m; v-- > 0;) {
        uint64_t p = point[v];
        uint64_t omp = or_gl_sub(1, p);
        for (size_t i = 0; i < cur; i++) {
            out[cur + i] = or_gl_mul(out[i], p);
            out[i] = or_gl_mul(out[i], omp);
        }
        cur *= 2;
    }
}

/* field_simd.rs:297-313 / mle.rs:152-172: bind variable 0 in place. */
void or_bind_first_half(uint64_t *evals, size_t len, uint64_t r)
{
    size_t half = len / 2;
    for (size_t i = 0; i < half; i++) {
        uint64_t a = evals[i], b = evals[i + half];
        evals[i] = or_gl_add(a, or_gl_mul(or_gl_sub(b, a), r));
    }
}

/* mle.rs:131-148 */
uint64_t or_mle_evaluate(const uint64_t *evals, size_t num_vars, const uint64_t *point)
{
    size_t len = (size_t)1 << num_vars;
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * len);
    memcpy(cur, evals, sizeof(uint64_t) * len);
    for (size_t v = 0; v < num_vars; v++) {
        or_bind_first_half(cur, len, point[v]);
        len /= 2;
    }
    uint64_t r = cur[0];
    free(cur);
    return r;
}

/* sumcheck.rs:286-306: Lagrange over nodes 0..n-1. */
uint64_t or_interpolate_at(const uint64_t *evals, size_t n, uint64_t r)
{
    uint64_t acc = 0;
    for (size_t i = 0; i < n; i++) {
        uint64_t weight = 1;
        for (size_t j = 0; j < n; j++) {
            if (i == j) continue;
            uint64_t num = or_gl_sub(r, (uint64_t)j);
            uint64_t den = or_gl_sub((uint64_t)i, (uint64_t)j);
            weight = or_gl_mul(weight, or_gl_mul(num, or_gl_inverse(den)));
        }
        acc = or_gl_add(acc, or_gl_mul(evals[i], weight));
    }
    return acc;
}

/* ------------------------------------------------------------------------ */
/* Sum-check: lattice-sumcheck/src/sumcheck.rs:62-371                        */
/* A virtual polynomial is nfactors tables of 2^num_vars canonical u64 plus  */
/* nterms terms (coeff, factor ids) in CSR form (term_off[nterms+1]).        */
/* ------------------------------------------------------------------------ */

/* virtual_poly.rs:98-104 */
static size_t vp_max_degree(size_t nterms, const uint32_t *term_off)
{
    size_t d = 0;
    for (size_t t = 0; t < nterms; t++) {
        size_t len = term_off[t + 1] - term_off[t];
        if (len > d) d = len;
    }
    return d;
}

size_t or_vp_max_degree(size_t nterms, const uint32_t *term_off) { return vp_max_degree(nterms, term_off); }

/* virtual_poly.rs:107-121: sum of P over the hypercube. */
uint64_t or_vp_sum_over_hypercube(size_t num_vars, size_t nfactors, const uint64_t *const *factors, size_t nterms,
                                  const uint64_t *coeffs, const uint32_t *term_off, const uint32_t *term_ids)
{
    (void)nfactors;
    size_t len = (size_t)1 << num_vars;
    uint64_t acc = 0;
    for (size_t t = 0; t < nterms; t++)
        for (size_t p = 0; p < len; p++) {
            uint64_t prod = coeffs[t];
            for (uint32_t k = term_off[t]; k < term_off[t + 1]; k++) prod = or_gl_mul(prod, factors[term_ids[k]][p]);
            acc = or_gl_add(acc, prod);
        }
    return acc;
}

/* sumcheck.rs:172-226: g(t), t = 0..d, from the current bound tables.
 * bind_half_slices (field_simd.rs:273-291) gives lo + (hi - lo)*t. */
void or_sumcheck_round_evals(size_t rem_vars, size_t nfactors, uint64_t *const *bound, size_t nterms,
                             const uint64_t *coeffs, const uint32_t *term_off, const uint32_t *term_ids, size_t d,
                             uint64_t *evals /* d+1 */)
{
    size_t points = (size_t)1 << (rem_vars - 1);
    uint64_t *vals = (uint64_t *)malloc(sizeof(uint64_t) * (nfactors ? nfactors : 1));
    for (size_t t = 0; t <= d; t++) {
        uint64_t acc = 0;
        for (size_t p = 0; p < points; p++) {
            for (size_t f = 0; f < nfactors; f++) {
                uint64_t lo = bound[f][p], hi = bound[f][p + points];
                vals[f] = t == 0 ? lo : t == 1 ? hi : or_gl_add(lo, or_gl_mul(or_gl_sub(hi, lo), (uint64_t)t));
            }
            for (size_t tt = 0; tt < nterms; tt++) {
                uint64_t prod = coeffs[tt];
                for (uint32_t k = term_off[tt]; k < term_off[tt + 1]; k++) prod = or_gl_mul(prod, vals[term_ids[k]]);
                acc = or_gl_add(acc, prod);
            }
        }
        evals[t] = acc;
    }
    free(vals);
}

/* sumcheck.rs:62-162.  rounds_out is num_vars*(d+1) (d = max degree, or 1 for
 * the zero polynomial); challenges_out num_vars; factor_claims_out nfactors.
 * *d_out receives the round width minus one. */
int or_sumcheck_prove(size_t num_vars, size_t nfactors, const uint64_t *const *factors, size_t nterms,
                      const uint64_t *coeffs, const uint32_t *term_off, const uint32_t *term_ids, uint64_t claim,
                      or_transcript *tr, uint64_t *rounds_out, uint64_t *challenges_out, uint64_t *final_claim_out,
                      uint64_t *factor_claims_out, size_t *d_out)
{
    static const uint8_t LROUND[] = "sumcheck-round";
    static const uint8_t LCHAL[] = "sumcheck-challenge";
    int st;
    if (nterms == 0) {
        if (claim != 0) return OR_SUMCHECK_CLAIM_MISMATCH;
        *d_out = 1;
        for (size_t r = 0; r < num_vars; r++) {
            uint64_t ev[2] = {0, 0};
            if ((st = or_transcript_append_field_slice(tr, LROUND, 14, ev, 2))) return st;
            if ((st = or_transcript_challenge_field(tr, LCHAL, 18, &challenges_out[r]))) return st;
            rounds_out[2 * r] = rounds_out[2 * r + 1] = 0;
        }
        *final_claim_out = 0;
        return OR_OK;
    }
    size_t d = vp_max_degree(nterms, term_off);
    *d_out = d;
    size_t len = (size_t)1 << num_vars;
    uint64_t **bound = (uint64_t **)malloc(sizeof(uint64_t *) * nfactors);
    for (size_t f = 0; f < nfactors; f++) {
        bound[f] = (uint64_t *)malloc(sizeof(uint64_t) * len);
        memcpy(bound[f], factors[f], sizeof(uint64_t) * len);
    }
    uint64_t current = claim;
    int status = OR_OK;
    for (size_t r = 0; r < num_vars; r++) {
        uint64_t *ev = rounds_out + r * (d + 1);
        or_sumcheck_round_evals(num_vars - r, nfactors, bound, nterms, coeffs, term_off, term_ids, d, ev);
        if ((status = or_transcript_append_field_slice(tr, LROUND, 14, ev, d + 1))) goto done;
        if ((status = or_transcript_challenge_field(tr, LCHAL, 18, &challenges_out[r]))) goto done;
        if (or_gl_add(ev[0], ev[1]) != current) {
            status = OR_SUMCHECK_CLAIM_MISMATCH;
            goto done;
        }
        current = or_interpolate_at(ev, d + 1, challenges_out[r]);
        for (size_t f = 0; f < nfactors; f++) or_bind_first_half(bound[f], len, challenges_out[r]);
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
    for (size_t f = 0; f < nfactors; f++) free(bound[f]);
    free(bound);
    return status;
}

/* sumcheck.rs:324-371.  rounds is num_vars rows of width round_len.
 * has_expected selects the Option<expected_final>. */
int or_sumcheck_verify(const uint64_t *rounds, size_t num_rounds, size_t round_len, size_t num_vars,
                       size_t max_degree, uint64_t claim, or_transcript *tr, int has_expected,
                       uint64_t expected_final, uint64_t *point_out, uint64_t *final_claim_out)
{
    static const uint8_t LROUND[] = "sumcheck-round";
    static const uint8_t LCHAL[] = "sumcheck-challenge";
    if (num_rounds != num_vars) return OR_SUMCHECK_BAD_ROUND_SHAPE;
    uint64_t current = claim;
    int st;
    for (size_t r = 0; r < num_rounds; r++) {
        const uint64_t *ev = rounds + r * round_len;
        /* The reference indexes evals[1]; a one-entry round cannot pass. */
        if (round_len == 0 || round_len > max_degree + 1 || round_len < 2) return OR_SUMCHECK_BAD_ROUND_SHAPE;
        if ((st = or_transcript_append_field_slice(tr, LROUND, 14, ev, round_len))) return st;
        uint64_t c;
        if ((st = or_transcript_challenge_field(tr, LCHAL, 18, &c))) return st;
        if (or_gl_add(ev[0], ev[1]) != current) return OR_SUMCHECK_ROUND_CHECK_FAILED;
        current = or_interpolate_at(ev, round_len, c);
        point_out[r] = c;
    }
    if (has_expected && current != expected_final) return OR_SUMCHECK_FINAL_CHECK_FAILED;
    *final_claim_out = current;
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Wide-modulus commitments (checker for the uneven commitment batch)        */
/* ------------------------------------------------------------------------ */
/*
 * The batch commits t_i = sum_j A_ij * s_j in Z_q[X]/(X^n+1) for moduli up
 * to 2^50, beyond the reference's Modulus32.  The semantics are those of
 * AjtaiPublicKey::commit (lattice-commitment/src/ajtai.rs:118-153) with the
 * ring product defined by negacyclic convolution (the schoolbook shape of
 * lattice-ring/src/ntt.rs:390-403), so the checker is that definition,
 * computed exactly in 128-bit integers.  A is [k][lda][n] (u32 or u64
 * residues, any representative), the witness is m elements of n balanced
 * int8 digits (ring.rs:232-243 from_signed), the output [k][n] canonical u64.
 */

static uint64_t or_mulmod64(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }

uint64_t or_powmod64(uint64_t b, uint64_t e, uint64_t q)
{
    uint64_t acc = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) acc = or_mulmod64(acc, b, q);
        b = or_mulmod64(b, b, q);
        e >>= 1;
    }
    return acc;
}

static uint64_t or_elem(const void *A, int elem_bytes, size_t idx)
{
    return elem_bytes == 4 ? ((const uint32_t *)A)[idx] : ((const uint64_t *)A)[idx];
}

int or_commit_wide_schoolbook(uint64_t q, uint32_t n, size_t k, size_t m, size_t lda, const void *A,
                              int elem_bytes, const int8_t *digits, uint64_t *out)
{
    if (q < 2 || (q >> 62) || n == 0 || lda < m || (elem_bytes != 4 && elem_bytes != 8)) return OR_BAD_ARGUMENT;
    /* |A mod q| < 2^62, |digit| <= 128: each product < 2^69, n*m terms fit i128 */
    __int128 *acc = (__int128 *)malloc(sizeof(__int128) * n);
    uint64_t *a = (uint64_t *)malloc(sizeof(uint64_t) * n);
    for (size_t i = 0; i < k; i++) {
        memset(acc, 0, sizeof(__int128) * n);
        for (size_t j = 0; j < m; j++) {
            const int8_t *s = digits + j * n;
            int nz = 0;
            for (uint32_t c = 0; c < n; c++) nz |= s[c] != 0;
            if (!nz) continue;
            for (uint32_t c = 0; c < n; c++) a[c] = or_elem(A, elem_bytes, (i * lda + j) * n + c) % q;
            for (uint32_t c = 0; c < n; c++) {
                if (!s[c]) continue;
                const __int128 sc = s[c];
                /* a(X) * s_c X^c: coefficient r gets a[r-c], wrapping with a sign flip */
                for (uint32_t r = 0; r < n; r++) {
                    if (r >= c)
                        acc[r] += (__int128)a[r - c] * sc;
                    else
                        acc[r] -= (__int128)a[r + n - c] * sc;
                }
            }
        }
        for (uint32_t r = 0; r < n; r++) {
            __int128 v = acc[r] % (__int128)q;
            if (v < 0) v += q;
            out[i * n + r] = (uint64_t)v;
        }
    }
    free(acc);
    free(a);
    return OR_OK;
}

/* Horner evaluation of an n-coefficient polynomial at x, mod q. */
static uint64_t or_eval_poly(uint64_t q, uint32_t n, const void *p, int elem_bytes, size_t off, uint64_t x)
{
    uint64_t v = 0;
    for (uint32_t c = n; c-- > 0;) v = (uint64_t)(((u128)v * x + or_elem(p, elem_bytes, off + c) % q) % q);
    return v;
}

/* Evaluation check of t = A s at a root z of X^n + 1 (z^n = -1 mod q, so
 * evaluation is a ring homomorphism Z_q[X]/(X^n+1) -> Z_q): counts the rows
 * i with t_i(z) != sum_j A_ij(z) s_j(z).  A wrong row passes with
 * probability <= (n-1)/q per random root (Schwartz-Zippel).  t is [k][n]
 * with elem_bytes-wide elements. */
size_t or_commit_wide_check_at(uint64_t q, uint32_t n, size_t k, size_t m, size_t lda, const void *A,
                               int elem_bytes, const int8_t *digits, const void *t, int t_bytes, uint64_t z)
{
    uint64_t *sz = (uint64_t *)malloc(sizeof(uint64_t) * (m ? m : 1));
    for (size_t j = 0; j < m; j++) {
        uint64_t v = 0;
        for (uint32_t c = n; c-- > 0;) {
            int64_t d = digits[j * n + c];
            uint64_t dr = d < 0 ? q - (uint64_t)(-d) % q : (uint64_t)d % q;
            v = (uint64_t)(((u128)v * z + dr) % q);
        }
        sz[j] = v;
    }
    size_t bad = 0;
    for (size_t i = 0; i < k; i++) {
        uint64_t lhs = or_eval_poly(q, n, t, t_bytes, i * n, z);
        uint64_t rhs = 0;
        for (size_t j = 0; j < m; j++) {
            if (!sz[j]) continue;
            rhs = (rhs + or_mulmod64(or_eval_poly(q, n, A, elem_bytes, (i * lda + j) * n, z), sz[j], q)) % q;
        }
        bad += lhs != rhs;
    }
    free(sz);
    return bad;
}

/* Checker of many commitments against one key, by evaluation at a root z of
 * X^n + 1 (the map a(X) -> a(z) is a ring homomorphism Z_q[X]/(X^n+1) -> Z_q,
 * so t_i = sum_j A_ij s_j implies t_i(z) = sum_j A_ij(z) s_j(z)).  The key's
 * evaluations are computed once; each witness then costs m*n + k*m steps.
 * Multiplication by the fixed z uses Shoup's precomputed quotient
 * floor(z * 2^64 / q) (q < 2^62). */
static inline uint64_t or_shoup_mul(uint64_t v, uint64_t z, uint64_t zs, uint64_t q)
{
    uint64_t hi = (uint64_t)(((u128)v * zs) >> 64);
    uint64_t r = v * z - hi * q;
    return r >= q ? r - q : r;
}

/* Az[i * m + j] = A_ij(z) mod q for i < k, j < m of a [k][lda][n] key. */
int or_eval_key_at(uint64_t q, uint32_t n, size_t k, size_t m, size_t lda, const void *A, int elem_bytes, uint64_t z,
                   uint64_t *Az)
{
    if (q < 2 || (q >> 62) || z >= q || lda < m || (elem_bytes != 4 && elem_bytes != 8)) return OR_BAD_ARGUMENT;
    const uint64_t zs = (uint64_t)(((u128)z << 64) / q);
    for (size_t i = 0; i < k; i++)
        for (size_t j = 0; j < m; j++) {
            const size_t off = (i * lda + j) * n;
            uint64_t v = 0;
            for (uint32_t c = n; c-- > 0;) {
                uint64_t a = or_elem(A, elem_bytes, off + c);
                if (a >= q) a %= q;
                v = or_shoup_mul(v, z, zs, q) + a;
                if (v >= q) v -= q;
            }
            Az[i * m + j] = v;
        }
    return OR_OK;
}

/* Number of rows i with t_i(z) != sum_{j<m} Az[i * ldz + j] * s_j(z), for one
 * witness of int8 digits [m][n] and its commitment t [k][n] (u32 or u64). */
size_t or_commit_check_at_evals(uint64_t q, uint32_t n, size_t k, size_t m, size_t ldz, const uint64_t *Az,
                                const int8_t *digits, const void *t, int t_bytes, uint64_t z)
{
    const uint64_t zs = (uint64_t)(((u128)z << 64) / q);
    uint64_t *sz = (uint64_t *)malloc(sizeof(uint64_t) * (m ? m : 1));
    for (size_t j = 0; j < m; j++) {
        uint64_t v = 0;
        for (uint32_t c = n; c-- > 0;) {
            const int64_t d = digits[j * n + c];
            const uint64_t dr = d < 0 ? q - (uint64_t)(-d) % q : (uint64_t)d % q;
            v = or_shoup_mul(v, z, zs, q) + dr;
            if (v >= q) v -= q;
        }
        sz[j] = v;
    }
    size_t bad = 0;
    for (size_t i = 0; i < k; i++) {
        uint64_t lhs = 0;
        for (uint32_t c = n; c-- > 0;) {
            lhs = or_shoup_mul(lhs, z, zs, q) + or_elem(t, t_bytes, i * n + c) % q;
            if (lhs >= q) lhs -= q;
        }
        u128 acc = 0;
        for (size_t j = 0; j < m; j++) {
            acc += (u128)Az[i * ldz + j] * sz[j]; /* < 2^124 per term */
            if ((j & 7) == 7) acc %= q;
        }
        bad += (uint64_t)(acc % q) != lhs;
    }
    free(sz);
    return bad;
}

/* ---- the ~2^50 quadratic-slot ring (crates/lattice-ring/src/modulus50.rs) ----
 * Status codes follow Mod50Error (:417-422): 40 LogNOutOfRange, 41
 * AdicityTooSmall, 42 GeneratorInvalid, 43 LengthMismatch. */

static uint64_t q50_pow(uint64_t b, uint64_t e, uint64_t q) /* modulus50.rs:168-180 */
{
    uint64_t acc = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) acc = (uint64_t)((u128)acc * b % q);
        b = (uint64_t)((u128)b * b % q);
        e >>= 1;
    }
    return acc;
}

static size_t q50_rev(size_t x, uint32_t bits) /* modulus50.rs:229-236 */
{
    size_t r = 0;
    for (uint32_t b = 0; b < bits; b++) r |= ((x >> b) & 1) << (bits - 1 - b);
    return r;
}

/* QuadSlotTables::new (modulus50.rs:241-300): zeta, the per-level twiddles
 * flattened (level d, block i at 2^d - 1 + i; 2^(log_n - 1) - 1 entries),
 * their inverses, and the n/2 leaf constants. */
int or_quad_tables(uint64_t q, uint32_t two_adicity, uint64_t generator, uint32_t log_n, uint64_t *zeta_out,
                   uint64_t *tw, uint64_t *tw_inv, uint64_t *gamma)
{
    if (log_n == 0 || log_n > 7) return 40;
    if (two_adicity < 7) return 41;
    const uint64_t zeta = q50_pow(generator, (q - 1) >> 7, q); /* root_of_unity(7) */
    if (q50_pow(zeta, 64, q) != q - 1) return 42;
    if (zeta_out) *zeta_out = zeta;
    for (uint32_t d = 0; d + 1 < log_n; d++)
        for (size_t i = 0; i < ((size_t)1 << d); i++) {
            const uint64_t e = (((uint64_t)(2 * q50_rev(i, d) + 1)) << (5 - d)) % 128;
            const uint64_t w = q50_pow(zeta, e, q);
            if (tw) tw[((size_t)1 << d) - 1 + i] = w;
            if (tw_inv) tw_inv[((size_t)1 << d) - 1 + i] = q50_pow(w, q - 2, q);
        }
    for (size_t p = 0; p < ((size_t)1 << (log_n - 1)); p++) {
        const uint64_t e = (((uint64_t)(2 * q50_rev(p, log_n - 1) + 1)) << (7 - log_n)) % 128;
        if (gamma) gamma[p] = q50_pow(zeta, e, q);
    }
    return 0;
}

static uint64_t q50_add(uint64_t a, uint64_t b, uint64_t q) { return a + b >= q ? a + b - q : a + b; }
static uint64_t q50_sub(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + (q - b); }
static uint64_t q50_mul(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }

/* forward (modulus50.rs:317-342): incomplete DIF, lo' = lo + w hi, hi' = lo - w hi */
int or_quad_forward(uint64_t q, uint32_t two_adicity, uint64_t generator, uint32_t log_n, uint64_t *a)
{
    uint64_t tw[64];
    const int st = or_quad_tables(q, two_adicity, generator, log_n, NULL, tw, NULL, NULL);
    if (st) return st;
    const size_t n = (size_t)1 << log_n;
    for (uint32_t d = 0; d + 1 < log_n; d++) {
        const size_t span = n >> d, half = span / 2;
        for (size_t i = 0; i < ((size_t)1 << d); i++) {
            const uint64_t w = tw[((size_t)1 << d) - 1 + i];
            for (size_t j = 0; j < half; j++) {
                const uint64_t lo = a[i * span + j], t = q50_mul(w, a[i * span + j + half], q);
                a[i * span + j] = q50_add(lo, t, q);
                a[i * span + j + half] = q50_sub(lo, t, q);
            }
        }
    }
    return 0;
}

/* inverse (modulus50.rs:346-375): lo = (lo' + hi') / 2, hi = (lo' - hi') / (2 w), levels reversed */
int or_quad_inverse(uint64_t q, uint32_t two_adicity, uint64_t generator, uint32_t log_n, uint64_t *a)
{
    uint64_t twi[64];
    const int st = or_quad_tables(q, two_adicity, generator, log_n, NULL, NULL, twi, NULL);
    if (st) return st;
    const size_t n = (size_t)1 << log_n;
    const uint64_t inv2 = q50_pow(2, q - 2, q);
    for (uint32_t d = log_n - 1; d-- > 0;) {
        const size_t span = n >> d, half = span / 2;
        for (size_t i = 0; i < ((size_t)1 << d); i++) {
            const uint64_t wi = twi[((size_t)1 << d) - 1 + i];
            for (size_t j = 0; j < half; j++) {
                const uint64_t lo = a[i * span + j], hi = a[i * span + j + half];
                a[i * span + j] = q50_mul(q50_add(lo, hi, q), inv2, q);
                a[i * span + j + half] = q50_mul(q50_mul(q50_sub(lo, hi, q), inv2, q), wi, q);
            }
        }
    }
    return 0;
}

/* slotwise_mul (modulus50.rs:380-402): (u0 + u1 X)(v0 + v1 X) mod (X^2 - gamma_p), Karatsuba */
int or_quad_slotwise_mul(uint64_t q, uint32_t two_adicity, uint64_t generator, uint32_t log_n, const uint64_t *a,
                         const uint64_t *b, uint64_t *out)
{
    uint64_t gamma[64];
    const int st = or_quad_tables(q, two_adicity, generator, log_n, NULL, NULL, NULL, gamma);
    if (st) return st;
    for (size_t p = 0; p < ((size_t)1 << (log_n - 1)); p++) {
        const uint64_t u0 = a[2 * p], u1 = a[2 * p + 1], v0 = b[2 * p], v1 = b[2 * p + 1];
        const uint64_t p0 = q50_mul(u0, v0, q), p1 = q50_mul(u1, v1, q);
        const uint64_t cross = q50_mul(q50_add(u0, u1, q), q50_add(v0, v1, q), q);
        out[2 * p] = q50_add(p0, q50_mul(gamma[p], p1, q), q);
        out[2 * p + 1] = q50_sub(cross, q50_add(p0, p1, q), q);
    }
    return 0;
}

/* mul (modulus50.rs:406-421): forward both, slotwise, inverse */
int or_quad_mul(uint64_t q, uint32_t two_adicity, uint64_t generator, uint32_t log_n, const uint64_t *a,
                const uint64_t *b, uint64_t *out)
{
    const size_t n = (size_t)1 << (log_n <= 7 ? log_n : 0);
    uint64_t ta[128], tb[128];
    int st = or_quad_tables(q, two_adicity, generator, log_n, NULL, NULL, NULL, NULL);
    if (st) return st;
    memcpy(ta, a, n * 8);
    memcpy(tb, b, n * 8);
    or_quad_forward(q, two_adicity, generator, log_n, ta);
    or_quad_forward(q, two_adicity, generator, log_n, tb);
    or_quad_slotwise_mul(q, two_adicity, generator, log_n, ta, tb, out);
    return or_quad_inverse(q, two_adicity, generator, log_n, out);
}

/* ---- witness packing (packing.rs:30-200, sparse.rs:30-54/118-144,
 * decomposition.rs:47-100) -------------------------------------------------
 * Status codes: PackingError 50 LengthMismatch, 51 LimbOutOfRange, 52
 * ModulusTooLarge, 53 ValueOutOfRange, 54 NotCoprime; DecompositionError 55
 * DigitCountMismatch, 56 DigitOutOfRange, 57 ValueOutOfRange, 58 ValueTooLarge. */

/* pack_field_elements (packing.rs:30-52): three 22-bit limbs per value, from_signed per element */
size_t or_pack_field_elements(uint32_t q, uint32_t n, const uint64_t *values, size_t count, uint32_t *out)
{
    const size_t limbs = 3 * count, elems = limbs ? (limbs + n - 1) / n : 1;
    for (size_t g = 0; g < elems * n; g++) {
        uint64_t l = 0;
        if (g < limbs) {
            const uint64_t raw = values[g / 3];
            l = g % 3 == 0 ? raw & 0x3FFFFF : g % 3 == 1 ? (raw >> 22) & 0x3FFFFF : raw >> 44;
        }
        out[g] = (uint32_t)(l % q);
    }
    return elems;
}

/* unpack_field_elements (packing.rs:57-91); returns the status, *got = values written */
int or_unpack_field_elements(uint32_t q, uint32_t n, const uint32_t *packed, size_t elems, uint64_t *out,
                             size_t *got)
{
    const int64_t half = q / 2, bound = (int64_t)1 << 22;
    const size_t total = elems * n;
    *got = 0;
    for (size_t g = 0; g < total; g++) {
        const int64_t c = packed[g], b = c <= half ? c : c - (int64_t)q;
        if (b <= -bound || b >= bound) return 51;
    }
    for (size_t i = 0; 3 * i + 3 <= total; i++) {
        int64_t b[3];
        for (int j = 0; j < 3; j++) {
            const int64_t c = packed[3 * i + j];
            b[j] = c <= half ? c : c - (int64_t)q;
        }
        const uint64_t raw = ((uint64_t)b[0] & 0x3FFFFF) | (((uint64_t)b[1] & 0x3FFFFF) << 22) | ((uint64_t)b[2] << 44);
        out[i] = raw >= GL_P ? raw - GL_P : raw; /* Goldilocks::from_u64 */
        *got = i + 1;
    }
    return 0;
}

/* pack_bits (sparse.rs:30-54): 31 bits per coefficient */
size_t or_pack_bits(uint32_t n, const uint8_t *bits, size_t nbits, uint32_t *out)
{
    size_t elems = (nbits + 31 * (size_t)n - 1) / (31 * (size_t)n);
    const size_t real = elems;
    if (elems == 0) elems = 1;
    for (size_t g = 0; g < elems * n; g++) {
        uint32_t v = 0;
        for (size_t b = 0; b < 31; b++) {
            const size_t idx = g * 31 + b;
            if (real && idx < nbits && bits[idx]) v |= 1u << b;
        }
        out[g] = v;
    }
    return elems;
}

/* unpack_bits (sparse.rs:57-71) */
void or_unpack_bits(uint32_t n, const uint32_t *elems, size_t nelems, uint8_t *bits, size_t nbits)
{
    for (size_t idx = 0; idx < nbits; idx++) {
        const size_t e = idx / (31 * (size_t)n), within = idx % (31 * (size_t)n), c = within / 31, b = within % 31;
        bits[idx] = e < nelems ? (elems[e * n + c] >> b) & 1 : 0;
    }
}

/* GadgetDecomposition::decompose (decomposition.rs:47-70) */
int or_gadget_decompose(uint64_t base, size_t num_digits, uint64_t x, int64_t *digits)
{
    u128 rem = x;
    const int64_t half = (int64_t)(base / 2);
    for (size_t i = 0; i < num_digits; i++) {
        const int64_t digit = (int64_t)(rem % base);
        const int64_t d = digit > half ? digit - (int64_t)base : digit;
        rem = (u128)(((__int128)rem - d) / (__int128)base);
        digits[i] = d;
    }
    return rem != 0 ? 58 : 0;
}

/* GadgetDecomposition::recompose (decomposition.rs:73-93) */
int or_gadget_recompose(uint64_t base, size_t num_digits, const int64_t *digits, size_t ndigits, uint64_t *x)
{
    if (ndigits != num_digits) return 55;
    __int128 acc = 0, power = 1;
    const int64_t half = (int64_t)(base / 2);
    for (size_t i = 0; i < ndigits; i++) {
        if (digits[i] < -half || digits[i] > half) return 56;
        acc += (__int128)digits[i] * power;
        power *= (__int128)base;
    }
    if (acc < 0 || acc > (__int128)UINT64_MAX) return 57;
    *x = (uint64_t)acc;
    return 0;
}

/* crt_pack (packing.rs:95-155): one combined residue in coefficient 0 per group */
int or_crt_pack(uint32_t q, uint32_t n, const uint64_t *values, size_t count, const uint64_t *moduli, size_t k,
                uint32_t *out, size_t *elems_out)
{
    if (k == 0) return 50;
    uint64_t product = 1;
    for (size_t i = 0; i < k; i++) {
        const uint64_t m = moduli[i];
        if (m == 0 || m >= q) return 52;
        if ((u128)product * m >= q) return 52;
        product *= m;
    }
    const size_t elems = count ? (count + k - 1) / k : 1;
    for (size_t e = 0; e < elems; e++) {
        uint64_t combined = 0, prod = 1;
        for (size_t slot = 0; slot < (k < n ? k : n); slot++) {
            const size_t idx = e * k + slot;
            if (idx >= count) break;
            const uint64_t m = moduli[slot], v = values[idx];
            if (v >= m) return 53;
            /* inverse of prod mod m by the extended Euclid algorithm */
            int64_t r0 = (int64_t)(prod % m), r1 = (int64_t)m, s0 = 1, s1 = 0;
            while (r1) {
                const int64_t qq = r0 / r1, t = r0 - qq * r1, ts = s0 - qq * s1;
                r0 = r1, r1 = t, s0 = s1, s1 = ts;
            }
            if (m <= 1 || r0 != 1) return 54;
            const uint64_t inv = (uint64_t)(((s0 % (int64_t)m) + (int64_t)m) % (int64_t)m);
            const uint64_t diff = (v + m - combined % m) % m, t = (uint64_t)((u128)diff * inv % m);
            combined += prod * t;
            prod *= m;
        }
        for (uint32_t c = 0; c < n; c++) out[e * n + c] = 0;
        out[e * n] = (uint32_t)(combined % q);
    }
    *elems_out = elems;
    return 0;
}

/* crt_unpack (packing.rs:158-185) */
int or_crt_unpack(uint32_t n, const uint32_t *packed, size_t elems, const uint64_t *moduli, size_t k, size_t expected,
                  uint64_t *out)
{
    size_t got = 0;
    for (size_t e = 0; e < elems && got < expected; e++)
        for (size_t i = 0; i < k && got < expected; i++) out[got++] = packed[e * n] % moduli[i];
    return got != expected ? 50 : 0;
}

/* ------------------------------------------------------------------------ */
/* Short ring-element challenges: lattice-core/src/short_challenge.rs        */
/* ------------------------------------------------------------------------ */

enum { OR_SC_FW = 0, OR_SC_BIASED = 1, OR_SC_SMALLSET = 2, OR_SC_FWSET = 3 };
enum { OR_SC_OK = 0, OR_SC_REJECTION = 60, OR_SC_INVALID = 61, OR_SC_GAMMA_CAP = 62 };

/* ceil_sqrt / ceil_sqrt_u128 (short_challenge.rs:281-340): min(ceil(sqrt(x)), 2^64 - 1),
 * by bisection on exact 128-bit squares. */
uint64_t or_ceil_sqrt_u128(uint64_t lo, uint64_t hi)
{
    const unsigned __int128 x = ((unsigned __int128)hi << 64) | lo;
    if (x == 0) return 0;
    /* smallest g with g^2 >= x, searched in [1, 2^64] */
    unsigned __int128 a = 1, b = (unsigned __int128)1 << 64;
    while (a < b) {
        const unsigned __int128 m = a + (b - a) / 2;
        const int ge = m >= ((unsigned __int128)1 << 64) ? 1 : m * m >= x;
        if (ge)
            b = m;
        else
            a = m + 1;
    }
    return a >= ((unsigned __int128)1 << 64) ? UINT64_MAX : (uint64_t)a;
}

/* stream (short_challenge.rs:343-352): SHAKE256("SHORTCHAL" || u32le|label| || label
 * || u32le|seed| || seed || u64le attempt). */
static void or_sc_stream(const char *label, const uint8_t *seed, size_t slen, uint8_t *out, size_t bytes)
{
    const size_t llen = strlen(label), blen = 9 + 4 + llen + 4 + slen + 8;
    uint8_t *buf = (uint8_t *)malloc(blen);
    memcpy(buf, "SHORTCHAL", 9);
    put_u32le(buf + 9, (uint32_t)llen);
    memcpy(buf + 13, label, llen);
    put_u32le(buf + 13 + llen, (uint32_t)slen);
    if (slen) memcpy(buf + 17 + llen, seed, slen);
    put_u64le(buf + 17 + llen + slen, 0);
    or_shake256(buf, blen, out, bytes);
    free(buf);
}

typedef struct {
    const uint8_t *s;
    size_t len, pos;
} or_cursor;

/* UniformCursor::next_index (short_challenge.rs:361-388); returns 0 and sets *ok = 0 when the stream runs dry. */
static uint64_t or_next_index(or_cursor *c, uint64_t range, int *ok)
{
    if (range == 0) {
        *ok = 0;
        return 0;
    }
    if (range <= 1) return 0;
    if (range > UINT32_MAX) {
        const uint64_t lim = (UINT64_MAX / range) * range;
        for (;;) {
            if (c->pos + 8 > c->len) {
                *ok = 0;
                return 0;
            }
            const uint64_t raw = get_u64le(c->s + c->pos);
            c->pos += 8;
            if (raw < lim) return raw % range;
        }
    }
    const uint64_t limit = (((uint64_t)UINT32_MAX + 1) / range) * range;
    for (;;) {
        if (c->pos + 4 > c->len) {
            *ok = 0;
            return 0;
        }
        const uint64_t raw = get_u32le(c->s + c->pos);
        c->pos += 4;
        if (raw < limit) return raw % range;
    }
}

static int or_sc_validate(size_t n, int family, size_t weight, int64_t amplitude, uint32_t permille, size_t nvalues)
{
    if (n == 0 || n > ((size_t)1 << 20)) return 0;
    switch (family) {
    case OR_SC_FW: return !(weight == 0 || weight > n || amplitude <= 0);
    case OR_SC_BIASED: return !(permille == 0 || permille > 1000);
    case OR_SC_SMALLSET: return !(nvalues == 0 || nvalues > 1024);
    case OR_SC_FWSET: return !(weight == 0 || weight > n || nvalues == 0 || nvalues > 1024);
    default: return 0;
    }
}

/* ShortChallengeSpec::sample (short_challenge.rs:183-205, samplers :392-497); *gamma = Γ_C. */
int or_short_challenge_sample(size_t n, int family, size_t weight, int64_t amplitude, uint32_t permille,
                              const int64_t *values, size_t nvalues, const uint8_t *seed, size_t slen, int64_t *out,
                              uint64_t *gamma)
{
    if (!or_sc_validate(n, family, weight, amplitude, permille, nvalues)) return OR_SC_INVALID;
    size_t bytes = 0;
    const char *label = "";
    switch (family) {
    case OR_SC_FW: bytes = weight * 24 + 32, label = "fw"; break;
    case OR_SC_BIASED: bytes = n * 6 + 32, label = "bias3"; break;
    case OR_SC_SMALLSET: bytes = n * 8 + 32, label = "smallset"; break;
    default: bytes = weight * 32 + 32, label = "fwset"; break;
    }
    uint8_t *s = (uint8_t *)malloc(bytes);
    or_sc_stream(label, seed, slen, s, bytes);
    or_cursor c = {s, bytes, 0};
    int ok = 1;
    for (size_t i = 0; i < n; i++) out[i] = 0;
    if (family == OR_SC_FW || family == OR_SC_FWSET) {
        uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * n); /* partial Fisher-Yates */
        for (size_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
        for (size_t slot = 0; slot < weight && ok; slot++) {
            const uint64_t pos = or_next_index(&c, n - slot, &ok);
            if (!ok) break;
            const uint32_t at = idx[pos];
            const uint32_t t = idx[pos];
            idx[pos] = idx[n - 1 - slot];
            idx[n - 1 - slot] = t;
            if (family == OR_SC_FW) {
                const uint64_t sign = or_next_index(&c, 2, &ok);
                if (ok) out[at] = sign == 0 ? amplitude : -amplitude;
            } else {
                const uint64_t vp = or_next_index(&c, nvalues, &ok);
                if (ok) out[at] = values[vp];
            }
        }
        free(idx);
    } else if (family == OR_SC_BIASED) {
        for (size_t i = 0; i < n && ok; i++) {
            if (c.pos + 4 > c.len) {
                ok = 0;
                break;
            }
            const uint32_t raw = get_u32le(c.s + c.pos);
            c.pos += 4;
            if ((raw & 0x3FF) >= permille) continue; /* zero */
            if (c.pos + 4 > c.len) {
                ok = 0;
                break;
            }
            const uint32_t sraw = get_u32le(c.s + c.pos);
            c.pos += 4;
            out[i] = (sraw & 1) == 0 ? 1 : -1;
        }
    } else {
        for (size_t i = 0; i < n && ok; i++) {
            const uint64_t vp = or_next_index(&c, nvalues, &ok);
            if (ok) out[i] = values[vp];
        }
    }
    free(s);
    if (!ok) return OR_SC_REJECTION;
    unsigned __int128 l2 = 0; /* certified_gamma (:271-277) */
    for (size_t i = 0; i < n; i++) l2 += (unsigned __int128)((__int128)out[i] * out[i]);
    *gamma = or_ceil_sqrt_u128((uint64_t)l2, (uint64_t)(l2 >> 64));
    return OR_SC_OK;
}

/* sample_rq2_challenge (extension.rs:180-216): spec.n must be 2n; even stream
 * positions -> c0, odd -> c1, each reduced mod q; g0, g1 the per-half Γ. */
int or_rq2_sample_challenge(uint32_t q, uint32_t n, size_t spec_n, int family, size_t weight, int64_t amplitude,
                            uint32_t permille, const int64_t *values, size_t nvalues, const uint8_t *seed, size_t slen,
                            uint32_t *out, uint64_t *g0, uint64_t *g1)
{
    if (spec_n != 2 * (size_t)n) return OR_SC_INVALID;
    int64_t *c = (int64_t *)malloc(sizeof(int64_t) * spec_n);
    uint64_t gam = 0;
    const int st =
        or_short_challenge_sample(spec_n, family, weight, amplitude, permille, values, nvalues, seed, slen, c, &gam);
    if (st == OR_SC_OK) {
        unsigned __int128 l0 = 0, l1 = 0;
        for (uint32_t i = 0; i < n; i++) {
            const int64_t a = c[2 * i], b = c[2 * i + 1];
            out[i] = (uint32_t)(((a % (int64_t)q) + q) % q);
            out[n + i] = (uint32_t)(((b % (int64_t)q) + q) % q);
            l0 += (unsigned __int128)((__int128)a * a);
            l1 += (unsigned __int128)((__int128)b * b);
        }
        *g0 = or_ceil_sqrt_u128((uint64_t)l0, (uint64_t)(l0 >> 64));
        *g1 = or_ceil_sqrt_u128((uint64_t)l1, (uint64_t)(l1 >> 64));
    }
    free(c);
    return st;
}

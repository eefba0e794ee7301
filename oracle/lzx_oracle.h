/* Types and entry points of the CPU oracle shared by its source files
 * (lzx_oracle.c: the restatement; lzx_baseline.c: the timed CPU baselines).
 * TEST INFRASTRUCTURE ONLY, like the rest of oracle/. */
#ifndef LZX_ORACLE_H
#define LZX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

typedef unsigned __int128 u128;

enum {
    OR_OK = 0,
    OR_NTT_LENGTH_MISMATCH = 1,
    OR_NTT_ADICITY_TOO_SMALL = 2,
    OR_AJTAI_DIMENSION_MISMATCH = 4,
    OR_AJTAI_NORM_EXCEEDED = 5,
    OR_AJTAI_COMMITMENT_MISMATCH = 6,
    OR_SUMCHECK_BAD_ROUND_SHAPE = 10,
    OR_SUMCHECK_ROUND_CHECK_FAILED = 11,
    OR_SUMCHECK_FINAL_CHECK_FAILED = 12,
    OR_SUMCHECK_CLAIM_MISMATCH = 13,
    OR_SUMCHECK_VIRTUAL_POLY = 14,
    OR_TRANSCRIPT_REJECTION_BUDGET = 20,
    OR_TRANSCRIPT_MESSAGE_TOO_LARGE = 21,
    OR_BAD_ARGUMENT = 30,
};

/* Modulus32: lattice-ring/src/modulus.rs:22-162 */
typedef struct {
    uint32_t q, two_adicity, generator;
    uint64_t barrett; /* floor((2^64-1)/q), modulus.rs:15-17 */
} or_mod32;

/* NttTables: lattice-ring/src/ntt.rs:44-106 */
typedef struct {
    or_mod32 mod;
    uint32_t log_n;
    uint32_t n;
    uint32_t psi, n_inv;
    uint32_t *pow_psi;       /* [n] */
    uint32_t *pow_psi_inv;   /* [n] */
    uint32_t *pow_omega;     /* [max(n/2,1)] */
    uint32_t *pow_omega_inv; /* [max(n/2,1)] */
} or_ntt;

/* KeccakSponge: lattice-core/src/keccak.rs */
typedef struct {
    uint64_t st[25];
    size_t pos;
    size_t rate;
    uint8_t suffix;
} or_sponge;

/* Transcript: lattice-core/src/transcript.rs:36-238 */
typedef struct {
    or_sponge sponge;
    uint64_t query_count;
    uint64_t query_budget;
} or_transcript;

uint32_t or_mul(const or_mod32 *m, uint32_t a, uint32_t b);
uint32_t or_add(const or_mod32 *m, uint32_t a, uint32_t b);
int or_ntt_forward(const or_ntt *t, uint32_t *a, size_t len);
int or_ntt_inverse(const or_ntt *t, uint32_t *a, size_t len);
uint64_t or_gl_from_u128(uint64_t hi, uint64_t lo);
uint64_t or_gl_add(uint64_t a, uint64_t b);
uint64_t or_gl_sub(uint64_t a, uint64_t b);
uint64_t or_gl_mul(uint64_t a, uint64_t b);
uint64_t or_interpolate_at(const uint64_t *evals, size_t n, uint64_t r);
int or_transcript_append_field_slice(or_transcript *t, const uint8_t *label, size_t llen, const uint64_t *v, size_t n);
int or_transcript_challenge_field(or_transcript *t, const uint8_t *label, size_t llen, uint64_t *out);
int or_sumcheck_prove(size_t num_vars, size_t nfactors, const uint64_t *const *factors, size_t nterms,
                      const uint64_t *coeffs, const uint32_t *term_off, const uint32_t *term_ids, uint64_t claim,
                      or_transcript *tr, uint64_t *rounds_out, uint64_t *challenges_out, uint64_t *final_claim_out,
                      uint64_t *factor_claims_out, size_t *d_out);

#endif

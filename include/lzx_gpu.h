/*
 * lzx_gpu.h -- C ABI of the MI355X (gfx950) lattice-commitment hot path.
 *
 * Drop-in boundary for these Rust items of idrees2516/lzx (paths relative to
 * the reference's crates/):
 *   - the negacyclic NTT            lattice-ring/src/ntt.rs, ring.rs
 *   - the ~2^50 quadratic-slot ring lattice-ring/src/modulus50.rs
 *   - the extension R_q[Y]/(Y^2+1)  lattice-ring/src/extension.rs
 *   - witness packing               lattice-ring/src/packing.rs,
 *                                   lattice-commitment/src/sparse.rs,
 *                                   lattice-core/src/decomposition.rs
 *   - the Ajtai commitment A*s      lattice-commitment/src/ajtai.rs
 *   - the Goldilocks sum-check      lattice-sumcheck/src/sumcheck.rs with the
 *     prover and its Fiat-Shamir     lattice-core/src/transcript.rs, keccak.rs,
 *     zerocheck, batching, split eq  zerocheck.rs, batch.rs, fastprover.rs
 *
 * Conventions
 *   - Every entry point returns lzx_status; values 1..29 and 40..62 map
 *     one-to-one onto the reference's error enums (noted per code), 30, 31 and
 *     100.. are ABI/device codes.
 *     Nothing unwinds across the ABI.
 *   - Pointers named *_dev are device (HBM) pointers; *_host are host pointers.
 *     `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Device entry points are stream-ordered and return before completion.
 *   - A batched device entry point whose batch is empty (no elements in and
 *     none out) returns LZX_OK without touching its data pointers, which may
 *     then be NULL.  (Packing an empty input still writes one element, as the
 *     reference's packers do.)
 *   - Residues are canonical: u32 in [0, q), Goldilocks u64 in [0, p).
 *   - Handles are immutable after creation and may be shared across host
 *     threads; calls on one handle from several threads only read it.
 */
#ifndef LZX_GPU_H
#define LZX_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t lzx_status;

enum {
    LZX_OK = 0,
    LZX_NTT_LENGTH_MISMATCH = 1,        /* NttError::LengthMismatch        ntt.rs:40 */
    LZX_NTT_ADICITY_TOO_SMALL = 2,      /* NttError::AdicityTooSmall       ntt.rs:41 */
    LZX_RING_LENGTH_MISMATCH = 3,       /* RingError::LengthMismatch       ring.rs:20 */
    LZX_AJTAI_DIMENSION_MISMATCH = 4,   /* AjtaiError::DimensionMismatch   ajtai.rs:45 */
    LZX_AJTAI_NORM_EXCEEDED = 5,        /* AjtaiError::NormExceeded        ajtai.rs:46 */
    LZX_AJTAI_COMMITMENT_MISMATCH = 6,  /* AjtaiError::CommitmentMismatch  ajtai.rs:47 */
    LZX_SUMCHECK_BAD_ROUND_SHAPE = 10,  /* SumcheckError::BadRoundShape    sumcheck.rs:34 */
    LZX_SUMCHECK_ROUND_CHECK_FAILED = 11, /* SumcheckError::RoundCheckFailed sumcheck.rs:39 */
    LZX_SUMCHECK_FINAL_CHECK_FAILED = 12, /* SumcheckError::FinalCheckFailed sumcheck.rs:43 */
    LZX_SUMCHECK_CLAIM_MISMATCH = 13,   /* SumcheckError::ClaimMismatch    sumcheck.rs:45 */
    LZX_SUMCHECK_VIRTUAL_POLY = 14,     /* SumcheckError::VirtualPoly      sumcheck.rs:30 */
    LZX_ZEROCHECK_FINAL_CLAIM_NONZERO = 15, /* ZerocheckError::FinalClaimNonZero zerocheck.rs:20 */
    LZX_BATCH_VARIABLE_COUNT_MISMATCH = 16, /* BatchError::VariableCountMismatch batch.rs:19 */
    LZX_BATCH_EMPTY = 17,               /* BatchError::EmptyBatch          batch.rs:20 */
    LZX_BATCH_CHALLENGE_COUNT = 18,     /* BatchError::ChallengeCount      batch.rs:21 */
    LZX_MOD50_LOG_N_OUT_OF_RANGE = 40,  /* Mod50Error::LogNOutOfRange    modulus50.rs:418 */
    LZX_MOD50_ADICITY_TOO_SMALL = 41,   /* Mod50Error::AdicityTooSmall   modulus50.rs:419 */
    LZX_MOD50_GENERATOR_INVALID = 42,   /* Mod50Error::GeneratorInvalid  modulus50.rs:420 */
    LZX_MOD50_LENGTH_MISMATCH = 43,     /* Mod50Error::LengthMismatch    modulus50.rs:421 */
    LZX_PACKING_LENGTH_MISMATCH = 50,   /* PackingError::LengthMismatch  packing.rs:210 */
    LZX_PACKING_LIMB_OUT_OF_RANGE = 51, /* PackingError::LimbOutOfRange  packing.rs:211 */
    LZX_PACKING_MODULUS_TOO_LARGE = 52, /* PackingError::ModulusTooLarge packing.rs:212 */
    LZX_PACKING_VALUE_OUT_OF_RANGE = 53, /* PackingError::ValueOutOfRange packing.rs:213 */
    LZX_PACKING_NOT_COPRIME = 54,       /* PackingError::NotCoprime      packing.rs:214 */
    LZX_DECOMP_DIGIT_COUNT_MISMATCH = 55, /* DecompositionError::DigitCountMismatch decomposition.rs:110 */
    LZX_DECOMP_DIGIT_OUT_OF_RANGE = 56, /* DecompositionError::DigitOutOfRange decomposition.rs:111 */
    LZX_DECOMP_VALUE_OUT_OF_RANGE = 57, /* DecompositionError::ValueOutOfRange decomposition.rs:112 */
    LZX_DECOMP_VALUE_TOO_LARGE = 58,    /* DecompositionError::ValueTooLarge   decomposition.rs:113 */
    LZX_SHORT_CHALLENGE_REJECTION_BUDGET = 60, /* ShortChallengeError::RejectionBudgetExceeded short_challenge.rs:84 */
    LZX_SHORT_CHALLENGE_INVALID_PARAMETERS = 61, /* ShortChallengeError::InvalidParameters short_challenge.rs:86 */
    LZX_SHORT_CHALLENGE_GAMMA_CAP_EXCEEDED = 62, /* ShortChallengeError::GammaCapExceeded short_challenge.rs:88 */
    LZX_TRANSCRIPT_REJECTION_BUDGET = 20, /* TranscriptError::RejectionBudgetExceeded transcript.rs:16 */
    LZX_TRANSCRIPT_MESSAGE_TOO_LARGE = 21, /* TranscriptError::MessageTooLarge transcript.rs:18 */
    LZX_INVALID_ARGUMENT = 30,          /* null handle/pointer or malformed shape */
    LZX_UNSUPPORTED = 31,               /* shape outside what the kernels implement */
    LZX_DEVICE_ERROR = 100,             /* HIP launch/runtime failure */
    LZX_DEVICE_OUT_OF_MEMORY = 101,
    LZX_NO_DEVICE = 102
};

/* Human-readable name of a status code (static storage). */
const char *lzx_status_string(lzx_status s);

/* Text of the most recent HIP error seen by this thread ("" if none). */
const char *lzx_last_device_error(void);

/* ---------------------------------------------------------------------------
 * Negacyclic NTT over Z_q[X]/(X^n + 1), n = 2^log_n, q an odd prime < 2^32.
 * Forward output is natural order: out[k] = sum_i a_i * psi^(i(2k+1)) with
 * psi = generator^((q-1) >> (log_n+1)), exactly NttTables::forward.
 * ------------------------------------------------------------------------- */
typedef struct lzx_ntt_plan lzx_ntt_plan;

/* Replaces NttTables::new(Modulus32::new(q, two_adicity, generator), log_n)
 * (ntt.rs:46, modulus.rs:37).  LZX_NTT_ADICITY_TOO_SMALL when
 * log_n + 1 > two_adicity.  Uploads the twiddle tables to the current device. */
lzx_status lzx_ntt_plan_create(uint32_t q, uint32_t two_adicity, uint32_t generator, uint32_t log_n,
                               lzx_ntt_plan **out);
void lzx_ntt_plan_destroy(lzx_ntt_plan *plan);

/* psi and n^{-1} of the plan (NttTables fields, ntt.rs:52-58). */
lzx_status lzx_ntt_plan_constants(const lzx_ntt_plan *plan, uint32_t *psi, uint32_t *n_inv);

/* The kernel class that a batched forward (inverse != 0: inverse) transform
 * of polys_per_plan polynomials per plan runs on 16-byte-aligned data, for
 * plans as passed to lzx_ntt_forward_multi.  No counterpart in the reference
 * (its NttTables has one scalar path); for tests and benchmarks that report
 * the kernel they measure.
 *   LZX_NTT_PATH_MFMA: the int8 matrix-core transform (n = 64 and
 *     2^20 < q < 2^31 - 2^16 for every plan; unavailable to plans created
 *     with LZX_NTT_MFMA=0);
 *   LZX_NTT_PATH_REGISTER: the register butterfly kernels (n <= 64; exact
 *     FP64 butterflies for q > 2^31);
 *   LZX_NTT_PATH_LDS: the LDS kernel (n > 64);
 *   LZX_NTT_PATH_AUTO: MFMA-capable and not yet measured.  The first such
 *     call for a (device, moduli, polys_per_plan, direction) times the MFMA
 *     and register classes on a scratch copy of its data and keeps the
 *     faster in the tuning cache (LZX_AUTOTUNE_CACHE); plans created with
 *     LZX_NTT_KERNEL=mfma or =register always run that class. */
#define LZX_NTT_PATH_REGISTER 0u
#define LZX_NTT_PATH_MFMA 1u
#define LZX_NTT_PATH_LDS 2u
#define LZX_NTT_PATH_AUTO 3u
lzx_status lzx_ntt_batch_path(const lzx_ntt_plan *const *plans, size_t nplans, size_t polys_per_plan,
                              uint32_t inverse, uint32_t *path);

/* In-place batched transforms of len/n polynomials laid out [poly][n] in
 * HBM.  Replace NttTables::forward/forward_fast (ntt.rs:153,118) and
 * inverse/inverse_fast (ntt.rs:220,187) applied to each polynomial.
 * LZX_NTT_LENGTH_MISMATCH when len is not a multiple of n. */
lzx_status lzx_ntt_forward_batch(const lzx_ntt_plan *plan, uint32_t *data_dev, size_t len, void *stream);
lzx_status lzx_ntt_inverse_batch(const lzx_ntt_plan *plan, uint32_t *data_dev, size_t len, void *stream);

/* Incomplete NTT (ntt.rs:257 forward_partial, ntt.rs:291 complete_partial),
 * batched.  LZX_NTT_LENGTH_MISMATCH when levels/already > log_n. */
lzx_status lzx_ntt_forward_partial_batch(const lzx_ntt_plan *plan, uint32_t *data_dev, size_t len, uint32_t levels,
                                         void *stream);
lzx_status lzx_ntt_complete_partial_batch(const lzx_ntt_plan *plan, uint32_t *data_dev, size_t len,
                                          uint32_t already, void *stream);

/* CRT batch: nplans moduli of one degree in ONE launch.  data_dev is laid
 * out [plan][poly][n] with polys_per_plan polynomials per modulus (the
 * reference applies NttTables::forward per element per modulus). */
lzx_status lzx_ntt_forward_multi(const lzx_ntt_plan *const *plans, size_t nplans, uint32_t *data_dev,
                                 size_t polys_per_plan, void *stream);
lzx_status lzx_ntt_inverse_multi(const lzx_ntt_plan *const *plans, size_t nplans, uint32_t *data_dev,
                                 size_t polys_per_plan, void *stream);

/* Synchronous host-buffer forms with the exact reference signatures
 * forward(&self, &mut [u32]) / inverse(&self, &mut [u32]): len must equal n
 * times the number of polynomials. */
lzx_status lzx_ntt_forward_host(const lzx_ntt_plan *plan, uint32_t *a_host, size_t len);
lzx_status lzx_ntt_inverse_host(const lzx_ntt_plan *plan, uint32_t *a_host, size_t len);

/* Negacyclic ring products out = a * b for len/n pairs (RingElement::mul,
 * ring.rs:309-340).  Device buffers, [poly][n]. */
lzx_status lzx_ring_mul_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, const uint32_t *b_dev,
                              uint32_t *out_dev, size_t len, void *stream);

/* ---------------------------------------------------------------------------
 * Quadratic extension ring R_q[Y]/(Y^2+1) (lattice-ring/src/extension.rs).
 * An element c0 + c1 Y is [c0 | c1], n canonical coefficients each; a batch
 * of count elements is [count][2][n] in device memory (the order of
 * Rq2::to_bytes). The plan names q and n. Outputs may alias or overlap
 * inputs in any way (an in-place norm writes [count][n] over the first half
 * of its input); overlapping calls stage their inputs in scratch first.
 * ------------------------------------------------------------------------- */

/* Rq2::mul (extension.rs:115-124): out = a * b, count elements. */
lzx_status lzx_rq2_mul_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, const uint32_t *b_dev,
                             uint32_t *out_dev, size_t count, void *stream);

/* Rq2::norm (extension.rs:142-147): out[count][n] = c0^2 + c1^2. */
lzx_status lzx_rq2_norm_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, uint32_t *out_dev, size_t count,
                              void *stream);

/* Rq2::inverse (extension.rs:151-167): ok_dev[i] = 1 and out[i] = a[i]^{-1}
 * when every NTT slot of N(a[i]) is non-zero, else ok_dev[i] = 0 (None) and
 * out[i] = 0. */
lzx_status lzx_rq2_inverse_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, uint32_t *out_dev,
                                 uint8_t *ok_dev, size_t count, void *stream);

/* Short ring-element challenges (lattice-core/src/short_challenge.rs:47-89):
 * ShortChallengeFamily as `family` with its parameters; unused fields are ignored. */
enum {
    LZX_SC_FIXED_WEIGHT = 0,           /* weight, amplitude */
    LZX_SC_BIASED_TERNARY = 1,         /* p_nonzero_permille */
    LZX_SC_SMALL_SET = 2,              /* values[nvalues] */
    LZX_SC_FIXED_WEIGHT_SMALL_SET = 3  /* weight, values[nvalues] */
};
typedef struct {
    size_t n;
    int32_t family;
    size_t weight;
    int64_t amplitude;
    uint32_t p_nonzero_permille;
    const int64_t *values;
    size_t nvalues;
} lzx_short_challenge_spec;

/* ShortChallengeSpec::sample (short_challenge.rs:183-205): coeffs_out[n]
 * signed coefficients and *gamma_out = Γ_C = ceil(sqrt(sum c_i^2)).  Host code. */
lzx_status lzx_short_challenge_sample(const lzx_short_challenge_spec *spec, const uint8_t *seed, size_t seed_len,
                                      int64_t *coeffs_out, uint64_t *gamma_out);

/* ShortChallengeSpec::sample_with_gamma_cap (short_challenge.rs:208-236):
 * tries seed || u64le(attempt) for attempt = 0 .. retries until Γ_C <= cap;
 * LZX_SHORT_CHALLENGE_GAMMA_CAP_EXCEEDED (with the last Γ in *gamma_out) otherwise. */
lzx_status lzx_short_challenge_sample_with_gamma_cap(const lzx_short_challenge_spec *spec, const uint8_t *seed,
                                                     size_t seed_len, uint64_t cap, size_t retries,
                                                     int64_t *coeffs_out, uint64_t *gamma_out);

/* sample_rq2_challenge (extension.rs:180-216) for RingConfig(q, log_n):
 * spec->n must be 2n; out_host[2n] = [c0 | c1] (even stream positions to c0,
 * odd to c1, reduced mod q); *g0, *g1 the per-half Γ.  Host code. */
lzx_status lzx_rq2_sample_challenge(uint32_t q, uint32_t log_n, const lzx_short_challenge_spec *spec,
                                    const uint8_t *seed, size_t seed_len, uint32_t *out_host, uint64_t *g0,
                                    uint64_t *g1);

/* ceil_sqrt_u128 (short_challenge.rs:293-340): min(ceil(sqrt(hi 2^64 + lo)), 2^64 - 1). */
uint64_t lzx_ceil_sqrt_u128(uint64_t lo, uint64_t hi);

/* Rq2::{add, sub, neg, conjugate} (extension.rs:91-139), coefficient-wise. */
lzx_status lzx_rq2_add_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, const uint32_t *b_dev,
                             uint32_t *out_dev, size_t count, void *stream);
lzx_status lzx_rq2_sub_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, const uint32_t *b_dev,
                             uint32_t *out_dev, size_t count, void *stream);
lzx_status lzx_rq2_neg_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, uint32_t *out_dev, size_t count,
                             void *stream);
lzx_status lzx_rq2_conjugate_batch(const lzx_ntt_plan *plan, const uint32_t *a_dev, uint32_t *out_dev,
                                   size_t count, void *stream);

/* ---------------------------------------------------------------------------
 * Ajtai commitment t = A * s over R_q (lattice-commitment/src/ajtai.rs).
 * The key keeps the NTT-domain matrix A^ resident in HBM ([i][j][n],
 * row-major k x m) and, for the int8 path, the signed 8-bit limb planes of
 * the centred coefficient matrix.
 * ------------------------------------------------------------------------- */
typedef struct lzx_ajtai_key lzx_ajtai_key;

enum { LZX_COMMIT_AUTO = 0, LZX_COMMIT_NTT = 1, LZX_COMMIT_INT8 = 2 };

/* Replaces AjtaiPublicKey::from_seed(AjtaiParams{ring, k, m, norm_bound},
 * seed) (ajtai.rs:87-103): A[i*m+j] = uniform_from_seed("ajtai-A", seed,
 * i*m+j) (ring.rs:102-180), A^ = forward NTT of every entry. */
lzx_status lzx_ajtai_key_from_seed(uint32_t q, uint32_t two_adicity, uint32_t generator, uint32_t log_n, size_t k,
                                   size_t m, uint32_t norm_bound, const uint8_t seed[32], lzx_ajtai_key **out);

/* The same key from an explicit coefficient matrix A (host, [k*m][n]). */
lzx_status lzx_ajtai_key_from_matrix(uint32_t q, uint32_t two_adicity, uint32_t generator, uint32_t log_n,
                                     size_t k, size_t m, uint32_t norm_bound, const uint32_t *A_host,
                                     lzx_ajtai_key **out);
void lzx_ajtai_key_destroy(lzx_ajtai_key *key);

/* Copy A (coefficient domain) or A^ (NTT domain) of the key to the host,
 * [k*m][n] each (AjtaiPublicKey::entry, ajtai.rs:106-112). */
lzx_status lzx_ajtai_key_matrix(const lzx_ajtai_key *key, uint32_t *A_host);
lzx_status lzx_ajtai_key_matrix_ntt(const lzx_ajtai_key *key, uint32_t *Ahat_host);

/* Batched commit of W witnesses (AjtaiPublicKey::commit, ajtai.rs:118-153):
 * s_dev [W][m][n] canonical residues, out_dev [W][k][n].  s_len is the
 * per-witness ring-element count and must equal m (DimensionMismatch). */
lzx_status lzx_ajtai_commit_batch(const lzx_ajtai_key *key, const uint32_t *s_dev, size_t s_len, size_t W,
                                  uint32_t *out_dev, void *stream);

/* Batched commit of small-digit witnesses: digits_dev [W][m][n] int8, any
 * value in [-128, 127] (coefficient c represents c mod q, ring.rs:232-243
 * from_signed).  path selects the NTT-domain kernels, the int8 MFMA
 * coefficient-domain kernels (n = 64, q < 2^32 - 2^25), or AUTO: the
 * faster of the two for the shape (device, q, n, k, m, W), measured on the
 * first AUTO call of that shape (which runs each path twice on the stream)
 * and kept for the process and in LZX_AUTOTUNE_CACHE when set; keys without
 * int8 support take the NTT path.  Results are identical on every path, for
 * every int8 digit. */
lzx_status lzx_ajtai_commit_digits_batch(const lzx_ajtai_key *key, const int8_t *digits_dev, size_t s_len,
                                         size_t W, uint32_t *out_dev, int32_t path, void *stream);

/* The path AUTO takes for W witnesses on this key: LZX_COMMIT_NTT or
 * LZX_COMMIT_INT8 once measured (or kept from LZX_AUTOTUNE_CACHE), else
 * LZX_COMMIT_AUTO (the next AUTO call measures).  No counterpart in the
 * reference; for callers that report or pre-tune their shapes. */
lzx_status lzx_ajtai_commit_path_choice(const lzx_ajtai_key *key, size_t W, int32_t *path_out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): lzx_ajtai_commit_digits_batch with path
 * LZX_COMMIT_INT8, bit-identical, but allocating nothing.  Its chunk partial
 * sums live in a fixed array of LZX_NOALLOC_SCRATCH_BYTES in the library's
 * device code, one per device, held until the process ends.  Limit:
 * ceil(m / 2032) * W * k * 256 bytes <= LZX_NOALLOC_SCRATCH_BYTES, else
 * LZX_UNSUPPORTED (the reason is in lzx_last_device_error()); n = 64 and
 * q < 2^32 - 2^25 as for the int8 path.  Every call on a device shares that
 * device's array: calls that overlap on one device (other streams, other
 * threads) race, so serialise them.  The array keeps the last call's
 * contents; each call writes every word it later reads. */
#define LZX_NOALLOC_SCRATCH_BYTES (16u << 20)
lzx_status lzx_ajtai_commit_digits_int8_noalloc(const lzx_ajtai_key *key, const int8_t *digits_dev, size_t s_len,
                                                size_t W, uint32_t *out_dev, void *stream);

/* Synchronous host forms: commit(&self, &[RingElement]) (ajtai.rs:118) and
 * verify_opening(&self, &AjtaiCommitment, &[RingElement]) (ajtai.rs:210):
 * norm check (NormExceeded) then recommit and compare (CommitmentMismatch). */
lzx_status lzx_ajtai_commit_host(const lzx_ajtai_key *key, const uint32_t *s_host, size_t s_len,
                                 uint32_t *out_host);

/* Replaces commit_naive_reference (ajtai.rs:158-180): every product A_ij * s_j
 * as its own ring multiplication from the coefficient-domain A (NTT of both,
 * pointwise, inverse NTT per product), summed over j in the coefficient
 * domain.  Equal to commit on every input; kept as the differential path.
 * Device form: s_dev [W][m][n], out_dev [W][k][n]; host form: one witness. */
lzx_status lzx_ajtai_commit_naive_batch(const lzx_ajtai_key *key, const uint32_t *s_dev, size_t s_len, size_t W,
                                        uint32_t *out_dev, void *stream);
lzx_status lzx_ajtai_commit_naive_host(const lzx_ajtai_key *key, const uint32_t *s_host, size_t s_len,
                                       uint32_t *out_host);

/* Replaces sample_small_secret(&RingConfig, m, bound, seed) (ajtai.rs:307-348)
 * on the device, one thread per ring element: m elements of n = 2^log_n
 * coefficients uniform in [-bound, bound], written as residues mod q to
 * out_dev [m][n] u32 and/or as int8 digits to digits_dev [m][n] (bound <= 127);
 * either pointer may be NULL, not both.  LZX_INVALID_ARGUMENT for
 * bound >= 2^31 (the reference's span 2*bound + 1 would exceed a u32 window
 * and its rejection loop would never end); LZX_UNSUPPORTED for seed_len > 95
 * (the framed XOF input must fit one SHAKE-256 block). */
lzx_status lzx_ajtai_sample_small_secret(uint32_t q, uint32_t log_n, size_t m, uint32_t bound, const uint8_t *seed,
                                         size_t seed_len, uint32_t *out_dev, int8_t *digits_dev, void *stream);
lzx_status lzx_ajtai_verify_opening_host(const lzx_ajtai_key *key, const uint32_t *t_host, const uint32_t *s_host,
                                         size_t s_len);

/* ---------------------------------------------------------------------------
 * Uneven commitment batch: many independent commitments t = A s, each on its
 * own ring Z_q[X]/(X^n + 1) with its own length m, in ONE persistent launch.
 * Every job has the semantics of AjtaiPublicKey::commit (ajtai.rs:118-153);
 * the reference has no batch form (its callers loop over commit).  Beyond the
 * reference's Modulus32 (modulus.rs:22-162), moduli up to 2^50 are accepted,
 * with u64 residues.  Results are the exact ring products, independent of
 * how the library computes them.
 * ------------------------------------------------------------------------- */
typedef struct lzx_wide_key lzx_wide_key;
typedef struct lzx_commit_batch lzx_commit_batch;

/* Commitment key over Z_q[X]/(X^n + 1), n = 2^log_n with 6 <= log_n <= 8.
 * q is an odd prime with 2n | q - 1 and 256 < q < 2^50; generator generates
 * Z_q^* (psi = generator^((q-1)/2n), as root_of_unity, modulus.rs:146-161).
 * A_dev is [k][m][n] residues (any representative): u32 elements when
 * q < 2^32, u64 otherwise.  The key keeps an NTT-domain copy built on the
 * device (stream-ordered); A_dev may be freed once `stream` has passed it.
 * Errors: NttAdicityTooSmall when 2n does not divide q - 1, InvalidArgument
 * when psi^n != -1 or a size is out of range. */
lzx_status lzx_wide_key_create(uint64_t q, uint64_t generator, uint32_t log_n, size_t k, size_t m,
                               const void *A_dev, void *stream, lzx_wide_key **out);
void lzx_wide_key_destroy(lzx_wide_key *key);
/* Element width in bytes (4 or 8) of A and of commitments under this key. */
uint32_t lzx_wide_key_elem_bytes(const lzx_wide_key *key);

typedef struct lzx_commit_job {
    const lzx_wide_key *key;
    const int8_t *digits_dev; /* [m][n] balanced digits (16-byte aligned); coefficient = digit mod q */
    size_t m;                 /* witness length; the key's first m columns are used (m <= key m) */
    void *out_dev;            /* [k][n] canonical residues of the key's element width */
} lzx_commit_job;

/* Plans a batch: device work queue (largest jobs first), per-class chunk
 * sizes (the autotune cache, else LZX_BATCH_CHUNK, else a default), job
 * descriptors and scratch.  The job array is copied; keys must outlive the
 * batch. */
lzx_status lzx_commit_batch_create(const lzx_commit_job *jobs, size_t njobs, lzx_commit_batch **out);
/* Runs the whole batch as one persistent kernel launch, replayed from a
 * hipGraph captured on first use.  Stream-ordered; calls on one batch must
 * not overlap. */
lzx_status lzx_commit_batch_run(lzx_commit_batch *b, void *stream);
/* Synchronizes `stream`, then reports a device-side failure of the last run:
 * a dependency wait past its bound gives LZX_DEVICE_ERROR. */
lzx_status lzx_commit_batch_check(lzx_commit_batch *b, void *stream);
/* Times candidate chunk sizes for each class present (runs that write the
 * jobs' outputs), keeps the fastest in the process-wide cache and in the
 * file named by LZX_AUTOTUNE_CACHE when set, and replans the batch. */
lzx_status lzx_commit_batch_autotune(lzx_commit_batch *b, void *stream);
/* Plan facts: work items, scratch bytes and the chunk (ring elements per work
 * item) of each class, class = 3 * kind + log_n - 6 with kind 0 for
 * q < 2^31, 1 for q < 2^32, 2 for q < 2^50 (0 when the class is absent). */
lzx_status lzx_commit_batch_info(const lzx_commit_batch *b, size_t *items, size_t *scratch_bytes,
                                 uint32_t chunks[9]);
void lzx_commit_batch_destroy(lzx_commit_batch *b);

/* ---------------------------------------------------------------------------
 * Fiat-Shamir transcript (lattice-core/src/transcript.rs), host side.
 * ------------------------------------------------------------------------- */
typedef struct lzx_transcript lzx_transcript;

/* Transcript::new(protocol, budget) (transcript.rs:46). */
lzx_status lzx_transcript_new(const uint8_t *protocol, size_t len, uint64_t query_budget, lzx_transcript **out);
lzx_status lzx_transcript_clone(const lzx_transcript *t, lzx_transcript **out);
void lzx_transcript_destroy(lzx_transcript *t);
/* append_message (transcript.rs:67); MessageTooLarge above 2^26 bytes. */
lzx_status lzx_transcript_append_message(lzx_transcript *t, const uint8_t *label, size_t label_len,
                                         const uint8_t *msg, size_t msg_len);
/* append_field_slice (transcript.rs:90). */
lzx_status lzx_transcript_append_field_slice(lzx_transcript *t, const uint8_t *label, size_t label_len,
                                             const uint64_t *values, size_t n);
/* challenge_fields (transcript.rs:109); RejectionBudgetExceeded past the budget. */
lzx_status lzx_transcript_challenge_fields(lzx_transcript *t, const uint8_t *label, size_t label_len, size_t n,
                                           uint64_t *out);
uint64_t lzx_transcript_query_count(const lzx_transcript *t);

/* ---------------------------------------------------------------------------
 * Sum-check over Goldilocks (lattice-sumcheck/src/sumcheck.rs).
 * A virtual polynomial P = sum_t c_t * prod_{f in ids_t} F_f is given as
 * nfactors device tables of 2^num_vars canonical u64 (variable 0 = most
 * significant index bit, mle.rs:1-7) and terms in CSR form: term t has
 * coefficient coeffs[t] and factor ids term_ids[term_off[t] .. term_off[t+1]).
 * Device limits (virtual_poly.rs:48-54 has none): at most 16 factors, 16
 * terms and degree 15 (the longest term); a polynomial beyond them is refused
 * with LZX_UNSUPPORTED before anything is enqueued.
 * ------------------------------------------------------------------------- */

/* sumcheck::prove(&vp, claim, &mut transcript) (sumcheck.rs:62-162).  The
 * factor tables are read, never written (the reference clones them).
 * Outputs (host): rounds_out [num_vars][degree+1], challenges_out
 * [num_vars], *final_claim, factor_claims_out [nfactors], *degree_out.
 * degree is the maximum term length, or 1 for the zero polynomial. */
lzx_status lzx_sumcheck_prove(size_t num_vars, size_t nfactors, const uint64_t *const *factors_dev, size_t nterms,
                              const uint64_t *coeffs, const uint32_t *term_off, const uint32_t *term_ids,
                              uint64_t claim, lzx_transcript *transcript, uint64_t *rounds_out,
                              uint64_t *challenges_out, uint64_t *final_claim, uint64_t *factor_claims_out,
                              size_t *degree_out, void *stream);

/* Same prover with the challenges supplied by the caller instead of a
 * transcript (kernel-level parity and timing; no Fiat-Shamir). */
lzx_status lzx_sumcheck_prove_with_challenges(size_t num_vars, size_t nfactors, const uint64_t *const *factors_dev,
                                              size_t nterms, const uint64_t *coeffs, const uint32_t *term_off,
                                              const uint32_t *term_ids, uint64_t claim,
                                              const uint64_t *challenges, uint64_t *rounds_out,
                                              uint64_t *final_claim, uint64_t *factor_claims_out,
                                              size_t *degree_out, void *stream);

/* SumcheckProof::verify (sumcheck.rs:324-371), host only.  rounds is
 * num_rounds rows of round_len values; expected_final is used when
 * has_expected != 0.  Writes point_out [num_rounds] and *final_claim. */
lzx_status lzx_sumcheck_verify(const uint64_t *rounds, size_t num_rounds, size_t round_len, size_t num_vars,
                               size_t max_degree, uint64_t claim, lzx_transcript *transcript, int32_t has_expected,
                               uint64_t expected_final, uint64_t *point_out, uint64_t *final_claim);

/* fastprover::prove_fast_with_eq (fastprover.rs:212-224; prove_fast :201-208
 * is this call with n_eq = 0): the same proof as lzx_sumcheck_prove, byte
 * for byte (the reference's fast prover is transcript-identical to
 * sumcheck::prove).  The factors named by eq_indices are eq(w, .) for the
 * host point w (num_vars canonical values) and their tables are never read,
 * so those factor pointers may be NULL.  When one eq factor appears exactly
 * once in every term the device never materialises it: each round's eq values
 * come from a prefix/suffix split of w, the tables stream one factor fewer and
 * the other factors are evaluated at d points instead of d + 1.  Otherwise the
 * eq tables are built from w on the device first.  The window and statistics
 * options of the reference's FastProverOpts have no effect on the result and
 * are not taken. */
lzx_status lzx_sumcheck_prove_eq(size_t num_vars, size_t nfactors, const uint64_t *const *factors_dev, size_t nterms,
                                 const uint64_t *coeffs, const uint32_t *term_off, const uint32_t *term_ids,
                                 uint64_t claim, lzx_transcript *transcript, const size_t *eq_indices, size_t n_eq,
                                 const uint64_t *w, uint64_t *rounds_out, uint64_t *challenges_out,
                                 uint64_t *final_claim, uint64_t *factor_claims_out, size_t *degree_out,
                                 void *stream);

/* zerocheck::prove (zerocheck.rs:26-54): r = challenge_fields("zerocheck-point",
 * num_vars) from the transcript; refuses with LZX_ZEROCHECK_FINAL_CLAIM_NONZERO
 * when P(r) != 0 (P does not vanish on the hypercube); otherwise runs the
 * sum-check of P(x) eq(r, x) with claim 0 (factors [P, eq(r, .)], one term,
 * degree 2).  point_out receives r (num_vars), rounds_out num_vars x 3 values,
 * factor_claims_out 2 values. */
lzx_status lzx_zerocheck_prove(size_t num_vars, const uint64_t *poly_dev, lzx_transcript *transcript,
                               uint64_t *point_out, uint64_t *rounds_out, uint64_t *challenges_out,
                               uint64_t *final_claim, uint64_t *factor_claims_out, void *stream);

/* zerocheck::verify (zerocheck.rs:58-84): draws r as the prover did, verifies
 * the degree-2 rounds against claim 0 (and the final claim against
 * expected_poly_eval when has_expected) and returns r in point_out. */
lzx_status lzx_zerocheck_verify(const uint64_t *rounds, size_t num_rounds, size_t round_len, size_t num_vars,
                                lzx_transcript *transcript, int32_t has_expected, uint64_t expected_poly_eval,
                                uint64_t *point_out);

/* Round-level sum-check session: the rounds of sumcheck::prove
 * (sumcheck.rs:100-143) one call at a time, with the transcript steps left to
 * the caller (absorb the round, draw r, check g(0)+g(1), interpolate).
 * create evaluates round 0 on the device; round_eval copies g_j(0..d) out
 * (round_len() = degree + 1 values; the zero polynomial gives 2 zeros);
 * fold(r) binds the round's variable to r and evaluates the next round;
 * final, once every variable is bound, returns the factor claims and
 * P(r) (sumcheck.rs:139-154).  round_eval/fold past the last variable and
 * final before it return LZX_SUMCHECK_BAD_ROUND_SHAPE.  The caller's tables
 * must stay valid until the first fold; the session owns its folded copies.
 * Calls on one session are sequential. */
typedef struct lzx_sumcheck_session lzx_sumcheck_session;
lzx_status lzx_sumcheck_session_create(size_t num_vars, size_t nfactors, const uint64_t *const *factors_dev,
                                       size_t nterms, const uint64_t *coeffs, const uint32_t *term_off,
                                       const uint32_t *term_ids, void *stream, lzx_sumcheck_session **out);
size_t lzx_sumcheck_session_round_len(const lzx_sumcheck_session *session);
lzx_status lzx_sumcheck_session_round_eval(lzx_sumcheck_session *session, uint64_t *evals_out);
lzx_status lzx_sumcheck_session_fold(lzx_sumcheck_session *session, uint64_t r);
lzx_status lzx_sumcheck_session_final(lzx_sumcheck_session *session, uint64_t *factor_claims_out,
                                      uint64_t *final_claim);
void lzx_sumcheck_session_destroy(lzx_sumcheck_session *session);

/* One claim of a batch (batch.rs:24-28 BatchClaim): a virtual polynomial in
 * the lzx_sumcheck_prove form (device tables of 2^num_vars canonical u64,
 * terms in CSR form) and its asserted sum. */
typedef struct {
    size_t num_vars;
    size_t nfactors;
    const uint64_t *const *factors_dev;
    size_t nterms;
    const uint64_t *coeffs;
    const uint32_t *term_off;
    const uint32_t *term_ids;
    uint64_t claimed_sum;
} lzx_batch_claim;

/* batch::prove_batch (batch.rs:33-80): draws rho_i = challenge_fields("batch-rho", k),
 * proves sum_i rho_i P_i with claim sum_i rho_i c_i on the device prover.
 * All claims share num_vars (else LZX_BATCH_VARIABLE_COUNT_MISMATCH).
 * Outputs: rhos_out[k], rounds_out[num_vars][degree+1], challenges_out[num_vars],
 * factor_claims_out[sum_i nfactors_i] in global factor order (claim 0's first).
 * Factor tables passed more than once (the same device pointer) are folded
 * once. At most 16 distinct tables, 16 terms and degree 15 in all (LZX_UNSUPPORTED). */
lzx_status lzx_sumcheck_prove_batch(const lzx_batch_claim *claims, size_t nclaims, lzx_transcript *transcript,
                                    uint64_t *rhos_out, uint64_t *rounds_out, uint64_t *challenges_out,
                                    uint64_t *final_claim, uint64_t *factor_claims_out, size_t *degree_out,
                                    void *stream);

/* batch::prove_batch_star (batch.rs:161-225): claims of different num_vars;
 * a claim over r < N = max num_vars variables is lifted by ignoring its last
 * N - r variables (f'[j 2^(N-r) + s] = f[j]; its claim scales by 2^(N-r)),
 * rho_i = challenge_fields("batch-star-rho", k). N rounds; *num_vars_out = N.
 * The per-claim point is challenges_out[0 .. num_vars_i). Padded tables are
 * built in device scratch and count towards the 16 distinct tables. */
lzx_status lzx_sumcheck_prove_batch_star(const lzx_batch_claim *claims, size_t nclaims, lzx_transcript *transcript,
                                         uint64_t *rhos_out, uint64_t *rounds_out, uint64_t *challenges_out,
                                         uint64_t *final_claim, uint64_t *factor_claims_out, size_t *num_vars_out,
                                         size_t *degree_out, void *stream);

/* batch::verify_batch (batch.rs:300-336): rho_i from "batch-rho", the rounds
 * verified against sum_i rho_i c_i over num_vars variables. */
lzx_status lzx_sumcheck_verify_batch(const uint64_t *rounds, size_t num_rounds, size_t round_len, size_t num_vars,
                                     size_t max_degree, const uint64_t *claimed_sums, size_t nclaims,
                                     lzx_transcript *transcript, int32_t has_expected, uint64_t expected_final,
                                     uint64_t *point_out);

/* batch::verify_batch_star (batch.rs:236-262) when rhos is null, else
 * verify_batch_star_with_rhos (batch.rs:268-296) with the caller's nrhos rhos:
 * the rounds are verified against sum_i rho_i 2^(N - num_vars_i) c_i over
 * N = max num_vars variables. */
lzx_status lzx_sumcheck_verify_batch_star(const uint64_t *rounds, size_t num_rounds, size_t round_len,
                                          const size_t *num_vars, size_t n_num_vars, size_t max_degree,
                                          const uint64_t *claimed_sums, size_t nclaims, const uint64_t *rhos,
                                          size_t nrhos, lzx_transcript *transcript, int32_t has_expected,
                                          uint64_t expected_final, uint64_t *point_out);

/* Multilinear helpers on device tables (field_simd.rs:391 eq_table,
 * virtual_poly.rs:107 sum_over_hypercube). */
lzx_status lzx_eq_table(const uint64_t *point_host, size_t num_vars, uint64_t *out_dev, void *stream);
lzx_status lzx_vp_sum_over_hypercube(size_t num_vars, size_t nfactors, const uint64_t *const *factors_dev,
                                     size_t nterms, const uint64_t *coeffs, const uint32_t *term_off,
                                     const uint32_t *term_ids, uint64_t *sum_out, void *stream);

/* Measurement aid with no counterpart in the reference: the device's
 * Goldilocks product rate (the sum-check kernels' VALU ceiling).  Runs
 * independent chains of 64-bit modular products (lazily reduced, as the
 * kernels' inner products are) on every CU at waves_per_simd (4 or 8) waves
 * per SIMD for about 5 ms and returns products per second. */
lzx_status lzx_gl_mul_rate(uint32_t waves_per_simd, double *products_per_s);

/* ---- witness packing feeding the commitment ----------------------------------
 * Device buffers throughout; ring elements are u32 [elems][n] residues. */

/* pack_field_elements (packing.rs:30-52): each canonical Goldilocks value
 * becomes three 22-bit limbs (low first) in consecutive coefficients, each
 * reduced mod q; out_dev holds lzx_pack_field_elements_len(count, n) =
 * max(1, ceil(3 count / n)) elements, zero-padded. */
size_t lzx_pack_field_elements_len(size_t count, uint32_t n);
lzx_status lzx_pack_field_elements(const uint64_t *values_dev, size_t count, uint32_t q, uint32_t n,
                                   uint32_t *out_dev, void *stream);
/* unpack_field_elements (packing.rs:57-91): out_dev receives floor(elems n / 3)
 * values; LZX_PACKING_LIMB_OUT_OF_RANGE if any balanced coefficient is
 * outside (-2^22, 2^22).  Synchronizes the stream. */
lzx_status lzx_unpack_field_elements(const uint32_t *packed_dev, size_t elems, uint32_t q, uint32_t n,
                                     uint64_t *out_dev, void *stream);
/* crt_pack (packing.rs:95-155): groups of k values share coefficient 0 of one
 * element (max(1, ceil(count / k)) elements); moduli on the host.
 * Synchronizes the stream. */
lzx_status lzx_crt_pack(const uint64_t *values_dev, size_t count, const uint64_t *moduli, size_t k, uint32_t q,
                        uint32_t n, uint32_t *out_dev, void *stream);
/* crt_unpack (packing.rs:158-185). */
lzx_status lzx_crt_unpack(const uint32_t *packed_dev, size_t elems, uint32_t n, const uint64_t *moduli, size_t k,
                          size_t expected, uint64_t *out_dev, void *stream);
/* pack_bits / unpack_bits (sparse.rs:30-71): 31 bits per coefficient, bit b of
 * coefficient g is bits[31 g + b]; bits are bytes (0 / non-zero).
 * lzx_pack_bits_len = max(1, ceil(nbits / (31 n))) elements. */
size_t lzx_pack_bits_len(size_t nbits, uint32_t n);
lzx_status lzx_pack_bits(const uint8_t *bits_dev, size_t nbits, uint32_t n, uint32_t *out_dev, void *stream);
lzx_status lzx_unpack_bits(const uint32_t *elems_dev, size_t elems, uint32_t n, uint8_t *bits_dev, size_t nbits,
                           void *stream);
/* commit_onehot_columns (sparse.rs:118-144): column c is the indicator of bit
 * positions[c] among lens[c], bit-packed and padded to m elements; s_dev
 * (u32 [m][n]) receives the sum of every column mod q and t_dev (u32 [k][n])
 * the commitment of s. */
lzx_status lzx_ajtai_commit_onehot(const lzx_ajtai_key *key, const uint64_t *positions, const uint64_t *lens,
                                   size_t ncols, uint32_t *s_dev, uint32_t *t_dev, void *stream);
/* GadgetDecomposition::decompose / recompose (decomposition.rs:47-93):
 * balanced base-b digits (a digit r > b/2 becomes r - b), [count][num_digits],
 * digit_bytes 1 (int8, b <= 255: feeds lzx_ajtai_commit_digits_batch) or 8
 * (int64).  Synchronize the stream. */
lzx_status lzx_gadget_decompose(const uint64_t *x_dev, size_t count, uint64_t base, uint32_t num_digits,
                                void *digits_dev, uint32_t digit_bytes, void *stream);
lzx_status lzx_gadget_recompose(const void *digits_dev, uint32_t digit_bytes, size_t count, uint64_t base,
                                uint32_t num_digits, uint64_t *x_dev, void *stream);

/* ---- the ~2^50 quadratic-slot ring (modulus50.rs:44-499) -------------------
 * QuadSlotTables / RingConfig50 over a prime q < 2^50 of two-adicity >= 7
 * (Q_50 = 2^50 - 2687, generator 3), n = 2^log_n with log_n in [1, 7].
 * Batched calls take len = (number of polynomials) x n canonical u64
 * residues in HBM, [poly][coeff]; len not a multiple of n is
 * LZX_MOD50_LENGTH_MISMATCH. */
typedef struct lzx_quad_plan lzx_quad_plan;

/* QuadSlotTables::new (modulus50.rs:241-300). */
lzx_status lzx_quad_plan_create(uint64_t q, uint32_t two_adicity, uint64_t generator, uint32_t log_n,
                                lzx_quad_plan **out);
void lzx_quad_plan_destroy(lzx_quad_plan *plan);
/* RingConfig50::slot_constants (modulus50.rs:446-449): the n/2 leaf gammas; and zeta. */
lzx_status lzx_quad_slot_constants(const lzx_quad_plan *plan, uint64_t *gamma_out, uint64_t *zeta_out);
/* QuadSlotTables::forward / inverse (modulus50.rs:317-375), in place. */
lzx_status lzx_quad_forward_batch(const lzx_quad_plan *plan, uint64_t *data_dev, size_t len, void *stream);
lzx_status lzx_quad_inverse_batch(const lzx_quad_plan *plan, uint64_t *data_dev, size_t len, void *stream);
/* QuadSlotTables::slotwise_mul (modulus50.rs:380-402) on slot-form inputs; out may alias a or b. */
lzx_status lzx_quad_slotwise_mul_batch(const lzx_quad_plan *plan, const uint64_t *a_dev, const uint64_t *b_dev,
                                       uint64_t *out_dev, size_t len, void *stream);
/* QuadSlotTables::mul / RingConfig50::mul (modulus50.rs:406-421, 451-454): ring products. */
lzx_status lzx_quad_mul_batch(const lzx_quad_plan *plan, const uint64_t *a_dev, const uint64_t *b_dev,
                              uint64_t *out_dev, size_t len, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LZX_GPU_H */

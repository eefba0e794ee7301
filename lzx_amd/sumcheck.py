"""Host-side mirror of lattice-sumcheck / lattice-core's prover surface.

  Transcript            crates/lattice-core/src/transcript.rs:36-238
  DenseMle              crates/lattice-core/src/mle.rs:12-260 (device tables)
  VirtualPolynomial     crates/lattice-sumcheck/src/virtual_poly.rs:48-160
  prove / verify        crates/lattice-sumcheck/src/sumcheck.rs:62-371
Factor tables live in HBM (torch int64 tensors holding canonical u64).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import device as dev
from ._lib import check, lib

GL_P = 0xFFFFFFFF00000001


class Transcript:
    """Transcript::new(protocol, budget) (transcript.rs:46)."""

    def __init__(self, protocol: bytes = b"", query_budget: int = 1 << 20, _h=None):
        self._destroy = lib().lzx_transcript_destroy  # bound now: lib() may be gone at interpreter exit
        if _h is not None:
            self._h = _h
            return
        self._h = C.c_void_p()
        check(lib().lzx_transcript_new(protocol, len(protocol), query_budget, C.byref(self._h)), "Transcript::new")

    @classmethod
    def new_default(cls, protocol: bytes) -> "Transcript":
        return cls(protocol, 1 << 20)

    def __del__(self):
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h and destroy:
            destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def clone(self) -> "Transcript":
        h = C.c_void_p()
        check(lib().lzx_transcript_clone(self._h, C.byref(h)))
        return Transcript(_h=h)

    def append_message(self, label: bytes, message: bytes) -> None:
        check(lib().lzx_transcript_append_message(self._h, label, len(label), message, len(message)),
              "Transcript::append_message")

    def append_field_slice(self, label: bytes, values) -> None:
        v = np.ascontiguousarray(values, dtype=np.uint64)
        check(lib().lzx_transcript_append_field_slice(self._h, label, len(label), v.ctypes.data, v.size),
              "Transcript::append_field_slice")

    def challenge_fields(self, label: bytes, n: int) -> np.ndarray:
        out = np.zeros(max(n, 1), dtype=np.uint64)
        check(lib().lzx_transcript_challenge_fields(self._h, label, len(label), n, out.ctypes.data),
              "Transcript::challenge_fields")
        return out[:n]

    def challenge_field(self, label: bytes) -> int:
        return int(self.challenge_fields(label, 1)[0])

    def query_count(self) -> int:
        return int(lib().lzx_transcript_query_count(self._h))


@dataclass
class VirtualPolynomial:
    """P = sum_t c_t prod_{f in ids_t} F_f over device factor tables."""

    num_vars: int
    factors: list = field(default_factory=list)  # torch int64 tensors, 2^num_vars entries
    terms: list = field(default_factory=list)    # (coeff, [factor ids])

    def add_factor(self, table) -> int:
        """Append a device table; None stands for an eq(w, .) factor that only
        prove_fast_with_eq supplies (never materialised)."""
        if table is not None and table.numel() != 1 << self.num_vars:
            raise ValueError("factor has the wrong number of evaluations")
        self.factors.append(table)
        return len(self.factors) - 1

    def add_term(self, coeff: int, ids) -> None:
        self.terms.append((int(coeff) % GL_P, list(ids)))

    def max_degree(self) -> int:
        return max((len(ids) for _, ids in self.terms), default=0)

    def _ffi(self):
        ptrs = (C.c_void_p * max(1, len(self.factors)))(*[t.data_ptr() if t is not None else None
                                                           for t in self.factors])
        coeffs = np.array([c for c, _ in self.terms] or [0], dtype=np.uint64)
        off = [0]
        ids = []
        for _, fs in self.terms:
            ids.extend(fs)
            off.append(len(ids))
        off = np.array(off, dtype=np.uint32)
        ids = np.array(ids or [0], dtype=np.uint32)
        return ptrs, coeffs, off, ids

    def sum_over_hypercube(self, stream=None) -> int:
        ptrs, coeffs, off, ids = self._ffi()
        out = C.c_uint64()
        check(lib().lzx_vp_sum_over_hypercube(self.num_vars, len(self.factors), ptrs, len(self.terms),
                                              coeffs.ctypes.data, off.ctypes.data, ids.ctypes.data, C.byref(out),
                                              dev.stream_handle(stream)), "sum_over_hypercube")
        return out.value


@dataclass
class SumcheckOutput:
    rounds: np.ndarray        # [num_vars][degree+1]
    challenges: np.ndarray    # [num_vars]
    final_claim: int
    factor_claims: np.ndarray  # [nfactors]


def _prove(vp: VirtualPolynomial, claim: int, transcript, challenges, stream):
    ptrs, coeffs, off, ids = vp._ffi()
    d = max(vp.max_degree(), 1)
    rounds = np.zeros(max(1, vp.num_vars * (d + 1)), dtype=np.uint64)
    chals = np.zeros(max(1, vp.num_vars), dtype=np.uint64)
    fclaims = np.zeros(max(1, len(vp.factors)), dtype=np.uint64)
    fin = C.c_uint64()
    deg = C.c_size_t()
    if transcript is not None:
        st = lib().lzx_sumcheck_prove(vp.num_vars, len(vp.factors), ptrs, len(vp.terms), coeffs.ctypes.data,
                                      off.ctypes.data, ids.ctypes.data, claim, transcript.handle,
                                      rounds.ctypes.data, chals.ctypes.data, C.byref(fin), fclaims.ctypes.data,
                                      C.byref(deg), dev.stream_handle(stream))
    else:
        given = np.ascontiguousarray(challenges, dtype=np.uint64)
        chals[: given.size] = given
        st = lib().lzx_sumcheck_prove_with_challenges(vp.num_vars, len(vp.factors), ptrs, len(vp.terms),
                                                      coeffs.ctypes.data, off.ctypes.data, ids.ctypes.data, claim,
                                                      given.ctypes.data, rounds.ctypes.data, C.byref(fin),
                                                      fclaims.ctypes.data, C.byref(deg), dev.stream_handle(stream))
    check(st, "sumcheck::prove")
    w = deg.value + 1
    # the zero polynomial's proof carries no factor claims (sumcheck.rs:86-91)
    return SumcheckOutput(rounds[: vp.num_vars * w].reshape(vp.num_vars, w), chals[: vp.num_vars], fin.value,
                          fclaims[: len(vp.factors) if vp.terms else 0])


def gl_mul_rate(waves_per_simd: int = 8) -> float:
    """Goldilocks products per second the device sustains (lzx_gl_mul_rate):
    the VALU ceiling the sum-check kernels are measured against."""
    out = C.c_double()
    check(lib().lzx_gl_mul_rate(waves_per_simd, C.byref(out)), "gl_mul_rate")
    return out.value


def prove(vp: VirtualPolynomial, claim: int, transcript: Transcript, stream=None) -> SumcheckOutput:
    """sumcheck::prove(&vp, claim, &mut transcript) (sumcheck.rs:62-162)."""
    return _prove(vp, claim, transcript, None, stream)


def prove_fast(vp: VirtualPolynomial, claim: int, transcript: Transcript, opts=None, stream=None) -> SumcheckOutput:
    """fastprover::prove_fast / prove_fast_with_opts (fastprover.rs:201-237): the
    reference's fast prover is transcript-identical to sumcheck::prove, so this
    is the device prover; the window / statistics options do not change the proof."""
    return _prove(vp, claim, transcript, None, stream)


def prove_fast_with_eq(vp: VirtualPolynomial, claim: int, transcript: Transcript, opts, eq_indices, w,
                       stream=None) -> SumcheckOutput:
    """fastprover::prove_fast_with_eq (fastprover.rs:212-224): the factors at
    eq_indices are eq(w, .) and are never read (they may be None)."""
    ptrs, coeffs, off, ids = vp._ffi()
    d = max(vp.max_degree(), 1)
    rounds = np.zeros(max(1, vp.num_vars * (d + 1)), dtype=np.uint64)
    chals = np.zeros(max(1, vp.num_vars), dtype=np.uint64)
    fclaims = np.zeros(max(1, len(vp.factors)), dtype=np.uint64)
    eqi = np.array(list(eq_indices) or [0], dtype=np.uintp)
    wv = np.ascontiguousarray(np.asarray(w, dtype=np.uint64).reshape(-1)[: vp.num_vars]) if len(w) else \
        np.zeros(1, dtype=np.uint64)
    if len(eq_indices) and wv.size < vp.num_vars:
        raise ValueError("w needs one coordinate per variable")
    fin = C.c_uint64()
    deg = C.c_size_t()
    check(lib().lzx_sumcheck_prove_eq(vp.num_vars, len(vp.factors), ptrs, len(vp.terms), coeffs.ctypes.data,
                                      off.ctypes.data, ids.ctypes.data, claim, transcript.handle, eqi.ctypes.data,
                                      len(eq_indices), wv.ctypes.data, rounds.ctypes.data, chals.ctypes.data,
                                      C.byref(fin), fclaims.ctypes.data, C.byref(deg), dev.stream_handle(stream)),
          "fastprover::prove_fast_with_eq")
    wd = deg.value + 1
    return SumcheckOutput(rounds[: vp.num_vars * wd].reshape(vp.num_vars, wd), chals[: vp.num_vars], fin.value,
                          fclaims[: len(vp.factors) if vp.terms else 0])


def prove_with_challenges(vp: VirtualPolynomial, claim: int, challenges, stream=None) -> SumcheckOutput:
    """The same prover with caller-supplied challenges (no Fiat-Shamir)."""
    return _prove(vp, claim, None, challenges, stream)


def verify(rounds, num_vars: int, max_degree: int, claim: int, transcript: Transcript, expected_final=None):
    """SumcheckProof::verify (sumcheck.rs:324-371): returns (point, final_claim)."""
    r = np.ascontiguousarray(rounds, dtype=np.uint64)
    nr = r.shape[0] if r.ndim == 2 else 0
    rl = r.shape[1] if r.ndim == 2 else 0
    point = np.zeros(max(1, nr), dtype=np.uint64)
    fin = C.c_uint64()
    check(lib().lzx_sumcheck_verify(r.ctypes.data, nr, rl, num_vars, max_degree, claim, transcript.handle,
                                    1 if expected_final is not None else 0, expected_final or 0, point.ctypes.data,
                                    C.byref(fin)), "SumcheckProof::verify")
    return point[:nr], fin.value


def eq_table(point, out=None, stream=None):
    """field_simd.rs:391-407 eq table on device (variable 0 = MSB)."""
    p = np.ascontiguousarray(point, dtype=np.uint64)
    if out is None:
        out = dev.empty(1 << p.size, np.uint64)
    check(lib().lzx_eq_table(p.ctypes.data, p.size, dev.ptr(out), dev.stream_handle(stream)), "eq_table")
    return out


@dataclass
class ZerocheckOutput:
    point: np.ndarray       # r, drawn as "zerocheck-point" (num_vars)
    output: SumcheckOutput  # the sum-check of P(x) eq(r, x) with claim 0


def zerocheck_prove(poly, num_vars: int, transcript: Transcript, stream=None) -> ZerocheckOutput:
    """zerocheck::prove(&poly, &mut transcript) (zerocheck.rs:26-54) on a device
    table of 2^num_vars canonical entries.  Raises LzxError with
    ZerocheckError::FinalClaimNonZero when P does not vanish on the hypercube."""
    if poly.numel() != 1 << num_vars:
        raise ValueError("poly has the wrong number of evaluations")
    point = np.zeros(max(1, num_vars), dtype=np.uint64)
    rounds = np.zeros(max(1, num_vars * 3), dtype=np.uint64)
    chals = np.zeros(max(1, num_vars), dtype=np.uint64)
    fclaims = np.zeros(2, dtype=np.uint64)
    fin = C.c_uint64()
    check(lib().lzx_zerocheck_prove(num_vars, dev.ptr(poly), transcript.handle, point.ctypes.data,
                                    rounds.ctypes.data, chals.ctypes.data, C.byref(fin), fclaims.ctypes.data,
                                    dev.stream_handle(stream)), "zerocheck::prove")
    out = SumcheckOutput(rounds[: num_vars * 3].reshape(num_vars, 3), chals[:num_vars], fin.value, fclaims)
    return ZerocheckOutput(point[:num_vars], out)


def zerocheck_verify(rounds, num_vars: int, transcript: Transcript, expected_poly_eval=None) -> np.ndarray:
    """zerocheck::verify (zerocheck.rs:58-84): returns the point r."""
    r = np.ascontiguousarray(rounds, dtype=np.uint64)
    nr = r.shape[0] if r.ndim == 2 else 0
    rl = r.shape[1] if r.ndim == 2 else 0
    point = np.zeros(max(1, num_vars), dtype=np.uint64)
    check(lib().lzx_zerocheck_verify(r.ctypes.data if r.size else None, nr, rl, num_vars, transcript.handle,
                                     1 if expected_poly_eval is not None else 0, expected_poly_eval or 0,
                                     point.ctypes.data), "zerocheck::verify")
    return point[:num_vars]


# ---- claim batching (crates/lattice-sumcheck/src/batch.rs:24-336) ------------------------

class _BatchClaimC(C.Structure):
    """lzx_batch_claim."""

    _fields_ = [("num_vars", C.c_size_t), ("nfactors", C.c_size_t), ("factors_dev", C.c_void_p),
                ("nterms", C.c_size_t), ("coeffs", C.c_void_p), ("term_off", C.c_void_p),
                ("term_ids", C.c_void_p), ("claimed_sum", C.c_uint64)]


@dataclass
class BatchClaim:
    """batch.rs:24-28: a virtual polynomial and its asserted sum."""

    poly: VirtualPolynomial
    claimed_sum: int


@dataclass
class PerClaimClaims:
    """batch.rs:115-126."""

    claimed_sum: int
    num_vars: int
    factor_claims: np.ndarray
    point: np.ndarray


@dataclass
class BatchStarOutput:
    """batch.rs:102-113: the single proof, the shared point, the rhos, the per-claim claims."""

    output: SumcheckOutput
    rhos: np.ndarray
    per_claim: list

    @property
    def challenges(self):
        return self.output.challenges

    @property
    def rounds(self):
        return self.output.rounds


def _batch_ffi(claims):
    arr = (_BatchClaimC * max(1, len(claims)))()
    keep = []
    for i, c in enumerate(claims):
        vp = c.poly
        if any(f is None for f in vp.factors):
            raise ValueError("batch claims need every factor table")
        ptrs, coeffs, off, ids = vp._ffi()
        keep += [ptrs, coeffs, off, ids]
        arr[i] = _BatchClaimC(vp.num_vars, len(vp.factors), C.cast(ptrs, C.c_void_p).value, len(vp.terms),
                              coeffs.ctypes.data, off.ctypes.data, ids.ctypes.data, int(c.claimed_sum) % GL_P)
    return arr, keep


def _prove_batch(claims, transcript, star, stream):
    arr, keep = _batch_ffi(claims)
    nv = max((c.poly.num_vars for c in claims), default=0)
    d = max([max(c.poly.max_degree(), 1) for c in claims] or [1])
    total_f = sum(len(c.poly.factors) for c in claims)
    rhos = np.zeros(max(1, len(claims)), dtype=np.uint64)
    rounds = np.zeros(max(1, nv * (d + 1)), dtype=np.uint64)
    chals = np.zeros(max(1, nv), dtype=np.uint64)
    fclaims = np.zeros(max(1, total_f), dtype=np.uint64)
    fin, deg, nvo = C.c_uint64(), C.c_size_t(), C.c_size_t(nv)
    if star:
        st = lib().lzx_sumcheck_prove_batch_star(arr, len(claims), transcript.handle, rhos.ctypes.data,
                                                 rounds.ctypes.data, chals.ctypes.data, C.byref(fin),
                                                 fclaims.ctypes.data, C.byref(nvo), C.byref(deg),
                                                 dev.stream_handle(stream))
    else:
        st = lib().lzx_sumcheck_prove_batch(arr, len(claims), transcript.handle, rhos.ctypes.data, rounds.ctypes.data,
                                            chals.ctypes.data, C.byref(fin), fclaims.ctypes.data, C.byref(deg),
                                            dev.stream_handle(stream))
    check(st, "batch::prove_batch_star" if star else "batch::prove_batch")
    del keep
    n, w = nvo.value, deg.value + 1
    has_terms = any(c.poly.terms for c in claims)
    out = SumcheckOutput(rounds[: n * w].reshape(n, w), chals[:n], fin.value,
                         fclaims[:total_f] if has_terms else np.zeros(0, dtype=np.uint64))
    return out, rhos[: len(claims)]


def prove_batch(claims, transcript: Transcript, stream=None):
    """batch::prove_batch(&claims, &mut t) (batch.rs:33-80): (SumcheckOutput, rhos);
    factor claims in global order (claim 0's factors first)."""
    return _prove_batch(claims, transcript, False, stream)


def prove_batch_star(claims, transcript: Transcript, stream=None) -> BatchStarOutput:
    """batch::prove_batch_star (batch.rs:161-225): claims of any num_vars, one
    sum-check over max num_vars rounds."""
    out, rhos = _prove_batch(claims, transcript, True, stream)
    per, base = [], 0
    for c in claims:
        nf = len(c.poly.factors)
        per.append(PerClaimClaims(int(c.claimed_sum) % GL_P, c.poly.num_vars, out.factor_claims[base: base + nf],
                                  out.challenges[: c.poly.num_vars]))
        base += nf
    return BatchStarOutput(out, rhos, per)


def _rounds_arg(rounds):
    r = np.ascontiguousarray(rounds, dtype=np.uint64)
    nr = r.shape[0] if r.ndim == 2 else 0
    rl = r.shape[1] if r.ndim == 2 else 0
    return r, nr, rl


def verify_batch(rounds, num_vars: int, max_degree: int, claimed_sums, transcript: Transcript,
                 expected_combined_final=None) -> np.ndarray:
    """batch::verify_batch (batch.rs:300-336): returns the point."""
    r, nr, rl = _rounds_arg(rounds)
    sums = np.array([int(x) % GL_P for x in claimed_sums] or [0], dtype=np.uint64)
    point = np.zeros(max(1, nr), dtype=np.uint64)
    check(lib().lzx_sumcheck_verify_batch(r.ctypes.data if r.size else None, nr, rl, num_vars, max_degree,
                                          sums.ctypes.data, len(claimed_sums), transcript.handle,
                                          1 if expected_combined_final is not None else 0,
                                          expected_combined_final or 0, point.ctypes.data), "batch::verify_batch")
    return point[:nr]


def verify_batch_star(rounds, num_vars, max_degree: int, claimed_sums, transcript: Transcript,
                      expected_combined_final=None, rhos=None) -> np.ndarray:
    """batch::verify_batch_star (batch.rs:236-262), or verify_batch_star_with_rhos
    (batch.rs:268-296) when rhos is given: returns the point."""
    r, nr, rl = _rounds_arg(rounds)
    nvs = np.array(list(num_vars) or [0], dtype=np.uintp)
    sums = np.array([int(x) % GL_P for x in claimed_sums] or [0], dtype=np.uint64)
    rh = None if rhos is None else np.array([int(x) for x in rhos] or [0], dtype=np.uint64)
    point = np.zeros(max(1, nr), dtype=np.uint64)
    check(lib().lzx_sumcheck_verify_batch_star(r.ctypes.data if r.size else None, nr, rl, nvs.ctypes.data,
                                               len(num_vars), max_degree, sums.ctypes.data, len(claimed_sums),
                                               rh.ctypes.data if rh is not None else None,
                                               0 if rhos is None else len(rhos), transcript.handle,
                                               1 if expected_combined_final is not None else 0,
                                               expected_combined_final or 0, point.ctypes.data),
          "batch::verify_batch_star")
    return point[:nr]


def verify_batch_star_with_rhos(rounds, num_vars, max_degree: int, claimed_sums, rhos, transcript: Transcript,
                                expected_combined_final=None) -> np.ndarray:
    """batch::verify_batch_star_with_rhos (batch.rs:268-296)."""
    return verify_batch_star(rounds, num_vars, max_degree, claimed_sums, transcript, expected_combined_final, rhos)


# ---- round-level session ------------------------------------------------------------

class SumcheckSession:
    """The prover's rounds one call at a time (sumcheck.rs:100-143), the
    transcript steps left to the caller: round_eval() gives g_j(0..d) for the
    current round, fold(r) binds its variable, final() the factor claims and
    P(r) once every variable is bound."""

    def __init__(self, vp: VirtualPolynomial, stream=None):
        if any(f is None for f in vp.factors):
            raise ValueError("a session needs every factor table")
        self._vp = vp  # the caller's tables are read until the first fold
        ptrs, coeffs, off, ids = vp._ffi()
        self._destroy = lib().lzx_sumcheck_session_destroy  # bound now: lib() may be gone at interpreter exit
        self._h = C.c_void_p()
        check(lib().lzx_sumcheck_session_create(vp.num_vars, len(vp.factors), ptrs, len(vp.terms), coeffs.ctypes.data,
                                                off.ctypes.data, ids.ctypes.data, dev.stream_handle(stream),
                                                C.byref(self._h)), "sumcheck session")
        self.width = int(lib().lzx_sumcheck_session_round_len(self._h))
        self.nfactors = len(vp.factors)

    def __del__(self):
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h and destroy:
            destroy(h)
            self._h = None

    def round_eval(self) -> np.ndarray:
        out = np.zeros(self.width, dtype=np.uint64)
        check(lib().lzx_sumcheck_session_round_eval(self._h, out.ctypes.data), "sumcheck session round_eval")
        return out

    def fold(self, r: int) -> None:
        check(lib().lzx_sumcheck_session_fold(self._h, int(r)), "sumcheck session fold")

    def final(self):
        """(factor_claims, P(r))."""
        claims = np.zeros(max(1, self.nfactors), dtype=np.uint64)
        fin = C.c_uint64()
        check(lib().lzx_sumcheck_session_final(self._h, claims.ctypes.data, C.byref(fin)), "sumcheck session final")
        return claims[: self.nfactors], fin.value


def interpolate_at(evals, r: int) -> int:
    """interpolate_at (sumcheck.rs:286-306): the degree-(len-1) polynomial through
    (i, evals[i]), i = 0.., evaluated at r, in the Goldilocks field."""
    n, r = len(evals), int(r) % GL_P
    acc = 0
    for i in range(n):
        num, den = 1, 1
        for j in range(n):
            if j != i:
                num = num * ((r - j) % GL_P) % GL_P
                den = den * ((i - j) % GL_P) % GL_P
        acc = (acc + int(evals[i]) * num % GL_P * pow(den, GL_P - 2, GL_P)) % GL_P
    return acc


def prove_with_session(vp: VirtualPolynomial, claim: int, transcript: Transcript, stream=None) -> SumcheckOutput:
    """sumcheck::prove (sumcheck.rs:62-162) driven round by round through a
    session: per round absorb "sumcheck-round", draw "sumcheck-challenge",
    check g(0) + g(1) against the running claim, interpolate, fold."""
    S = SumcheckSession(vp, stream)
    m = vp.num_vars
    rounds = np.zeros((m, S.width), dtype=np.uint64)
    chals = np.zeros(m, dtype=np.uint64)
    current = int(claim) % GL_P
    if not vp.terms and current != 0:
        check(13, "sumcheck::prove")  # SumcheckError::ClaimMismatch
    for j in range(m):
        ev = S.round_eval()
        transcript.append_field_slice(b"sumcheck-round", ev)
        r = transcript.challenge_field(b"sumcheck-challenge")
        if (int(ev[0]) + int(ev[1])) % GL_P != current:
            check(13, "sumcheck::prove")  # SumcheckError::ClaimMismatch
        current = interpolate_at(ev, r)
        S.fold(r)
        rounds[j], chals[j] = ev, r
    claims, fin = S.final() if vp.terms else (np.zeros(0, dtype=np.uint64), 0)
    if vp.terms and fin != current:
        check(12, "sumcheck::prove")  # SumcheckError::FinalCheckFailed
    return SumcheckOutput(rounds, chals, fin, claims)

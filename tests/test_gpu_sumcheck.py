"""GPU parity of the Goldilocks sum-check prover and its host transcript.

Mirrors crates/lattice-sumcheck/src/sumcheck.rs:402-550 and
crates/lattice-core/src/transcript.rs:244-316, against the oracle's
restatement (same transcript flow) and the brute-force golden vectors.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
VEC = json.loads((Path(__file__).resolve().parent / "golden" / "vectors.json").read_text())
P = O.GL_P


@pytest.fixture(scope="module")
def sc():
    from lzx_amd import device as D
    from lzx_amd import sumcheck as S
    D.require_gpu()
    return S, D


def _vp(S, D, nv, tables, terms):
    vp = S.VirtualPolynomial(nv)
    for t in tables:
        vp.add_factor(D.to_device(np.ascontiguousarray(t, dtype=np.uint64)))
    for c, ids in terms:
        vp.add_term(c, ids)
    return vp


def test_transcript_matches_oracle(sc):
    S, D = sc
    rng = np.random.default_rng(1)
    a, b = S.Transcript(b"lzx-proto"), O.Transcript(b"lzx-proto")
    for step in range(40):
        op = step % 4
        if op == 0:
            msg = rng.integers(0, 256, int(rng.integers(0, 300)), dtype=np.uint8).tobytes()
            a.append_message(b"lbl", msg)
            b.append_message(b"lbl", msg)
        elif op == 1:
            v = rng.integers(0, P, 7, dtype=np.uint64)
            a.append_field_slice(b"fs", v)
            b.append_field_slice(b"fs", v)
        else:
            n = int(rng.integers(1, 20))
            assert a.challenge_fields(b"chal", n).tolist() == b.challenge_fields(b"chal", n)[1].tolist()
    from lzx_amd._lib import LzxError
    t = S.Transcript(b"budget", 2)
    t.challenge_field(b"a")
    t.challenge_field(b"b")
    with pytest.raises(LzxError) as e:
        t.challenge_field(b"c")
    assert e.value.code == 20


def test_prove_with_challenges_matches_brute_force(sc):
    S, D = sc
    for v in VEC["sumcheck"]:
        vp = _vp(S, D, v["num_vars"], v["factors"], [(c, fs) for c, fs in v["terms"]])
        assert vp.sum_over_hypercube() == v["claim"]
        out = S.prove_with_challenges(vp, v["claim"], v["challenges"])
        assert out.rounds.tolist() == v["rounds"]
        assert out.factor_claims.tolist() == v["factor_claims"]


SHAPES = [
    # (num_vars, nfactors, terms)
    (1, 3, [(3, [0, 1]), (5, [1, 2, 0]), (11, [2])]),
    (5, 3, [(3, [0, 1]), (5, [1, 2, 0]), (11, [2])]),
    (10, 4, [(1, [0, 1, 2]), (P - 1, [0, 3])]),            # eq*(a*b - c)
    (9, 2, [(1, [0, 1])]),                                   # zerocheck shape
    (11, 5, [(7, [0, 0, 1]), (P - 2, [2, 3, 4, 1]), (1, [4])]),  # repeated factor, generic kernel
    (12, 1, [(1, [0])]),
    # the compile-time shapes of the kernels, with fold rounds (num_vars >= 13)
    (14, 4, [(1, [0, 1, 2]), (P - 1, [0, 3])]),             # eq*(a*b - c)
    (14, 2, [(1, [0, 1])]),                                   # f*e
    (3, 4, [(5, [0, 1]), (P - 9, [2, 3])]),                  # rho_0 f_0 e_0 + rho_1 f_1 e_1
    (15, 4, [(12345678901, [0, 1]), (1, [3, 2])]),
    (13, 4, [(1, [1, 2, 0]), (P - 1, [3, 0])]),              # eq*(a*b - c), factors in another order
    (13, 4, [(1, [0, 1, 2]), (P - 1, [0, 2])]),              # -c's factor also in the other term: generic
]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"nv{s[0]}_f{s[1]}_t{len(s[2])}")
def test_prove_matches_oracle_transcript(sc, shape):
    S, D = sc
    nv, nf, terms = shape
    rng = np.random.default_rng(nv * 31 + nf)
    tables = [rng.integers(0, P, 1 << nv, dtype=np.uint64) for _ in range(nf)]
    claim = O.vp_sum_over_hypercube(nv, tables, terms)
    vp = _vp(S, D, nv, tables, terms)
    assert vp.sum_over_hypercube() == claim
    out = S.prove(vp, claim, S.Transcript(b"lzx-sumcheck-test"))
    st, want = O.sumcheck_prove(nv, tables, terms, claim, O.Transcript(b"lzx-sumcheck-test"))
    assert st == 0
    assert out.rounds.tolist() == want["rounds"].tolist()
    assert out.challenges.tolist() == want["challenges"].tolist()
    assert out.final_claim == want["final_claim"]
    assert out.factor_claims.tolist() == want["factor_claims"].tolist()
    # the factor tables are untouched (the reference clones them)
    for t, dt in zip(tables, vp.factors):
        assert np.array_equal(D.to_host(dt, np.uint64), t)
    point, fin = S.verify(out.rounds, nv, vp.max_degree(), claim, S.Transcript(b"lzx-sumcheck-test"),
                          out.final_claim)
    assert point.tolist() == out.challenges.tolist() and fin == out.final_claim


def test_errors_and_zero_polynomial(sc):
    S, D = sc
    from lzx_amd._lib import LzxError
    nv = 6
    rng = np.random.default_rng(3)
    tables = [rng.integers(0, P, 1 << nv, dtype=np.uint64) for _ in range(3)]
    terms = [(3, [0, 1]), (5, [1, 2, 0]), (11, [2])]
    vp = _vp(S, D, nv, tables, terms)
    claim = vp.sum_over_hypercube()
    with pytest.raises(LzxError) as e:
        S.prove(vp, (claim + 1) % P, S.Transcript(b"t"))
    assert e.value.code == 13
    out = S.prove(vp, claim, S.Transcript(b"t"))
    with pytest.raises(LzxError) as e:
        S.verify(out.rounds, nv, 3, (claim + 1) % P, S.Transcript(b"t"))
    assert e.value.code == 11
    with pytest.raises(LzxError) as e:
        S.verify(out.rounds, nv, 3, claim, S.Transcript(b"t"), (out.final_claim + 1) % P)
    assert e.value.code == 12
    with pytest.raises(LzxError) as e:
        S.verify(out.rounds[:-1], nv, 3, claim, S.Transcript(b"t"))
    assert e.value.code == 10
    with pytest.raises(LzxError) as e:
        S.verify(out.rounds, nv, 3, P, S.Transcript(b"t"))
    assert e.value.code == 30
    z = S.prove(S.VirtualPolynomial(3), 0, S.Transcript(b"t"))
    assert z.rounds.shape == (3, 2) and not z.rounds.any()
    S.verify(z.rounds, 3, 1, 0, S.Transcript(b"t"), 0)
    # non-canonical Goldilocks words (p encodes the same element as 0) are
    # refused before anything is absorbed: round values, claim, final claim
    for i in range(z.rounds.size):
        r2 = z.rounds.copy()
        r2.reshape(-1)[i] = P
        with pytest.raises(LzxError) as e:
            S.verify(r2, 3, 1, 0, S.Transcript(b"t"), 0)
        assert e.value.code == 30
    for claim_, fin_ in [(P, 0), (0, P)]:
        with pytest.raises(LzxError) as e:
            S.verify(z.rounds, 3, 1, claim_, S.Transcript(b"t"), fin_)
        assert e.value.code == 30
    bad = S.VirtualPolynomial(nv)
    bad.add_factor(D.to_device(tables[0]))
    bad.add_term(1, [])
    with pytest.raises(LzxError) as e:
        S.prove(bad, 0, S.Transcript(b"t"))
    assert e.value.code == 14


def test_eq_table_matches_oracle(sc):
    S, D = sc
    rng = np.random.default_rng(8)
    for m in (0, 1, 5, 13):
        pt = rng.integers(0, P, m, dtype=np.uint64)
        assert np.array_equal(D.to_host(S.eq_table(pt), np.uint64), O.eq_table(pt))


def _device_prove(S, D, nv, protocol, prior, tables, terms, budget=1 << 20):
    """A device prove (its transcript runs in the close and tail kernels)
    after `prior` host appends; returns (output, transcript afterwards)."""
    t = S.Transcript(protocol, budget)
    for label, msg in prior:
        t.append_message(label, msg)
    vp = _vp(S, D, nv, tables, terms)
    return S.prove(vp, O.vp_sum_over_hypercube(nv, tables, terms), t), t


def test_reference_transcript_properties_through_the_device_sponge(sc):
    """transcript.rs:244-316's five properties, on the sponge the device runs
    between rounds: 14 variables use both the per-round close kernel (halves
    above 2048) and the one-block tail, and the device hands its state back to
    the host transcript, which continues it."""
    S, D = sc
    from lzx_amd._lib import LzxError
    nv = 14
    rng = np.random.default_rng(244)
    tables = [rng.integers(0, P, 1 << nv, dtype=np.uint64) for _ in range(2)]
    terms = [(1, [0, 1]), (9, [1])]
    base, t_base = _device_prove(S, D, nv, b"test-proto", [(b"a", b"hello")], tables, terms)
    # deterministic_and_label_sensitive: the same history gives the same proof;
    # another label or protocol changes every challenge
    again, _ = _device_prove(S, D, nv, b"test-proto", [(b"a", b"hello")], tables, terms)
    assert again.rounds.tolist() == base.rounds.tolist() and again.challenges.tolist() == base.challenges.tolist()
    for proto, prior in [(b"test-proto", [(b"other", b"hello")]), (b"other-proto", [(b"a", b"hello")])]:
        o, _ = _device_prove(S, D, nv, proto, prior, tables, terms)
        assert o.rounds[0].tolist() == base.rounds[0].tolist()  # round 0 precedes any challenge
        assert all(x != y for x, y in zip(o.challenges.tolist(), base.challenges.tolist()))
    # label_concatenation_unambiguous: ("ab", "c") and ("a", "bc") frame differently
    ab, _ = _device_prove(S, D, nv, b"amb", [(b"ab", b"c")], tables, terms)
    a_bc, _ = _device_prove(S, D, nv, b"amb", [(b"a", b"bc")], tables, terms)
    assert all(x != y for x, y in zip(ab.challenges.tolist(), a_bc.challenges.tolist()))
    # resumable_squeeze_keeps_isolation: after the device rounds the host keeps
    # squeezing the same sponge; a fresh replay reproduces the sequence, and
    # so does the oracle's transcript after its own prove
    seq = [t_base.challenge_field(lbl) for lbl in (b"r1", b"r2", b"r3")]
    assert len(set(seq)) == 3
    _, t_again = _device_prove(S, D, nv, b"test-proto", [(b"a", b"hello")], tables, terms)
    assert [t_again.challenge_field(lbl) for lbl in (b"r1", b"r2", b"r3")] == seq
    ot = O.Transcript(b"test-proto")
    ot.append_message(b"a", b"hello")
    O.sumcheck_prove(nv, tables, terms, O.vp_sum_over_hypercube(nv, tables, terms), ot)
    assert [ot.challenge_field(lbl)[1] for lbl in (b"r1", b"r2", b"r3")] == seq
    # query_budget_enforced: one round's challenge past the budget fails on the device
    with pytest.raises(LzxError) as e:
        _device_prove(S, D, nv, b"budget", [], tables, terms, budget=nv - 1)
    assert e.value.code == 20
    _, t_full = _device_prove(S, D, nv, b"budget", [], tables, terms, budget=nv)
    with pytest.raises(LzxError) as e:
        t_full.challenge_field(b"d")
    assert e.value.code == 20
    # many_challenges_distinct_and_canonical
    ch = base.challenges.tolist() + ab.challenges.tolist()
    assert len(set(ch)) == len(ch) and max(ch) < P
    many = S.Transcript(b"many").challenge_fields(b"vec", 1000).tolist()
    assert len(many) == 1000 and len(set(many)) == 1000 and max(many) < P


def test_full_workload_matches_oracle(sc):
    """2^24-entry tables, VP eq(r,x)*(a*b - c), degree 3, transcript
    challenges: every round message, challenge and factor claim bit-exact
    against the oracle."""
    S, D = sc
    from lzx_amd import workloads as W
    nv = 24
    N = 1 << nv
    a = W.uniform_gl(0xC4A, N)
    b = W.uniform_gl(0xC4B, N)
    c = W.uniform_gl(0xC4C, N)
    r = W.uniform_gl(0xC4E, nv)
    eq_dev = S.eq_table(r)
    eq = D.to_host(eq_dev, np.uint64)
    terms = [(1, [0, 1, 2]), (P - 1, [0, 3])]
    vp = S.VirtualPolynomial(nv)
    vp.add_factor(eq_dev)
    for t in (a, b, c):
        vp.add_factor(D.to_device(t))
    for cf, ids in terms:
        vp.add_term(cf, ids)
    claim = vp.sum_over_hypercube()
    assert claim == O.vp_sum_over_hypercube(nv, [eq, a, b, c], terms)
    out = S.prove(vp, claim, S.Transcript(b"lzx-mi355x-sumcheck"))
    st, want = O.sumcheck_prove(nv, [eq, a, b, c], terms, claim, O.Transcript(b"lzx-mi355x-sumcheck"))
    assert st == 0
    assert out.rounds.tolist() == want["rounds"].tolist()
    assert out.challenges.tolist() == want["challenges"].tolist()
    assert out.factor_claims.tolist() == want["factor_claims"].tolist()
    assert out.final_claim == want["final_claim"]


@pytest.mark.parametrize("nv", [1, 4, 11, 14])
def test_zerocheck_matches_oracle(sc, nv):
    # zerocheck.rs:26-84: point, rounds, challenges, claims and the transcript
    # state after the proof, across the close-kernel and tail paths
    S, D = sc
    z = np.zeros(1 << nv, dtype=np.uint64)
    t_dev, t_or = S.Transcript(b"lzx-zerocheck-test"), O.Transcript(b"lzx-zerocheck-test")
    t_dev.append_message(b"ctx", b"zc")
    t_or.append_message(b"ctx", b"zc")
    got = S.zerocheck_prove(D.to_device(z), nv, t_dev)
    st, r, want = O.zerocheck_prove(nv, z, t_or)
    assert st == 0
    assert got.point.tolist() == r.tolist()
    assert got.output.rounds.tolist() == want["rounds"].tolist()
    assert got.output.challenges.tolist() == want["challenges"].tolist()
    assert got.output.factor_claims.tolist() == want["factor_claims"].tolist()
    assert got.output.final_claim == want["final_claim"]
    assert t_dev.challenge_field(b"ident") == t_or.challenge_field(b"ident")[1]
    # the device proof verifies (zerocheck::verify), expected P(r') = 0
    tv = S.Transcript(b"lzx-zerocheck-test")
    tv.append_message(b"ctx", b"zc")
    assert S.zerocheck_verify(got.output.rounds, nv, tv, 0).tolist() == r.tolist()


def test_zerocheck_refuses_nonzero_and_rejects_tampering(sc):
    from lzx_amd._lib import LzxError
    S, D = sc
    nv = 6
    p = np.zeros(1 << nv, dtype=np.uint64)
    p[3] = 1
    with pytest.raises(LzxError) as e:
        S.zerocheck_prove(D.to_device(p), nv, S.Transcript(b"lzx-zerocheck-test"))
    assert e.value.code == O.ZEROCHECK_FINAL_CLAIM_NONZERO
    out = S.zerocheck_prove(D.to_device(np.zeros(1 << nv, dtype=np.uint64)), nv, S.Transcript(b"lzx-zerocheck-test"))
    bad = out.output.rounds.copy()
    bad[0, 0] = 1
    with pytest.raises(LzxError):
        S.zerocheck_verify(bad, nv, S.Transcript(b"lzx-zerocheck-test"), None)


# ---- fastprover::prove_fast_with_eq (fastprover.rs:212-224) ----------------------
# The reference's fast prover is transcript-identical to sumcheck::prove with
# the eq factor materialised (tests/fastprover.rs:140-170), so the oracle is the
# oracle prover over eq_table(w).

EQ_SHAPES = [
    # (num_vars, nfactors, eq index, terms); ids refer to factors, eq appears once per term -> split
    (1, 2, 1, [(1, [0, 1])]),
    (7, 2, 1, [(1, [0, 1])]),
    (12, 2, 1, [(1, [0, 1])]),
    (16, 2, 1, [(1, [0, 1])]),
    (5, 4, 0, [(1, [0, 1, 2]), (P - 1, [0, 3])]),
    (13, 4, 0, [(1, [0, 1, 2]), (P - 1, [0, 3])]),
    (18, 4, 0, [(1, [0, 1, 2]), (P - 1, [0, 3])]),
    (10, 3, 1, [(3, [0, 1, 2]), (5, [1])]),                 # eq alone in a term: degree-0 remainder
    (9, 5, 2, [(7, [0, 2, 1, 1]), (P - 2, [3, 2, 4])]),     # a squared factor beside eq, NF = 8 kernels
    # eq twice in a term or missing from a term: materialised from w
    (8, 3, 1, [(1, [0, 1, 1])]),
    (8, 3, 1, [(1, [0, 1]), (1, [2])]),
]


@pytest.mark.parametrize("shape", EQ_SHAPES, ids=lambda s: f"nv{s[0]}_f{s[1]}_eq{s[2]}_{len(s[3])}t")
def test_prove_fast_with_eq_matches_oracle(sc, shape):
    S, D = sc
    nv, nf, e, terms = shape
    rng = np.random.default_rng(1000 + nv * 7 + nf)
    w = rng.integers(0, P, nv, dtype=np.uint64)
    tabs = [rng.integers(0, P, 1 << nv, dtype=np.uint64) for _ in range(nf)]
    tabs[e] = O.eq_table(w)
    claim = O.vp_sum_over_hypercube(nv, tabs, terms)
    t_dev, t_or = S.Transcript(b"sc-eq"), O.Transcript(b"sc-eq")
    st, want = O.sumcheck_prove(nv, tabs, terms, claim, t_or)
    assert st == 0
    vp = S.VirtualPolynomial(nv)
    for f, t in enumerate(tabs):
        vp.add_factor(None if f == e else D.to_device(t))  # the eq table is never handed over
    for c, ids in terms:
        vp.add_term(c, ids)
    got = S.prove_fast_with_eq(vp, claim, t_dev, None, [e], w)
    assert got.rounds.tolist() == want["rounds"].tolist()
    assert got.challenges.tolist() == want["challenges"].tolist()
    assert got.factor_claims.tolist() == want["factor_claims"].tolist()
    assert got.final_claim == want["final_claim"]
    assert t_dev.challenge_field(b"ident") == t_or.challenge_field(b"ident")[1]


def test_prove_fast_equals_prove_and_rejects_bad_claims(sc):
    from lzx_amd._lib import LzxError
    S, D = sc
    nv = 9
    rng = np.random.default_rng(77)
    w = rng.integers(0, P, nv, dtype=np.uint64)
    a = rng.integers(0, P, 1 << nv, dtype=np.uint64)
    tabs = [a, O.eq_table(w)]
    vp = _vp(S, D, nv, tabs, [(1, [0, 1])])
    claim = O.vp_sum_over_hypercube(nv, tabs, [(1, [0, 1])])
    x = S.prove(vp, claim, S.Transcript(b"pf"))
    y = S.prove_fast(vp, claim, S.Transcript(b"pf"))
    z = S.prove_fast_with_eq(vp, claim, S.Transcript(b"pf"), None, [1], w)
    for o in (y, z):
        assert o.rounds.tolist() == x.rounds.tolist() and o.factor_claims.tolist() == x.factor_claims.tolist()
    with pytest.raises(LzxError) as e:  # a wrong claim fails the first round check (sumcheck.rs:120-123)
        S.prove_fast_with_eq(vp, (claim + 1) % P, S.Transcript(b"pf"), None, [1], w)
    assert e.value.code == 13
    with pytest.raises(LzxError) as e:  # w must be canonical field elements
        S.prove_fast_with_eq(vp, claim, S.Transcript(b"pf"), None, [1], np.full(nv, P, dtype=np.uint64))
    assert e.value.code == 30


@pytest.mark.parametrize("nv", [3, 12, 17])
def test_zerocheck_nonzero_table_with_vanishing_point(sc, nv):
    # r is fixed by the transcript before P is seen, so a P that is non-zero on
    # the hypercube but has P(r) = 0 passes zerocheck::prove's refusal check
    # (zerocheck.rs:37-45) and yields non-trivial rounds; one entry off and
    # both provers refuse.  Pins the device's eq-free evaluation of P(r).
    S, D = sc
    probe = O.Transcript(b"zc-vanish")
    st, r = probe.challenge_fields(b"zerocheck-point", nv)
    assert st == 0
    eq = O.eq_table(r)
    rng = np.random.default_rng(nv)
    p = rng.integers(0, P, 1 << nv, dtype=np.uint64)
    k = int(rng.integers(0, 1 << nv))
    p[k] = 0
    rest = O.vp_sum_over_hypercube(nv, [p, eq], [(1, [0, 1])])
    p[k] = (P - rest) * O.gl_inverse(int(eq[k])) % P
    assert O.mle_evaluate(p, r) == 0
    t_dev, t_or = S.Transcript(b"zc-vanish"), O.Transcript(b"zc-vanish")
    got = S.zerocheck_prove(D.to_device(p), nv, t_dev)
    st, r2, want = O.zerocheck_prove(nv, p, t_or)
    assert st == 0 and got.point.tolist() == r2.tolist() == r.tolist()
    assert want["rounds"].any()
    assert got.output.rounds.tolist() == want["rounds"].tolist()
    assert got.output.factor_claims.tolist() == want["factor_claims"].tolist()
    assert got.output.final_claim == want["final_claim"]
    assert t_dev.challenge_field(b"ident") == t_or.challenge_field(b"ident")[1]
    from lzx_amd._lib import LzxError
    p[k] = (int(p[k]) + 1) % P
    st, _, _ = O.zerocheck_prove(nv, p, O.Transcript(b"zc-vanish"))
    assert st == O.ZEROCHECK_FINAL_CLAIM_NONZERO
    with pytest.raises(LzxError) as e:
        S.zerocheck_prove(D.to_device(p), nv, S.Transcript(b"zc-vanish"))
    assert e.value.code == O.ZEROCHECK_FINAL_CLAIM_NONZERO


# ---- round-level session -------------------------------------------------------------

SESSION_SHAPES = [(1, 2, [(1, [0, 1])]), (5, 3, [(3, [0, 1]), (5, [1, 2, 0]), (11, [2])]),
                  (13, 4, [(1, [0, 1, 2]), (P - 1, [0, 3])]), (16, 2, [(7, [0, 0, 1])]),
                  (9, 8, [(i + 1, [i, (i + 3) % 8]) for i in range(8)])]


@pytest.mark.parametrize("nv,nf,terms", SESSION_SHAPES)
def test_session_rounds_match_oracle(sc, nv, nf, terms):
    S, D = sc
    rng = np.random.default_rng(nv * 7 + nf)
    tables = [rng.integers(0, P, 1 << nv, dtype=np.uint64) for _ in range(nf)]
    claim = O.vp_sum_over_hypercube(nv, tables, terms)
    st, ref = O.sumcheck_prove(nv, tables, terms, claim, O.Transcript(b"lzx-session"))
    assert st == 0
    sess = S.SumcheckSession(_vp(S, D, nv, tables, terms))
    for j in range(nv):
        assert sess.round_eval().tolist() == ref["rounds"][j].tolist(), j
        sess.fold(int(ref["challenges"][j]))
    claims, fin = sess.final()
    assert claims.tolist() == ref["factor_claims"].tolist() and fin == ref["final_claim"]


@pytest.mark.parametrize("nv,nf,terms", SESSION_SHAPES[:4] + [(0, 2, [(4, [0, 1])]), (4, 2, [])])
def test_prove_with_session_equals_device_transcript_prove(sc, nv, nf, terms):
    S, D = sc
    rng = np.random.default_rng(nv + 100)
    tables = [rng.integers(0, P, 1 << nv, dtype=np.uint64) for _ in range(nf)]
    claim = O.vp_sum_over_hypercube(nv, tables, terms) if terms else 0
    ta, tb = S.Transcript(b"lzx-session-eq"), S.Transcript(b"lzx-session-eq")
    a = S.prove_with_session(_vp(S, D, nv, tables, terms), claim, ta)
    b = S.prove(_vp(S, D, nv, tables, terms), claim, tb)
    assert a.rounds.tolist() == b.rounds.tolist() and a.challenges.tolist() == b.challenges.tolist()
    assert a.final_claim == b.final_claim and a.factor_claims.tolist() == b.factor_claims.tolist()
    assert ta.challenge_field(b"next") == tb.challenge_field(b"next")  # both transcripts end in one state


def test_session_errors(sc):
    from lzx_amd._lib import LzxError
    S, D = sc
    rng = np.random.default_rng(3)
    tables = [rng.integers(0, P, 8, dtype=np.uint64) for _ in range(2)]
    sess = S.SumcheckSession(_vp(S, D, 3, tables, [(1, [0, 1])]))
    with pytest.raises(LzxError) as e:  # variables left to bind
        sess.final()
    assert e.value.code == 10
    for r in (5, 6, 7):
        sess.round_eval()
        sess.fold(r)
    for call in (sess.round_eval, lambda: sess.fold(1)):
        with pytest.raises(LzxError) as e:
            call()
        assert e.value.code == 10
    with pytest.raises(LzxError) as e:  # a challenge outside the field
        S.SumcheckSession(_vp(S, D, 3, tables, [(1, [0, 1])])).fold(P)
    assert e.value.code == 30
    tampered = 1 + O.vp_sum_over_hypercube(3, tables, [(1, [0, 1])])
    with pytest.raises(LzxError) as e:  # sumcheck.rs:120-123
        S.prove_with_session(_vp(S, D, 3, tables, [(1, [0, 1])]), tampered, S.Transcript(b"x"))
    assert e.value.code == 13


def test_prove_graph_replays_match_oracle(sc):
    """From the second prove with the same tables the launches replay from a
    captured graph; every replay, interleaved with other proves and host
    reads, gives the oracle's proof."""
    S, D = sc
    cases = []
    for nv, terms in ((15, [(1, [0, 1, 2]), (P - 1, [0, 3])]), (9, [(3, [0, 1]), (5, [1, 2, 0]), (11, [2])])):
        nf = 1 + max(max(ids) for _, ids in terms)
        rng = np.random.default_rng(7000 + nv)
        tables = [rng.integers(0, P, 1 << nv, dtype=np.uint64) for _ in range(nf)]
        claim = O.vp_sum_over_hypercube(nv, tables, terms)
        st, want = O.sumcheck_prove(nv, tables, terms, claim, O.Transcript(b"sc-graph"))
        assert st == 0
        cases.append((_vp(S, D, nv, tables, terms), claim, want, tables))
    for rep in range(4):
        for vp, claim, want, tables in cases:
            out = S.prove(vp, claim, S.Transcript(b"sc-graph"))
            assert out.rounds.tolist() == want["rounds"].tolist(), rep
            assert out.challenges.tolist() == want["challenges"].tolist(), rep
            assert out.factor_claims.tolist() == want["factor_claims"].tolist(), rep
            assert out.final_claim == want["final_claim"], rep
            assert np.array_equal(D.to_host(vp.factors[0], np.uint64), tables[0])  # inputs untouched
    # a wrong claim on a replayed shape is still refused
    from lzx_amd._lib import LzxError
    vp, claim, _, _ = cases[0]
    with pytest.raises(LzxError) as e:
        S.prove(vp, (claim + 1) % P, S.Transcript(b"sc-graph"))
    assert e.value.code == 13  # SumcheckError::ClaimMismatch


def test_gl_mul_rate_probe(sc):
    S, _ = sc
    from lzx_amd._lib import LzxError
    rate = S.gl_mul_rate(8)
    assert 1e10 < rate < 1e14  # products per second on one device
    with pytest.raises(LzxError):
        S.gl_mul_rate(3)


# crates/lattice-memory/src/onehot_check.rs:143-159: the raf-evaluation
# sum-check is one term eq_j * w * ra_1 * ... * ra_d (2 + d factors, degree 2 + d)
@pytest.mark.parametrize("nv,d", [(6, 7), (14, 7), (13, 13), (11, 10)])
def test_onehot_raf_shape_matches_oracle(sc, nv, d):
    S, D = sc
    rng = np.random.default_rng(31 * nv + d)
    nf = 2 + d
    tables = [rng.integers(0, P, 1 << nv, dtype=np.uint64) for _ in range(nf)]
    tables[0] = O.eq_table(rng.integers(0, P, nv, dtype=np.uint64))
    terms = [(1, list(range(nf)))]
    claim = O.vp_sum_over_hypercube(nv, tables, terms)
    vp = _vp(S, D, nv, tables, terms)
    assert vp.sum_over_hypercube() == claim
    out = S.prove(vp, claim, S.Transcript(b"onehot-raf"))
    st, want = O.sumcheck_prove(nv, tables, terms, claim, O.Transcript(b"onehot-raf"))
    assert st == 0
    assert out.rounds.tolist() == want["rounds"].tolist()
    assert out.challenges.tolist() == want["challenges"].tolist()
    assert out.factor_claims.tolist() == want["factor_claims"].tolist()
    assert out.final_claim == want["final_claim"]


def test_virtual_polynomial_beyond_the_device_limits_is_refused(sc):
    S, D = sc
    from lzx_amd._lib import LzxError
    nv = 3
    rng = np.random.default_rng(5)
    for nf, terms in ((17, [(1, list(range(17)))]), (2, [(1, [0] * 9 + [1] * 7)])):  # 17 factors; degree 16
        tables = [rng.integers(0, P, 1 << nv, dtype=np.uint64) for _ in range(nf)]
        vp = _vp(S, D, nv, tables, terms)
        t = S.Transcript(b"too-wide")
        with pytest.raises(LzxError) as e:
            S.prove(vp, 0, t)
        assert e.value.code == 31  # LZX_UNSUPPORTED
        # refused before anything ran: the caller's transcript is untouched (a CPU fallback continues from it)
        assert t.challenge_field(b"after") == S.Transcript(b"too-wide").challenge_field(b"after")
        tb = S.Transcript(b"too-wide")
        with pytest.raises(LzxError) as e:
            S.prove_batch([S.BatchClaim(vp, 0), S.BatchClaim(vp, 0)], tb)
        assert e.value.code == 31
        assert tb.challenge_field(b"after") == S.Transcript(b"too-wide").challenge_field(b"after")

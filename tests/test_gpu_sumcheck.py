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
    z = S.prove(S.VirtualPolynomial(3), 0, S.Transcript(b"t"))
    assert z.rounds.shape == (3, 2) and not z.rounds.any()
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

"""GPU parity of claim batching (crates/lattice-sumcheck/src/batch.rs:33-336):
the device prove_batch / prove_batch_star against the oracle's restatement
(same transcript flow; pinned in test_oracle.py by the reference's cases and
the per-claim round closed form), and the reference's own batch tests."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
P = O.GL_P


@pytest.fixture(scope="module")
def sc():
    from lzx_amd import device as D
    from lzx_amd import sumcheck as S
    D.require_gpu()
    return S, D


def _spec(nv, rng, nf=2, terms=((2, [0, 1]), (9, [0]))):
    return (nv, [rng.integers(0, P, 1 << nv, dtype=np.uint64) for _ in range(nf)], [(c, list(i)) for c, i in terms])


def _device_claims(S, D, specs, shared=None):
    """BatchClaims over device copies; `shared` maps (claim, factor) -> (claim, factor) whose tensor to reuse."""
    tens, claims = {}, []
    for ci, (nv, facs, terms) in enumerate(specs):
        vp = S.VirtualPolynomial(nv)
        for fi, f in enumerate(facs):
            src = (shared or {}).get((ci, fi))
            t = tens[src] if src is not None else D.to_device(f)
            tens[(ci, fi)] = t
            vp.add_factor(t)
        for c, ids in terms:
            vp.add_term(c, ids)
        claims.append(S.BatchClaim(vp, O.vp_sum_over_hypercube(nv, facs, terms)))
    return claims


def _oracle_claims(specs):
    return [(nv, facs, terms, O.vp_sum_over_hypercube(nv, facs, terms)) for nv, facs, terms in specs]


@pytest.mark.parametrize("nv,k", [(1, 1), (5, 3), (11, 2), (16, 4)])
def test_prove_batch_matches_oracle(sc, nv, k):
    S, D = sc
    rng = np.random.default_rng(nv * 10 + k)
    specs = [_spec(nv, rng) for _ in range(k)]
    out, rhos = S.prove_batch(_device_claims(S, D, specs), S.Transcript(b"lzx-batch-test"))
    st, ref = O.sumcheck_prove_batch(_oracle_claims(specs), O.Transcript(b"lzx-batch-test"))
    assert st == 0
    assert rhos.tolist() == ref["rhos"].tolist()
    assert out.rounds.tolist() == ref["rounds"].tolist()
    assert out.challenges.tolist() == ref["challenges"].tolist()
    assert out.factor_claims.tolist() == ref["factor_claims"].tolist()
    assert out.final_claim == ref["final_claim"]
    sums = [c[3] for c in _oracle_claims(specs)]
    pt = S.verify_batch(out.rounds, nv, 2, sums, S.Transcript(b"lzx-batch-test"), out.final_claim)
    assert pt.tolist() == out.challenges.tolist()


@pytest.mark.parametrize("nvs", [(3, 5, 5, 4), (0, 2), (7, 12, 9), (18, 14, 20), (6, 13, 9, 11, 14, 8)])
def test_prove_batch_star_matches_oracle(sc, nvs):
    S, D = sc
    rng = np.random.default_rng(sum(nvs))
    specs = [_spec(nv, rng) for nv in nvs]
    out = S.prove_batch_star(_device_claims(S, D, specs), S.Transcript(b"lzx-batch-star-test"))
    st, ref = O.sumcheck_prove_batch(_oracle_claims(specs), O.Transcript(b"lzx-batch-star-test"), star=True)
    assert st == 0
    assert out.rhos.tolist() == ref["rhos"].tolist()
    assert out.rounds.tolist() == ref["rounds"].tolist()
    assert out.challenges.tolist() == ref["challenges"].tolist()
    assert out.output.factor_claims.tolist() == ref["factor_claims"].tolist()
    assert out.output.final_claim == ref["final_claim"]
    assert out.rounds.shape[0] == max(nvs)
    for pc, nv in zip(out.per_claim, nvs):
        assert pc.num_vars == nv and pc.point.tolist() == out.challenges[:nv].tolist() and len(pc.factor_claims) == 2
    sums = [c[3] for c in _oracle_claims(specs)]
    pt = S.verify_batch_star(out.rounds, list(nvs), 2, sums, S.Transcript(b"lzx-batch-star-test"),
                             out.output.final_claim)
    assert pt.tolist() == out.challenges.tolist()


def test_batch_terminal_from_per_claim_openings(sc):
    # batch.rs:481-514: the combined final claim is sum_j rho_j P_j(r_j), r_j the truncated point
    S, D = sc
    rng = np.random.default_rng(5)
    specs = [_spec(nv, rng) for nv in (3, 5, 5)]
    out = S.prove_batch_star(_device_claims(S, D, specs), S.Transcript(b"lzx-batch-star-test"))
    expected = 0
    for rho, (nv, facs, terms), pc in zip(out.rhos, specs, out.per_claim):
        vals = [O.mle_evaluate(f, pc.point) for f in facs]
        assert pc.factor_claims.tolist() == vals
        pj = 0
        for c, ids in terms:
            p = c
            for i in ids:
                p = O.gl_mul(p, vals[i])
            pj = O.gl_add(pj, p)
        expected = O.gl_add(expected, O.gl_mul(int(rho), pj))
    sums = [c[3] for c in _oracle_claims(specs)]
    S.verify_batch_star(out.rounds, [3, 5, 5], 2, sums, S.Transcript(b"lzx-batch-star-test"), expected)
    vt = S.Transcript(b"lzx-batch-star-test")
    rh = vt.challenge_fields(b"batch-star-rho", 3)
    assert rh.tolist() == out.rhos.tolist()
    S.verify_batch_star_with_rhos(out.rounds, [3, 5, 5], 2, sums, rh, vt, expected)


def test_shared_tables_fold_once(sc):
    # the same device table in several claims: one fold, the same proof as distinct copies
    S, D = sc
    rng = np.random.default_rng(6)
    g = rng.integers(0, P, 1 << 9, dtype=np.uint64)
    specs = []
    for _ in range(5):
        nv, facs, terms = _spec(9, rng)
        specs.append((nv, [facs[0], g], terms))
    shared = {(ci, 1): (0, 1) for ci in range(1, 5)}  # 10 factors, 6 distinct tables
    out, rhos = S.prove_batch(_device_claims(S, D, specs, shared), S.Transcript(b"shared"))
    st, ref = O.sumcheck_prove_batch(_oracle_claims(specs), O.Transcript(b"shared"))
    assert st == 0 and out.rounds.tolist() == ref["rounds"].tolist()
    assert out.factor_claims.tolist() == ref["factor_claims"].tolist()


def test_batch_errors(sc):
    from lzx_amd._lib import LzxError
    S, D = sc
    rng = np.random.default_rng(7)
    with pytest.raises(LzxError) as e:  # batch.rs:412-437
        S.prove_batch(_device_claims(S, D, [_spec(3, rng), _spec(4, rng)]), S.Transcript(b"t"))
    assert e.value.code == 16
    with pytest.raises(LzxError) as e:
        S.prove_batch([], S.Transcript(b"t"))
    assert e.value.code == 17
    with pytest.raises(LzxError) as e:
        S.prove_batch_star([], S.Transcript(b"t"))
    assert e.value.code == 17
    with pytest.raises(LzxError) as e:  # 18 distinct tables exceed the device prover's 16
        S.prove_batch(_device_claims(S, D, [_spec(3, rng) for _ in range(9)]), S.Transcript(b"t"))
    assert e.value.code == 31
    specs = [_spec(4, rng) for _ in range(2)]
    out, _ = S.prove_batch(_device_claims(S, D, specs), S.Transcript(b"lzx-batch-test"))
    bad = [O.gl_add(c[3], 1) if i == 0 else c[3] for i, c in enumerate(_oracle_claims(specs))]
    with pytest.raises(LzxError) as e:  # batch.rs:391-410
        S.verify_batch(out.rounds, 4, 2, bad, S.Transcript(b"lzx-batch-test"))
    assert e.value.code == 11
    tampered = out.rounds.copy()
    tampered[0, 0] = O.gl_add(int(tampered[0, 0]), 1)
    sums = [c[3] for c in _oracle_claims(specs)]
    with pytest.raises(LzxError) as e:  # batch.rs:535-556 (star form below)
        S.verify_batch(tampered, 4, 2, sums, S.Transcript(b"lzx-batch-test"))
    assert e.value.code == 11
    with pytest.raises(LzxError) as e:
        S.verify_batch_star(out.rounds, [4, 4, 4], 2, sums, S.Transcript(b"x"))
    assert e.value.code == 16
    with pytest.raises(LzxError) as e:
        S.verify_batch_star_with_rhos(out.rounds, [4, 4], 2, sums, [1], S.Transcript(b"x"))
    assert e.value.code == 18
    with pytest.raises(LzxError) as e:  # a budget too small for the rhos: BatchError::ChallengeCount
        S.prove_batch(_device_claims(S, D, specs), S.Transcript(b"t", 0))
    assert e.value.code == 18

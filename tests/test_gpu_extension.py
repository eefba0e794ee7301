"""GPU parity of the quadratic extension ring R_q[Y]/(Y^2+1)
(crates/lattice-ring/src/extension.rs:27-443) against the oracle's
restatement (pinned in test_oracle.py by schoolbook products and the
reference's cases)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ext():
    from lzx_amd import device as D
    from lzx_amd import extension as E
    from lzx_amd import ring as R
    D.require_gpu()
    return E, D, R


MODS = {"q32": (3221225473, 30, 5), "q31": (2013265921, 27, 31), "q12289": (12289, 12, 11)}


def _setup(R, name, log_n):
    q, adic, g = MODS[name]
    ring = R.RingConfig(R.Modulus32(q, adic, g), log_n)
    return ring, O.Ntt(q, adic, g, log_n), q


@pytest.mark.parametrize("name", ["q32", "q31"])
@pytest.mark.parametrize("log_n,count", [(0, 5), (1, 7), (2, 33), (3, 64), (4, 31), (5, 100), (6, 1000), (7, 9),
                                         (10, 3)])
def test_rq2_mul_norm_inverse_match_oracle(ext, name, log_n, count):
    E, D, R = ext
    ring, ntt, q = _setup(R, name, log_n)
    n = 1 << log_n
    rng = np.random.default_rng(log_n * 100 + count)
    a = rng.integers(0, q, (count, 2 * n), dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, q, (count, 2 * n), dtype=np.uint64).astype(np.uint32)
    da, db = D.to_device(a), D.to_device(b)
    got = D.to_host(E.mul_batch(ring, da, db), np.uint32).reshape(count, 2 * n)
    assert np.array_equal(got, ntt.rq2_mul(a, b))
    got = D.to_host(E.norm_batch(ring, da), np.uint32).reshape(count, n)
    assert np.array_equal(got, ntt.rq2_norm(a))
    ok, inv = E.inverse_batch(ring, da)
    wok, winv = ntt.rq2_inverse(a)
    assert D.to_host(ok, np.uint8).tolist() == wok.tolist()
    assert np.array_equal(D.to_host(inv, np.uint32).reshape(count, 2 * n), winv)
    # in place: out aliases the first operand
    E.mul_batch(ring, da, db, out=da)
    assert np.array_equal(D.to_host(da, np.uint32).reshape(count, 2 * n), ntt.rq2_mul(a, b))


@pytest.mark.parametrize("name", ["q32", "q31"])
@pytest.mark.parametrize("log_n,count", [(6, 1000), (4, 257), (8, 40)])
def test_rq2_in_place_norm_and_overlaps(ext, name, log_n, count):
    """The norm's [count][n] output written over its own [count][2][n] input
    (many tiles, so another tile's unread input would be overwritten by a
    fused in-place pass), in-place inverse, and a product whose output
    overlaps its input shifted by one element."""
    E, D, R = ext
    ring, ntt, q = _setup(R, name, log_n)
    n = 1 << log_n
    rng = np.random.default_rng(log_n * 7 + count)
    a = rng.integers(0, q, (count, 2 * n), dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, q, (count, 2 * n), dtype=np.uint64).astype(np.uint32)
    da = D.to_device(a.reshape(-1))
    E.norm_batch(ring, da, out=da)
    assert np.array_equal(D.to_host(da, np.uint32)[: count * n].reshape(count, n), ntt.rq2_norm(a))
    da = D.to_device(a)
    ok, _ = E.inverse_batch(ring, da, out=da)
    wok, winv = ntt.rq2_inverse(a)
    assert D.to_host(ok, np.uint8).tolist() == wok.tolist()
    assert np.array_equal(D.to_host(da, np.uint32).reshape(count, 2 * n), winv)
    # out = a shifted by one element, over a buffer of count + 1 elements
    buf = D.to_device(np.concatenate([a, np.zeros((1, 2 * n), np.uint32)]).reshape(-1))
    db = D.to_device(b)
    E.mul_batch(ring, buf[: count * 2 * n], db, out=buf[2 * n:])
    got = D.to_host(buf, np.uint32).reshape(count + 1, 2 * n)[1:]
    assert np.array_equal(got, ntt.rq2_mul(a, b))


@pytest.mark.parametrize("log_n", [4, 6, 8])
def test_rq2_inverse_flags_zero_norm_slots(ext, log_n):
    # over q = 12289 a norm slot vanishes often enough to see both outcomes in one batch
    E, D, R = ext
    ring, ntt, q = _setup(R, "q12289", log_n)
    n = 1 << log_n
    rng = np.random.default_rng(log_n)
    a = rng.integers(0, q, (600, 2 * n), dtype=np.uint64).astype(np.uint32)
    ev = rng.integers(1, q, n, dtype=np.uint64).astype(np.uint32)
    ev[3] = 0
    a[0, :n] = ntt.inverse(ev)[1]  # c0 with a zero slot, c1 = 0: N has a zero slot
    a[0, n:] = 0
    ok, inv = E.inverse_batch(ring, D.to_device(a))
    wok, winv = ntt.rq2_inverse(a)
    okh = D.to_host(ok, np.uint8)
    assert okh.tolist() == wok.tolist() and okh[0] == 0 and okh.sum() > 0
    assert np.array_equal(D.to_host(inv, np.uint32).reshape(600, 2 * n), winv)
    assert not D.to_host(inv, np.uint32).reshape(600, 2 * n)[okh == 0].any()


@pytest.mark.parametrize("log_n", [2, 6])
def test_rq2_inverse_flags_default_modulus(ext, log_n):
    """q = 3*2^30 + 1 (the FP64 fused pass): elements whose norm N = c0^2 + c1^2
    vanishes in one slot (c1's evaluation there is i times c0's, i^2 = -1),
    zero elements and ordinary ones in one batch."""
    E, D, R = ext
    ring, ntt, q = _setup(R, "q32", log_n)
    n = 1 << log_n
    rng = np.random.default_rng(40 + log_n)
    count = 100
    a = rng.integers(0, q, (count, 2 * n), dtype=np.uint64).astype(np.uint32)
    i4 = pow(5, (q - 1) // 4, q)  # 5 generates the group mod q, so this is a square root of -1
    assert i4 * i4 % q == q - 1
    for e in range(0, count, 7):  # one vanishing slot per chosen element
        e0 = rng.integers(1, q, n, dtype=np.uint64)
        e1 = rng.integers(1, q, n, dtype=np.uint64)
        j = int(rng.integers(0, n))
        e1[j] = e0[j] * i4 % q
        a[e, :n] = ntt.inverse(e0.astype(np.uint32))[1]
        a[e, n:] = ntt.inverse(e1.astype(np.uint32))[1]
    a[3] = 0
    ok, inv = E.inverse_batch(ring, D.to_device(a))
    wok, winv = ntt.rq2_inverse(a)
    okh = D.to_host(ok, np.uint8)
    assert okh.tolist() == wok.tolist()
    assert okh[0] == 0 and okh[3] == 0 and okh[7] == 0 and okh[1] == 1
    assert np.array_equal(D.to_host(inv, np.uint32).reshape(count, 2 * n), winv)
    norm = D.to_host(E.norm_batch(ring, D.to_device(a)), np.uint32).reshape(count, n)
    assert np.array_equal(norm, ntt.rq2_norm(a))


def test_rq2_linear_ops(ext):
    E, D, R = ext
    ring, _ntt, q = _setup(R, "q32", 6)
    rng = np.random.default_rng(2)
    a = rng.integers(0, q, (77, 128), dtype=np.uint64)
    b = rng.integers(0, q, (77, 128), dtype=np.uint64)
    da, db = D.to_device(a.astype(np.uint32)), D.to_device(b.astype(np.uint32))
    h = lambda t: D.to_host(t, np.uint32).reshape(77, 128).astype(np.uint64)
    assert np.array_equal(h(E.add_batch(ring, da, db)), (a + b) % q)
    assert np.array_equal(h(E.sub_batch(ring, da, db)), (a + q - b) % q)
    assert np.array_equal(h(E.neg_batch(ring, da)), (q - a) % q)
    conj = a.copy()
    conj[:, 64:] = (q - conj[:, 64:]) % q
    assert np.array_equal(h(E.conjugate_batch(ring, da)), conj)


def test_rq2_reference_cases(ext):
    # extension.rs:263-391 through the element API
    E, D, R = ext
    ring, ntt, q = _setup(R, "q32", 6)
    rng = np.random.default_rng(7)
    el = lambda: E.Rq2(R.RingElement(ring, rng.integers(0, q, 64, dtype=np.uint64).astype(np.uint32)),
                       R.RingElement(ring, rng.integers(0, q, 64, dtype=np.uint64).astype(np.uint32)))
    y2 = E.Rq2.y(ring).square()
    assert y2.c0.coeffs[0] == q - 1 and not y2.c0.coeffs[1:].any() and y2.c1.is_zero()
    a, b, c = el(), el(), el()
    assert a.mul(b) == b.mul(a)
    assert a.mul(b).mul(c) == a.mul(b.mul(c))
    assert a.mul(b.add(c)) == a.mul(b).add(a.mul(c))
    assert a.mul(E.Rq2.one(ring)) == a and a.mul(E.Rq2.zero(ring)).is_zero() and a.add(a.neg()).is_zero()
    assert a.conjugate().conjugate() == a
    zc = a.mul(a.conjugate())
    assert zc.c1.is_zero() and zc.c0 == a.norm()
    assert a.mul(b).norm() == a.norm().mul(b.norm())
    inv = a.inverse()
    assert inv is not None and a.mul(inv) == E.Rq2.one(ring)
    assert E.Rq2.zero(ring).inverse() is None
    assert a.to_bytes() == a.c0.to_bytes() + a.c1.to_bytes()
    s0 = a.c0.mul(b.c0).coeffs.astype(np.int64) - a.c1.mul(b.c1).coeffs
    s1 = a.c0.mul(b.c1).coeffs.astype(np.uint64) + a.c1.mul(b.c0).coeffs
    m = a.mul(b)
    assert m.c0.coeffs.tolist() == (s0 % q).tolist() and m.c1.coeffs.tolist() == (s1 % q).tolist()
    from lzx_amd._lib import LzxError
    with pytest.raises(LzxError) as e:
        E.Rq2(a.c0, R.RingElement(R.RingConfig(R.Modulus32.Q_32, 5), np.zeros(32, dtype=np.uint32)))
    assert e.value.code == 3


def test_rq2_full_batch_sampled(ext):
    # 2^17 elements at n = 64 (64 MiB per operand): 512 sampled rows against the oracle
    E, D, R = ext
    ring, ntt, q = _setup(R, "q32", 6)
    count = 1 << 17
    rng = np.random.default_rng(17)
    a = rng.integers(0, q, (count, 128), dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, q, (count, 128), dtype=np.uint64).astype(np.uint32)
    got = D.to_host(E.mul_batch(ring, D.to_device(a), D.to_device(b)), np.uint32).reshape(count, 128)
    rows = np.sort(rng.choice(count, 512, replace=False))
    rows[-1] = count - 1
    assert np.array_equal(got[rows], ntt.rq2_mul(a[rows], b[rows]))


# ---- short challenges and sample_rq2_challenge (library host code against the oracle) ----

SPECS = [(256, "fw", dict(weight=23, amplitude=1)), (64, "fw", dict(weight=63, amplitude=3)),
         (1 << 16, "fw", dict(weight=300, amplitude=5)), (512, "bt", dict(permille=500)),
         (100, "bt", dict(permille=1000)), (64, "ss", dict(values=[-2, -1, 0, 1, 2])),
         (1000, "ss", dict(values=list(range(-500, 501)))), (256, "fws", dict(weight=128, values=[-1, 1, 2])),
         (7, "fws", dict(weight=7, values=[9]))]


def _spec(SC, n, fam, kw):
    f = {"fw": lambda: SC.FixedWeight(kw["weight"], kw["amplitude"]), "bt": lambda: SC.BiasedTernary(kw["permille"]),
         "ss": lambda: SC.SmallSet(tuple(kw["values"])),
         "fws": lambda: SC.FixedWeightSmallSet(kw["weight"], tuple(kw["values"]))}[fam]()
    return SC.ShortChallengeSpec(n, f)


OFAM = {"fw": O.SC_FIXED_WEIGHT, "bt": O.SC_BIASED_TERNARY, "ss": O.SC_SMALL_SET, "fws": O.SC_FIXED_WEIGHT_SMALL_SET}


@pytest.mark.parametrize("n,fam,kw", SPECS)
def test_short_challenge_sample_matches_oracle(n, fam, kw):
    from lzx_amd import short_challenge as SC
    from lzx_amd._lib import LzxError
    spec = _spec(SC, n, fam, kw)
    for seed in (b"", b"seed-a", bytes(range(200))):
        st, want, g = O.short_challenge_sample(n, OFAM[fam], seed, **kw)
        if st:  # e.g. the biased ternary at p = 1/2 can outrun its n * 6 + 32 byte stream
            with pytest.raises(LzxError) as e:
                spec.sample(seed)
            assert e.value.code == st
            continue
        c = spec.sample(seed)
        assert c.coefficients.tolist() == want.tolist() and c.gamma_c == g


def test_short_challenge_cap_errors_and_bookkeeping():
    # short_challenge.rs:668-886 through the library
    from lzx_amd import short_challenge as SC
    from lzx_amd._lib import LzxError
    pk = SC.pikkufold_spec()
    c = pk.sample(b"seed-a")
    assert c.hamming_weight() == 23 and c.l_inf() == 1 and c.gamma_c == 5 and pk.gamma_c_worst_case() == 5
    assert pk.entropy_bits() > 100.0
    with pytest.raises(LzxError) as e:
        pk.sample_with_gamma_cap(b"x", 4, 4)
    assert e.value.code == 62
    assert pk.sample_with_gamma_cap(b"x", 5, 4).gamma_c == 5
    lf = SC.latticefold_plus_spec()
    st, want, g = O.short_challenge_sample_with_gamma_cap(64, O.SC_SMALL_SET, b"t", 10, 200, values=[-2, -1, 0, 1, 2])
    if st == 0:
        got = lf.sample_with_gamma_cap(b"t", 10, 200)
        assert got.coefficients.tolist() == want.tolist() and got.gamma_c == g <= 10
    assert lf.gamma_c_worst_case() == 16 and abs(lf.entropy_bits() - 64 * np.log2(5)) < 1e-9
    assert abs(SC.cyclo_spec().entropy_bits() - 768.0) < 1.0 and SC.protogalattice_spec(256).entropy_bits() > 128
    assert SC.symphony_spec(512).entropy_bits() == 512 * np.log2(5) and SC.rokoko_spec(1024).entropy_bits() > 100
    pk.require_entropy(90.0)
    with pytest.raises(LzxError):
        pk.require_entropy(200.0)
    for bad in (SC.ShortChallengeSpec(0, SC.FixedWeight(1, 1)), SC.ShortChallengeSpec(8, SC.FixedWeight(9, 1)),
                SC.ShortChallengeSpec(8, SC.SmallSet(())), SC.ShortChallengeSpec(8, SC.BiasedTernary(1001))):
        with pytest.raises(LzxError) as e:
            bad.sample(b"s")
        assert e.value.code == 61
        with pytest.raises(LzxError):
            bad.validate()
    M = (1 << 64) - 1
    for x in (0, 1, 2, 3, 24, 25, 26, 1 << 96, (1 << 96) + 1, M * M, M * M + 1, (1 << 128) - 1):
        assert SC.ceil_sqrt(x) == O.ceil_sqrt_u128(x)


def test_rq2_sample_challenge_matches_oracle(ext):
    # extension.rs:393-441 through the library, against the oracle
    E, D, R = ext
    from lzx_amd import short_challenge as SC
    from lzx_amd._lib import LzxError
    ring, _ntt, q = _setup(R, "q32", 6)
    spec = SC.ShortChallengeSpec(128, SC.FixedWeight(23, 1))
    z, g0, g1 = E.sample_rq2_challenge(ring, spec, b"chal-seed")
    st, want, w0, w1 = O.rq2_sample_challenge(q, 64, 128, O.SC_FIXED_WEIGHT, b"chal-seed", weight=23, amplitude=1)
    assert st == 0 and np.concatenate([z.c0.coeffs, z.c1.coeffs]).tolist() == want.tolist() and (g0, g1) == (w0, w1)
    assert int(np.count_nonzero(z.c0.coeffs)) + int(np.count_nonzero(z.c1.coeffs)) == 23 and g0 >= 1 and g1 >= 1
    with pytest.raises(LzxError) as e:
        E.sample_rq2_challenge(ring, SC.ShortChallengeSpec(64, SC.FixedWeight(8, 1)), b"x")
    assert e.value.code == 61
    sym = SC.symphony_spec(128)
    z, _, _ = E.sample_rq2_challenge(ring, sym, b"sym-seed")
    for c in np.concatenate([z.c0.coeffs, z.c1.coeffs]).astype(np.int64):
        assert min(c, q - c) <= 2

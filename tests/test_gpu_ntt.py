"""GPU parity of the batched negacyclic NTT against the oracle (bit-exact).

Mirrors the reference's NTT tests (crates/lattice-ring/src/ntt.rs:364-479):
round trip, schoolbook product, partial+complete == full, fast == generic,
bad geometry; plus the batched/CRT layouts and the full 4-prime x 2^16 x d=64
workload.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

MODULI = [O.Q_32, O.Q_12289, *O.CRT_PRIMES]


def _rng_poly(rng, q, count):
    return rng.integers(0, q, size=count, dtype=np.uint64).astype(np.uint32)


@pytest.fixture(scope="module")
def ring():
    from lzx_amd import ring as R
    from lzx_amd import device as D
    D.require_gpu()
    return R


@pytest.mark.parametrize("mod", MODULI, ids=lambda m: str(m[0]))
@pytest.mark.parametrize("log_n", [1, 2, 3, 4, 5, 6, 7, 8, 10])
def test_forward_inverse_batch_matches_oracle(ring, mod, log_n):
    q, adic, g = mod
    if log_n + 1 > adic:
        pytest.skip("adicity")
    from lzx_amd import device as D
    rng = np.random.default_rng(1000 + log_n)
    polys = 131  # not a multiple of the 64-polynomial tile
    a = _rng_poly(rng, q, polys << log_n)
    tab = ring.NttTables(ring.Modulus32(q, adic, g), log_n)
    ot = O.Ntt(q, adic, g, log_n)
    st, want = ot.forward(a)
    assert st == 0
    t = D.to_device(a)
    tab.forward_batch(t)
    got = D.to_host(t, np.uint32)
    assert np.array_equal(got, want), f"forward q={q} log_n={log_n}"
    tab.inverse_batch(t)
    back = D.to_host(t, np.uint32)
    assert np.array_equal(back, a), "inverse(forward(a)) != a"
    e = _rng_poly(rng, q, polys << log_n)
    st, want_inv = ot.inverse(e)
    t = D.to_device(e)
    tab.inverse_batch(t)
    assert np.array_equal(D.to_host(t, np.uint32), want_inv), "inverse vs oracle"


@pytest.mark.parametrize("log_n", [1, 3, 5, 6, 8])
def test_partial_then_complete(ring, log_n):
    from lzx_amd import device as D
    q, adic, g = O.Q_32
    rng = np.random.default_rng(7 + log_n)
    polys = 37
    a = _rng_poly(rng, q, polys << log_n)
    tab = ring.NttTables(ring.Modulus32(q, adic, g), log_n)
    ot = O.Ntt(q, adic, g, log_n)
    _, full = ot.forward(a)
    for levels in range(log_n + 1):
        st, want_part = ot.forward_partial(a, levels)
        assert st == 0
        t = D.to_device(a)
        tab.forward_partial_batch(t, levels)
        part = D.to_host(t, np.uint32)
        assert np.array_equal(part, want_part), f"partial levels={levels}"
        tab.complete_partial_batch(t, levels)
        assert np.array_equal(D.to_host(t, np.uint32), full), f"complete after {levels}"


def test_crt_multi_full_workload(ring):
    """4 CRT primes x 2^16 polynomials of degree 64 in one launch, forward
    bit-exact vs the oracle and inverse round trip."""
    from lzx_amd import device as D
    from lzx_amd import workloads as W
    polys = 1 << 16
    data = W.ntt_batch_input(ring.CRT_PRIMES, polys, 64)
    tabs = [ring.NttTables(m, 6) for m in ring.CRT_PRIMES]
    t = D.to_device(data)
    ring.forward_multi(tabs, t, polys)
    got = D.to_host(t, np.uint32).reshape(data.shape)
    for i, m in enumerate(ring.CRT_PRIMES):
        ot = O.Ntt(m.q, m.two_adicity, m.generator, 6)
        _, want = ot.forward(data[i].reshape(-1))
        assert np.array_equal(got[i].reshape(-1), want), f"prime {m.q}"
    ring.inverse_multi(tabs, t, polys)
    assert np.array_equal(D.to_host(t, np.uint32).reshape(data.shape), data)


# the int8-MFMA transform's modulus range is (2^20, 2^31 - 2^16): its smallest
# and largest NTT-friendly primes (q = 1 mod 128, primitive roots 7 and 5)
MFMA_EDGE = [(1049089, 9, 7), (2147415553, 9, 5)]


def _patterns(rng, q, polys):
    n = polys * 64
    yield "random", rng.integers(0, q, size=n, dtype=np.uint64).astype(np.uint32)
    yield "zeros", np.zeros(n, dtype=np.uint32)
    yield "q-1", np.full(n, q - 1, dtype=np.uint32)
    alt = np.zeros(n, dtype=np.uint32)
    alt[1::2] = q - 1
    yield "alternating", alt
    yield "small", rng.integers(0, 256, size=n, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("mod", [*O.CRT_PRIMES, *MFMA_EDGE], ids=lambda m: str(m[0]))
def test_mfma_transform_edges(ring, mod, monkeypatch):
    """The int8-MFMA transform (n = 64): extreme residues and partial 32-
    polynomial tiles, forward and inverse bit-exact vs the oracle, at the CRT
    primes and at both ends of the modulus range."""
    from lzx_amd import device as D
    q, adic, g = mod
    monkeypatch.setenv("LZX_NTT_KERNEL", "mfma")
    tab = ring.NttTables(ring.Modulus32(q, adic, g), 6)
    assert tab.path() == "mfma" and tab.path(inverse=True) == "mfma"
    ot = O.Ntt(q, adic, g, 6)
    rng = np.random.default_rng(q % 1000)
    for polys in (1, 31, 32, 33, 65, 1000):
        for name, a in _patterns(rng, q, polys):
            _, want = ot.forward(a)
            t = D.to_device(a)
            tab.forward_batch(t)
            assert np.array_equal(D.to_host(t, np.uint32), want), f"forward {name} polys={polys}"
            _, want_inv = ot.inverse(a)
            t = D.to_device(a)
            tab.inverse_batch(t)
            assert np.array_equal(D.to_host(t, np.uint32), want_inv), f"inverse {name} polys={polys}"


def test_mfma_multi_plan_ragged(ring, monkeypatch):
    """One launch over the 4 CRT primes with 1000 polynomials each (plan
    boundaries inside blocks, a partial tile per plan): forward and inverse vs
    the oracle."""
    from lzx_amd import device as D
    polys = 1000
    rng = np.random.default_rng(5)
    monkeypatch.setenv("LZX_NTT_KERNEL", "mfma")
    tabs = [ring.NttTables(m, 6) for m in ring.CRT_PRIMES]
    assert ring.batch_path(tabs, polys) == "mfma"
    data = np.stack([rng.integers(0, m.q, size=polys * 64, dtype=np.uint64).astype(np.uint32) for m in ring.CRT_PRIMES])
    t = D.to_device(data)
    ring.forward_multi(tabs, t, polys)
    got = D.to_host(t, np.uint32).reshape(data.shape)
    for i, m in enumerate(ring.CRT_PRIMES):
        _, want = O.Ntt(m.q, m.two_adicity, m.generator, 6).forward(data[i])
        assert np.array_equal(got[i], want), f"forward prime {m.q}"
    t = D.to_device(data)
    ring.inverse_multi(tabs, t, polys)
    got = D.to_host(t, np.uint32).reshape(data.shape)
    for i, m in enumerate(ring.CRT_PRIMES):
        _, want = O.Ntt(m.q, m.two_adicity, m.generator, 6).inverse(data[i])
        assert np.array_equal(got[i], want), f"inverse prime {m.q}"


@pytest.mark.parametrize("env", [("LZX_NTT_MFMA", "0"), ("LZX_NTT_KERNEL", "register")], ids=lambda e: e[0])
@pytest.mark.parametrize("mod", [O.CRT_PRIMES[0], O.CRT_PRIMES[3]], ids=lambda m: str(m[0]))
def test_butterfly_kernel_when_mfma_disabled(ring, mod, env, monkeypatch):
    """Plans created without MFMA tables (LZX_NTT_MFMA=0) or pinned to the
    register class (LZX_NTT_KERNEL=register) run the register butterfly
    kernel (both lazy Shoup classes: q above and below 2^30), bit-exact vs
    the oracle."""
    from lzx_amd import device as D
    q, adic, g = mod
    monkeypatch.setenv(*env)
    tab = ring.NttTables(ring.Modulus32(q, adic, g), 6)
    assert tab.path() == "register"
    ot = O.Ntt(q, adic, g, 6)
    rng = np.random.default_rng(9)
    for polys in (1, 63, 64, 1000):
        for name, a in _patterns(rng, q, polys):
            _, want = ot.forward(a)
            t = D.to_device(a)
            tab.forward_batch(t)
            assert np.array_equal(D.to_host(t, np.uint32), want), f"forward {name} polys={polys}"
            tab.inverse_batch(t)
            assert np.array_equal(D.to_host(t, np.uint32), a), f"round trip {name} polys={polys}"


def test_ntt_path_selection(ring):
    """Kernel class per plan: MFMA-capable (so "auto" until measured) for
    n = 64 inside its modulus range, the register kernels otherwise (q above
    2^31, small q, n < 64), LDS above 64."""
    assert ring.NttTables(ring.CRT_PRIMES[0], 6).path(polys=12345) == "auto"
    assert ring.NttTables(ring.Modulus32(*O.Q_32), 6).path() == "register"
    assert ring.NttTables(ring.Modulus32(*O.Q_12289), 6).path() == "register"
    assert ring.NttTables(ring.CRT_PRIMES[0], 5).path() == "register"
    assert ring.NttTables(ring.CRT_PRIMES[0], 7).path() == "lds"


def test_auto_kernel_choice_measures_on_a_copy(ring):
    """The first AUTO call of a shape times both classes on a scratch copy:
    the caller's data is transformed exactly once (forward vs the oracle,
    then inverse back), the choice is kept per direction, and later calls
    reuse it."""
    from lzx_amd import device as D
    q, adic, g = O.CRT_PRIMES[1]
    polys = 777
    tab = ring.NttTables(ring.Modulus32(q, adic, g), 6)
    assert tab.path(polys) == "auto" and tab.path(polys, inverse=True) == "auto"
    rng = np.random.default_rng(77)
    a = rng.integers(0, q, size=polys * 64, dtype=np.uint64).astype(np.uint32)
    _, want = O.Ntt(q, adic, g, 6).forward(a)
    t = D.to_device(a)
    tab.forward_batch(t)
    assert np.array_equal(D.to_host(t, np.uint32), want)
    assert tab.path(polys) in ("mfma", "register") and tab.path(polys, inverse=True) == "auto"
    tab.inverse_batch(t)
    assert np.array_equal(D.to_host(t, np.uint32), a)
    assert tab.path(polys, inverse=True) in ("mfma", "register")
    tab.forward_batch(t)
    assert np.array_equal(D.to_host(t, np.uint32), want)


@pytest.mark.parametrize("mod", [O.Q_32, O.CRT_PRIMES[0]], ids=lambda m: str(m[0]))
def test_lds_staged_register_kernel(ring, mod, monkeypatch):
    """The LDS-staged register kernel (plans created with LZX_NTT_STAGE=lds):
    the FP64 class for q > 2^31, Shoup below, every register size, a partial
    last tile."""
    from lzx_amd import device as D
    q, adic, g = mod
    monkeypatch.setenv("LZX_NTT_STAGE", "lds")
    for log_n in (2, 3, 4, 5, 6):
        rng = np.random.default_rng(77 + log_n)
        polys = 77
        a = _rng_poly(rng, q, polys << log_n)
        tab = ring.NttTables(ring.Modulus32(q, adic, g), log_n)
        ot = O.Ntt(q, adic, g, log_n)
        _, want = ot.forward(a)
        t = D.to_device(a)
        tab.forward_batch(t)
        assert np.array_equal(D.to_host(t, np.uint32), want), f"forward log_n={log_n}"
        tab.inverse_batch(t)
        assert np.array_equal(D.to_host(t, np.uint32), a), f"round trip log_n={log_n}"


def test_host_signature_and_errors(ring):
    from lzx_amd._lib import LzxError
    q, adic, g = O.Q_32
    tab = ring.NttTables(ring.Modulus32(q, adic, g), 5)
    a = _rng_poly(np.random.default_rng(3), q, 32)
    b = a.copy()
    tab.forward(b)
    _, want = O.Ntt(q, adic, g, 5).forward(a)
    assert np.array_equal(b, want)
    tab.inverse(b)
    assert np.array_equal(b, a)
    with pytest.raises(LzxError) as ei:
        tab.forward(np.zeros(48, dtype=np.uint32))
    assert ei.value.code == ring.LZX_NTT_LENGTH_MISMATCH
    # ntt.rs:470-479: adicity 12 -> max log_n = 11
    ring.NttTables(ring.Modulus32(*O.Q_12289), 11)
    with pytest.raises(LzxError) as ei:
        ring.NttTables(ring.Modulus32(*O.Q_12289), 12)
    assert ei.value.code == ring.LZX_NTT_ADICITY_TOO_SMALL


@pytest.mark.parametrize("mod", [O.Q_32, O.Q_2013265921], ids=lambda m: str(m[0]))
def test_ring_mul_matches_schoolbook(ring, mod):
    from lzx_amd import device as D
    q, adic, g = mod
    log_n = 5
    n = 1 << log_n
    rng = np.random.default_rng(11)
    polys = 9
    a = _rng_poly(rng, q, polys * n)
    b = _rng_poly(rng, q, polys * n)
    tab = ring.NttTables(ring.Modulus32(q, adic, g), log_n)
    out = D.empty(polys * n, np.uint32)
    tab.ring_mul_batch(D.to_device(a), D.to_device(b), out)
    got = D.to_host(out, np.uint32).reshape(polys, n)
    for i in range(polys):
        want = O.negacyclic_schoolbook(q, a[i * n:(i + 1) * n], b[i * n:(i + 1) * n])
        assert np.array_equal(got[i], want)

"""GPU parity of the Ajtai commitment (NTT-domain and int8-MFMA paths).

Mirrors crates/lattice-commitment/src/ajtai.rs:368-570: commit/open/verify,
tamper rejection, norm bound, deterministic derivation, cached == naive
across the reference's four shapes, linearity, zero elements; plus the
batched device layouts and the full d=64, k=16, m=2^16, W=64 workload.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
VEC = json.loads((Path(__file__).resolve().parent / "golden" / "vectors.json").read_text())


@pytest.fixture(scope="module")
def mods():
    from lzx_amd import commitment as CM
    from lzx_amd import device as D
    from lzx_amd import ring as R
    D.require_gpu()
    return CM, D, R


def _params(mods, log_n, k, m, bound, mod=O.Q_32):
    CM, D, R = mods
    return CM.AjtaiParams(R.RingConfig(R.Modulus32(*mod), log_n), k, m, bound)


def test_key_derivation_matches_oracle(mods):
    CM, D, R = mods
    for log_n, k, m in [(3, 2, 2), (6, 3, 12)]:
        pk = CM.AjtaiPublicKey.from_seed(_params(mods, log_n, k, m, 16), bytes([7]) * 32)
        A, Ah = O.ajtai_from_seed(O.Q_32, log_n, k, m, bytes([7]) * 32)
        assert np.array_equal(pk.matrix(), A)
        assert np.array_equal(pk.matrix_ntt(), Ah)
        pk2 = CM.AjtaiPublicKey.from_seed(_params(mods, log_n, k, m, 16), bytes([8]) * 32)
        assert not np.array_equal(pk2.matrix()[0], A[0])


def test_device_derivation_through_the_rejection_fallback(mods):
    """from_seed derives A on the device.  At q = 2^31 + 129 half of all words
    are rejected, so about half of the elements run past the first
    n + max(n, 16) words into ring.rs:155-175's re-squeeze loop; the oracle
    restates that loop literally (a fresh XOF per further word)."""
    CM, D, R = mods
    q = (2147483777, 7, 3)
    for log_n, k, m in [(6, 2, 40), (3, 3, 50), (4, 1, 7)]:
        pk = CM.AjtaiPublicKey.from_seed(_params(mods, log_n, k, m, 16, mod=q), bytes([0x5A]) * 32)
        A, Ah = O.ajtai_from_seed(q, log_n, k, m, bytes([0x5A]) * 32)
        assert np.array_equal(pk.matrix(), A), (log_n, k, m)
        assert np.array_equal(pk.matrix_ntt(), Ah), (log_n, k, m)


def test_reference_shapes_match_golden_commitments(mods):
    """ajtai.rs:490-506 shapes; expected values from schoolbook (gen_golden.py)."""
    CM, D, R = mods
    for v in VEC["ajtai"]:
        pk = CM.AjtaiPublicKey.from_seed(_params(mods, v["log_n"], v["k"], v["m"], v["norm_bound"]),
                                         bytes([v["key_seed_byte"]]) * 32)
        s = O.sample_small_secret(v["q"], v["log_n"], v["m"], v["density"], v["secret_seed"].encode())
        t = pk.commit(s)
        assert t.rows_array.tolist() == v["commitment"], (v["log_n"], v["k"], v["m"])
        pk.verify_opening(t, s)


def test_open_verify_tamper_and_norm(mods):
    """ajtai.rs:368-412."""
    CM, D, R = mods
    from lzx_amd._lib import LzxError
    q = O.Q_32[0]
    pk = CM.AjtaiPublicKey.from_seed(_params(mods, 4, 2, 3, 64), bytes([2]) * 32)
    s = O.sample_small_secret(q, 4, 3, 8, b"secret")
    t = pk.commit(s)
    pk.verify_opening(t, s)
    s2 = O.sample_small_secret(q, 4, 3, 8, b"other")
    with pytest.raises(LzxError) as e:
        pk.verify_opening(t, s2)
    assert e.value.code == 6
    bad = CM.AjtaiCommitment(t.rows_array.copy(), pk.params.ring)
    bad.rows_array[0, 0] = (bad.rows_array[0, 0] + 1) % q
    with pytest.raises(LzxError) as e:
        pk.verify_opening(bad, s)
    assert e.value.code == 6
    pk8 = CM.AjtaiPublicKey.from_seed(_params(mods, 4, 2, 3, 8), bytes([3]) * 32)
    big = np.zeros((3, 16), np.uint32)
    big[:, 0] = 1000
    with pytest.raises(LzxError) as e:
        pk8.verify_opening(pk8.commit(big), big)
    assert e.value.code == 5
    with pytest.raises(LzxError) as e:
        pk.commit(s[:2])
    assert e.value.code == 4


@pytest.mark.parametrize("shape", [(6, 3, 37, 11), (6, 16, 200, 9), (6, 5, 64, 3), (6, 1, 9, 1), (5, 4, 20, 6),
                                   (7, 2, 10, 3)])
def test_batched_commit_paths_match_oracle(mods, shape):
    """Canonical-residue batch, digits via NTT path and digits via int8 MFMA
    (d = 64 only) all equal the oracle's cached-NTT commit per witness."""
    CM, D, R = mods
    log_n, k, m, W = shape
    n = 1 << log_n
    q = O.Q_32[0]
    pk = CM.AjtaiPublicKey.from_seed(_params(mods, log_n, k, m, 8), bytes([log_n + k]) * 32)
    A, Ah = O.ajtai_from_seed(O.Q_32, log_n, k, m, bytes([log_n + k]) * 32)
    rng = np.random.default_rng(k * 1000 + m)
    digits = rng.integers(-8, 8, size=(W, m, n)).astype(np.int8)
    digits[:, ::5] = 0  # whole zero ring elements (zero-skip)
    resid = np.mod(digits.astype(np.int64), q).astype(np.uint32)
    ot = O.Ntt(*O.Q_32, log_n)
    want = np.stack([O.ajtai_commit(ot, k, m, Ah, resid[w])[1] for w in range(W)])
    out = D.empty(W * k * n, np.uint32)
    pk.commit_batch(D.to_device(resid), W, out)
    assert np.array_equal(D.to_host(out, np.uint32).reshape(W, k, n), want), "residue batch"
    dd = D.to_device(digits)
    paths = [CM.COMMIT_NTT] + ([CM.COMMIT_INT8, CM.COMMIT_AUTO] if n == 64 else [])
    for path in paths:
        out = D.empty(W * k * n, np.uint32)
        pk.commit_digits_batch(dd, W, out, path=path)
        assert np.array_equal(D.to_host(out, np.uint32).reshape(W, k, n), want), f"digits path {path}"


SMALL_Q = O.CRT_PRIMES[0]  # q < 2^31: the Shoup register transform on the witness side


@pytest.mark.parametrize("mac", ["mfma", "valu"])
@pytest.mark.parametrize("shape", [(6, 9, 77, 17), (4, 3, 40, 130), (5, 16, 33, 5), (6, 8, 1024, 1), (6, 16, 31, 8)])
def test_ntt_path_multiply_accumulate_matches_oracle(mods, shape, mac, monkeypatch):
    """The NTT-domain path's multiply-accumulate, on int8 MFMA (the per-
    coefficient limb-split product) and on VALU (LZX_AJTAI_MAC=valu), equals
    the oracle's cached-NTT commit (ajtai.rs:118-153): k leaving a row block
    partly empty, m not a multiple of 32, W leaving a column block partly
    empty, more than one column group (W = 130), digit and residue witnesses,
    at q > 2^31 (FP64 witness transform) and q < 2^31."""
    CM, D, R = mods
    if mac == "valu":
        monkeypatch.setenv("LZX_AJTAI_MAC", "valu")
    else:
        monkeypatch.delenv("LZX_AJTAI_MAC", raising=False)
    log_n, k, m, W = shape
    n = 1 << log_n
    for mod in (O.Q_32, SMALL_Q):
        q = mod[0]
        seed = bytes([log_n + 3 * k + (q & 7)]) * 32
        pk = CM.AjtaiPublicKey.from_seed(_params(mods, log_n, k, m, 8, mod=mod), seed)
        _, Ah = O.ajtai_from_seed(mod, log_n, k, m, seed)
        rng = np.random.default_rng(m * 7 + W + q % 1000)
        digits = rng.integers(-128, 128, size=(W, m, n)).astype(np.int8)
        digits[:, ::7] = 0
        resid = np.mod(digits.astype(np.int64), q).astype(np.uint32)
        ot = O.Ntt(*mod, log_n)
        want = np.stack([O.ajtai_commit(ot, k, m, Ah, resid[w])[1] for w in range(W)])
        out = D.empty(W * k * n, np.uint32)
        pk.commit_digits_batch(D.to_device(digits), W, out, path=CM.COMMIT_NTT)
        assert np.array_equal(D.to_host(out, np.uint32).reshape(W, k, n), want), (mod, "digits")
        out = D.empty(W * k * n, np.uint32)
        pk.commit_batch(D.to_device(resid), W, out)
        assert np.array_equal(D.to_host(out, np.uint32).reshape(W, k, n), want), (mod, "residues")


def test_ntt_path_limb_extremes(mods):
    """NTT-domain operands at the limb split's extremes: the key's A^ entries
    and the witness's s^ entries are drawn from {0, 1, (q-1)/2, (q+1)/2, q-2,
    q-1} (A and s are their inverse transforms), the ring elements split
    over several chunks; the commit equals the oracle's."""
    CM, D, R = mods
    q = O.Q_32[0]
    k, m, W = 16, 2093, 9
    rng = np.random.default_rng(5150)
    ext = np.array([0, 1, (q - 1) // 2, (q + 1) // 2, q - 2, q - 1], dtype=np.uint64)
    ot = O.Ntt(*O.Q_32, 6)
    Ah = ext[rng.integers(0, ext.size, size=(k * m, 64))].astype(np.uint32)
    A = ot.inverse(Ah)[1]
    pk = CM.AjtaiPublicKey.from_matrix(_params(mods, 6, k, m, 8), A)
    assert np.array_equal(pk.matrix_ntt(), Ah)
    sh = ext[rng.integers(0, ext.size, size=(W * m, 64))].astype(np.uint32)
    s = ot.inverse(sh)[1].reshape(W, m, 64)
    want = np.stack([O.ajtai_commit(ot, k, m, Ah, s[w])[1] for w in range(W)])
    out = D.empty(W * k * 64, np.uint32)
    pk.commit_batch(D.to_device(s), W, out)
    assert np.array_equal(D.to_host(out, np.uint32).reshape(W, k, 64), want)


def test_int8_extreme_digits_and_limbs(mods):
    """Digits at +-127 and A entries at the centred extremes (0, (q-1)/2,
    (q+1)/2, q-1) exercise the limb split and int32 fold-back bound."""
    CM, D, R = mods
    q = O.Q_32[0]
    k, m, W = 4, 2100, 2  # m above one 2048-element fold-back chunk
    rng = np.random.default_rng(77)
    ext = np.array([0, 1, (q - 1) // 2, (q + 1) // 2, q - 1, q - 2], dtype=np.uint64)
    A = ext[rng.integers(0, ext.size, size=(k * m, 64))].astype(np.uint32)
    pk = CM.AjtaiPublicKey.from_matrix(_params(mods, 6, k, m, 127), A)
    digits = rng.choice(np.array([-127, 127, -1, 0], dtype=np.int8), size=(W, m, 64))
    ot = O.Ntt(*O.Q_32, 6)
    _, Ah = None, np.stack([ot.forward(A[i])[1] for i in range(k * m)])
    resid = np.mod(digits.astype(np.int64), q).astype(np.uint32)
    want = np.stack([O.ajtai_commit(ot, k, m, Ah, resid[w])[1] for w in range(W)])
    for path in (CM.COMMIT_NTT, CM.COMMIT_INT8):
        out = D.empty(W * k * 64, np.uint32)
        pk.commit_digits_batch(D.to_device(digits), W, out, path=path)
        assert np.array_equal(D.to_host(out, np.uint32).reshape(W, k, 64), want), f"path {path}"


def test_int8_digit_minus128_exact(mods):
    """Digit -128 is a valid int8 digit: every path, AUTO included, gives the
    oracle's cached-NTT commit (ajtai.rs:118-153).  Batches mix elements that
    are all -128, sparse -128 among +-127, and 16-element runs with no -128
    at all (both sides of the kernel's per-batch split); m spans several
    int32 fold-back chunks; A entries sit at the centred extremes."""
    CM, D, R = mods
    q = O.Q_32[0]
    k, m, W = 4, 4100, 3
    rng = np.random.default_rng(128)
    ext = np.array([0, 1, (q - 1) // 2, (q + 1) // 2, q - 1, q - 2], dtype=np.uint64)
    A = ext[rng.integers(0, ext.size, size=(k * m, 64))].astype(np.uint32)
    A[: k * m // 2] = rng.integers(0, q, size=(k * m // 2, 64), dtype=np.uint64).astype(np.uint32)
    pk = CM.AjtaiPublicKey.from_matrix(_params(mods, 6, k, m, 128), A)
    digits = rng.choice(np.array([-128, 127, -127, 1, 0, -1], dtype=np.int8), size=(W, m, 64))
    digits[:, 32:64] = rng.choice(np.array([127, -127, 5, 0], dtype=np.int8), size=(W, 32, 64))  # no -128
    digits[0, 100:140] = -128
    digits[1, 2031:2034] = -128  # across a chunk boundary
    digits[2, 7] = 0
    digits[2, 7, 63] = -128
    ot = O.Ntt(*O.Q_32, 6)
    Ah = np.stack([ot.forward(A[i])[1] for i in range(k * m)])
    resid = np.mod(digits.astype(np.int64), q).astype(np.uint32)
    want = np.stack([O.ajtai_commit(ot, k, m, Ah, resid[w])[1] for w in range(W)])
    dd = D.to_device(digits)
    for path in (CM.COMMIT_NTT, CM.COMMIT_INT8, CM.COMMIT_AUTO):
        out = D.empty(W * k * 64, np.uint32)
        pk.commit_digits_batch(dd, W, out, path=path)
        assert np.array_equal(D.to_host(out, np.uint32).reshape(W, k, 64), want), f"path {path}"
    # one witness of nothing but -128
    full = np.full((1, m, 64), -128, np.int8)
    res1 = np.mod(full[0].astype(np.int64), q).astype(np.uint32)
    want1 = O.ajtai_commit(ot, k, m, Ah, res1)[1]
    out = D.empty(k * 64, np.uint32)
    pk.commit_digits_batch(D.to_device(full), 1, out, path=CM.COMMIT_INT8)
    assert np.array_equal(D.to_host(out, np.uint32).reshape(k, 64), want1)


def test_full_workload_int8_equals_ntt_path(mods):
    """d=64, k=16, m=2^16, W=64, A from seed [0x43;32], digits in [-8,8) with
    P(0)=17/32: the two device paths agree bit for bit; witness 0 equals the
    oracle's cached-NTT commit exactly, and all 64 witnesses pass the
    oracle's evaluation check against the key; linearity holds across
    witnesses."""
    CM, D, R = mods
    from lzx_amd import workloads as Wk
    q = O.Q_32[0]
    k, m, W = 16, 1 << 16, 64
    pk = CM.AjtaiPublicKey.from_seed(_params(mods, 6, k, m, 8), bytes([0x43]) * 32)
    digits = Wk.small_digits(0xC3, W * m * 64).reshape(W, m, 64)
    dd = D.to_device(digits)
    o_ntt = D.empty(W * k * 64, np.uint32)
    o_i8 = D.empty(W * k * 64, np.uint32)
    pk.commit_digits_batch(dd, W, o_ntt, path=CM.COMMIT_NTT)
    pk.commit_digits_batch(dd, W, o_i8, path=CM.COMMIT_INT8)
    a = D.to_host(o_ntt, np.uint32).reshape(W, k, 64)
    b = D.to_host(o_i8, np.uint32).reshape(W, k, 64)
    assert np.array_equal(a, b)
    # oracle spot check of witness 0
    Ah = pk.matrix_ntt()
    A = pk.matrix()
    ot = O.Ntt(*O.Q_32, 6)
    for idx in (0, 12345, k * m - 1):
        assert np.array_equal(ot.forward(A[idx])[1], Ah[idx])
    s0 = np.mod(digits[0].astype(np.int64), q).astype(np.uint32)
    _, want0 = O.ajtai_commit(ot, k, m, Ah, s0)
    assert np.array_equal(a[0], want0)
    # every witness against the key by evaluation at a root z of X^64 + 1
    psi = O.powmod64(O.Q_32[2], (q - 1) // 128, q)
    z = O.powmod64(psi, 77, q)
    Az = O.eval_key_at(q, A.reshape(k, m, 64), z)
    bad = [w for w in range(W) if O.commit_check_at_evals(q, Az, digits[w], a[w], z) != 0]
    assert not bad, bad
    # linearity: commit(s1 + s2) == commit(s1) + commit(s2) (digits stay small)
    s12 = (digits[1].astype(np.int16) + digits[2]).astype(np.int8)[None]
    o12 = D.empty(k * 64, np.uint32)
    pk.commit_digits_batch(D.to_device(np.ascontiguousarray(s12)), 1, o12, path=CM.COMMIT_INT8)
    got = D.to_host(o12, np.uint32).reshape(k, 64)
    assert np.array_equal(got, ((a[1].astype(np.uint64) + a[2]) % q).astype(np.uint32))


def test_c1_plain_commit_every_entry_point(mods):
    """The plain Ajtai workload exactly as stated: Q_32, d = 64, k = 8,
    m = 1024, A from seed [0x42; 32], one witness from splitmix64(0xC1) in
    [-8, 8).  The reference-signature host commit, the naive host path, the
    residue batch, the digit batch on every path and the allocation-free
    entry point all equal the first-principles schoolbook vector."""
    CM, D, R = mods
    from lzx_amd import workloads as Wk
    from lzx_amd._lib import lib
    v = VEC["c1"]
    q, k, m = v["q"], v["k"], v["m"]
    want = np.array(v["commitment"], dtype=np.uint32)
    pk = CM.AjtaiPublicKey.from_seed(_params(mods, 6, k, m, 8), bytes([v["key_seed_byte"]]) * 32)
    digits = Wk.small_digits(v["witness_seed"], m * 64, zero_half=False).reshape(1, m, 64)
    s = Wk.digits_to_residues(digits[0], q)
    t = pk.commit(s)
    assert np.array_equal(t.rows_array, want)
    assert np.array_equal(pk.commit_naive_reference(s).rows_array, want)
    pk.verify_opening(t, s)
    out = D.empty(k * 64, np.uint32)
    pk.commit_batch(D.to_device(s), 1, out)
    assert np.array_equal(D.to_host(out, np.uint32).reshape(k, 64), want)
    dd = D.to_device(digits)
    for path in (CM.COMMIT_NTT, CM.COMMIT_INT8, CM.COMMIT_AUTO):
        out = D.empty(k * 64, np.uint32)
        pk.commit_digits_batch(dd, 1, out, path=path)
        assert np.array_equal(D.to_host(out, np.uint32).reshape(k, 64), want), f"path {path}"
    out = D.empty(k * 64, np.uint32)
    assert lib().lzx_ajtai_commit_digits_int8_noalloc(pk.handle, D.ptr(dd), m, 1, D.ptr(out), D.stream_handle()) == 0
    assert np.array_equal(D.to_host(out, np.uint32).reshape(k, 64), want)


def test_commit_naive_equals_commit(mods):
    """commit_naive_reference == commit (ajtai.rs:490-506's differential test)
    on the reference's four shapes plus batched witnesses, through the device
    naive path (coefficient-domain A, one ring product per (i, j))."""
    CM, D, R = mods
    q = O.Q_32[0]
    for v in VEC["ajtai"]:
        pk = CM.AjtaiPublicKey.from_seed(_params(mods, v["log_n"], v["k"], v["m"], v["norm_bound"]),
                                         bytes([v["key_seed_byte"]]) * 32)
        s = O.sample_small_secret(v["q"], v["log_n"], v["m"], v["density"], v["secret_seed"].encode())
        assert pk.commit_naive_reference(s).rows_array.tolist() == v["commitment"]
    for log_n, k, m, W in [(6, 4, 37, 3), (5, 3, 200, 2), (8, 2, 20, 2), (2, 5, 9, 4)]:
        n = 1 << log_n
        pk = CM.AjtaiPublicKey.from_seed(_params(mods, log_n, k, m, 8), bytes([log_n * 3 + k]) * 32)
        rng = np.random.default_rng(m)
        s = rng.integers(0, q, size=(W, m, n), dtype=np.uint64).astype(np.uint32)
        s[:, 1::4] = 0
        ds = D.to_device(s)
        a = D.empty(W * k * n, np.uint32)
        b = D.empty(W * k * n, np.uint32)
        pk.commit_batch(ds, W, a)
        pk.commit_naive_batch(ds, W, b)
        assert np.array_equal(D.to_host(a, np.uint32), D.to_host(b, np.uint32)), (log_n, k, m, W)


def test_sample_small_secret_matches_oracle(mods):
    """sample_small_secret (ajtai.rs:307-348) on the device: the golden
    vectors (hashlib), the oracle over several spans, seed lengths and ring
    degrees, int8 digits beside residues, and the argument checks."""
    CM, D, R = mods
    from lzx_amd._lib import LzxError
    for v in VEC["small_secret"]:
        ring = R.RingConfig(R.Modulus32(v["q"], 30, 5), v["log_n"])
        got = CM.sample_small_secret(ring, v["m"], v["bound"], v["seed"].encode())
        assert D.to_host(got, np.uint32).tolist() == v["coeffs"]
    cases = [(O.Q_32, 6, 300, 8, bytes(range(32))), (O.Q_32, 6, 5, 127, b""), (O.Q_32, 3, 50, 1, b"x" * 95),
             (O.Q_12289, 5, 77, 3000, b"secret"), (O.Q_2013265921, 8, 9, (1 << 31) - 1, b"wide"),
             (O.Q_32, 0, 10, 2, b"n1"), (O.Q_32, 10, 4, 100, b"big-n")]
    for mod, log_n, m, bound, seed in cases:
        ring = R.RingConfig(R.Modulus32(*mod), log_n)
        want = O.sample_small_secret(mod[0], log_n, m, bound, seed)
        got = CM.sample_small_secret(ring, m, bound, seed)
        assert np.array_equal(D.to_host(got, np.uint32), want.reshape(m, -1)), (log_n, m, bound)
        if bound <= 127:
            res, dig = CM.sample_small_secret(ring, m, bound, seed, out=D.empty((m, 1 << log_n), np.uint32),
                                              digits=D.empty((m, 1 << log_n), np.int8))
            d = D.to_host(dig, np.int8).astype(np.int64)
            assert np.array_equal(np.mod(d, mod[0]).astype(np.uint32), want.reshape(m, -1))
            assert np.array_equal(D.to_host(res, np.uint32), want.reshape(m, -1))
    ring = R.RingConfig(R.Modulus32.Q_32, 4)
    for bound, seed, digits, code in [(1 << 31, b"s", False, 30), (200, b"s", True, 30), (3, b"s" * 96, False, 31)]:
        with pytest.raises(LzxError) as e:
            CM.sample_small_secret(ring, 2, bound, seed, digits=D.empty((2, 16), np.int8) if digits else None)
        assert e.value.code == code


def test_dimension_mismatch_batch(mods):
    CM, D, R = mods
    from lzx_amd._lib import LzxError
    pk = CM.AjtaiPublicKey.from_seed(_params(mods, 6, 2, 8, 8), bytes([1]) * 32)
    out = D.empty(2 * 64, np.uint32)
    with pytest.raises(LzxError) as e:
        pk.commit_digits_batch(D.to_device(np.zeros((1, 7, 64), np.int8)), 1, out, s_len=7)
    assert e.value.code == 4


def test_auto_path_is_measured_and_kept(mods):
    """LZX_COMMIT_AUTO times both digit paths on its first call of a shape,
    keeps the faster, and gives the same commitment as either path."""
    CM, D, R = mods
    k, m, W = 3, 96, 2
    pk = CM.AjtaiPublicKey.from_seed(CM.AjtaiParams(R.RingConfig(R.Modulus32.Q_32, 6), k, m, 8), bytes([0x5A]) * 32)
    digits = D.to_device(np.random.default_rng(77).integers(-128, 128, size=(W, m, 64)).astype(np.int8))
    assert pk.commit_path_choice(W) == CM.COMMIT_AUTO  # nothing measured for this shape yet
    outs = []
    for path in (CM.COMMIT_AUTO, CM.COMMIT_NTT, CM.COMMIT_INT8, CM.COMMIT_AUTO):
        o = D.empty(W * k * 64, np.uint32)
        pk.commit_digits_batch(digits, W, o, path=path)
        outs.append(D.to_host(o, np.uint32))
    assert all(np.array_equal(outs[0], x) for x in outs[1:])
    assert pk.commit_path_choice(W) in (CM.COMMIT_NTT, CM.COMMIT_INT8)
    assert pk.commit_path_choice(W + 1) == CM.COMMIT_AUTO  # kept per W
    # a key without the int8 path always takes the NTT one
    pk32 = CM.AjtaiPublicKey.from_seed(CM.AjtaiParams(R.RingConfig(R.Modulus32.Q_32, 5), 2, 8, 8), bytes([1]) * 32)
    assert pk32.commit_path_choice(1) == CM.COMMIT_NTT

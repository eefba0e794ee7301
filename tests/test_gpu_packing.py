"""GPU parity of witness packing (packing.rs, sparse.rs, decomposition.rs) against
the oracle's restatement (pinned by the reference's own cases in test_oracle.py)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
Q32 = 3221225473


@pytest.fixture(scope="module")
def pk():
    from lzx_amd import commitment as CM
    from lzx_amd import device as D
    from lzx_amd import packing as PK
    from lzx_amd import ring as R
    D.require_gpu()
    return PK, D, R, CM


@pytest.mark.parametrize("log_n,count", [(3, 0), (3, 5), (4, 40), (6, 10000), (4, 1)])
def test_field_element_packing_matches_oracle(pk, log_n, count):
    PK, D, R, _ = pk
    cfg = R.RingConfig(R.Modulus32.Q_32, log_n)
    n = 1 << log_n
    rng = np.random.default_rng(count + n)
    vals = rng.integers(0, O.GL_P, count, dtype=np.uint64)
    if count >= 5:
        vals[:5] = [0, 1, O.GL_P - 1, O.GL_P // 2, 0x7FFFFFFF00000000]
    got = PK.pack_field_elements(cfg, vals)
    want = O.pack_field_elements(Q32, n, vals)
    assert np.array_equal(D.to_host(got, np.uint32).reshape(want.shape), want)
    back = D.to_host(PK.unpack_field_elements(cfg, got), np.uint64)
    st, wback = O.unpack_field_elements(Q32, n, want)
    assert st == 0 and back.tolist() == wback.tolist() and back[:count].tolist() == vals.tolist()


def test_field_unpacking_rejects_oversized_limbs(pk):
    from lzx_amd._lib import LzxError
    PK, D, R, _ = pk
    cfg = R.RingConfig(R.Modulus32.Q_32, 3)
    for bad_val in (1 << 22, Q32 - (1 << 22)):
        bad = np.zeros(8 * 3, dtype=np.uint32)
        bad[17] = bad_val  # anywhere, padding included (packing.rs:64-75)
        with pytest.raises(LzxError) as e:
            PK.unpack_field_elements(cfg, bad)
        assert e.value.code == 51
    ok = np.zeros(8, dtype=np.uint32)
    ok[0], ok[1] = (1 << 22) - 1, Q32 - ((1 << 22) - 1)
    assert O.unpack_field_elements(Q32, 8, ok)[0] == 0
    PK.unpack_field_elements(cfg, ok)


@pytest.mark.parametrize("n,nbits", [(8, 0), (8, 1), (16, 496), (16, 501), (64, 123457)])
def test_bit_packing_matches_oracle(pk, n, nbits):
    PK, D, R, _ = pk
    cfg = R.RingConfig(R.Modulus32.Q_32, int(np.log2(n)))
    rng = np.random.default_rng(nbits + 1)
    bits = rng.integers(0, 2, nbits, dtype=np.uint8)
    got = PK.pack_bits(cfg, bits)
    want = O.pack_bits(n, bits)
    assert np.array_equal(D.to_host(got, np.uint32).reshape(want.shape), want)
    back = D.to_host(PK.unpack_bits(cfg, got, nbits + 40), np.uint8)  # bits past the data read as 0
    assert back.tolist() == O.unpack_bits(n, want, nbits + 40).tolist()


def test_crt_packing_matches_oracle(pk):
    from lzx_amd._lib import LzxError
    PK, D, R, _ = pk
    cfg = R.RingConfig(R.Modulus32.Q_32, 4)
    moduli = [7, 11, 13, 17]
    rng = np.random.default_rng(3)
    vals = np.array([int(rng.integers(0, moduli[i % 4])) for i in range(1001)], dtype=np.uint64)
    got = PK.crt_pack(cfg, vals, moduli)
    st, want = O.crt_pack(Q32, 16, vals, moduli)
    assert st == 0 and np.array_equal(D.to_host(got, np.uint32).reshape(want.shape), want)
    back = D.to_host(PK.crt_unpack(cfg, got, moduli, 1001), np.uint64)
    assert back.tolist() == vals.tolist()
    for vs, ms, code in (([7, 0], [7, 11], 53), ([1, 1, 1, 1], [1000000, 1000003, 1000019, 999999], 52),
                         ([1, 1], [6, 4], 54)):
        assert O.crt_pack(Q32, 16, vs, ms)[0] == code
        with pytest.raises(LzxError) as e:
            PK.crt_pack(cfg, np.array(vs, dtype=np.uint64), ms)
        assert e.value.code == code
    with pytest.raises(LzxError) as e:  # more values than the elements carry
        PK.crt_unpack(cfg, got, moduli, 4 * 251 + 1)
    assert e.value.code == 50


@pytest.mark.parametrize("base,nd,db", [(256, 9, 8), (16, 17, 1), (10, 21, 8), (2, 65, 1), (255, 9, 1), (4, 33, 8),
                                        (2, 125, 1), (1 << 32, 3, 8)])
def test_gadget_decomposition_matches_oracle(pk, base, nd, db):
    PK, D, R, _ = pk
    g = PK.GadgetDecomposition(base, nd)
    rng = np.random.default_rng(base)
    xs = np.concatenate([np.array([0, 1, 127, 128, 129, 255, 256, 65535, 2**32 - 1, 2**64 - 2, 2**64 - 1],
                                  dtype=np.uint64), rng.integers(0, 2**63, 2000, dtype=np.uint64)])
    dg = g.decompose_batch(xs, digit_bytes=db)
    got = D.to_host(dg, np.int8 if db == 1 else np.int64).reshape(len(xs), nd)
    for i, x in enumerate(xs.tolist()):
        st, want = O.gadget_decompose(base, nd, x)
        assert st == 0 and got[i].tolist() == want.tolist()
    back = D.to_host(g.recompose_batch(dg, digit_bytes=db), np.uint64)
    assert back.tolist() == xs.tolist()


def test_gadget_errors(pk):
    from lzx_amd._lib import LzxError
    PK, D, R, _ = pk
    g = PK.GadgetDecomposition(256, 2)
    with pytest.raises(LzxError) as e:  # decomposition.rs:66-68
        g.decompose_batch(np.array([70000], dtype=np.uint64))
    assert e.value.code == 58
    g9 = PK.GadgetDecomposition(256, 9)
    digits = np.zeros(9, dtype=np.int64)
    digits[0] = 200
    with pytest.raises(LzxError) as e:  # decomposition.rs:80-86
        g9.recompose(digits)
    assert e.value.code == 56
    with pytest.raises(LzxError) as e:
        g9.recompose([0] * 7)
    assert e.value.code == 55
    with pytest.raises(LzxError) as e:  # sum above 2^64 - 1
        PK.GadgetDecomposition(2**32, 3).recompose([0, 0, 2**31])
    assert e.value.code == 57


def test_commit_onehot_columns_matches_oracle(pk):
    from lzx_amd._lib import LzxError
    PK, D, R, CM = pk
    q, log_n, k, m = Q32, 6, 4, 8
    seed = bytes([0x61]) * 32
    params = CM.AjtaiParams(R.RingConfig(R.Modulus32.Q_32, log_n), k, m, 8)
    key = CM.AjtaiPublicKey.from_seed(params, seed)
    rng = np.random.default_rng(7)
    cols = []
    for _ in range(300):
        ln = int(rng.integers(1, 31 * 64 * m))
        cols.append(PK.OnehotColumn(ln, int(rng.integers(0, ln + 5)), 1))  # a few positions past len: empty
    t, s = PK.commit_onehot_columns(key, cols)
    want_s = np.zeros((m, 64), dtype=np.uint64)
    for c in cols:
        pkd = O.pack_bits(64, c.indicator())
        want_s[: pkd.shape[0]] = (want_s[: pkd.shape[0]] + pkd) % q
    want_s = want_s.astype(np.uint32)
    assert np.array_equal(D.to_host(s, np.uint32).reshape(m, 64), want_s)
    _, Ah = O.ajtai_from_seed(O.Q_32, log_n, k, m, seed, want_A=False)
    _, want_t = O.ajtai_commit(O.Ntt(*O.Q_32, log_n), k, m, Ah, want_s.reshape(-1))
    assert np.array_equal(t.rows_array, want_t.reshape(k, 64))
    with pytest.raises(LzxError) as e:  # a column needing more than m elements (sparse.rs:91-96)
        PK.commit_onehot_columns(key, [PK.OnehotColumn(31 * 64 * m + 1, 0, 1)])
    assert e.value.code == 4


def test_commit_onehot_reference_cases(pk):
    """sparse.rs:229-293 on the device: Hamming weight, one element per spread column, homomorphism, overflow."""
    from lzx_amd._lib import LzxError
    PK, D, R, CM = pk
    q = Q32
    params = CM.AjtaiParams(R.RingConfig(R.Modulus32.Q_32, 5), 2, 8, 1 << 23)
    key = CM.AjtaiPublicKey.from_seed(params, bytes([77]) * 32)
    t, s = PK.commit_onehot_columns(key, [PK.OnehotColumn(256, p, 5) for p in (3, 100, 200)])
    sh = D.to_host(s, np.uint32).reshape(8, 32)
    assert int(np.count_nonzero(sh.any(axis=1))) == 1
    assert sum(bin(int(c)).count("1") for c in sh[0]) == 3
    key.verify_opening(t, sh)
    spread = [PK.OnehotColumn(992 * 8, i * 992, 1) for i in range(8)]
    t2, s2 = PK.commit_onehot_columns(key, spread)
    s2h = D.to_host(s2, np.uint32).reshape(8, 32)
    assert int(np.count_nonzero(s2h.any(axis=1))) == 8
    key.verify_opening(t2, s2h)
    acc = key.commit(np.zeros((8, 32), dtype=np.uint32)).rows_array.astype(np.uint64)
    for c in spread:
        tc, sc = PK.commit_onehot_columns(key, [c])
        key.verify_opening(tc, D.to_host(sc, np.uint32).reshape(8, 32))
        acc = (acc + tc.rows_array) % q
    assert np.array_equal(acc.astype(np.uint32), t2.rows_array)
    small = CM.AjtaiPublicKey.from_seed(CM.AjtaiParams(R.RingConfig(R.Modulus32.Q_32, 2), 1, 2, 1 << 23),
                                        bytes([77]) * 32)
    with pytest.raises(LzxError) as e:  # 400 bits need 4 elements of 124 bits; m = 2
        PK.commit_onehot_columns(small, [PK.OnehotColumn(400, 3, 1)])
    assert e.value.code == 4

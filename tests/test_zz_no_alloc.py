"""GPU test of lzx_ajtai_commit_digits_int8_noalloc, the allocation-free form
of lzx_ajtai_commit_digits_batch(path = LZX_COMMIT_INT8), called through the
C ABI.  Its output must equal the pooled entry point's bit for bit on two
calls in a row (the second smaller, so stale scratch contents would show),
and it must refuse a shape whose partial sums exceed its fixed array."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
LZX_UNSUPPORTED = 31
LIMIT = 16 << 20  # LZX_NOALLOC_SCRATCH_BYTES


@pytest.fixture(scope="module")
def env():
    from lzx_amd import commitment as CM
    from lzx_amd import device as D
    from lzx_amd import ring as R
    from lzx_amd._lib import lib
    D.require_gpu()
    return CM, D, R, lib()


def _key(env, k, m, seed):
    CM, D, R, L = env
    params = CM.AjtaiParams(R.RingConfig(R.Modulus32.Q_32, 6), k, m, 128)
    return CM.AjtaiPublicKey.from_seed(params, bytes([seed]) * 32)


def _digits(rng, W, m):
    d = rng.integers(-8, 8, size=(W, m, 64)).astype(np.int8)
    d[:, ::7] = 0
    d[:, 3, :5] = -128
    d[:, 5] = rng.choice(np.array([-128, 127, -127], dtype=np.int8), size=(W, 64))
    return d


def _both(env, pk, digits):
    """(pooled int8 path, no-alloc path) outputs as [W][k][64] u32."""
    CM, D, R, L = env
    W, m, _ = digits.shape
    k = pk.params.k
    dd = D.to_device(digits)
    ref = D.empty(W * k * 64, np.uint32)
    pk.commit_digits_batch(dd, W, ref, path=CM.COMMIT_INT8)
    got = D.empty(W * k * 64, np.uint32)
    got.fill_(-1)
    st = L.lzx_ajtai_commit_digits_int8_noalloc(pk._h, D.ptr(dd), m, W, D.ptr(got), D.stream_handle())
    assert st == 0, L.lzx_status_string(st)
    return (D.to_host(ref, np.uint32).reshape(W, k, 64), D.to_host(got, np.uint32).reshape(W, k, 64))


def test_noalloc_equals_pooled_int8_commit(env):
    CM, D, R, L = env
    rng = np.random.default_rng(0x7A)
    q = O.Q_32[0]
    # first call: k = 16, m = 3000 (two fold-back chunks at least), W = 40
    pk1 = _key(env, 16, 3000, 0x51)
    d1 = _digits(rng, 40, 3000)
    ref, got = _both(env, pk1, d1)
    assert np.array_equal(got, ref)
    Ah = pk1.matrix_ntt()
    ot = O.Ntt(*O.Q_32, 6)
    for w in (0, 39):
        s = np.mod(d1[w].astype(np.int64), q).astype(np.uint32)
        assert np.array_equal(got[w], O.ajtai_commit(ot, 16, 3000, Ah, s)[1])
    # second call, smaller input on another key: k = 4, m = 300, W = 3
    pk2 = _key(env, 4, 300, 0x52)
    d2 = _digits(rng, 3, 300)
    ref, got = _both(env, pk2, d2)
    assert np.array_equal(got, ref)
    # and a smaller W on the first key after the larger call
    d3 = _digits(rng, 2, 3000)
    ref, got = _both(env, pk1, d3)
    assert np.array_equal(got, ref)


def test_noalloc_refuses_oversized_partials(env):
    CM, D, R, L = env
    k, m = 64, 4096
    W = LIMIT // (3 * k * 256) + 1  # ceil(4096 / 2032) = 3 chunks of W * k * 256 bytes
    assert 3 * W * k * 256 > LIMIT and 3 * (W - 1) * k * 256 <= LIMIT
    pk = _key(env, k, m, 0x53)
    dd = D.to_device(np.zeros((W, m, 64), np.int8))
    out = D.empty(W * k * 64, np.uint32)
    out.fill_(7)
    st = L.lzx_ajtai_commit_digits_int8_noalloc(pk._h, D.ptr(dd), m, W, D.ptr(out), D.stream_handle())
    assert st == LZX_UNSUPPORTED
    assert b"LZX_NOALLOC_SCRATCH_BYTES" in L.lzx_last_device_error()
    assert (D.to_host(out, np.uint32) == 7).all()
    # one witness fewer fits, and matches the pooled path
    rng = np.random.default_rng(5)
    d = _digits(rng, W - 1, m)
    ref, got = _both(env, pk, d)
    assert np.array_equal(got, ref)
    # the dimension check comes first, as in the pooled entry point
    st = L.lzx_ajtai_commit_digits_int8_noalloc(pk._h, D.ptr(dd), m - 1, 1, D.ptr(out), D.stream_handle())
    assert st == 4

"""GPU parity of the ~2^50 quadratic-slot ring (crates/lattice-ring/src/modulus50.rs)
against the oracle's restatement (itself pinned by schoolbook products and the
slot-reduction closed form in test_oracle.py)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
Q, ADIC, GEN = O.Q50


@pytest.fixture(scope="module")
def r50():
    from lzx_amd import device as D
    from lzx_amd import ring50 as R
    D.require_gpu()
    return R, D


# (log_n, polynomials): whole and partial 128-value tiles for every n
SHAPES = [(1, 1), (1, 77), (2, 33), (3, 1), (3, 50), (4, 9), (5, 5), (5, 203), (6, 3), (6, 1001), (7, 1), (7, 517)]


@pytest.mark.parametrize("log_n,polys", SHAPES)
def test_transforms_and_products_match_oracle(r50, log_n, polys):
    R, D = r50
    n = 1 << log_n
    rng = np.random.default_rng(log_n * 1000 + polys)
    a = rng.integers(0, Q, polys * n, dtype=np.uint64)
    b = rng.integers(0, Q, polys * n, dtype=np.uint64)
    T = R.QuadSlotTables(R.Modulus50.Q_50, log_n)
    assert T.leaf_gamma.tolist() == O.quad_tables(Q, ADIC, GEN, log_n)[4].tolist()
    ta = D.to_device(a)
    T.forward_batch(ta)
    _, fa = O.quad_forward(Q, ADIC, GEN, log_n, a)
    assert np.array_equal(D.to_host(ta, np.uint64), fa)
    _, fb = O.quad_forward(Q, ADIC, GEN, log_n, b)
    out = D.empty(polys * n, np.uint64)
    T.slotwise_mul_batch(ta, D.to_device(fb), out)
    _, sm = O.quad_slotwise_mul(Q, ADIC, GEN, log_n, fa, fb)
    assert np.array_equal(D.to_host(out, np.uint64), sm)
    T.inverse_batch(out)
    _, prod = O.quad_mul(Q, ADIC, GEN, log_n, a, b)
    assert np.array_equal(D.to_host(out, np.uint64), prod)
    out2 = D.empty(polys * n, np.uint64)
    T.mul_batch(D.to_device(a), D.to_device(b), out2)  # the fused forward-slot-inverse kernel
    assert np.array_equal(D.to_host(out2, np.uint64), prod)
    T.inverse_batch(ta)
    assert np.array_equal(D.to_host(ta, np.uint64), a)


@pytest.mark.parametrize("log_n", range(1, 8))
def test_extreme_residues_every_path(r50, log_n):
    """Inputs drawn from the largest and middle residues (q-1, q-2, (q+-1)/2, 0, 1)
    push the kernels' intermediate ranges to their bounds on every path."""
    R, D = r50
    n = 1 << log_n
    polys = 300
    rng = np.random.default_rng(50 + log_n)
    pool = np.array([Q - 1, Q - 2, (Q - 1) // 2, (Q + 1) // 2, 0, 1], dtype=np.uint64)
    a = pool[rng.integers(0, len(pool), polys * n)]
    b = pool[rng.integers(0, 2, polys * n)]  # q-1 and q-2 only
    a[:n] = Q - 1
    b[:n] = Q - 1
    T = R.QuadSlotTables(R.Modulus50.Q_50, log_n)
    ta = D.to_device(a)
    T.forward_batch(ta)
    _, fa = O.quad_forward(Q, ADIC, GEN, log_n, a)
    assert np.array_equal(D.to_host(ta, np.uint64), fa)
    ti = D.to_device(a)  # a as slot-form input to the inverse
    T.inverse_batch(ti)
    _, back = O.quad_inverse(Q, ADIC, GEN, log_n, a)
    assert np.array_equal(D.to_host(ti, np.uint64), back)
    out = D.empty(polys * n, np.uint64)
    T.slotwise_mul_batch(D.to_device(a), D.to_device(b), out)
    _, sm = O.quad_slotwise_mul(Q, ADIC, GEN, log_n, a, b)
    assert np.array_equal(D.to_host(out, np.uint64), sm)
    T.mul_batch(D.to_device(a), D.to_device(b), out)
    _, prod = O.quad_mul(Q, ADIC, GEN, log_n, a, b)
    assert np.array_equal(D.to_host(out, np.uint64), prod)


def test_edge_values_and_ring_identities(r50):
    R, D = r50
    ring = R.RingConfig50(7)
    n = 128
    one = np.zeros(n, dtype=np.uint64)
    one[0] = 1
    x = np.zeros(n, dtype=np.uint64)
    x[1] = 1
    xn1 = np.zeros(n, dtype=np.uint64)
    xn1[n - 1] = 1
    top = np.full(n, Q - 1, dtype=np.uint64)  # every coefficient -1: the largest residues
    assert ring.mul(x, xn1).tolist() == [Q - 1] + [0] * (n - 1)
    assert ring.mul(top, one).tolist() == top.tolist()
    _, want = O.quad_mul(Q, ADIC, GEN, 7, top, top)
    assert ring.mul(top, top).tolist() == want.tolist()
    v = ring.from_signed([-3, -2, -1, 0, 1, 2, 3, 4, 5])
    assert int(v[0]) == Q - 3 and ring.infinity_norm(v) == 5


def test_errors(r50):
    from lzx_amd._lib import LzxError
    R, D = r50
    for log_n, code in ((0, 40), (8, 40)):
        with pytest.raises(LzxError) as e:
            R.QuadSlotTables(R.Modulus50.Q_50, log_n)
        assert e.value.code == code
    with pytest.raises(LzxError) as e:
        R.QuadSlotTables(R.Modulus50(Q, 5, GEN), 3)
    assert e.value.code == 41
    with pytest.raises(LzxError) as e:
        R.QuadSlotTables(R.Modulus50(Q, 7, 1), 3)
    assert e.value.code == 42
    T = R.QuadSlotTables(R.Modulus50.Q_50, 3)
    with pytest.raises(LzxError) as e:
        T.forward_batch(D.to_device(np.zeros(12, dtype=np.uint64)))
    assert e.value.code == 43


def test_large_batch_properties(r50):
    # 2^16 polynomials of n = 128: products against the identity and commutativity
    R, D = r50
    polys, n = 1 << 16, 128
    rng = np.random.default_rng(9)
    a = D.to_device(rng.integers(0, Q, polys * n, dtype=np.uint64))
    b = D.to_device(rng.integers(0, Q, polys * n, dtype=np.uint64))
    T = R.QuadSlotTables(R.Modulus50.Q_50, 7)
    ab, ba = D.empty(polys * n, np.uint64), D.empty(polys * n, np.uint64)
    T.mul_batch(a, b, ab)
    T.mul_batch(b, a, ba)
    assert bool((ab == ba).all())
    one = np.zeros(polys * n, dtype=np.uint64)
    one[::n] = 1
    ida = D.empty(polys * n, np.uint64)
    T.mul_batch(a, D.to_device(one), ida)
    assert bool((ida == a).all())
    idx = rng.integers(0, polys, 4)
    hab = D.to_host(ab, np.uint64).reshape(polys, n)
    ha, hb = D.to_host(a, np.uint64).reshape(polys, n), D.to_host(b, np.uint64).reshape(polys, n)
    for p in idx:
        assert hab[p].tolist() == O.quad_mul(Q, ADIC, GEN, 7, ha[p], hb[p])[1].tolist()

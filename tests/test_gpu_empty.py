"""Empty batches through every batched device entry point: each returns without
error and without touching its (possibly null) buffers.  An empty torch tensor
has a null data pointer, so these also pin the C-ABI's rule that a zero-length
batch may pass null pointers."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from lzx_amd import device as D
    from lzx_amd import extension as E
    from lzx_amd import ring as R
    from lzx_amd import ring50 as R50
    D.require_gpu()
    return D, E, R, R50


def test_empty_ntt_and_ring_products(mods):
    D, _, R, _ = mods
    T = R.NttTables(R.Modulus32.Q_32, 6)
    empty = D.empty(0, np.uint32)
    T.forward_batch(empty)
    T.inverse_batch(empty)
    T.ring_mul_batch(empty, empty, D.empty(0, np.uint32))
    R.forward_multi([T, T], empty, 0)
    D.synchronize()


def test_empty_extension_ring(mods):
    D, E, R, _ = mods
    ring = R.RingConfig(R.Modulus32.Q_32, 6)
    a = D.empty((0, 2, 64), np.uint32)
    assert E.mul_batch(ring, a, a).numel() == 0
    assert E.norm_batch(ring, a).numel() == 0
    ok, inv = E.inverse_batch(ring, a)
    assert ok.numel() == 0 and inv.numel() == 0
    for fn in (E.add_batch, E.sub_batch):
        assert fn(ring, a, a).numel() == 0
    for fn in (E.neg_batch, E.conjugate_batch):
        assert fn(ring, a).numel() == 0
    D.synchronize()


def test_empty_quadratic_slot_ring(mods):
    D, _, _, R50 = mods
    T = R50.QuadSlotTables(R50.Modulus50.Q_50, 7)
    e = D.empty(0, np.uint64)
    T.forward_batch(e)
    T.inverse_batch(e)
    T.slotwise_mul_batch(e, e, D.empty(0, np.uint64))
    T.mul_batch(e, e, D.empty(0, np.uint64))
    D.synchronize()

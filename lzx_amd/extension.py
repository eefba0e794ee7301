"""Host-side mirror of the quadratic extension ring R_q[Y]/(Y^2+1)
(crates/lattice-ring/src/extension.rs:27-176); kernels in csrc/rq2.hip.

An element c0 + c1 Y is [c0 | c1] (n canonical u32 coefficients each); the
batched forms take device tensors of count elements laid out [count][2][n].
The element-wise `Rq2` methods upload, run the batched kernels and download.
"""
from __future__ import annotations

import numpy as np

from . import device as dev
from ._lib import LzxError, check, lib
from .ring import RingConfig, RingElement

LZX_RING_LENGTH_MISMATCH = 3


def _plan(ring: RingConfig):
    return ring.tables.handle


def _count(ring: RingConfig, t) -> int:
    per = 2 * ring.n()
    if t.numel() % per:
        raise LzxError(LZX_RING_LENGTH_MISMATCH, "Rq2 batch")
    return t.numel() // per


def mul_batch(ring: RingConfig, a, b, out=None, stream=None):
    """Rq2::mul (extension.rs:115-124) over a batch; returns out."""
    out = dev.empty(tuple(a.shape), np.uint32) if out is None else out
    check(lib().lzx_rq2_mul_batch(_plan(ring), dev.ptr(a), dev.ptr(b), dev.ptr(out), _count(ring, a),
                                  dev.stream_handle(stream)), "Rq2::mul")
    return out


def norm_batch(ring: RingConfig, a, out=None, stream=None):
    """Rq2::norm (extension.rs:142-147): [count][n] = c0^2 + c1^2."""
    count = _count(ring, a)
    out = dev.empty((count, ring.n()), np.uint32) if out is None else out
    check(lib().lzx_rq2_norm_batch(_plan(ring), dev.ptr(a), dev.ptr(out), count, dev.stream_handle(stream)),
          "Rq2::norm")
    return out


def inverse_batch(ring: RingConfig, a, out=None, stream=None):
    """Rq2::inverse (extension.rs:151-167): (ok [count] u8 device tensor, out);
    ok[i] = 0 is the reference's None (out[i] is then zero)."""
    count = _count(ring, a)
    out = dev.empty(tuple(a.shape), np.uint32) if out is None else out
    ok = dev.empty(max(1, count), np.int8)
    check(lib().lzx_rq2_inverse_batch(_plan(ring), dev.ptr(a), dev.ptr(out), dev.ptr(ok), count,
                                      dev.stream_handle(stream)), "Rq2::inverse")
    return ok[:count], out


def _linear(fn, what, ring, a, b, out, stream):
    out = dev.empty(tuple(a.shape), np.uint32) if out is None else out
    args = [_plan(ring), dev.ptr(a)] + ([dev.ptr(b)] if b is not None else []) + [dev.ptr(out), _count(ring, a),
                                                                                   dev.stream_handle(stream)]
    check(getattr(lib(), fn)(*args), what)
    return out


def add_batch(ring, a, b, out=None, stream=None):
    return _linear("lzx_rq2_add_batch", "Rq2::add", ring, a, b, out, stream)


def sub_batch(ring, a, b, out=None, stream=None):
    return _linear("lzx_rq2_sub_batch", "Rq2::sub", ring, a, b, out, stream)


def neg_batch(ring, a, out=None, stream=None):
    return _linear("lzx_rq2_neg_batch", "Rq2::neg", ring, a, None, out, stream)


def conjugate_batch(ring, a, out=None, stream=None):
    return _linear("lzx_rq2_conjugate_batch", "Rq2::conjugate", ring, a, None, out, stream)


class Rq2:
    """Rq2 (extension.rs:27-176): c0 + c1 Y with Y^2 = -1."""

    def __init__(self, c0: RingElement, c1: RingElement):
        """Rq2::new (extension.rs:54-63): RingError::LengthMismatch unless both halves share q and n."""
        if c0.config.modulus.q != c1.config.modulus.q or c0.config.log_n != c1.config.log_n:
            raise LzxError(LZX_RING_LENGTH_MISMATCH, "Rq2::new")
        self.c0, self.c1 = c0, c1

    @property
    def ring(self) -> RingConfig:
        return self.c0.config

    @classmethod
    def zero(cls, ring: RingConfig) -> "Rq2":
        return cls(ring.zero(), ring.zero())

    @classmethod
    def one(cls, ring: RingConfig) -> "Rq2":
        c0 = np.zeros(ring.n(), dtype=np.uint32)
        c0[0] = 1
        return cls(RingElement(ring, c0), ring.zero())

    @classmethod
    def y(cls, ring: RingConfig) -> "Rq2":
        c1 = np.zeros(ring.n(), dtype=np.uint32)
        c1[0] = 1
        return cls(ring.zero(), RingElement(ring, c1))

    def is_zero(self) -> bool:
        return self.c0.is_zero() and self.c1.is_zero()

    def _array(self) -> np.ndarray:
        return np.concatenate([self.c0.coeffs, self.c1.coeffs])

    @classmethod
    def _from(cls, ring: RingConfig, arr) -> "Rq2":
        n = ring.n()
        a = np.asarray(arr, dtype=np.uint32).reshape(-1)
        return cls(RingElement(ring, a[:n].copy()), RingElement(ring, a[n:].copy()))

    def _check(self, other: "Rq2", what: str) -> None:
        if other.ring.modulus.q != self.ring.modulus.q or other.ring.log_n != self.ring.log_n:
            raise LzxError(LZX_RING_LENGTH_MISMATCH, what)

    def _binary(self, other: "Rq2", fn) -> "Rq2":
        r = fn(self.ring, dev.to_device(self._array()), dev.to_device(other._array()))
        return Rq2._from(self.ring, dev.to_host(r, np.uint32))

    def add(self, other: "Rq2") -> "Rq2":
        self._check(other, "Rq2::add")
        return self._binary(other, add_batch)

    def sub(self, other: "Rq2") -> "Rq2":
        self._check(other, "Rq2::sub")
        return self._binary(other, sub_batch)

    def mul(self, other: "Rq2") -> "Rq2":
        self._check(other, "Rq2::mul")
        return self._binary(other, mul_batch)

    def square(self) -> "Rq2":
        return self.mul(self)

    def neg(self) -> "Rq2":
        return Rq2._from(self.ring, dev.to_host(neg_batch(self.ring, dev.to_device(self._array())), np.uint32))

    def conjugate(self) -> "Rq2":
        return Rq2._from(self.ring, dev.to_host(conjugate_batch(self.ring, dev.to_device(self._array())), np.uint32))

    def norm(self) -> RingElement:
        r = norm_batch(self.ring, dev.to_device(self._array()))
        return RingElement(self.ring, dev.to_host(r, np.uint32).reshape(-1))

    def inverse(self):
        """Some(Rq2) or None (extension.rs:151-167)."""
        ok, r = inverse_batch(self.ring, dev.to_device(self._array()))
        if not int(dev.to_host(ok, np.uint8)[0]):
            return None
        return Rq2._from(self.ring, dev.to_host(r, np.uint32))

    def to_bytes(self) -> bytes:
        """extension.rs:170-175: c0's bytes then c1's."""
        return self.c0.to_bytes() + self.c1.to_bytes()

    def __eq__(self, other) -> bool:
        return isinstance(other, Rq2) and self.c0 == other.c0 and self.c1 == other.c1


def sample_rq2_challenge(ring: RingConfig, spec, seed: bytes):
    """sample_rq2_challenge (extension.rs:180-216): (Rq2, g0, g1); spec.n must be 2n."""
    import ctypes as C

    n = ring.n()
    s = spec._c()
    out = np.zeros(2 * n, dtype=np.uint32)
    g0, g1 = C.c_uint64(), C.c_uint64()
    check(lib().lzx_rq2_sample_challenge(ring.modulus.q, ring.log_n, C.byref(s), seed, len(seed), out.ctypes.data,
                                         C.byref(g0), C.byref(g1)), "sample_rq2_challenge")
    return Rq2._from(ring, out), g0.value, g1.value

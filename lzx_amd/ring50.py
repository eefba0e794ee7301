"""Host-side mirror of lattice-ring's ~2^50 quadratic-slot ring.

  Modulus50        crates/lattice-ring/src/modulus50.rs:55-198
  QuadSlotTables   crates/lattice-ring/src/modulus50.rs:209-405
  RingConfig50     crates/lattice-ring/src/modulus50.rs:410-499

The transforms and products run batched on the device (torch int64 tensors
holding canonical residues, [poly][coeff]); single-element calls on host
arrays upload, run and download.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import device as dev
from ._lib import check, lib

Q_50 = 1125899906839937       # 2^50 - 2687 (modulus50.rs:44)
Q_50_TWO_ADICITY = 7
Q_50_GENERATOR = 3


@dataclass(frozen=True)
class Modulus50:
    """A ~2^50 prime with two-adicity >= 7 (host arithmetic for setup and checks)."""

    q: int = Q_50
    two_adicity: int = Q_50_TWO_ADICITY
    generator: int = Q_50_GENERATOR

    def reduce_u128(self, x: int) -> int:
        return x % self.q

    def reduce_i64(self, x: int) -> int:
        return x % self.q

    def mul(self, a: int, b: int) -> int:
        return a * b % self.q

    def add(self, a: int, b: int) -> int:
        return (a + b) % self.q

    def sub(self, a: int, b: int) -> int:
        return (a - b) % self.q

    def neg(self, a: int) -> int:
        return (-a) % self.q

    def pow(self, base: int, exp: int) -> int:
        return pow(base % self.q, exp, self.q)

    def inv(self, a: int):
        return None if a % self.q == 0 else pow(a, self.q - 2, self.q)

    def root_of_unity(self, k: int):
        return None if k > self.two_adicity else pow(self.generator, (self.q - 1) >> k, self.q)


Modulus50.Q_50 = Modulus50()


class QuadSlotTables:
    """QuadSlotTables::new(modulus, log_n) (modulus50.rs:241-300) as a device plan."""

    def __init__(self, modulus: Modulus50, log_n: int):
        self.modulus, self.log_n = modulus, log_n
        self._h = C.c_void_p()
        self._destroy = lib().lzx_quad_plan_destroy  # bound now: lib() may be gone at interpreter exit
        check(lib().lzx_quad_plan_create(modulus.q, modulus.two_adicity, modulus.generator, log_n,
                                         C.byref(self._h)), "QuadSlotTables::new")
        g = np.zeros(max(1, self.slots()), dtype=np.uint64)
        z = C.c_uint64()
        check(lib().lzx_quad_slot_constants(self._h, g.ctypes.data, C.byref(z)))
        self.zeta = z.value
        self.leaf_gamma = g[: self.slots()]

    def __del__(self):
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h and destroy:
            destroy(h)
            self._h = None

    def n(self) -> int:
        return 1 << self.log_n

    def slots(self) -> int:
        return self.n() // 2

    # batched device forms: tensors of len = polys * n
    def forward_batch(self, t, stream=None):
        check(lib().lzx_quad_forward_batch(self._h, dev.ptr(t), t.numel(), dev.stream_handle(stream)),
              "QuadSlotTables::forward")

    def inverse_batch(self, t, stream=None):
        check(lib().lzx_quad_inverse_batch(self._h, dev.ptr(t), t.numel(), dev.stream_handle(stream)),
              "QuadSlotTables::inverse")

    def slotwise_mul_batch(self, a, b, out, stream=None):
        if a.numel() != b.numel() or out.numel() != a.numel():
            raise ValueError("operand lengths differ")
        check(lib().lzx_quad_slotwise_mul_batch(self._h, dev.ptr(a), dev.ptr(b), dev.ptr(out), a.numel(),
                                                dev.stream_handle(stream)), "QuadSlotTables::slotwise_mul")

    def mul_batch(self, a, b, out, stream=None):
        if a.numel() != b.numel() or out.numel() != a.numel():
            raise ValueError("operand lengths differ")
        check(lib().lzx_quad_mul_batch(self._h, dev.ptr(a), dev.ptr(b), dev.ptr(out), a.numel(),
                                       dev.stream_handle(stream)), "QuadSlotTables::mul")

    # single-element host forms (the reference's signatures)
    def _host(self, fn, *arrays):
        ins = [dev.to_device(np.ascontiguousarray(x, dtype=np.uint64)) for x in arrays]
        out = ins[0] if len(ins) == 1 else dev.empty(ins[0].numel(), np.uint64)
        fn(*ins, out) if len(ins) > 1 else fn(ins[0])
        return dev.to_host(out, np.uint64)

    def forward(self, a):
        return self._host(self.forward_batch, a)

    def inverse(self, a):
        return self._host(self.inverse_batch, a)

    def slotwise_mul(self, a, b):
        return self._host(self.slotwise_mul_batch, a, b)

    def mul(self, a, b):
        return self._host(self.mul_batch, a, b)


class RingConfig50:
    """RingConfig50::new / with_modulus (modulus50.rs:425-438)."""

    def __init__(self, log_n: int, modulus: Modulus50 = Modulus50.Q_50):
        self.modulus, self.log_n = modulus, log_n
        self.tables = QuadSlotTables(modulus, log_n)

    def n(self) -> int:
        return 1 << self.log_n

    def slot_constants(self) -> np.ndarray:
        return self.tables.leaf_gamma

    def mul(self, a, b):
        return self.tables.mul(a, b)

    def add(self, a, b):
        a, b = np.asarray(a, dtype=np.uint64), np.asarray(b, dtype=np.uint64)
        if a.size != self.n() or b.size != self.n():
            raise ValueError("Mod50Error::LengthMismatch")
        return np.array([(int(x) + int(y)) % self.modulus.q for x, y in zip(a, b)], dtype=np.uint64)

    def sub(self, a, b):
        a, b = np.asarray(a, dtype=np.uint64), np.asarray(b, dtype=np.uint64)
        if a.size != self.n() or b.size != self.n():
            raise ValueError("Mod50Error::LengthMismatch")
        return np.array([(int(x) - int(y)) % self.modulus.q for x, y in zip(a, b)], dtype=np.uint64)

    def from_signed(self, signed) -> np.ndarray:
        out = np.zeros(self.n(), dtype=np.uint64)
        for i, s in enumerate(list(signed)[: self.n()]):
            out[i] = int(s) % self.modulus.q
        return out

    def infinity_norm(self, a) -> int:
        half = self.modulus.q // 2
        return max((int(c) if int(c) <= half else self.modulus.q - int(c) for c in a), default=0)

"""Host-side mirror of the witness-packing surface (device kernels in csrc/packing.hip).

  pack_field_elements / unpack_field_elements  crates/lattice-ring/src/packing.rs:30-91
  crt_pack / crt_unpack                        crates/lattice-ring/src/packing.rs:95-185
  pack_bits / unpack_bits, OnehotColumn,
  commit_onehot_columns                        crates/lattice-commitment/src/sparse.rs:30-144
  GadgetDecomposition                          crates/lattice-core/src/decomposition.rs:25-105

Batched forms take and return torch device tensors ([elems][n] u32 ring
elements, u64 values, i8 / i64 digits); the reference's element-wise helpers
upload, run and download.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import device as dev
from ._lib import check, lib
from .commitment import AjtaiCommitment, AjtaiPublicKey
from .ring import RingConfig

LIMB_BITS = 22
LIMBS_PER_VALUE = 3
BITS_PER_COEFF = 31


def _dev_u64(x):
    return x if hasattr(x, "data_ptr") else dev.to_device(np.ascontiguousarray(x, dtype=np.uint64))


def pack_field_elements(config: RingConfig, values, stream=None):
    """packing.rs:30-52: device tensor [elems][n] of u32 (three 22-bit limbs per value)."""
    v = _dev_u64(values)
    n = config.n()
    elems = lib().lzx_pack_field_elements_len(v.numel(), n)
    out = dev.empty((elems, n), np.uint32)
    check(lib().lzx_pack_field_elements(dev.ptr(v) if v.numel() else None, v.numel(), config.modulus.q, n,
                                        dev.ptr(out), dev.stream_handle(stream)), "pack_field_elements")
    return out


def unpack_field_elements(config: RingConfig, packed, stream=None):
    """packing.rs:57-91: every complete limb triple back to a Goldilocks value
    (PackingError::LimbOutOfRange when a balanced limb leaves (-2^22, 2^22))."""
    p = packed if hasattr(packed, "data_ptr") else dev.to_device(np.ascontiguousarray(packed, dtype=np.uint32))
    n = config.n()
    total = p.numel()
    out = dev.empty(max(1, total // 3), np.uint64)
    check(lib().lzx_unpack_field_elements(dev.ptr(p), total // n, config.modulus.q, n, dev.ptr(out),
                                          dev.stream_handle(stream)), "unpack_field_elements")
    return out[: total // 3]


def crt_pack(config: RingConfig, values, moduli, stream=None):
    """packing.rs:95-155: groups of len(moduli) values in coefficient 0 of one element each."""
    v = _dev_u64(values)
    m = np.ascontiguousarray(moduli, dtype=np.uint64)
    n = config.n()
    elems = max(1, -(-v.numel() // max(1, m.size)))
    out = dev.empty((elems, n), np.uint32)
    check(lib().lzx_crt_pack(dev.ptr(v) if v.numel() else None, v.numel(), m.ctypes.data if m.size else None,
                             m.size, config.modulus.q, n, dev.ptr(out), dev.stream_handle(stream)), "crt_pack")
    return out


def crt_unpack(config: RingConfig, packed, moduli, expected_values: int, stream=None):
    """packing.rs:158-185."""
    m = np.ascontiguousarray(moduli, dtype=np.uint64)
    n = config.n()
    out = dev.empty(max(1, expected_values), np.uint64)
    check(lib().lzx_crt_unpack(dev.ptr(packed), packed.numel() // n, n, m.ctypes.data if m.size else None, m.size,
                               expected_values, dev.ptr(out), dev.stream_handle(stream)), "crt_unpack")
    return out[:expected_values]


def pack_bits(ring: RingConfig, bits, stream=None):
    """sparse.rs:30-54: 31 bits per coefficient; device tensor [elems][n] u32."""
    b = bits if hasattr(bits, "data_ptr") else dev.to_device(np.ascontiguousarray(bits, dtype=np.uint8))
    n = ring.n()
    elems = lib().lzx_pack_bits_len(b.numel(), n)
    out = dev.empty((elems, n), np.uint32)
    check(lib().lzx_pack_bits(dev.ptr(b) if b.numel() else None, b.numel(), n, dev.ptr(out),
                              dev.stream_handle(stream)), "pack_bits")
    return out


def unpack_bits(ring: RingConfig, elems, num_bits: int, stream=None):
    """sparse.rs:57-71."""
    n = ring.n()
    out = dev.empty(max(1, num_bits), np.int8)
    check(lib().lzx_unpack_bits(dev.ptr(elems), elems.numel() // n, n, dev.ptr(out), num_bits,
                                dev.stream_handle(stream)), "unpack_bits")
    return out[:num_bits]


@dataclass
class OnehotColumn:
    """sparse.rs:75-107: a length-`len` indicator with its one bit at `position`."""

    len: int
    position: int
    value: int = 0

    def indicator(self) -> np.ndarray:
        bits = np.zeros(self.len, dtype=np.uint8)
        if self.position < self.len:
            bits[self.position] = 1
        return bits


def commit_onehot_columns(pk: AjtaiPublicKey, columns, stream=None):
    """sparse.rs:118-144: (commitment, packed witness s as a device tensor [m][n] u32)."""
    ring = pk.params.ring
    n, m, k = ring.n(), pk.params.m, pk.params.k
    pos = np.array([c.position for c in columns] or [0], dtype=np.uint64)
    lens = np.array([c.len for c in columns] or [0], dtype=np.uint64)
    s = dev.empty((m, n), np.uint32)
    t = dev.empty((k, n), np.uint32)
    check(lib().lzx_ajtai_commit_onehot(pk.handle, pos.ctypes.data, lens.ctypes.data, len(columns), dev.ptr(s),
                                        dev.ptr(t), dev.stream_handle(stream)), "commit_onehot_columns")
    return AjtaiCommitment(dev.to_host(t, np.uint32).reshape(k, n), ring), s


class GadgetDecomposition:
    """decomposition.rs:25-105: balanced base-b digits (a digit r > b/2 becomes r - b)."""

    def __init__(self, base: int, num_digits: int):
        self.base, self.num_digits = base, num_digits

    @classmethod
    def power_of_two(cls, bits: int, log_base: int) -> "GadgetDecomposition":
        return cls(1 << log_base, -(-(bits + 1) // log_base))

    def decompose_batch(self, x, digit_bytes: int = 8, stream=None):
        """[count][num_digits] digits on the device (int8 for bases <= 255 feeds the int8 commit)."""
        xs = _dev_u64(x)
        out = dev.empty((max(1, xs.numel()), self.num_digits), np.int8 if digit_bytes == 1 else np.uint64)
        check(lib().lzx_gadget_decompose(dev.ptr(xs) if xs.numel() else None, xs.numel(), self.base,
                                         self.num_digits, dev.ptr(out), digit_bytes, dev.stream_handle(stream)),
              "GadgetDecomposition::decompose")
        return out[: xs.numel()]

    def recompose_batch(self, digits, digit_bytes: int = 8, stream=None):
        count = digits.numel() // max(1, self.num_digits)
        out = dev.empty(max(1, count), np.uint64)
        check(lib().lzx_gadget_recompose(dev.ptr(digits), digit_bytes, count, self.base, self.num_digits,
                                         dev.ptr(out), dev.stream_handle(stream)), "GadgetDecomposition::recompose")
        return out[:count]

    def decompose(self, x: int) -> list:
        d = self.decompose_batch(np.array([x], dtype=np.uint64))
        return [int(v) for v in dev.to_host(d, np.int64).reshape(-1)]

    def recompose(self, digits) -> int:
        if len(digits) != self.num_digits:
            check(55, "GadgetDecomposition::recompose")  # DecompositionError::DigitCountMismatch
        t = dev.to_device(np.ascontiguousarray(digits, dtype=np.int64).view(np.uint64))
        return int(dev.to_host(self.recompose_batch(t), np.uint64)[0])

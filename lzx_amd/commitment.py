"""Host-side mirror of lattice-commitment's Ajtai surface over the HIP kernels.

  AjtaiParams        crates/lattice-commitment/src/ajtai.rs:32-41
  AjtaiPublicKey     crates/lattice-commitment/src/ajtai.rs:54-258
  AjtaiCommitment    crates/lattice-commitment/src/ajtai.rs:262-295
Errors surface as LzxError with the C-ABI code of the AjtaiError variant.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import device as dev
from ._lib import check, lib
from .ring import Modulus32, RingConfig, RingElement

COMMIT_AUTO, COMMIT_NTT, COMMIT_INT8 = 0, 1, 2


@dataclass
class AjtaiParams:
    ring: RingConfig
    k: int
    m: int
    norm_bound: int


class AjtaiCommitment:
    """k ring elements (ajtai.rs:262-295)."""

    def __init__(self, rows: np.ndarray, ring: RingConfig):
        self.rows_array = np.ascontiguousarray(rows, dtype=np.uint32)
        self.ring = ring

    @property
    def rows(self):
        return [RingElement(self.ring, r) for r in self.rows_array]

    def to_bytes(self) -> bytes:
        return self.rows_array.astype("<u4").tobytes()

    def __eq__(self, other):
        return isinstance(other, AjtaiCommitment) and np.array_equal(self.rows_array, other.rows_array)


class AjtaiPublicKey:
    """AjtaiPublicKey with A and the cached NTT-domain matrix resident in HBM."""

    def __init__(self, params: AjtaiParams, handle, seed=None):
        self.params = params
        self._h = handle
        self._destroy = lib().lzx_ajtai_key_destroy  # bound now: lib() may be gone at interpreter exit
        self.seed = seed

    @classmethod
    def from_seed(cls, params: AjtaiParams, seed: bytes) -> "AjtaiPublicKey":
        """ajtai.rs:87-103."""
        if len(seed) != 32:
            raise ValueError("seed must be 32 bytes")
        m: Modulus32 = params.ring.modulus
        h = C.c_void_p()
        check(lib().lzx_ajtai_key_from_seed(m.q, m.two_adicity, m.generator, params.ring.log_n, params.k,
                                            params.m, params.norm_bound, seed, C.byref(h)),
              "AjtaiPublicKey::from_seed")
        return cls(params, h, seed)

    @classmethod
    def from_matrix(cls, params: AjtaiParams, A: np.ndarray) -> "AjtaiPublicKey":
        m: Modulus32 = params.ring.modulus
        A = np.ascontiguousarray(A, dtype=np.uint32)
        h = C.c_void_p()
        check(lib().lzx_ajtai_key_from_matrix(m.q, m.two_adicity, m.generator, params.ring.log_n, params.k,
                                              params.m, params.norm_bound, A.ctypes.data, C.byref(h)),
              "AjtaiPublicKey::from_matrix")
        return cls(params, h)

    def __del__(self):
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h and destroy:
            destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def _shape(self):
        return (self.params.k * self.params.m, self.params.ring.n())

    def matrix(self) -> np.ndarray:
        out = np.zeros(self._shape(), dtype=np.uint32)
        check(lib().lzx_ajtai_key_matrix(self._h, out.ctypes.data))
        return out

    def matrix_ntt(self) -> np.ndarray:
        out = np.zeros(self._shape(), dtype=np.uint32)
        check(lib().lzx_ajtai_key_matrix_ntt(self._h, out.ctypes.data))
        return out

    def entry(self, i: int, j: int):
        if i < self.params.k and j < self.params.m:
            return RingElement(self.params.ring, self.matrix()[i * self.params.m + j])
        return None

    # -- reference signatures (host) -----------------------------------------
    def commit(self, s) -> AjtaiCommitment:
        """commit(&[RingElement]) (ajtai.rs:118-153)."""
        arr = _witness_array(s, self.params.ring.n())
        out = np.zeros((self.params.k, self.params.ring.n()), dtype=np.uint32)
        check(lib().lzx_ajtai_commit_host(self._h, arr.ctypes.data, arr.shape[0], out.ctypes.data),
              "AjtaiPublicKey::commit")
        return AjtaiCommitment(out, self.params.ring)

    def commit_naive_reference(self, s) -> AjtaiCommitment:
        """commit_naive_reference(&[RingElement]) (ajtai.rs:158-180): one ring
        product per (i, j), summed in the coefficient domain."""
        arr = _witness_array(s, self.params.ring.n())
        out = np.zeros((self.params.k, self.params.ring.n()), dtype=np.uint32)
        check(lib().lzx_ajtai_commit_naive_host(self._h, arr.ctypes.data, arr.shape[0], out.ctypes.data),
              "AjtaiPublicKey::commit_naive_reference")
        return AjtaiCommitment(out, self.params.ring)

    def verify_opening(self, commitment: AjtaiCommitment, s) -> None:
        """verify_opening (ajtai.rs:210-230); raises LzxError on failure."""
        arr = _witness_array(s, self.params.ring.n())
        check(lib().lzx_ajtai_verify_opening_host(self._h, commitment.rows_array.ctypes.data, arr.ctypes.data,
                                                  arr.shape[0]), "AjtaiPublicKey::verify_opening")

    # -- batched device forms ---------------------------------------------------
    def commit_batch(self, s_dev, W: int, out_dev, s_len=None, stream=None) -> None:
        """s_dev: [W][m][n] residues (int32 view), out_dev: [W][k][n]."""
        check(lib().lzx_ajtai_commit_batch(self._h, dev.ptr(s_dev), self.params.m if s_len is None else s_len, W,
                                           dev.ptr(out_dev), dev.stream_handle(stream)), "lzx_ajtai_commit_batch")

    def commit_digits_batch(self, digits_dev, W: int, out_dev, path: int = COMMIT_AUTO, s_len=None,
                            stream=None) -> None:
        """digits_dev: [W][m][n] int8 small digits, out_dev: [W][k][n]."""
        check(lib().lzx_ajtai_commit_digits_batch(self._h, dev.ptr(digits_dev),
                                                  self.params.m if s_len is None else s_len, W, dev.ptr(out_dev),
                                                  path, dev.stream_handle(stream)),
              "lzx_ajtai_commit_digits_batch")

    def commit_path_choice(self, W: int) -> int:
        """The path COMMIT_AUTO takes for W witnesses: COMMIT_NTT or COMMIT_INT8
        once measured, COMMIT_AUTO before the first AUTO call of the shape."""
        out = C.c_int32()
        check(lib().lzx_ajtai_commit_path_choice(self._h, W, C.byref(out)), "lzx_ajtai_commit_path_choice")
        return int(out.value)

    def commit_naive_batch(self, s_dev, W: int, out_dev, s_len=None, stream=None) -> None:
        """commit_naive_reference over [W][m][n] residues in HBM."""
        check(lib().lzx_ajtai_commit_naive_batch(self._h, dev.ptr(s_dev), self.params.m if s_len is None else s_len,
                                                 W, dev.ptr(out_dev), dev.stream_handle(stream)),
              "lzx_ajtai_commit_naive_batch")


def sample_small_secret(ring: RingConfig, m: int, bound: int, seed: bytes, out=None, digits=None, stream=None):
    """sample_small_secret(&RingConfig, m, bound, seed) (ajtai.rs:307-348) on
    the device.  Writes residues [m][n] into `out` and/or int8 digits into
    `digits` (device tensors; allocated when both are None, residues then)."""
    n = ring.n()
    if out is None and digits is None:
        out = dev.empty((m, n), np.uint32)
    mod: Modulus32 = ring.modulus
    check(lib().lzx_ajtai_sample_small_secret(mod.q, ring.log_n, m, bound, seed, len(seed),
                                              dev.ptr(out) if out is not None else None,
                                              dev.ptr(digits) if digits is not None else None,
                                              dev.stream_handle(stream)), "sample_small_secret")
    return out if digits is None else (out, digits)


def _witness_array(s, n):
    if isinstance(s, np.ndarray):
        arr = np.ascontiguousarray(s, dtype=np.uint32).reshape(-1, n)
    else:
        arr = np.ascontiguousarray(np.stack([e.coeffs for e in s]) if len(s) else np.zeros((0, n)), dtype=np.uint32)
    return arr

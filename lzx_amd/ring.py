"""Host-side mirror of lattice-ring's NTT surface over the HIP kernels.

Names, argument meaning and errors follow the reference (paths relative to
the reference's crates/):
  Modulus32            lattice-ring/src/modulus.rs:22-162
  NttTables            lattice-ring/src/ntt.rs:16-342
  RingConfig           lattice-ring/src/ring.rs:11-44
  RingElement          lattice-ring/src/ring.rs:183-447
Errors surface as LzxError whose .code is the C-ABI status mapped from
NttError / RingError.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import device as dev
from ._lib import LzxError, check, lib

LZX_NTT_LENGTH_MISMATCH = 1
LZX_NTT_ADICITY_TOO_SMALL = 2
LZX_RING_LENGTH_MISMATCH = 3


@dataclass(frozen=True)
class Modulus32:
    """(q, two_adicity, generator) of an odd prime q < 2^32 (modulus.rs:22-44)."""

    q: int
    two_adicity: int
    generator: int


Modulus32.Q_32 = Modulus32(3221225473, 30, 5)                # modulus.rs:48
Modulus32.Q_12289 = Modulus32(12289, 12, 11)                 # modulus.rs:52
Modulus32.Q_2013265921 = Modulus32(2013265921, 27, 31)       # modulus.rs:55
CRT_PRIMES = [Modulus32(2013265921, 27, 31), Modulus32(1811939329, 26, 13),
              Modulus32(2113929217, 25, 5), Modulus32(998244353, 23, 3)]


class NttTables:
    """NttTables::new(modulus, log_n) (ntt.rs:46) backed by a device plan."""

    def __init__(self, modulus: Modulus32, log_n: int):
        self.modulus = modulus
        self.log_n = log_n
        self._h = C.c_void_p()
        self._destroy = lib().lzx_ntt_plan_destroy  # bound now: lib() may be gone at interpreter exit
        check(lib().lzx_ntt_plan_create(modulus.q, modulus.two_adicity, modulus.generator, log_n,
                                        C.byref(self._h)), "NttTables::new")

    def __del__(self):
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h and destroy:
            destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def n(self) -> int:
        return 1 << self.log_n

    def constants(self):
        psi, ninv = C.c_uint32(), C.c_uint32()
        check(lib().lzx_ntt_plan_constants(self._h, C.byref(psi), C.byref(ninv)))
        return psi.value, ninv.value

    def path(self, polys: int = 1, inverse: bool = False) -> str:
        """The kernel class of this plan's batched transform of `polys`
        polynomials (see batch_path)."""
        return batch_path([self], polys, inverse)

    # -- exact reference signatures: in place on a host &mut [u32] ---------
    def _host(self, fn, a: np.ndarray, what: str):
        if a.dtype != np.uint32 or not a.flags.c_contiguous:
            raise TypeError("expected a C-contiguous uint32 array")
        check(fn(self._h, a.ctypes.data, a.size), what)

    def forward(self, a: np.ndarray) -> None:
        self._host(lib().lzx_ntt_forward_host, a, "NttTables::forward")

    forward_fast = forward

    def inverse(self, a: np.ndarray) -> None:
        self._host(lib().lzx_ntt_inverse_host, a, "NttTables::inverse")

    inverse_fast = inverse

    def forward_partial(self, a: np.ndarray, levels: int) -> None:
        self._via_device(a, lambda t: self.forward_partial_batch(t, levels))

    def complete_partial(self, a: np.ndarray, already: int) -> None:
        self._via_device(a, lambda t: self.complete_partial_batch(t, already))

    def _via_device(self, a: np.ndarray, fn) -> None:
        if a.size % self.n():
            raise LzxError(LZX_NTT_LENGTH_MISMATCH, "NttTables")
        t = dev.to_device(a)
        fn(t)
        a[...] = dev.to_host(t, np.uint32).reshape(a.shape)

    # -- batched device forms (torch tensors holding uint32 as int32) ------
    def forward_batch(self, data, stream=None) -> None:
        check(lib().lzx_ntt_forward_batch(self._h, dev.ptr(data), data.numel(), dev.stream_handle(stream)),
              "lzx_ntt_forward_batch")

    def inverse_batch(self, data, stream=None) -> None:
        check(lib().lzx_ntt_inverse_batch(self._h, dev.ptr(data), data.numel(), dev.stream_handle(stream)),
              "lzx_ntt_inverse_batch")

    def forward_partial_batch(self, data, levels: int, stream=None) -> None:
        check(lib().lzx_ntt_forward_partial_batch(self._h, dev.ptr(data), data.numel(), levels,
                                                  dev.stream_handle(stream)), "forward_partial")

    def complete_partial_batch(self, data, already: int, stream=None) -> None:
        check(lib().lzx_ntt_complete_partial_batch(self._h, dev.ptr(data), data.numel(), already,
                                                   dev.stream_handle(stream)), "complete_partial")

    def ring_mul_batch(self, a, b, out, stream=None) -> None:
        check(lib().lzx_ring_mul_batch(self._h, dev.ptr(a), dev.ptr(b), dev.ptr(out), a.numel(),
                                       dev.stream_handle(stream)), "lzx_ring_mul_batch")


def _plan_array(tables):
    arr = (C.c_void_p * len(tables))(*[t.handle.value for t in tables])
    return arr


PATHS = {0: "register", 1: "mfma", 2: "lds", 3: "auto"}


def batch_path(tables, polys_per_plan: int, inverse: bool = False) -> str:
    """The kernel class a batched transform of these plans runs: "mfma" (the
    int8 matrix-core transform), "register" (butterflies in VGPRs), "lds"
    (n > 64), or "auto" (MFMA-capable, not yet measured for this shape: the
    first call times both classes and keeps the faster)."""
    v = C.c_uint32()
    check(lib().lzx_ntt_batch_path(_plan_array(tables), len(tables), polys_per_plan, int(inverse), C.byref(v)))
    return PATHS[v.value]


def forward_multi(tables, data, polys_per_plan: int, stream=None) -> None:
    """CRT batch: data laid out [plan][poly][n], one launch per arithmetic class."""
    check(lib().lzx_ntt_forward_multi(_plan_array(tables), len(tables), dev.ptr(data), polys_per_plan,
                                      dev.stream_handle(stream)), "lzx_ntt_forward_multi")


def inverse_multi(tables, data, polys_per_plan: int, stream=None) -> None:
    check(lib().lzx_ntt_inverse_multi(_plan_array(tables), len(tables), dev.ptr(data), polys_per_plan,
                                      dev.stream_handle(stream)), "lzx_ntt_inverse_multi")


class RingConfig:
    """RingConfig::new(modulus, log_n) (ring.rs:33-40)."""

    def __init__(self, modulus: Modulus32, log_n: int):
        self.modulus = modulus
        self.log_n = log_n
        self.tables = NttTables(modulus, log_n)

    def n(self) -> int:
        return 1 << self.log_n

    def zero(self):
        return RingElement(self, np.zeros(self.n(), dtype=np.uint32))


class RingElement:
    """RingElement (ring.rs:183-447): natural-order canonical coefficients."""

    def __init__(self, config: RingConfig, coeffs):
        self.config = config
        self.coeffs = np.ascontiguousarray(coeffs, dtype=np.uint32)

    @classmethod
    def from_signed(cls, config: RingConfig, signed):
        """ring.rs:232-243: reduce balanced integers, pad to n."""
        q = config.modulus.q
        c = np.zeros(config.n(), dtype=np.uint32)
        s = np.asarray(signed, dtype=np.int64)
        c[: s.size] = np.mod(s, q).astype(np.uint32)
        return cls(config, c)

    def is_zero(self) -> bool:
        return not self.coeffs.any()

    def to_ntt(self) -> np.ndarray:
        a = self.coeffs.copy()
        self.config.tables.forward(a)
        return a

    @classmethod
    def from_ntt(cls, config: RingConfig, evals) -> "RingElement":
        e = np.ascontiguousarray(evals, dtype=np.uint32).copy()
        if e.size != config.n():
            raise LzxError(LZX_RING_LENGTH_MISMATCH, "RingElement::from_ntt")
        config.tables.inverse(e)
        return cls(config, e)

    def mul(self, other: "RingElement") -> "RingElement":
        if other.config.modulus != self.config.modulus or other.config.log_n != self.config.log_n:
            raise LzxError(LZX_RING_LENGTH_MISMATCH, "RingElement::mul")
        a = dev.to_device(self.coeffs)
        b = dev.to_device(other.coeffs)
        out = dev.empty(self.config.n(), np.uint32)
        self.config.tables.ring_mul_batch(a, b, out)
        return RingElement(self.config, dev.to_host(out, np.uint32))

    def to_bytes(self) -> bytes:
        return self.coeffs.astype("<u4").tobytes()

    def __eq__(self, other) -> bool:
        return (isinstance(other, RingElement) and self.config.modulus == other.config.modulus
                and self.config.log_n == other.config.log_n and np.array_equal(self.coeffs, other.coeffs))

"""Short ring-element challenges (crates/lattice-core/src/short_challenge.rs:47-690).

Sampling runs in the library (host code: one SHAKE-256 stream per sample);
the spec bookkeeping (validation, Γ bounds, entropy) is restated here.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import LzxError, check, lib

LZX_SC_FIXED_WEIGHT, LZX_SC_BIASED_TERNARY, LZX_SC_SMALL_SET, LZX_SC_FIXED_WEIGHT_SMALL_SET = 0, 1, 2, 3
LZX_SHORT_CHALLENGE_INVALID_PARAMETERS = 61


class SpecC(C.Structure):
    """lzx_short_challenge_spec."""

    _fields_ = [("n", C.c_size_t), ("family", C.c_int32), ("weight", C.c_size_t), ("amplitude", C.c_int64),
                ("p_nonzero_permille", C.c_uint32), ("values", C.c_void_p), ("nvalues", C.c_size_t)]


@dataclass(frozen=True)
class FixedWeight:
    weight: int
    amplitude: int


@dataclass(frozen=True)
class BiasedTernary:
    p_nonzero_permille: int


@dataclass(frozen=True)
class SmallSet:
    values: tuple


@dataclass(frozen=True)
class FixedWeightSmallSet:
    weight: int
    values: tuple


def ceil_sqrt(x: int) -> int:
    """short_challenge.rs:281-340: min(ceil(sqrt(x)), 2^64 - 1) for x < 2^128."""
    return int(lib().lzx_ceil_sqrt_u128(x & ((1 << 64) - 1), x >> 64))


def _ln_factorial(x: float) -> float:
    return 0.0 if x <= 1.0 else x * math.log(x) - x + 0.5 * math.log(2.0 * math.pi * x)


def _ln_binom(n: float, k: float) -> float:
    return _ln_factorial(n) - _ln_factorial(k) - _ln_factorial(max(n - k, 0.0))


@dataclass
class ShortChallenge:
    """short_challenge.rs:73-78, 238-268."""

    coefficients: np.ndarray
    family: object
    gamma_c: int

    def l_inf(self) -> int:
        return int(np.abs(self.coefficients).max()) if self.coefficients.size else 0

    def l2_squared(self) -> int:
        return sum(int(c) * int(c) for c in self.coefficients)

    def hamming_weight(self) -> int:
        return int(np.count_nonzero(self.coefficients))


@dataclass
class ShortChallengeSpec:
    """ShortChallengeSpec (short_challenge.rs:65-236)."""

    n: int
    family: object
    _keep: list = field(default_factory=list, repr=False, compare=False)

    def _c(self) -> SpecC:
        f = self.family
        vals = getattr(f, "values", ())
        arr = np.ascontiguousarray(list(vals) or [0], dtype=np.int64)
        self._keep = [arr]
        fam = {FixedWeight: LZX_SC_FIXED_WEIGHT, BiasedTernary: LZX_SC_BIASED_TERNARY, SmallSet: LZX_SC_SMALL_SET,
               FixedWeightSmallSet: LZX_SC_FIXED_WEIGHT_SMALL_SET}[type(f)]
        return SpecC(self.n, fam, getattr(f, "weight", 0), getattr(f, "amplitude", 0),
                     getattr(f, "p_nonzero_permille", 0), arr.ctypes.data if vals else None, len(vals))

    def validate(self) -> None:
        """short_challenge.rs:97-124 (raises LzxError InvalidParameters)."""
        f, n = self.family, self.n
        bad = n == 0 or n > 1 << 20
        if isinstance(f, FixedWeight):
            bad |= f.weight == 0 or f.weight > n or f.amplitude <= 0
        elif isinstance(f, BiasedTernary):
            bad |= f.p_nonzero_permille == 0 or f.p_nonzero_permille > 1000
        elif isinstance(f, SmallSet):
            bad |= len(f.values) == 0 or len(f.values) > 1024
        else:
            bad |= f.weight == 0 or f.weight > n or len(f.values) == 0 or len(f.values) > 1024
        if bad:
            raise LzxError(LZX_SHORT_CHALLENGE_INVALID_PARAMETERS, "ShortChallengeSpec::validate")

    def amplitude_max(self) -> int:
        f = self.family
        if isinstance(f, FixedWeight):
            return f.amplitude
        if isinstance(f, BiasedTernary):
            return 1
        return max((abs(v) for v in f.values), default=0)

    def gamma_c_worst_case(self) -> int:
        """short_challenge.rs:142-151: B * ceil(sqrt(k)), k the most non-zero coefficients."""
        f = self.family
        k = min(f.weight, self.n) if isinstance(f, (FixedWeight, FixedWeightSmallSet)) else self.n
        return max(self.amplitude_max(), 1) * ceil_sqrt(k)

    def entropy_bits(self) -> float:
        """short_challenge.rs:153-171 (Stirling ln n!, as the reference computes it)."""
        f, n = self.family, float(self.n)
        if isinstance(f, FixedWeight):
            return _ln_binom(n, float(f.weight)) / math.log(2) + f.weight
        if isinstance(f, BiasedTernary):
            p = f.p_nonzero_permille / 1000.0
            q = 1.0 - p
            h = 0.0
            if 0.0 < p / 2.0 < 1.0:
                h -= (p / 2.0) * math.log2(p / 2.0) * 2.0
            if 0.0 < q < 1.0:
                h -= q * math.log2(q)
            return n * h
        if isinstance(f, SmallSet):
            return n * math.log2(len(f.values))
        return _ln_binom(n, float(f.weight)) / math.log(2) + f.weight * math.log2(len(f.values))

    def require_entropy(self, min_bits: float) -> None:
        if self.entropy_bits() < min_bits:
            raise LzxError(LZX_SHORT_CHALLENGE_INVALID_PARAMETERS, "ShortChallengeSpec::require_entropy")

    def sample(self, seed: bytes) -> ShortChallenge:
        """short_challenge.rs:183-205."""
        spec = self._c()
        out = np.zeros(max(1, self.n), dtype=np.int64)
        g = C.c_uint64()
        check(lib().lzx_short_challenge_sample(C.byref(spec), seed, len(seed), out.ctypes.data, C.byref(g)),
              "ShortChallengeSpec::sample")
        return ShortChallenge(out[: self.n], self.family, g.value)

    def sample_with_gamma_cap(self, seed: bytes, cap: int, retries: int) -> ShortChallenge:
        """short_challenge.rs:208-236."""
        spec = self._c()
        out = np.zeros(max(1, self.n), dtype=np.int64)
        g = C.c_uint64()
        check(lib().lzx_short_challenge_sample_with_gamma_cap(C.byref(spec), seed, len(seed), cap, retries,
                                                              out.ctypes.data, C.byref(g)),
              "ShortChallengeSpec::sample_with_gamma_cap")
        return ShortChallenge(out[: self.n], self.family, g.value)


# paper profiles (short_challenge.rs:553-660)
def pikkufold_spec() -> ShortChallengeSpec:
    return ShortChallengeSpec(256, FixedWeight(23, 1))


def cyclo_spec() -> ShortChallengeSpec:
    return ShortChallengeSpec(512, BiasedTernary(500))


def latticefold_plus_spec() -> ShortChallengeSpec:
    return ShortChallengeSpec(64, SmallSet((-2, -1, 0, 1, 2)))


def protogalattice_spec(n: int) -> ShortChallengeSpec:
    return ShortChallengeSpec(n, FixedWeightSmallSet(min(128, n), (-1, 1, 2)))


def symphony_spec(n: int) -> ShortChallengeSpec:
    return ShortChallengeSpec(n, SmallSet((0, 1, -1, 2, -2)))


def rokoko_spec(n: int) -> ShortChallengeSpec:
    return ShortChallengeSpec(n, FixedWeight(min(22, n), 1))


def hyperwolf_spec(n: int) -> ShortChallengeSpec:
    return ShortChallengeSpec(n, FixedWeight(min(10, n), 1))

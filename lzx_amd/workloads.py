"""Seeded synthetic inputs for the hot-path workloads (numpy, vectorised).

All streams are splitmix64 (state += 0x9E3779B97F4A7C15, then the standard
two-multiply finaliser).  Each workload documents its exact rule so that any
implementation can regenerate identical inputs.
"""
from __future__ import annotations

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
GL_P = 0xFFFFFFFF00000001


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Outputs start .. start+count-1 of the splitmix64 stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        idx = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * GAMMA
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def uniform_u32_mod(seed: int, count: int, q: int) -> np.ndarray:
    """Low 32 bits of each splitmix64 output, keeping those < q, first `count` kept."""
    out = np.empty(0, dtype=np.uint32)
    start = 0
    while out.size < count:
        need = count - out.size
        batch = int(need * (2**32 / q) * 1.02) + 64
        cand = (splitmix64(seed, batch, start) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        start += batch
        out = np.concatenate([out, cand[cand < q]])
    return out[:count]


def uniform_gl(seed: int, count: int) -> np.ndarray:
    """splitmix64 outputs < p (Goldilocks), first `count` kept."""
    out = np.empty(0, dtype=np.uint64)
    start = 0
    while out.size < count:
        need = count - out.size
        cand = splitmix64(seed, need + 16, start)
        start += need + 16
        out = np.concatenate([out, cand[cand < np.uint64(GL_P)]])
    return out[:count]


def small_digits(seed: int, count: int, zero_half: bool = True) -> np.ndarray:
    """Witness digits in [-8, 8): x = splitmix64; if zero_half and x is even
    the digit is 0, else ((x >> 60) - 8).  P(0) = 17/32 with zero_half."""
    x = splitmix64(seed, count)
    d = ((x >> np.uint64(60)).astype(np.int16) - 8).astype(np.int8)
    if zero_half:
        d[(x & np.uint64(1)) == 0] = 0
    return d


def digits_to_residues(d: np.ndarray, q: int) -> np.ndarray:
    """Balanced small integers -> canonical residues (ring.rs:232-243 from_signed)."""
    return np.mod(d.astype(np.int64), q).astype(np.uint32)


# ---- workload definitions ---------------------------------------------------

def ntt_batch_input(moduli, polys: int, n: int = 64, seed_base: int = 0xC2) -> np.ndarray:
    """Batched NTT input, [prime][poly][n] u32, prime i from seed seed_base + i."""
    return np.stack([uniform_u32_mod(seed_base + i, polys * n, m.q).reshape(polys, n)
                     for i, m in enumerate(moduli)])

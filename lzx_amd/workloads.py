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


def uneven_batch_jobs(seed: int = 0xC5, njobs: int = 512, moduli=None):
    """The uneven commitment batch: job i takes outputs 3i, 3i+1, 3i+2 of the
    splitmix64 stream `seed`: u = (x0 >> 11) / 2^53, m = floor(2^(8 + 8u));
    log_n = 6 + x1 mod 3 (n in {64, 128, 256}); modulus = moduli[x2 mod 3].
    Returns a list of (m, log_n, (q, generator))."""
    if moduli is None:
        from .commit_batch import WIDE_MODULI as moduli  # noqa: WPS433
    x = splitmix64(seed, 3 * njobs).reshape(njobs, 3)
    u = (x[:, 0] >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    m = np.floor(np.exp2(8.0 + 8.0 * u)).astype(np.int64)
    log_n = 6 + (x[:, 1] % np.uint64(3)).astype(np.int64)
    mi = (x[:, 2] % np.uint64(3)).astype(np.int64)
    return [(int(m[i]), int(log_n[i]), moduli[int(mi[i])]) for i in range(njobs)]


def small_digits_torch(seed: int, count: int, device="cuda", start: int = 0):
    """small_digits(seed, count) computed on the device (same splitmix64 rule;
    int64 arithmetic wraps, logical shifts emulated with masks)."""
    import torch  # noqa: WPS433
    g = torch.tensor(np.array([int(GAMMA)], dtype=np.uint64).view(np.int64)[0], device=device)
    m1 = torch.tensor(np.array([int(M1)], dtype=np.uint64).view(np.int64)[0], device=device)
    m2 = torch.tensor(np.array([int(M2)], dtype=np.uint64).view(np.int64)[0], device=device)
    s0 = torch.tensor(np.array([seed & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64).view(np.int64)[0], device=device)

    def lsr(v, k):
        return (v >> k) & ((1 << (64 - k)) - 1)

    idx = torch.arange(start + 1, start + count + 1, dtype=torch.int64, device=device)
    z = s0 + idx * g
    z = (z ^ lsr(z, 30)) * m1
    z = (z ^ lsr(z, 27)) * m2
    z = z ^ lsr(z, 31)
    d = (lsr(z, 60) - 8).to(torch.int8)
    return torch.where((z & 1) == 0, torch.zeros_like(d), d)

"""Uneven commitment batch: many commitments t = A s of mixed ring degree,
modulus and length, computed by one persistent launch (liblzx_gpu.so).

Each job is AjtaiPublicKey::commit (crates/lattice-commitment/src/ajtai.rs:
118-153) on its own ring Z_q[X]/(X^n + 1), n in {64, 128, 256}; moduli may
exceed the reference's Modulus32 up to 2^50 (u64 residues).  Keys hold A in
HBM ([k][m][n]) and a job commits with the key's first m columns.  The batch
plans a device work queue once and replays it as a hipGraph; `autotune()`
picks the chunk size (ring elements per work item) per modulus class and ring
degree and caches it (file named by LZX_AUTOTUNE_CACHE).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import device as D
from ._lib import check, lib

# (q, generator): the Modulus32 default Q_32, the reference's 2013265921, and a
# 50-bit NTT prime q = (2^24 - 7) * 2^26 + 1 with generator 5.
Q_32 = (3221225473, 5)
Q_2013265921 = (2013265921, 31)
Q_50 = (1125899437080577, 5)
WIDE_MODULI = [Q_32, Q_2013265921, Q_50]

CLASS_KINDS = ("q<2^31", "q<2^32", "q<2^50")


def class_of(q: int, log_n: int) -> int:
    kind = 0 if q < (1 << 31) else 1 if q < (1 << 32) else 2
    return 3 * kind + log_n - 6


class _Job(C.Structure):
    _fields_ = [("key", C.c_void_p), ("digits_dev", C.c_void_p), ("m", C.c_size_t), ("out_dev", C.c_void_p)]


class WideKey:
    """Commitment key over Z_q[X]/(X^n + 1) with A resident in HBM."""

    def __init__(self, q: int, generator: int, log_n: int, k: int, m: int, A_dev, stream=None):
        self.q, self.generator, self.log_n, self.k, self.m = q, generator, log_n, k, m
        self.n = 1 << log_n
        h = C.c_void_p()
        check(lib().lzx_wide_key_create(q, generator, log_n, k, m, D.ptr(A_dev) if m else None,
                                        D.stream_handle(stream), C.byref(h)), "lzx_wide_key_create")
        self.h = h
        self.elem_bytes = lib().lzx_wide_key_elem_bytes(h)
        self.dtype = np.uint32 if self.elem_bytes == 4 else np.uint64

    @classmethod
    def from_host(cls, q: int, generator: int, A: np.ndarray, stream=None):
        """A: (k, m, n) residues; uploaded as u32 when q < 2^32, else u64."""
        k, m, n = A.shape
        dt = np.uint32 if q < (1 << 32) else np.uint64
        A_dev = D.to_device(np.ascontiguousarray(A, dtype=dt))
        key = cls(q, generator, int(n).bit_length() - 1, k, m, A_dev, stream)
        D.synchronize()
        return key

    def close(self):
        if getattr(self, "h", None):
            lib().lzx_wide_key_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CommitBatch:
    """jobs: sequence of (key, digits_dev [m][n] int8, m, out_dev [k][n])."""

    def __init__(self, jobs):
        self._keep = list(jobs)
        arr = (_Job * max(1, len(self._keep)))()
        for i, (key, digits, m, out) in enumerate(self._keep):
            arr[i].key = key.h.value
            arr[i].digits_dev = digits.data_ptr() if digits is not None else None
            arr[i].m = m
            arr[i].out_dev = out.data_ptr()
        h = C.c_void_p()
        check(lib().lzx_commit_batch_create(C.cast(arr, C.c_void_p), len(self._keep), C.byref(h)),
              "lzx_commit_batch_create")
        self.h = h

    def run(self, stream=None):
        check(lib().lzx_commit_batch_run(self.h, D.stream_handle(stream)), "lzx_commit_batch_run")

    def check(self, stream=None):
        check(lib().lzx_commit_batch_check(self.h, D.stream_handle(stream)), "lzx_commit_batch_check")

    def autotune(self, stream=None):
        check(lib().lzx_commit_batch_autotune(self.h, D.stream_handle(stream)), "lzx_commit_batch_autotune")

    def info(self):
        items, scratch = C.c_size_t(), C.c_size_t()
        chunks = (C.c_uint32 * 9)()
        check(lib().lzx_commit_batch_info(self.h, C.byref(items), C.byref(scratch), chunks), "lzx_commit_batch_info")
        return {"items": items.value, "scratch_bytes": scratch.value, "chunks": list(chunks)}

    def close(self):
        if getattr(self, "h", None):
            lib().lzx_commit_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def uneven_batch_workload(seed: int = 0xC5, njobs: int = 512, k: int = 16, stream=None):
    """The 512-job uneven batch on the device (synthetic, seeded).

    Jobs follow workloads.uneven_batch_jobs(seed).  One key per (modulus, n)
    class with k = 16 rows and as many columns as the class's longest job;
    A uniform in [0, q) from torch's generator seeded with seed * 16 + class.
    Job i's witness is workloads.small_digits(seed * 65536 + i, m * n) (digits
    in [-8, 8), P(0) = 17/32), computed on the device; W = 1 per job.
    Returns a dict with the batch, keys, per-job tensors and byte counts."""
    import torch  # noqa: WPS433

    from . import workloads as W

    spec = W.uneven_batch_jobs(seed, njobs)
    classes = {}
    for m, log_n, (q, g) in spec:
        c = class_of(q, log_n)
        classes.setdefault(c, [q, g, log_n, 0])
        classes[c][3] = max(classes[c][3], m)
    keys, A_dev = {}, {}
    for c, (q, g, log_n, mmax) in sorted(classes.items()):
        n = 1 << log_n
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed * 16 + c)
        A = torch.randint(0, q, (k, mmax, n), dtype=torch.int64, device="cuda", generator=gen)
        if q < (1 << 32):
            A = A.to(torch.int32)  # bit pattern of the u32 residue
        keys[c] = WideKey(q, g, log_n, k, mmax, A, stream)
        A_dev[c] = A
    jobs, digits, outs = [], [], []
    key_bytes = digit_bytes = out_bytes = 0
    for i, (m, log_n, (q, g)) in enumerate(spec):
        c = class_of(q, log_n)
        key, n = keys[c], 1 << log_n
        d = W.small_digits_torch(seed * 65536 + i, m * n).reshape(m, n)
        out = D.empty(k * n, key.dtype)
        jobs.append((key, d, m, out))
        digits.append(d)
        outs.append(out)
        key_bytes += k * m * n * key.elem_bytes
        digit_bytes += m * n
        out_bytes += k * n * key.elem_bytes
    D.synchronize()
    return {"spec": spec, "keys": keys, "A": A_dev, "jobs": jobs, "batch": CommitBatch(jobs),
            "ring_macs": sum(k * m for m, _, _ in spec), "key_bytes": key_bytes, "digit_bytes": digit_bytes,
            "out_bytes": out_bytes}

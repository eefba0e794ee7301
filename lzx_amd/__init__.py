"""lzx_amd: MI355X (gfx950) lattice-commitment hot path of LZX.

A thin Python mirror of the reference's Rust interface over the C ABI of
liblzx_gpu.so (include/lzx_gpu.h): negacyclic NTT (lzx_amd.ring), Ajtai
commitment (lzx_amd.commitment) and the Goldilocks sum-check
(lzx_amd.sumcheck).  All compute runs in hand-written HIP kernels; there is
no CPU fallback.
"""
from ._lib import LzxError, lib  # noqa: F401

__all__ = ["LzxError", "lib"]

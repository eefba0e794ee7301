"""ctypes binding of liblzx_gpu.so (the C ABI declared in include/lzx_gpu.h).

There is no fallback: if the library is missing, loading raises, so nothing in
this package can silently run a CPU path in place of the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liblzx_gpu.so"

_lib = None

vp = C.c_void_p
sz = C.c_size_t
u8p = C.c_char_p
u32, u64, i32 = C.c_uint32, C.c_uint64, C.c_int32
st_t = C.c_int32

# name -> (restype, argtypes); mirrors include/lzx_gpu.h one-to-one.
PROTOTYPES = {
    "lzx_status_string": (C.c_char_p, [st_t]),
    "lzx_last_device_error": (C.c_char_p, []),
    # NTT
    "lzx_ntt_plan_create": (st_t, [u32, u32, u32, u32, C.POINTER(vp)]),
    "lzx_ntt_plan_destroy": (None, [vp]),
    "lzx_ntt_plan_constants": (st_t, [vp, C.POINTER(u32), C.POINTER(u32)]),
    "lzx_ntt_batch_path": (st_t, [C.POINTER(vp), sz, sz, u32, C.POINTER(u32)]),
    "lzx_ntt_forward_batch": (st_t, [vp, vp, sz, vp]),
    "lzx_ntt_inverse_batch": (st_t, [vp, vp, sz, vp]),
    "lzx_ntt_forward_partial_batch": (st_t, [vp, vp, sz, u32, vp]),
    "lzx_ntt_complete_partial_batch": (st_t, [vp, vp, sz, u32, vp]),
    "lzx_ntt_forward_multi": (st_t, [C.POINTER(vp), sz, vp, sz, vp]),
    "lzx_ntt_inverse_multi": (st_t, [C.POINTER(vp), sz, vp, sz, vp]),
    "lzx_ntt_forward_host": (st_t, [vp, vp, sz]),
    "lzx_ntt_inverse_host": (st_t, [vp, vp, sz]),
    "lzx_ring_mul_batch": (st_t, [vp, vp, vp, vp, sz, vp]),
    # Quadratic extension ring R_q[Y]/(Y^2+1)
    "lzx_rq2_mul_batch": (st_t, [vp, vp, vp, vp, sz, vp]),
    "lzx_rq2_norm_batch": (st_t, [vp, vp, vp, sz, vp]),
    "lzx_rq2_inverse_batch": (st_t, [vp, vp, vp, vp, sz, vp]),
    "lzx_short_challenge_sample": (st_t, [vp, vp, sz, vp, C.POINTER(u64)]),
    "lzx_short_challenge_sample_with_gamma_cap": (st_t, [vp, vp, sz, u64, sz, vp, C.POINTER(u64)]),
    "lzx_rq2_sample_challenge": (st_t, [u32, u32, vp, vp, sz, vp, C.POINTER(u64), C.POINTER(u64)]),
    "lzx_ceil_sqrt_u128": (u64, [u64, u64]),
    "lzx_rq2_add_batch": (st_t, [vp, vp, vp, vp, sz, vp]),
    "lzx_rq2_sub_batch": (st_t, [vp, vp, vp, vp, sz, vp]),
    "lzx_rq2_neg_batch": (st_t, [vp, vp, vp, sz, vp]),
    "lzx_rq2_conjugate_batch": (st_t, [vp, vp, vp, sz, vp]),
    # Quadratic-slot ring over ~2^50 moduli
    "lzx_quad_plan_create": (st_t, [u64, u32, u64, u32, C.POINTER(vp)]),
    "lzx_quad_plan_destroy": (None, [vp]),
    "lzx_quad_slot_constants": (st_t, [vp, vp, C.POINTER(u64)]),
    "lzx_quad_forward_batch": (st_t, [vp, vp, sz, vp]),
    "lzx_quad_inverse_batch": (st_t, [vp, vp, sz, vp]),
    "lzx_quad_slotwise_mul_batch": (st_t, [vp, vp, vp, vp, sz, vp]),
    "lzx_quad_mul_batch": (st_t, [vp, vp, vp, vp, sz, vp]),
    # Ajtai
    "lzx_ajtai_key_from_seed": (st_t, [u32, u32, u32, u32, sz, sz, u32, u8p, C.POINTER(vp)]),
    "lzx_ajtai_key_from_matrix": (st_t, [u32, u32, u32, u32, sz, sz, u32, vp, C.POINTER(vp)]),
    "lzx_ajtai_key_destroy": (None, [vp]),
    "lzx_ajtai_key_matrix": (st_t, [vp, vp]),
    "lzx_ajtai_key_matrix_ntt": (st_t, [vp, vp]),
    "lzx_ajtai_commit_batch": (st_t, [vp, vp, sz, sz, vp, vp]),
    "lzx_ajtai_commit_digits_batch": (st_t, [vp, vp, sz, sz, vp, i32, vp]),
    "lzx_ajtai_commit_path_choice": (st_t, [vp, sz, C.POINTER(i32)]),
    "lzx_ajtai_commit_digits_int8_noalloc": (st_t, [vp, vp, sz, sz, vp, vp]),
    "lzx_ajtai_commit_host": (st_t, [vp, vp, sz, vp]),
    "lzx_ajtai_commit_naive_batch": (st_t, [vp, vp, sz, sz, vp, vp]),
    "lzx_ajtai_commit_naive_host": (st_t, [vp, vp, sz, vp]),
    "lzx_ajtai_sample_small_secret": (st_t, [u32, u32, sz, u32, u8p, sz, vp, vp, vp]),
    "lzx_ajtai_verify_opening_host": (st_t, [vp, vp, vp, sz]),
    # Witness packing
    "lzx_pack_field_elements_len": (sz, [sz, u32]),
    "lzx_pack_field_elements": (st_t, [vp, sz, u32, u32, vp, vp]),
    "lzx_unpack_field_elements": (st_t, [vp, sz, u32, u32, vp, vp]),
    "lzx_crt_pack": (st_t, [vp, sz, vp, sz, u32, u32, vp, vp]),
    "lzx_crt_unpack": (st_t, [vp, sz, u32, vp, sz, sz, vp, vp]),
    "lzx_pack_bits_len": (sz, [sz, u32]),
    "lzx_pack_bits": (st_t, [vp, sz, u32, vp, vp]),
    "lzx_unpack_bits": (st_t, [vp, sz, u32, vp, sz, vp]),
    "lzx_ajtai_commit_onehot": (st_t, [vp, vp, vp, sz, vp, vp, vp]),
    "lzx_gadget_decompose": (st_t, [vp, sz, u64, u32, vp, u32, vp]),
    "lzx_gadget_recompose": (st_t, [vp, u32, sz, u64, u32, vp, vp]),
    # Uneven commitment batch
    "lzx_wide_key_create": (st_t, [u64, u64, u32, sz, sz, vp, vp, C.POINTER(vp)]),
    "lzx_wide_key_destroy": (None, [vp]),
    "lzx_wide_key_elem_bytes": (u32, [vp]),
    "lzx_commit_batch_create": (st_t, [vp, sz, C.POINTER(vp)]),
    "lzx_commit_batch_run": (st_t, [vp, vp]),
    "lzx_commit_batch_check": (st_t, [vp, vp]),
    "lzx_commit_batch_autotune": (st_t, [vp, vp]),
    "lzx_commit_batch_info": (st_t, [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(u32)]),
    "lzx_commit_batch_destroy": (None, [vp]),
    # Transcript
    "lzx_transcript_new": (st_t, [u8p, sz, u64, C.POINTER(vp)]),
    "lzx_transcript_clone": (st_t, [vp, C.POINTER(vp)]),
    "lzx_transcript_destroy": (None, [vp]),
    "lzx_transcript_append_message": (st_t, [vp, u8p, sz, u8p, sz]),
    "lzx_transcript_append_field_slice": (st_t, [vp, u8p, sz, vp, sz]),
    "lzx_transcript_challenge_fields": (st_t, [vp, u8p, sz, sz, vp]),
    "lzx_transcript_query_count": (u64, [vp]),
    # Sum-check
    "lzx_sumcheck_prove": (st_t, [sz, sz, C.POINTER(vp), sz, vp, vp, vp, u64, vp, vp, vp, C.POINTER(u64), vp,
                                  C.POINTER(sz), vp]),
    "lzx_sumcheck_prove_with_challenges": (st_t, [sz, sz, C.POINTER(vp), sz, vp, vp, vp, u64, vp, vp,
                                                  C.POINTER(u64), vp, C.POINTER(sz), vp]),
    "lzx_sumcheck_verify": (st_t, [vp, sz, sz, sz, sz, u64, vp, i32, u64, vp, C.POINTER(u64)]),
    "lzx_eq_table": (st_t, [vp, sz, vp, vp]),
    "lzx_sumcheck_prove_eq": (st_t, [sz, sz, C.POINTER(vp), sz, vp, vp, vp, u64, vp, vp, sz, vp, vp, vp,
                                     C.POINTER(u64), vp, C.POINTER(sz), vp]),
    "lzx_zerocheck_prove": (st_t, [sz, vp, vp, vp, vp, vp, C.POINTER(u64), vp, vp]),
    "lzx_zerocheck_verify": (st_t, [vp, sz, sz, sz, vp, i32, u64, vp]),
    "lzx_sumcheck_session_create": (st_t, [sz, sz, C.POINTER(vp), sz, vp, vp, vp, vp, C.POINTER(vp)]),
    "lzx_sumcheck_session_round_len": (sz, [vp]),
    "lzx_sumcheck_session_round_eval": (st_t, [vp, vp]),
    "lzx_sumcheck_session_fold": (st_t, [vp, u64]),
    "lzx_sumcheck_session_final": (st_t, [vp, vp, C.POINTER(u64)]),
    "lzx_sumcheck_session_destroy": (None, [vp]),
    "lzx_sumcheck_prove_batch": (st_t, [vp, sz, vp, vp, vp, vp, C.POINTER(u64), vp, C.POINTER(sz), vp]),
    "lzx_sumcheck_prove_batch_star": (st_t, [vp, sz, vp, vp, vp, vp, C.POINTER(u64), vp, C.POINTER(sz),
                                             C.POINTER(sz), vp]),
    "lzx_sumcheck_verify_batch": (st_t, [vp, sz, sz, sz, sz, vp, sz, vp, i32, u64, vp]),
    "lzx_sumcheck_verify_batch_star": (st_t, [vp, sz, sz, vp, sz, sz, vp, sz, vp, sz, vp, i32, u64, vp]),
    "lzx_vp_sum_over_hypercube": (st_t, [sz, sz, C.POINTER(vp), sz, vp, vp, vp, C.POINTER(u64), vp]),
    "lzx_gl_mul_rate": (st_t, [u32, C.POINTER(C.c_double)]),
}


class LzxError(RuntimeError):
    """A non-zero lzx_status, carrying the code and the reference enum name."""

    def __init__(self, code: int, what: str = ""):
        self.code = code
        name = lib().lzx_status_string(code).decode()
        dev = lib().lzx_last_device_error().decode() if code >= 100 else ""
        super().__init__(f"{what}: {name} ({code}){' ' + dev if dev else ''}")


def lib():
    """Load liblzx_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m lzx_amd._build` (hipcc, gfx950). "
            "There is no CPU fallback."
        )
    L = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | getattr(os, "RTLD_GLOBAL", 0))
    for name, (res, args) in PROTOTYPES.items():
        f = getattr(L, name, None)
        if f is None:  # completeness is asserted by tests/test_abi.py
            continue
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(code: int, what: str = "lzx call") -> None:
    if code != 0:
        raise LzxError(code, what)

"""ctypes view of the CPU oracle (oracle/liblzx_oracle.so) plus closed forms.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product package (lzx_amd) never imports it.

Functions mirror the reference (idrees2516/lzx) items cited in
oracle/lzx_oracle.c; the closed forms at the bottom are independent
big-integer restatements of the mathematics used to pin the C code.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB_PATH = HERE / "liblzx_oracle.so"
_lib = None

GL_P = 0xFFFFFFFF00000001

# Moduli used by the workloads: (q, two_adicity, generator).
Q_32 = (3221225473, 30, 5)            # modulus.rs:48
Q_12289 = (12289, 12, 11)             # modulus.rs:52
Q_2013265921 = (2013265921, 27, 31)   # modulus.rs:55
CRT_PRIMES = [(2013265921, 27, 31), (1811939329, 26, 13), (2113929217, 25, 5), (998244353, 23, 3)]
Q_50_BUILD = (1125899437080577, 26, 5)  # ~50-bit prime chosen by this build (not in the reference)

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u8p = C.c_char_p


class OrMod32(C.Structure):
    _fields_ = [("q", C.c_uint32), ("two_adicity", C.c_uint32), ("generator", C.c_uint32), ("barrett", C.c_uint64)]


def lib():
    """Load (building first if absent) the oracle shared library."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        from oracle import build as _b  # noqa: WPS433
        _b.build()
    L = C.CDLL(str(_LIB_PATH))
    vp = C.c_void_p
    sz = C.c_size_t
    u32, u64, i64 = C.c_uint32, C.c_uint64, C.c_int64
    sig = {
        "or_mod32_new": (OrMod32, [u32, u32, u32]),
        "or_reduce_u64": (u32, [C.POINTER(OrMod32), u64]),
        "or_reduce_i64": (u32, [C.POINTER(OrMod32), i64]),
        "or_mul": (u32, [C.POINTER(OrMod32), u32, u32]),
        "or_add": (u32, [C.POINTER(OrMod32), u32, u32]),
        "or_sub": (u32, [C.POINTER(OrMod32), u32, u32]),
        "or_neg": (u32, [C.POINTER(OrMod32), u32]),
        "or_pow": (u32, [C.POINTER(OrMod32), u32, u64]),
        "or_inv": (u32, [C.POINTER(OrMod32), u32]),
        "or_root_of_unity": (u32, [C.POINTER(OrMod32), u32]),
        "or_ntt_new": (C.c_int, [u32, u32, u32, u32, C.POINTER(vp)]),
        "or_ntt_free": (None, [vp]),
        "or_ntt_psi": (u32, [vp]),
        "or_ntt_n_inv": (u32, [vp]),
        "or_ntt_forward": (C.c_int, [vp, _u32p, sz]),
        "or_ntt_inverse": (C.c_int, [vp, _u32p, sz]),
        "or_ntt_forward_partial": (C.c_int, [vp, _u32p, sz, u32]),
        "or_ntt_complete_partial": (C.c_int, [vp, _u32p, sz, u32]),
        "or_ntt_forward_batch": (C.c_int, [vp, _u32p, sz]),
        "or_ntt_inverse_batch": (C.c_int, [vp, _u32p, sz]),
        "or_ring_mul": (C.c_int, [vp, _u32p, _u32p, _u32p]),
        "or_rq2_mul": (C.c_int, [vp, _u32p, _u32p, _u32p]),
        "or_rq2_norm": (C.c_int, [vp, _u32p, _u32p]),
        "or_rq2_inverse": (C.c_int, [vp, _u32p, _u32p]),
        "or_keccak_f1600": (None, [_u64p]),
        "or_shake256": (None, [_u8p, sz, C.c_void_p, sz]),
        "or_sha3_256": (None, [_u8p, sz, C.c_void_p]),
        "or_xof": (None, [_u8p, sz, _u8p, sz, C.c_void_p, sz]),
        "or_ceil_sqrt_u128": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "or_short_challenge_sample": (C.c_int, [sz, C.c_int, sz, C.c_int64, C.c_uint32, vp, sz, vp, sz, vp,
                                                C.POINTER(C.c_uint64)]),
        "or_rq2_sample_challenge": (C.c_int, [C.c_uint32, C.c_uint32, sz, C.c_int, sz, C.c_int64, C.c_uint32, vp,
                                              sz, vp, sz, vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "or_uniform_from_seed": (None, [u32, u32, _u8p, sz, _u8p, sz, u64, _u32p]),
        "or_sample_small_secret": (None, [u32, u32, sz, u32, _u8p, sz, _u32p]),
        "or_ajtai_from_seed": (C.c_int, [u32, u32, u32, u32, sz, sz, _u8p, vp, vp]),
        "or_ajtai_commit": (C.c_int, [vp, sz, sz, _u32p, _u32p, sz, _u32p]),
        "or_ajtai_commit_naive": (C.c_int, [vp, sz, sz, _u32p, _u32p, sz, _u32p]),
        "or_negacyclic_schoolbook": (None, [u32, u32, _u32p, _u32p, _u32p]),
        "or_gl_from_u64": (u64, [u64]),
        "or_gl_from_u128": (u64, [u64, u64]),
        "or_gl_add": (u64, [u64, u64]),
        "or_gl_sub": (u64, [u64, u64]),
        "or_gl_neg": (u64, [u64]),
        "or_gl_mul": (u64, [u64, u64]),
        "or_gl_pow": (u64, [u64, u64]),
        "or_gl_inverse": (u64, [u64]),
        "or_transcript_new": (vp, [_u8p, sz, u64]),
        "or_transcript_free": (None, [vp]),
        "or_transcript_clone": (vp, [vp]),
        "or_transcript_append_message": (C.c_int, [vp, _u8p, sz, _u8p, sz]),
        "or_transcript_append_field_slice": (C.c_int, [vp, _u8p, sz, _u64p, sz]),
        "or_transcript_challenge_fields": (C.c_int, [vp, _u8p, sz, sz, _u64p]),
        "or_eq_table": (None, [_u64p, sz, _u64p]),
        "or_bind_first_half": (None, [_u64p, sz, u64]),
        "or_mle_evaluate": (u64, [_u64p, sz, _u64p]),
        "or_interpolate_at": (u64, [_u64p, sz, u64]),
        "or_vp_max_degree": (sz, [sz, _u32p]),
        "or_vp_sum_over_hypercube": (u64, [sz, sz, C.POINTER(C.c_void_p), sz, _u64p, _u32p, _u32p]),
        "or_sumcheck_prove": (C.c_int, [sz, sz, C.POINTER(C.c_void_p), sz, _u64p, _u32p, _u32p, u64, vp,
                                        _u64p, _u64p, C.POINTER(u64), _u64p, C.POINTER(sz)]),
        "or_sumcheck_verify": (C.c_int, [_u64p, sz, sz, sz, sz, u64, vp, C.c_int, u64, _u64p, C.POINTER(u64)]),
        "or_sc_prove_ref": (C.c_int, [sz, sz, C.POINTER(C.c_void_p), sz, _u64p, _u32p, _u32p, u64, vp,
                                      _u64p, _u64p, C.POINTER(u64), _u64p, C.POINTER(sz), C.c_int, C.c_int]),
        "or_cpu_has_avx512f": (C.c_int, []),
        "or_ntt_roundtrip_par": (None, [vp, _u32p, sz, C.c_int]),
        "or_ajtai_commit_par": (C.c_int, [vp, sz, sz, _u32p, _u32p, sz, _u32p, C.c_int]),
        "or_powmod64": (u64, [u64, u64, u64]),
        "or_commit_wide_schoolbook": (C.c_int, [u64, u32, sz, sz, sz, vp, C.c_int, vp, _u64p]),
        "or_commit_wide_check_at": (sz, [u64, u32, sz, sz, sz, vp, C.c_int, vp, vp, C.c_int, u64]),
        "or_eval_key_at": (C.c_int, [u64, u32, sz, sz, sz, vp, C.c_int, u64, vp]),
        "or_commit_check_at_evals": (sz, [u64, u32, sz, sz, sz, vp, vp, vp, C.c_int, u64]),
        "or_quad_tables": (C.c_int, [u64, u32, u64, u32, C.POINTER(u64), vp, vp, vp]),
        "or_quad_forward": (C.c_int, [u64, u32, u64, u32, _u64p]),
        "or_quad_inverse": (C.c_int, [u64, u32, u64, u32, _u64p]),
        "or_quad_slotwise_mul": (C.c_int, [u64, u32, u64, u32, _u64p, _u64p, _u64p]),
        "or_quad_mul": (C.c_int, [u64, u32, u64, u32, _u64p, _u64p, _u64p]),
        "or_pack_field_elements": (sz, [u32, u32, _u64p, sz, _u32p]),
        "or_unpack_field_elements": (C.c_int, [u32, u32, _u32p, sz, _u64p, C.POINTER(sz)]),
        "or_pack_bits": (sz, [u32, C.c_void_p, sz, _u32p]),
        "or_unpack_bits": (None, [u32, _u32p, sz, C.c_void_p, sz]),
        "or_gadget_decompose": (C.c_int, [u64, sz, u64, C.c_void_p]),
        "or_gadget_recompose": (C.c_int, [u64, sz, C.c_void_p, sz, C.POINTER(u64)]),
        "or_crt_pack": (C.c_int, [u32, u32, _u64p, sz, _u64p, sz, _u32p, C.POINTER(sz)]),
        "or_crt_unpack": (C.c_int, [u32, _u32p, sz, _u64p, sz, sz, _u64p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


# ---------------------------------------------------------------------------
# Modulus32 / NTT
# ---------------------------------------------------------------------------

class Mod32:
    """Modulus32 (modulus.rs:22-162)."""

    def __init__(self, q, two_adicity=1, generator=1):
        self.m = lib().or_mod32_new(q, two_adicity, generator)
        self.q = q

    def reduce_u64(self, x):
        return lib().or_reduce_u64(C.byref(self.m), x)

    def reduce_i64(self, x):
        return lib().or_reduce_i64(C.byref(self.m), x)

    def mul(self, a, b):
        return lib().or_mul(C.byref(self.m), a, b)

    def add(self, a, b):
        return lib().or_add(C.byref(self.m), a, b)

    def sub(self, a, b):
        return lib().or_sub(C.byref(self.m), a, b)

    def inv(self, a):
        return lib().or_inv(C.byref(self.m), a)

    def root_of_unity(self, k):
        return lib().or_root_of_unity(C.byref(self.m), k)


class Ntt:
    """NttTables (ntt.rs:16-342) restated; operates on uint32 numpy arrays."""

    def __init__(self, q, two_adicity, generator, log_n):
        h = C.c_void_p()
        st = lib().or_ntt_new(q, two_adicity, generator, log_n, C.byref(h))
        self.status = st
        self.h = h if st == 0 else None
        self.q, self.log_n, self.n = q, log_n, 1 << log_n

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.or_ntt_free(self.h)
            self.h = None

    @property
    def psi(self):
        return lib().or_ntt_psi(self.h)

    @property
    def n_inv(self):
        return lib().or_ntt_n_inv(self.h)

    def forward(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint32).copy()
        st = lib().or_ntt_forward_batch(self.h, a, a.size // self.n) if a.size % self.n == 0 else 1
        return st, a

    def inverse(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint32).copy()
        st = lib().or_ntt_inverse_batch(self.h, a, a.size // self.n) if a.size % self.n == 0 else 1
        return st, a

    def forward_partial(self, a, levels):
        a = np.ascontiguousarray(a, dtype=np.uint32).copy()
        flat = a.reshape(-1, self.n) if a.size % self.n == 0 else None
        if flat is None:
            return 1, a
        for row in flat:
            st = lib().or_ntt_forward_partial(self.h, row, self.n, levels)
            if st:
                return st, a
        return 0, a

    def complete_partial(self, a, already):
        a = np.ascontiguousarray(a, dtype=np.uint32).copy()
        flat = a.reshape(-1, self.n) if a.size % self.n == 0 else None
        if flat is None:
            return 1, a
        for row in flat:
            st = lib().or_ntt_complete_partial(self.h, row, self.n, already)
            if st:
                return st, a
        return 0, a

    def ring_mul(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        out = np.zeros_like(a)
        for i in range(a.size // self.n):
            o = np.zeros(self.n, dtype=np.uint32)
            lib().or_ring_mul(self.h, np.ascontiguousarray(a.reshape(-1, self.n)[i]),
                              np.ascontiguousarray(b.reshape(-1, self.n)[i]), o)
            out.reshape(-1, self.n)[i] = o
        return out

    # R_q[Y]/(Y^2+1) (extension.rs:115-167): elements are rows [c0 | c1] of 2n
    def rq2_mul(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint32).reshape(-1, 2 * self.n)
        b = np.ascontiguousarray(b, dtype=np.uint32).reshape(-1, 2 * self.n)
        out = np.zeros_like(a)
        for i in range(a.shape[0]):
            o = np.zeros(2 * self.n, dtype=np.uint32)
            lib().or_rq2_mul(self.h, np.ascontiguousarray(a[i]), np.ascontiguousarray(b[i]), o)
            out[i] = o
        return out

    def rq2_norm(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint32).reshape(-1, 2 * self.n)
        out = np.zeros((a.shape[0], self.n), dtype=np.uint32)
        for i in range(a.shape[0]):
            o = np.zeros(self.n, dtype=np.uint32)
            lib().or_rq2_norm(self.h, np.ascontiguousarray(a[i]), o)
            out[i] = o
        return out

    def rq2_inverse(self, a):
        """(ok[count], out[count][2n]); rows with ok = 0 (no inverse) are zero."""
        a = np.ascontiguousarray(a, dtype=np.uint32).reshape(-1, 2 * self.n)
        out = np.zeros_like(a)
        ok = np.zeros(a.shape[0], dtype=np.uint8)
        for i in range(a.shape[0]):
            o = np.zeros(2 * self.n, dtype=np.uint32)
            ok[i] = lib().or_rq2_inverse(self.h, np.ascontiguousarray(a[i]), o)
            if ok[i]:
                out[i] = o
        return ok, out


# ---------------------------------------------------------------------------
# Keccak / XOF / sampling / Ajtai
# ---------------------------------------------------------------------------

def shake256(data: bytes, n: int) -> bytes:
    buf = C.create_string_buffer(max(n, 1))
    lib().or_shake256(data, len(data), buf, n)
    return buf.raw[:n]


def sha3_256(data: bytes) -> bytes:
    buf = C.create_string_buffer(32)
    lib().or_sha3_256(data, len(data), buf)
    return buf.raw


def keccak_f1600(state):
    st = np.ascontiguousarray(state, dtype=np.uint64).copy()
    lib().or_keccak_f1600(st)
    return st


def xof(domain: bytes, msg: bytes, n: int) -> bytes:
    buf = C.create_string_buffer(max(n, 1))
    lib().or_xof(domain, len(domain), msg, len(msg), buf, n)
    return buf.raw[:n]


def uniform_from_seed(q, log_n, domain: bytes, seed: bytes, index: int):
    out = np.zeros(1 << log_n, dtype=np.uint32)
    lib().or_uniform_from_seed(q, log_n, domain, len(domain), seed, len(seed), index, out)
    return out


def sample_small_secret(q, log_n, m, bound, seed: bytes):
    out = np.zeros(m << log_n, dtype=np.uint32)
    lib().or_sample_small_secret(q, log_n, m, bound, seed, len(seed), out)
    return out.reshape(m, 1 << log_n)


def ajtai_from_seed(mod, log_n, k, m, seed: bytes, want_A=True, want_Ahat=True):
    """AjtaiPublicKey::from_seed (ajtai.rs:87-103): returns (A, Ahat), each (k*m, n)."""
    q, adic, g = mod
    n = 1 << log_n
    A = np.zeros((k * m, n), dtype=np.uint32) if want_A else None
    Ah = np.zeros((k * m, n), dtype=np.uint32) if want_Ahat else None
    st = lib().or_ajtai_from_seed(q, adic, g, log_n, k, m, seed,
                                  A.ctypes.data if want_A else None, Ah.ctypes.data if want_Ahat else None)
    assert st == 0, st
    return A, Ah


def ajtai_commit(ntt: Ntt, k, m, Ahat, s):
    """AjtaiPublicKey::commit (ajtai.rs:118-153); s is (m, n) canonical."""
    s = np.ascontiguousarray(s, dtype=np.uint32)
    out = np.zeros((k, ntt.n), dtype=np.uint32)
    st = lib().or_ajtai_commit(ntt.h, k, m, np.ascontiguousarray(Ahat, dtype=np.uint32), s.reshape(-1),
                               s.shape[0] if s.ndim > 1 else s.size // ntt.n, out.reshape(-1))
    return st, out


def cpu_has_avx512f() -> bool:
    return bool(lib().or_cpu_has_avx512f())


def ntt_roundtrip_par(ntt: Ntt, data, polys, threads):
    """Forward then inverse of `polys` polynomials in place over `threads` threads."""
    lib().or_ntt_roundtrip_par(ntt.h, data, polys, threads)


def ajtai_commit_par(ntt: Ntt, k, m, Ahat, s, threads):
    """ajtai.rs:118-153 over `threads` OpenMP threads; same values as ajtai_commit."""
    s = np.ascontiguousarray(s, dtype=np.uint32).reshape(-1)
    out = np.zeros((k, ntt.n), dtype=np.uint32)
    st = lib().or_ajtai_commit_par(ntt.h, k, m, np.ascontiguousarray(Ahat, dtype=np.uint32).reshape(-1), s,
                                   s.size // ntt.n, out.reshape(-1), threads)
    return st, out


def ajtai_commit_naive(ntt: Ntt, k, m, A, s):
    """commit_naive_reference (ajtai.rs:158-180)."""
    s = np.ascontiguousarray(s, dtype=np.uint32)
    out = np.zeros((k, ntt.n), dtype=np.uint32)
    st = lib().or_ajtai_commit_naive(ntt.h, k, m, np.ascontiguousarray(A, dtype=np.uint32), s.reshape(-1),
                                     s.shape[0], out.reshape(-1))
    return st, out


def negacyclic_schoolbook(q, a, b):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    out = np.zeros_like(a)
    lib().or_negacyclic_schoolbook(q, a.size, a, b, out)
    return out


def powmod64(b, e, q):
    return lib().or_powmod64(b, e, q)


def commit_wide_schoolbook(q, A, digits, m=None):
    """t = A s over Z_q[X]/(X^n+1), exact schoolbook (ajtai.rs:118-153 semantics).
    A: (k, lda, n) u32/u64, digits: (>= m, n) int8.  Returns (k, n) uint64."""
    A = np.ascontiguousarray(A)
    assert A.dtype in (np.uint32, np.uint64) and A.ndim == 3
    k, lda, n = A.shape
    digits = np.ascontiguousarray(digits, dtype=np.int8)
    m = digits.shape[0] if m is None else m
    out = np.zeros((k, n), dtype=np.uint64)
    st = lib().or_commit_wide_schoolbook(q, n, k, m, lda, A.ctypes.data, A.itemsize, digits.ctypes.data, out)
    assert st == 0, st
    return out


def commit_wide_check_at(q, A, digits, t, z, m=None):
    """Number of rows i with t_i(z) != sum_j A_ij(z) s_j(z) mod q (z a root of X^n + 1)."""
    A = np.ascontiguousarray(A)
    k, lda, n = A.shape
    digits = np.ascontiguousarray(digits, dtype=np.int8)
    m = digits.shape[0] if m is None else m
    t = np.ascontiguousarray(t)
    assert t.shape == (k, n) and t.dtype in (np.uint32, np.uint64)
    return lib().or_commit_wide_check_at(q, n, k, m, lda, A.ctypes.data, A.itemsize, digits.ctypes.data,
                                         t.ctypes.data, t.itemsize, z)


def eval_key_at(q, A, z, m=None):
    """A_ij(z) mod q for a (k, lda, n) key (u32/u64), j < m: (k, m) uint64."""
    A = np.ascontiguousarray(A)
    assert A.dtype in (np.uint32, np.uint64) and A.ndim == 3
    k, lda, n = A.shape
    m = lda if m is None else m
    out = np.zeros((k, m), dtype=np.uint64)
    st = lib().or_eval_key_at(q, n, k, m, lda, A.ctypes.data, A.itemsize, z, out.ctypes.data)
    assert st == 0, st
    return out


def commit_check_at_evals(q, Az, digits, t, z, m=None):
    """Rows of t (k, n) with t_i(z) != sum_{j<m} Az[i, j] s_j(z), s = int8 digits (>= m, n)."""
    Az = np.ascontiguousarray(Az, dtype=np.uint64)
    digits = np.ascontiguousarray(digits, dtype=np.int8)
    t = np.ascontiguousarray(t)
    k, n = t.shape
    m = digits.shape[0] if m is None else m
    assert digits.shape[1] == n and Az.shape[0] == k and Az.shape[1] >= m and t.dtype in (np.uint32, np.uint64)
    return lib().or_commit_check_at_evals(q, n, k, m, Az.shape[1], Az.ctypes.data, digits.ctypes.data,
                                          t.ctypes.data, t.itemsize, z)


# ---------------------------------------------------------------------------
# Goldilocks / transcript / sum-check
# ---------------------------------------------------------------------------

def gl_add(a, b): return lib().or_gl_add(a, b)
def gl_sub(a, b): return lib().or_gl_sub(a, b)
def gl_mul(a, b): return lib().or_gl_mul(a, b)
def gl_neg(a): return lib().or_gl_neg(a)
def gl_inverse(a): return lib().or_gl_inverse(a)
def gl_from_u128(x): return lib().or_gl_from_u128(x >> 64, x & (2**64 - 1))


class Transcript:
    """Transcript (transcript.rs:36-238)."""

    def __init__(self, protocol: bytes = b"", budget: int = 1 << 20, _h=None):
        self.h = _h if _h is not None else lib().or_transcript_new(protocol, len(protocol), budget)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_transcript_free(self.h)
            self.h = None

    def clone(self):
        return Transcript(_h=lib().or_transcript_clone(self.h))

    def append_message(self, label: bytes, msg: bytes):
        return lib().or_transcript_append_message(self.h, label, len(label), msg, len(msg))

    def append_field_slice(self, label: bytes, vals):
        v = np.ascontiguousarray(vals, dtype=np.uint64)
        return lib().or_transcript_append_field_slice(self.h, label, len(label), v, v.size)

    def challenge_fields(self, label: bytes, n: int):
        out = np.zeros(max(n, 1), dtype=np.uint64)
        st = lib().or_transcript_challenge_fields(self.h, label, len(label), n, out)
        return st, out[:n]

    def challenge_field(self, label: bytes):
        st, out = self.challenge_fields(label, 1)
        return st, int(out[0])


def eq_table(point):
    p = np.ascontiguousarray(point, dtype=np.uint64)
    out = np.zeros(1 << p.size, dtype=np.uint64)
    lib().or_eq_table(p, p.size, out)
    return out


def mle_evaluate(evals, point):
    e = np.ascontiguousarray(evals, dtype=np.uint64)
    p = np.ascontiguousarray(point, dtype=np.uint64)
    return lib().or_mle_evaluate(e, p.size, p if p.size else np.zeros(1, np.uint64))


def bind_first_half(evals, r):
    e = np.ascontiguousarray(evals, dtype=np.uint64).copy()
    lib().or_bind_first_half(e, e.size, r)
    return e[: e.size // 2]


def interpolate_at(evals, r):
    e = np.ascontiguousarray(evals, dtype=np.uint64)
    return lib().or_interpolate_at(e, e.size, r)


def _terms_csr(terms):
    coeffs = np.array([c for c, _ in terms] or [0], dtype=np.uint64)
    off = [0]
    ids = []
    for _, fs in terms:
        ids.extend(fs)
        off.append(len(ids))
    return coeffs, np.array(off, dtype=np.uint32), np.array(ids or [0], dtype=np.uint32)


def vp_sum_over_hypercube(num_vars, factors, terms):
    coeffs, off, ids = _terms_csr(terms)
    facs = [np.ascontiguousarray(f, dtype=np.uint64) for f in factors]
    ptrs = (C.c_void_p * max(len(facs), 1))(*[f.ctypes.data for f in facs])
    return lib().or_vp_sum_over_hypercube(num_vars, len(facs), ptrs, len(terms), coeffs, off, ids)


def sumcheck_prove(num_vars, factors, terms, claim, transcript: Transcript, simd=None, threads=1):
    """sumcheck::prove (sumcheck.rs:62-162).  Returns (status, dict).
    simd=None runs the plain restatement; simd=0/1 the reference's
    single-binding round structure in its scalar (LZX_NO_SIMD=1) or AVX-512
    form (the CPU baseline), over `threads` OpenMP threads."""
    coeffs, off, ids = _terms_csr(terms)
    facs = [np.ascontiguousarray(f, dtype=np.uint64) for f in factors]
    ptrs = (C.c_void_p * max(len(facs), 1))(*[f.ctypes.data for f in facs])
    d = max((len(fs) for _, fs in terms), default=1)
    rounds = np.zeros(max(num_vars * (d + 1), 1), dtype=np.uint64)
    chals = np.zeros(max(num_vars, 1), dtype=np.uint64)
    fclaims = np.zeros(max(len(facs), 1), dtype=np.uint64)
    fin = C.c_uint64(0)
    dout = C.c_size_t(0)
    if simd is None:
        st = lib().or_sumcheck_prove(num_vars, len(facs), ptrs, len(terms), coeffs, off, ids, claim, transcript.h,
                                     rounds, chals, C.byref(fin), fclaims, C.byref(dout))
    else:
        st = lib().or_sc_prove_ref(num_vars, len(facs), ptrs, len(terms), coeffs, off, ids, claim, transcript.h,
                                   rounds, chals, C.byref(fin), fclaims, C.byref(dout), int(simd), int(threads))
    w = dout.value + 1
    return st, {
        "rounds": rounds[: num_vars * w].reshape(num_vars, w),
        "challenges": chals[:num_vars],
        "final_claim": fin.value,
        "factor_claims": fclaims[: len(facs)],
        "degree": dout.value,
    }


def sumcheck_verify(rounds, num_vars, max_degree, claim, transcript: Transcript, expected=None):
    r = np.ascontiguousarray(rounds, dtype=np.uint64)
    nr = r.shape[0] if r.ndim == 2 else 0
    rl = r.shape[1] if r.ndim == 2 else 0
    point = np.zeros(max(nr, 1), dtype=np.uint64)
    fin = C.c_uint64(0)
    st = lib().or_sumcheck_verify(r.reshape(-1) if r.size else np.zeros(1, np.uint64), nr, rl, num_vars,
                                  max_degree, claim, transcript.h, 1 if expected is not None else 0,
                                  expected or 0, point, C.byref(fin))
    return st, point[:nr], fin.value


# ---------------------------------------------------------------------------
# Closed forms (pure Python big integers; small sizes only)
# ---------------------------------------------------------------------------

def closed_psi(q, g, log_n):
    """psi = g^((q-1) >> (log_n+1)) (modulus.rs:153-155 with ntt.rs:52-54)."""
    return pow(g, (q - 1) >> (log_n + 1), q)


def closed_ntt(a, q, g, log_n):
    """Direct evaluation out[k] = sum_i a_i psi^(i(2k+1)) mod q."""
    n = 1 << log_n
    psi = closed_psi(q, g, log_n)
    return [sum(int(a[i]) * pow(psi, i * (2 * k + 1), q) for i in range(n)) % q for k in range(n)]


def closed_negacyclic(a, b, q):
    n = len(a)
    out = [0] * n
    for i in range(n):
        for j in range(n):
            v = int(a[i]) * int(b[j])
            if i + j < n:
                out[i + j] += v
            else:
                out[i + j - n] -= v
    return [x % q for x in out]


def splitmix64(seed):
    """splitmix64 stream (the seeded generator the workloads use for inputs)."""
    state = seed & (2**64 - 1)
    while True:
        state = (state + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        yield z ^ (z >> 31)


# ---------------------------------------------------------------------------
# zerocheck (crates/lattice-sumcheck/src/zerocheck.rs:26-84), composed from the
# oracle's transcript, eq table, MLE evaluation and sum-check.
# ---------------------------------------------------------------------------

ZEROCHECK_FINAL_CLAIM_NONZERO = 15


def zerocheck_prove(num_vars, poly, transcript: Transcript):
    """zerocheck::prove (zerocheck.rs:26-54).  Returns (status, point, sumcheck dict)."""
    st, r = transcript.challenge_fields(b"zerocheck-point", num_vars)  # :31-34
    if st:
        return st, None, None
    r = np.asarray(r, dtype=np.uint64)[:num_vars]
    eq = eq_table(r)                                   # :37
    if mle_evaluate(poly, r) != 0:                     # :38-45
        return ZEROCHECK_FINAL_CLAIM_NONZERO, r, None
    st, out = sumcheck_prove(num_vars, [poly, eq], [(1, [0, 1])], 0, transcript)  # :46-54
    return st, r, out


def zerocheck_verify(rounds, num_vars, transcript: Transcript, expected_poly_eval=None):
    """zerocheck::verify (zerocheck.rs:58-84).  Returns (status, r)."""
    st, r = transcript.challenge_fields(b"zerocheck-point", num_vars)
    if st:
        return st, None
    st, _point, _fin = sumcheck_verify(rounds, num_vars, 2, 0, transcript, expected_poly_eval)
    return st, (np.asarray(r, dtype=np.uint64)[:num_vars] if st == 0 else None)


# ---------------------------------------------------------------------------
# Claim batching (crates/lattice-sumcheck/src/batch.rs:33-336), composed from
# the transcript and sumcheck_prove / sumcheck_verify above
# ---------------------------------------------------------------------------

BATCH_VARIABLE_COUNT_MISMATCH, BATCH_EMPTY, BATCH_CHALLENGE_COUNT = 16, 17, 18


def gl_pow2(k):
    """pow2 (batch.rs:150-157): 1 doubled k times."""
    v = 1
    for _ in range(k):
        v = (v + v) % GL_P
    return v


def pad_factor(f, k):
    """pad_factor (batch.rs:130-148): every entry repeated 2^k times consecutively."""
    return np.repeat(np.ascontiguousarray(f, dtype=np.uint64), 1 << k)


def sumcheck_prove_batch(claims, transcript: Transcript, star=False):
    """prove_batch (batch.rs:33-80) or prove_batch_star (:161-225).  claims are
    (num_vars, factors, terms, claimed_sum).  Returns (status, dict) with
    rhos, num_vars and sumcheck_prove's outputs."""
    if not claims:
        return BATCH_EMPTY, None
    nvs = [c[0] for c in claims]
    n = max(nvs) if star else nvs[0]
    if not star and any(v != n for v in nvs):                                   # :40-46
        return BATCH_VARIABLE_COUNT_MISMATCH, None
    st, rhos = transcript.challenge_fields(b"batch-star-rho" if star else b"batch-rho", len(claims))
    if st:
        return BATCH_CHALLENGE_COUNT, None
    factors, terms, combined = [], [], 0
    for ci, (nv, facs, tms, claimed) in enumerate(claims):
        rho, k = int(rhos[ci]), n - nv
        combined = (combined + rho * gl_pow2(k) % GL_P * (int(claimed) % GL_P)) % GL_P   # :58, :185-187
        base = len(factors)
        factors += [pad_factor(f, k) for f in facs]
        terms += [(rho * (int(c) % GL_P) % GL_P, [base + i for i in ids]) for c, ids in tms]
    st, out = sumcheck_prove(n, factors, terms, combined, transcript)
    if st:
        return st, None
    out["rhos"] = np.asarray(rhos, dtype=np.uint64)[: len(claims)]
    out["num_vars"] = n
    out["combined_claim"] = combined
    return 0, out


def sumcheck_verify_batch(rounds, num_vars, max_degree, claimed_sums, transcript: Transcript, expected=None):
    """verify_batch (batch.rs:300-336).  Returns (status, point)."""
    st, rhos = transcript.challenge_fields(b"batch-rho", len(claimed_sums))
    if st:
        return BATCH_CHALLENGE_COUNT, None
    comb = sum(int(r) * (int(c) % GL_P) for r, c in zip(rhos, claimed_sums)) % GL_P
    st, point, _ = sumcheck_verify(rounds, num_vars, max_degree, comb, transcript, expected)
    return st, point


def sumcheck_verify_batch_star(rounds, num_vars, max_degree, claimed_sums, transcript: Transcript, expected=None,
                               rhos=None):
    """verify_batch_star (batch.rs:236-262); with rhos, verify_batch_star_with_rhos (:268-296)."""
    if not num_vars or len(num_vars) != len(claimed_sums):
        return BATCH_VARIABLE_COUNT_MISMATCH, None
    if rhos is None:
        st, rhos = transcript.challenge_fields(b"batch-star-rho", len(claimed_sums))
        if st:
            return BATCH_CHALLENGE_COUNT, None
    if len(rhos) != len(claimed_sums):
        return BATCH_CHALLENGE_COUNT, None
    n = max(num_vars)
    comb = sum(int(r) * gl_pow2(n - v) % GL_P * (int(c) % GL_P) for r, v, c in zip(rhos, num_vars, claimed_sums))
    st, point, _ = sumcheck_verify(rounds, n, max_degree, comb % GL_P, transcript, expected)
    return st, point


# ---------------------------------------------------------------------------
# The ~2^50 quadratic-slot ring (crates/lattice-ring/src/modulus50.rs:44-499)
# ---------------------------------------------------------------------------

Q50 = (1125899906839937, 7, 3)  # 2^50 - 2687, two-adicity 7, generator 3 (modulus50.rs:44-53)


def quad_tables(q, two_adicity, generator, log_n):
    """QuadSlotTables::new: (status, zeta, twiddles, inverse twiddles, leaf gammas)."""
    n = 1 << min(max(log_n, 1), 7)
    zeta = C.c_uint64(0)
    tw = np.zeros(n, dtype=np.uint64)
    twi = np.zeros(n, dtype=np.uint64)
    gam = np.zeros(n, dtype=np.uint64)
    st = lib().or_quad_tables(q, two_adicity, generator, log_n, C.byref(zeta), tw.ctypes.data, twi.ctypes.data,
                              gam.ctypes.data)
    return st, zeta.value, tw[: n // 2 - 1], twi[: n // 2 - 1], gam[: n // 2]


def _quad_each(fn, q, two_adicity, generator, log_n, *arrays):
    n = 1 << log_n
    ins = [np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, n) for a in arrays]
    out = np.zeros_like(ins[0])
    for p in range(out.shape[0]):
        if len(ins) == 1:
            row = ins[0][p].copy()
            st = fn(q, two_adicity, generator, log_n, row)
            out[p] = row
        else:
            row = np.zeros(n, dtype=np.uint64)
            st = fn(q, two_adicity, generator, log_n, ins[0][p].copy(), ins[1][p].copy(), row)
            out[p] = row
        if st:
            return st, None
    return 0, out.reshape(-1)


def quad_forward(q, two_adicity, generator, log_n, a):
    return _quad_each(lib().or_quad_forward, q, two_adicity, generator, log_n, a)


def quad_inverse(q, two_adicity, generator, log_n, a):
    return _quad_each(lib().or_quad_inverse, q, two_adicity, generator, log_n, a)


def quad_slotwise_mul(q, two_adicity, generator, log_n, a, b):
    return _quad_each(lib().or_quad_slotwise_mul, q, two_adicity, generator, log_n, a, b)


def quad_mul(q, two_adicity, generator, log_n, a, b):
    return _quad_each(lib().or_quad_mul, q, two_adicity, generator, log_n, a, b)


# ---------------------------------------------------------------------------
# Witness packing (packing.rs:30-200, sparse.rs:30-71, decomposition.rs:47-93)
# ---------------------------------------------------------------------------

def pack_field_elements(q, n, values):
    v = np.ascontiguousarray(values, dtype=np.uint64)
    out = np.zeros(max(1, -(-3 * v.size // n)) * n, dtype=np.uint32)
    e = lib().or_pack_field_elements(q, n, v if v.size else np.zeros(1, np.uint64), v.size, out)
    return out[: e * n].reshape(e, n)


def unpack_field_elements(q, n, packed):
    pk = np.ascontiguousarray(packed, dtype=np.uint32).reshape(-1)
    out = np.zeros(max(1, pk.size // 3), dtype=np.uint64)
    got = C.c_size_t(0)
    st = lib().or_unpack_field_elements(q, n, pk, pk.size // n, out, C.byref(got))
    return st, out[: got.value]


def pack_bits(n, bits):
    b = np.ascontiguousarray(bits, dtype=np.uint8)
    e = max(1, -(-b.size // (31 * n)))
    out = np.zeros(e * n, dtype=np.uint32)
    b1 = b if b.size else np.zeros(1, np.uint8)
    lib().or_pack_bits(n, b1.ctypes.data, b.size, out)
    return out.reshape(e, n)


def unpack_bits(n, elems, nbits):
    el = np.ascontiguousarray(elems, dtype=np.uint32).reshape(-1)
    out = np.zeros(max(1, nbits), dtype=np.uint8)
    lib().or_unpack_bits(n, el if el.size else np.zeros(1, np.uint32), el.size // n, out.ctypes.data, nbits)
    return out[:nbits]


def gadget_decompose(base, num_digits, x):
    d = np.zeros(max(1, num_digits), dtype=np.int64)
    st = lib().or_gadget_decompose(base, num_digits, x, d.ctypes.data)
    return st, d[:num_digits]


def gadget_recompose(base, num_digits, digits):
    d = np.ascontiguousarray(digits, dtype=np.int64)
    x = C.c_uint64(0)
    st = lib().or_gadget_recompose(base, num_digits, d.ctypes.data if d.size else None, d.size, C.byref(x))
    return st, x.value


def crt_pack(q, n, values, moduli):
    v = np.ascontiguousarray(values, dtype=np.uint64)
    m = np.ascontiguousarray(moduli, dtype=np.uint64)
    e_cap = max(1, -(-v.size // max(1, m.size)))
    out = np.zeros(e_cap * n, dtype=np.uint32)
    got = C.c_size_t(0)
    st = lib().or_crt_pack(q, n, v if v.size else np.zeros(1, np.uint64), v.size,
                           m if m.size else np.zeros(1, np.uint64), m.size, out, C.byref(got))
    return st, out[: got.value * n].reshape(got.value, n)


def crt_unpack(n, packed, moduli, expected):
    pk = np.ascontiguousarray(packed, dtype=np.uint32).reshape(-1)
    m = np.ascontiguousarray(moduli, dtype=np.uint64)
    out = np.zeros(max(1, expected), dtype=np.uint64)
    st = lib().or_crt_unpack(n, pk if pk.size else np.zeros(1, np.uint32), pk.size // n, m, m.size, expected, out)
    return st, out[:expected]


# ---------------------------------------------------------------------------
# Short ring-element challenges (crates/lattice-core/src/short_challenge.rs)
# ---------------------------------------------------------------------------

SC_FIXED_WEIGHT, SC_BIASED_TERNARY, SC_SMALL_SET, SC_FIXED_WEIGHT_SMALL_SET = 0, 1, 2, 3
SC_REJECTION_BUDGET, SC_INVALID_PARAMETERS, SC_GAMMA_CAP_EXCEEDED = 60, 61, 62


def ceil_sqrt_u128(x: int) -> int:
    return lib().or_ceil_sqrt_u128(x & ((1 << 64) - 1), x >> 64)


def _spec_args(family, weight=0, amplitude=0, permille=0, values=()):
    vals = np.ascontiguousarray(list(values) or [0], dtype=np.int64)
    return family, weight, amplitude, permille, vals, len(values)


def short_challenge_sample(n, family, seed: bytes, weight=0, amplitude=0, permille=0, values=()):
    """ShortChallengeSpec::sample (short_challenge.rs:183-205): (status, coefficients, gamma_c)."""
    fam, w, a, pm, vals, nv = _spec_args(family, weight, amplitude, permille, values)
    out = np.zeros(max(1, n), dtype=np.int64)
    g = C.c_uint64(0)
    sb = (C.c_uint8 * max(1, len(seed))).from_buffer_copy(seed or b"\0")
    st = lib().or_short_challenge_sample(n, fam, w, a, pm, vals.ctypes.data, nv, C.cast(sb, C.c_void_p), len(seed),
                                         out.ctypes.data, C.byref(g))
    return st, out[:n], g.value


def short_challenge_sample_with_gamma_cap(n, family, seed: bytes, cap, retries, **kw):
    """sample_with_gamma_cap (short_challenge.rs:208-236): (status, coefficients, gamma_c)."""
    if cap == 0:
        return SC_INVALID_PARAMETERS, None, 0
    last = 0
    for attempt in range(retries + 1):
        st, c, g = short_challenge_sample(n, family, seed + attempt.to_bytes(8, "little"), **kw)
        if st:
            return st, None, 0
        last = g
        if g <= cap:
            return 0, c, g
    return SC_GAMMA_CAP_EXCEEDED, None, last


def rq2_sample_challenge(q, n, spec_n, family, seed: bytes, weight=0, amplitude=0, permille=0, values=()):
    """sample_rq2_challenge (extension.rs:180-216): (status, [c0 | c1] u32, g0, g1)."""
    fam, w, a, pm, vals, nv = _spec_args(family, weight, amplitude, permille, values)
    out = np.zeros(2 * n, dtype=np.uint32)
    g0, g1 = C.c_uint64(0), C.c_uint64(0)
    sb = (C.c_uint8 * max(1, len(seed))).from_buffer_copy(seed or b"\0")
    st = lib().or_rq2_sample_challenge(q, n, spec_n, fam, w, a, pm, vals.ctypes.data, nv, C.cast(sb, C.c_void_p),
                                       len(seed), out.ctypes.data, C.byref(g0), C.byref(g1))
    return st, out, g0.value, g1.value

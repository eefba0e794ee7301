"""CPU tests of the C ABI: the library loads and exports every entry point
that include/lzx_gpu.h declares, with the Python prototypes in sync.
No compute calls are made (there is no GPU here)."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = (ROOT / "include" / "lzx_gpu.h").read_text()


def declared_functions():
    body = re.sub(r"/\*.*?\*/", "", HEADER, flags=re.S)
    return sorted(set(re.findall(r"\b(lzx_[a-z0-9_]+)\s*\(", body)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ["lzx_ntt_plan_create", "lzx_ntt_forward_batch", "lzx_ntt_inverse_batch",
                 "lzx_ajtai_key_from_seed", "lzx_ajtai_commit_batch", "lzx_ajtai_commit_digits_batch",
                 "lzx_sumcheck_prove", "lzx_transcript_challenge_fields"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from lzx_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("liblzx_gpu.so not built (run __graft_entry__.build())")
    L = C.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert set(declared_functions()) == set(_lib.PROTOTYPES), "ctypes prototypes out of sync with the header"


def test_status_strings_name_reference_enums():
    from lzx_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("liblzx_gpu.so not built")
    L = _lib.lib()
    assert L.lzx_status_string(1) == b"NttError::LengthMismatch"
    assert L.lzx_status_string(4) == b"AjtaiError::DimensionMismatch"
    assert L.lzx_status_string(13) == b"SumcheckError::ClaimMismatch"
    assert L.lzx_status_string(20) == b"TranscriptError::RejectionBudgetExceeded"


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from lzx_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "absent.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.lib()

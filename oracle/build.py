"""Build recipe for the CPU oracle: oracle/liblzx_oracle.so (gcc, plain C).

The oracle is test infrastructure: tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg load it; the product path never does.
The reference itself (Rust) cannot be built in this image, so there is no
oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "lzx_oracle.c"
LIB = HERE / "liblzx_oracle.so"

# x86-64-v2 keeps the library loadable on any host CPU the GPU box may have.
CFLAGS = ["-O3", "-march=x86-64-v2", "-std=gnu11", "-fPIC", "-shared", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def build(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", *CFLAGS, "-o", str(tmp), str(SRC)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)

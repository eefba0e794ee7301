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
SRCS = [HERE / "lzx_oracle.c", HERE / "lzx_baseline.c"]
LIB = HERE / "liblzx_oracle.so"

# x86-64-v2 keeps the library loadable on any host CPU the GPU box may have;
# the AVX-512 baseline kernels are compiled for avx512f per function and run
# only where the CPU reports it (checked at run time, as the reference does).
# OpenMP serves the labelled all-cores baseline column.
CFLAGS = ["-O3", "-march=x86-64-v2", "-std=gnu11", "-fPIC", "-shared", "-fopenmp", "-Wall", "-Wextra",
          "-Wno-unused-parameter"]


def build(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", *CFLAGS, "-I", str(HERE), "-o", str(tmp), *map(str, SRCS)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)

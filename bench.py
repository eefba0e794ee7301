"""Benchmark of the lattice-commitment hot path on one MI355X.

Headline workload (value): batched negacyclic NTT, 4 CRT primes (~31-bit,
each = 1 mod 128) x 2^16 polynomials of degree 64, u32 residues resident in
HBM; one step = forward pass + inverse pass over the whole batch.
GB/s counts the algorithmic bytes of each pass (read + write of every
coefficient): 2 * 4 * 2^16 * 64 * 4 B = 134,217,728 B per pass.

Prints ONE JSON line (rank 0).  --gpus above 1 is refused: this build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Ajtai ring-MACs/s (int8-MFMA vs NTT path, bit-exact) + batched NTT GB/s"
HBM_PEAK_GBS = 8000.0
POLYS = 1 << 16
LOG_N = 6
N = 1 << LOG_N


def _events(torch, stream, k):
    return [torch.cuda.Event(enable_timing=True) for _ in range(k)]


def bench_ntt(steps: int, warmup: int):
    import torch

    from lzx_amd import device as D
    from lzx_amd import ring as R
    from lzx_amd import workloads as W

    moduli = R.CRT_PRIMES
    tabs = [R.NttTables(m, LOG_N) for m in moduli]
    host = W.ntt_batch_input(moduli, POLYS, N)
    data = D.to_device(host)
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        R.forward_multi(tabs, data, POLYS, stream)
        R.inverse_multi(tabs, data, POLYS, stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for e0, e1, e2 in ev:
        e0.record(stream)
        R.forward_multi(tabs, data, POLYS, stream)
        e1.record(stream)
        R.inverse_multi(tabs, data, POLYS, stream)
        e2.record(stream)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = start.elapsed_time(end)
    fwd_ms = statistics.mean(a.elapsed_time(b) for a, b, _ in ev)
    inv_ms = statistics.mean(b.elapsed_time(c) for _, b, c in ev)
    # round trip must restore the input exactly after every step
    ok = bool(np.array_equal(D.to_host(data, np.uint32).reshape(host.shape), host))
    bytes_pass = 2 * len(moduli) * POLYS * N * 4
    return {
        "ms_per_step": total_ms / steps,
        "wall_s": wall,
        "fwd_ms": fwd_ms,
        "inv_ms": inv_ms,
        "bytes_per_pass": bytes_pass,
        "gbs": 2 * bytes_pass / (total_ms / steps) / 1e6,
        "fwd_gbs": bytes_pass / fwd_ms / 1e6,
        "inv_gbs": bytes_pass / inv_ms / 1e6,
        "roundtrip_exact": ok,
    }


def cpu_baseline_ntt(polys: int = POLYS, reps: int = 5):
    """The oracle's scalar forward+inverse (restating ntt.rs:118-216) on one
    core over the same 4 x polys x 64 input; median of reps (harness.rs:24-41)."""
    from lzx_amd import ring as R
    from lzx_amd import workloads as W
    from oracle import oracle as O

    host = W.ntt_batch_input(R.CRT_PRIMES, polys, N)
    ots = [O.Ntt(m.q, m.two_adicity, m.generator, LOG_N) for m in R.CRT_PRIMES]
    L = O.lib()
    times = []
    for r in range(reps + 1):
        buf = host.copy()
        t0 = time.perf_counter()
        for i, ot in enumerate(ots):
            row = np.ascontiguousarray(buf[i].reshape(-1))
            L.or_ntt_forward_batch(ot.h, row, polys)
            L.or_ntt_inverse_batch(ot.h, row, polys)
        dt = time.perf_counter() - t0
        if r:
            times.append(dt)
    med = statistics.median(times)
    bytes_step = 2 * 2 * len(R.CRT_PRIMES) * polys * N * 4
    return {"value": bytes_step / med / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"oracle scalar forward+inverse, 4 primes x {polys} polys x d=64, median of {reps} reps "
                      f"({med:.3f} s each)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the path would shard polynomials/primes with no exchange", file=sys.stderr)
        sys.exit(2)
    import torch

    from lzx_amd import device as D
    D.require_gpu()
    r = bench_ntt(args.steps, args.warmup)
    line = {
        "metric": METRIC,
        "value": round(r["gbs"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (splitmix64-seeded residues)",
        "config": {"workload": "batched negacyclic NTT forward+inverse, 4 CRT primes x 2^16 polys x d=64, u32",
                   "polys_per_prime": POLYS, "degree": N, "primes": [m.q for m in __import__(
                       "lzx_amd.ring", fromlist=["CRT_PRIMES"]).CRT_PRIMES], "roundtrip_exact": r["roundtrip_exact"]},
        "roofline": {"bound": "hbm", "kernel": "ntt_reg_kernel<6,false,FWD>",
                     "achieved": round(r["fwd_gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["fwd_gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                     "fwd_us": round(r["fwd_ms"] * 1e3, 3), "inv_us": round(r["inv_ms"] * 1e3, 3),
                     "inv_gbs": round(r["inv_gbs"], 2)},
        "cpu_baseline": None if args.no_cpu else cpu_baseline_ntt(),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

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
# HBM bytes per forward launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (each scaled by the reading for a 64 MiB copy kernel), per kernel
# class of the plans; the rocprofv3 summaries are kept under profiles/.
NTT_FWD_KERNEL = {"mfma": "ntt_mfma_kernel<NTT_FWD>", "register": "ntt_lazy_kernel<6,NTT_FWD>"}
NTT_FWD_TRAFFIC_BYTES = {"mfma": (66657 + 65536) * 1024, "register": (65812 + 65536) * 1024}
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
    # the timed region: K steps back to back, bracketed by one pair of events
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(steps):
        R.forward_multi(tabs, data, POLYS, stream)
        R.inverse_multi(tabs, data, POLYS, stream)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = start.elapsed_time(end)
    # per-launch durations for the roofline: K forward launches back to back
    # between one event pair, then K inverse launches likewise (each launch is
    # a full transform of the buffer, so K inverses restore the input), queued
    # behind a spin kernel so that the first one does not start on an idle GPU
    fwd_ms, inv_ms = [], []
    for _ in range(3):
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        _queue_ahead(torch, stream)
        e0.record(stream)
        for _ in range(steps):
            R.forward_multi(tabs, data, POLYS, stream)
        e1.record(stream)
        for _ in range(steps):
            R.inverse_multi(tabs, data, POLYS, stream)
        e2.record(stream)
        torch.cuda.synchronize()
        fwd_ms.append(e0.elapsed_time(e1) / steps)
        inv_ms.append(e1.elapsed_time(e2) / steps)
    fwd_ms, inv_ms = statistics.median(fwd_ms), statistics.median(inv_ms)
    # round trip must restore the input exactly after every step
    ok = bool(np.array_equal(D.to_host(data, np.uint32).reshape(host.shape), host))
    bytes_pass = 2 * len(moduli) * POLYS * N * 4
    return {
        "path": R.batch_path(tabs, POLYS),
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


def host_cpu():
    """Model name, logical CPUs, and the OpenMP thread count the all-cores
    columns use (OMP_NUM_THREADS when set: the box's CPU share)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS") or nproc)
    return {"model": model, "nproc": nproc, "threads": max(1, min(threads, nproc))}


def _median_s(fn, reps, warmup=1):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def cpu_baseline_ntt(polys: int = POLYS, reps: int = 5):
    """The oracle's scalar forward+inverse (restating ntt.rs:118-216, which has
    no SIMD) over the same 4 x polys x 64 input: one core, median of reps after
    a warmup (harness.rs:24-41), plus a labelled all-cores column (OpenMP over
    polynomials)."""
    from lzx_amd import ring as R
    from lzx_amd import workloads as W
    from oracle import oracle as O

    cpu = host_cpu()
    host = W.ntt_batch_input(R.CRT_PRIMES, polys, N)
    ots = [O.Ntt(m.q, m.two_adicity, m.generator, LOG_N) for m in R.CRT_PRIMES]
    L = O.lib()
    buf = host.copy()

    def one_core():
        for i, ot in enumerate(ots):
            row = buf[i].reshape(-1)
            L.or_ntt_forward_batch(ot.h, row, polys)
            L.or_ntt_inverse_batch(ot.h, row, polys)

    def all_cores():
        for i, ot in enumerate(ots):
            O.ntt_roundtrip_par(ot, buf[i].reshape(-1), polys, cpu["threads"])

    med = _median_s(one_core, reps)
    med_all = _median_s(all_cores, reps)
    ok = bool(np.array_equal(buf, host))
    bytes_step = 2 * 2 * len(R.CRT_PRIMES) * polys * N * 4
    return {"value": bytes_step / med / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"the full workload: oracle scalar forward+inverse, 4 primes x {polys} polys x d=64, "
                      f"median of {reps} reps ({med:.3f} s each), -O3 -march=x86-64-v2",
            "cpu": cpu["model"], "nproc": cpu["nproc"], "roundtrip_exact": ok,
            "all_cores": {"value": bytes_step / med_all / 1e9, "cores": cpu["threads"],
                          "how": "OpenMP over polynomials (the reference is single-threaded)"}}


def _median_gpu_ms(torch, fn, reps, warmup=1):
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        _queue_ahead(torch, stream)
        a.record(stream)
        fn()
        b.record(stream)
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b))
    return statistics.median(times)


def achievable_copy_gbs(torch, nbytes: int = 134217728, reps: int = 20) -> float:
    """Achievable HBM rate on this GPU: a device-to-device copy of nbytes
    (read + write counted), median of reps, HIP events on the launch stream."""
    a = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream()
    b.copy_(a)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        _queue_ahead(torch, stream)
        e0.record(stream)
        b.copy_(a)
        e1.record(stream)
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1) * 1e-3)
    return 2 * nbytes / statistics.median(ts) / 1e9


def achievable_int8_tops(torch, dim: int = 8192, reps: int = 10):
    """Achievable dense int8 matrix rate of the library GEMM (torch._int_mm on
    dim^3 int8 operands); (TOP/s, note)."""
    try:
        a = torch.randint(-128, 127, (dim, dim), dtype=torch.int8, device="cuda")
        b = torch.randint(-128, 127, (dim, dim), dtype=torch.int8, device="cuda")
        torch._int_mm(a, b)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            _queue_ahead(torch, stream)
            e0.record(stream)
            torch._int_mm(a, b)
            e1.record(stream)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        return 2 * dim ** 3 / statistics.median(ts) / 1e12, f"torch._int_mm {dim}^3"
    except Exception as e:  # the library GEMM is optional here; report why it is missing
        return None, f"torch._int_mm unavailable: {type(e).__name__}"


def bench_ajtai(reps: int = 5):
    """Ajtai commit, d=64, k=16, m=2^16, W=64 small-digit witnesses (digits in
    [-8, 8), P(0) = 17/32), A from seed [0x43; 32], all legs bit-identical:
      - plain_ntt_path: the NTT-domain commit with the multiply-accumulate on
        the VALU (LZX_AJTAI_MAC=valu), the plain drop-in;
      - ntt_path: the same NTT-domain commit with the multiply-accumulate (and
        the digit transform) on int8 MFMA, 2^36 int8 MACs;
      - int8_path: the coefficient-domain Toeplitz product on int8 MFMA, 2^40
        int8 MACs (2.2e12 int8 ops).
    2^26 ring-MACs per commit."""
    import torch

    from lzx_amd import commitment as CM
    from lzx_amd import device as D
    from lzx_amd import ring as R
    from lzx_amd import workloads as W

    k, m, nw = 16, 1 << 16, 64
    t0 = time.perf_counter()
    pk = CM.AjtaiPublicKey.from_seed(CM.AjtaiParams(R.RingConfig(R.Modulus32.Q_32, 6), k, m, 8), bytes([0x43]) * 32)
    key_s = time.perf_counter() - t0
    digits = D.to_device(W.small_digits(0xC3, nw * m * 64).reshape(nw, m, 64))
    o0 = D.empty(nw * k * 64, np.uint32)
    o1 = D.empty(nw * k * 64, np.uint32)
    o2 = D.empty(nw * k * 64, np.uint32)
    prev = os.environ.get("LZX_AJTAI_MAC")
    os.environ["LZX_AJTAI_MAC"] = "valu"  # read by the library on every commit call
    try:
        ms_plain = _median_gpu_ms(torch, lambda: pk.commit_digits_batch(digits, nw, o0, path=CM.COMMIT_NTT), reps)
    finally:
        if prev is None:
            del os.environ["LZX_AJTAI_MAC"]
        else:
            os.environ["LZX_AJTAI_MAC"] = prev
    ms_ntt = _median_gpu_ms(torch, lambda: pk.commit_digits_batch(digits, nw, o1, path=CM.COMMIT_NTT), reps)
    ms_i8 = _median_gpu_ms(torch, lambda: pk.commit_digits_batch(digits, nw, o2, path=CM.COMMIT_INT8), reps)
    same = bool(torch.equal(o1, o2)) and bool(torch.equal(o0, o1))
    o3 = D.empty(nw * k * 64, np.uint32)
    ms_auto = _median_gpu_ms(torch, lambda: pk.commit_digits_batch(digits, nw, o3, path=CM.COMMIT_AUTO), reps)
    auto_choice = {CM.COMMIT_NTT: "ntt_path", CM.COMMIT_INT8: "int8_path"}.get(pk.commit_path_choice(nw), "not measured")
    rmacs = k * m * nw
    i8_ops = 2.0 * rmacs * 64 * 64 * 4
    nm_ops = 2.0 * rmacs * 64 * 16
    return {"workload": "Ajtai commit d=64 k=16 m=2^16 W=64, q=3221225473, digits in [-8,8)",
            "ring_macs_per_commit": rmacs, "key_from_seed_s": round(key_s, 3),
            "plain_ntt_path": {"ms": round(ms_plain, 4), "ring_macs_per_s": rmacs / (ms_plain * 1e-3),
                               "multiply_accumulate": "VALU (LZX_AJTAI_MAC=valu)"},
            "ntt_path": {"ms": round(ms_ntt, 4), "ring_macs_per_s": rmacs / (ms_ntt * 1e-3),
                         "multiply_accumulate": "int8 MFMA, NTT domain (1024 int8 MACs per ring-MAC)",
                         "int8_tops": round(nm_ops / (ms_ntt * 1e-3) / 1e12, 1)},
            "int8_path": {"ms": round(ms_i8, 4), "ring_macs_per_s": rmacs / (ms_i8 * 1e-3),
                          "multiply_accumulate": "int8 MFMA, coefficient domain (16384 int8 MACs per ring-MAC)",
                          "int8_tops": round(i8_ops / (ms_i8 * 1e-3) / 1e12, 1),
                          "int8_mfma_frac_of_5000_tops": round(i8_ops / (ms_i8 * 1e-3) / 5e15, 4)},
            "int8_over_plain_ntt": round(ms_plain / ms_i8, 3), "ntt_mfma_over_plain_ntt": round(ms_plain / ms_ntt, 3),
            "int8_over_ntt": round(ms_ntt / ms_i8, 3), "paths_bit_identical": same,
            "auto": {"ms": round(ms_auto, 4), "choice": auto_choice, "equals_paths": bool(torch.equal(o1, o3)),
                     "faster_path": "int8_path" if ms_i8 <= ms_ntt else "ntt_path"},
            "int8_achievable": _int8_achievable(torch, i8_ops / (ms_i8 * 1e-3) / 1e12)}


def _int8_achievable(torch, tops):
    peak, note = achievable_int8_tops(torch)
    return {"tops": None if peak is None else round(peak, 1), "how": note,
            "int8_path_frac": None if peak is None else round(tops / peak, 4)}


def bench_sumcheck(reps: int = 3):
    """Sum-check prove of eq(r,x)*(a(x)*b(x) - c(x)) over 2^24-entry Goldilocks
    tables with the Fiat-Shamir transcript (run on the device; one host sync).
    Algorithmic bytes with 4 materialised tables: sum_j 4*8*(2^(24-j) + 2^(23-j)).
    The default path replays a captured graph (from the second call on); the
    direct path enqueues kernel by kernel.  The round-0 evaluation and the first
    fold are timed on their own (each call includes a small block-total
    reduction and a readback, so the fractions are lower bounds) against the
    HBM peak and against the measured Goldilocks product rate."""
    import os

    import torch

    from lzx_amd import device as D
    from lzx_amd import sumcheck as S
    from lzx_amd import workloads as W

    nv = 24
    N = 1 << nv
    P = S.GL_P
    vp = S.VirtualPolynomial(nv)
    vp.add_factor(S.eq_table(W.uniform_gl(0xC4E, nv)))
    for seed in (0xC4A, 0xC4B, 0xC4C):
        vp.add_factor(D.to_device(W.uniform_gl(seed, N)))
    vp.add_term(1, [0, 1, 2])
    vp.add_term(P - 1, [0, 3])
    claim = vp.sum_over_hypercube()
    S.prove(vp, claim, S.Transcript(b"lzx-mi355x-sumcheck"))
    torch.cuda.synchronize()
    prove = lambda: S.prove(vp, claim, S.Transcript(b"lzx-mi355x-sumcheck"))
    sec = _median_s(prove, max(reps, 5), warmup=1)
    out = prove()
    os.environ["LZX_SC_GRAPH"] = "0"
    try:
        sec_direct = _median_s(prove, max(reps, 5), warmup=1)
    finally:
        del os.environ["LZX_SC_GRAPH"]
    alg_bytes = sum(4 * 8 * ((1 << (nv - j)) + (1 << (nv - 1 - j))) for j in range(nv))
    # round 0 alone (sc_eval_kernel: 4 tables read, 8 products per pair) and
    # the first fold (sc_fold_eval_kernel: 4 tables read, half written, 8 fold
    # products + 8 point products per pair of the next round)
    t_eval = _median_s(vp.sum_over_hypercube, 5, warmup=1)
    sess_t = []
    for _ in range(3):
        sess = S.SumcheckSession(vp)
        sess.round_eval()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sess.fold(12345)
        sess_t.append(time.perf_counter() - t0)
        del sess
    t_fold = statistics.median(sess_t)
    rate = S.gl_mul_rate(8)
    eval_bytes, fold_bytes = 4 * 8 * N, 4 * 8 * N + 4 * 8 * (N // 2)
    eval_prods, fold_prods = 8 * (N // 2), 8 * (N // 2) + 8 * (N // 4)

    def rl(b, prods, t):
        return {"bound": "hbm", "achieved": round(b / t / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(b / t / 1e9 / HBM_PEAK_GBS, 3), "traffic": None,
                "valu": {"goldilocks_products_per_s": round(prods / t / 1e12, 3),
                         "measured_rate": round(rate / 1e12, 3), "unit": "T/s", "frac": round(prods / t / rate, 3)}}

    # fastprover::prove_fast_with_eq on the same polynomial: eq(r, .) is never
    # materialised (3 streamed tables); the proof is the same, byte for byte
    r = W.uniform_gl(0xC4E, nv)
    vq = S.VirtualPolynomial(nv)
    vq.add_factor(None)
    for f in vp.factors[1:]:
        vq.add_factor(f)
    vq.terms = list(vp.terms)
    fast_call = lambda: S.prove_fast_with_eq(vq, claim, S.Transcript(b"lzx-mi355x-sumcheck"), None, [0], r)
    fast = fast_call()
    torch.cuda.synchronize()
    sec_eq = _median_s(fast_call, max(reps, 5), warmup=1)
    alg_eq = sum(3 * 8 * ((1 << (nv - j)) + (1 << (nv - 1 - j))) for j in range(nv))
    return {"workload": "sum-check prove, 2^24-entry tables, eq*(a*b-c), degree 3, transcript challenges",
            "ms_per_prove": round(sec * 1e3, 3), "path": "captured graph (replayed)",
            "ms_per_prove_direct": round(sec_direct * 1e3, 3), "alg_bytes": alg_bytes,
            "gbs": round(alg_bytes / sec / 1e9, 1), "final_claim": int(out.final_claim),
            "kernels": {"round0_eval": dict(us=round(t_eval * 1e6, 1), **rl(eval_bytes, eval_prods, t_eval)),
                        "first_fold_eval": dict(us=round(t_fold * 1e6, 1), **rl(fold_bytes, fold_prods, t_fold))},
            "split_eq": {"entry": "prove_fast_with_eq (eq never materialised)", "ms_per_prove": round(sec_eq * 1e3, 3),
                         "alg_bytes": alg_eq, "gbs": round(alg_eq / sec_eq / 1e9, 1),
                         "same_proof": bool(fast.rounds.tolist() == out.rounds.tolist()
                                            and fast.final_claim == out.final_claim)}}


def bench_ring50(reps: int = 10):
    """The ~2^50 quadratic-slot ring (q = 2^50 - 2687): 2^16 polynomials of
    n = 128, u64 residues.  Forward transform in place (read + write 64 MiB) and
    the fused ring product (forward both, slotwise Karatsuba, inverse: read
    2 x 64 MiB, write 64 MiB)."""
    import torch

    from lzx_amd import device as D
    from lzx_amd import ring50 as R

    polys, n = 1 << 16, 128
    rng = np.random.default_rng(0x50)
    a = D.to_device(rng.integers(0, R.Q_50, polys * n, dtype=np.uint64))
    b = D.to_device(rng.integers(0, R.Q_50, polys * n, dtype=np.uint64))
    out = D.empty(polys * n, np.uint64)
    T = R.QuadSlotTables(R.Modulus50.Q_50, 7)
    words = polys * n

    t_fwd = _timed_events(torch, lambda: T.forward_batch(a), reps)
    t_mul = _timed_events(torch, lambda: T.mul_batch(a, b, out), reps)
    return {"workload": "quadratic-slot ring q=2^50-2687, 2^16 polys x n=128, u64",
            "forward_us": round(t_fwd * 1e6, 2), "forward_gbs": round(2 * 8 * words / t_fwd / 1e9, 1),
            "mul_us": round(t_mul * 1e6, 2), "mul_gbs": round(3 * 8 * words / t_mul / 1e9, 1),
            "mul_polys_per_s": round(polys / t_mul, 1)}


def bench_ntt_q32(reps: int = 10):
    """The register NTT at the reference's default modulus q = 3*2^30 + 1
    (above 2^31, so the kernels run exact FP64 butterflies): 2^18 polynomials
    of d = 64, forward and inverse per call (134 MB each).  Every forward call
    is matched by an inverse call, so the data must come back unchanged."""
    import torch

    from lzx_amd import device as D
    from lzx_amd import ring as R

    polys = 1 << 18
    rng = np.random.default_rng(0x32)
    host = rng.integers(0, R.Modulus32.Q_32.q, polys * N, dtype=np.uint32)
    data = D.to_device(host)
    T = R.NttTables(R.Modulus32.Q_32, LOG_N)
    t_fwd = _timed_events(torch, lambda: T.forward_batch(data), reps)
    t_inv = _timed_events(torch, lambda: T.inverse_batch(data), reps)
    ok = bool(np.array_equal(D.to_host(data, np.uint32).reshape(-1), host))
    bytes_pass = 2 * polys * N * 4
    return {"workload": "batched negacyclic NTT, q=3221225473, 2^18 polys x d=64, u32",
            "fwd_us": round(t_fwd * 1e6, 2), "fwd_gbs": round(bytes_pass / t_fwd / 1e9, 1),
            "inv_us": round(t_inv * 1e6, 2), "inv_gbs": round(bytes_pass / t_inv / 1e9, 1),
            "roundtrip_exact": ok}


def bench_ntt_large(reps: int = 10):
    """The headline transform on a 4x larger batch (4 CRT primes x 2^18
    polynomials of d = 64, 537 MB per pass): the per-launch ramp-up and drain
    weigh 4x less, so this is the kernels' streaming rate.  Both kernel
    classes are timed on the same data, each pinned at plan creation
    (LZX_NTT_KERNEL): the int8-MFMA transform and the register butterfly
    kernel."""
    import torch

    from lzx_amd import device as D
    from lzx_amd import ring as R
    from lzx_amd import workloads as W

    polys = 1 << 18
    host = W.ntt_batch_input(R.CRT_PRIMES, polys, N)
    data = D.to_device(host)
    bytes_pass = 2 * len(R.CRT_PRIMES) * polys * N * 4
    out = {"workload": "batched negacyclic NTT, 4 CRT primes x 2^18 polys x d=64, u32"}
    prev = os.environ.get("LZX_NTT_KERNEL")
    for name in ("mfma", "register"):
        os.environ["LZX_NTT_KERNEL"] = name  # read at plan creation
        tabs = [R.NttTables(m, LOG_N) for m in R.CRT_PRIMES]
        t_fwd = _timed_events(torch, lambda: R.forward_multi(tabs, data, polys), reps)
        t_inv = _timed_events(torch, lambda: R.inverse_multi(tabs, data, polys), reps)
        out[name] = {"path": R.batch_path(tabs, polys), "fwd_us": round(t_fwd * 1e6, 2),
                     "fwd_gbs": round(bytes_pass / t_fwd / 1e9, 1), "fwd_hbm_frac": round(bytes_pass / t_fwd / 8e12, 4),
                     "inv_us": round(t_inv * 1e6, 2), "inv_gbs": round(bytes_pass / t_inv / 1e9, 1)}
    if prev is None:
        os.environ.pop("LZX_NTT_KERNEL", None)
    else:
        os.environ["LZX_NTT_KERNEL"] = prev
    out["roundtrip_exact"] = bool(np.array_equal(D.to_host(data, np.uint32).reshape(host.shape), host))
    return out


def _queue_ahead(torch, stream):
    """A ~1 ms spin kernel on the stream ahead of a timed region: the host
    queues the timed launches while it runs, so the first timed launch starts
    from a busy queue, as in a pipeline of such calls, not from an idle GPU
    waking up (which added about 2 us per launch to 50 back-to-back 25 us NTT
    launches; with the spin kernel ahead, event times match the kernel trace)."""
    with torch.cuda.stream(stream):
        torch.cuda._sleep(2_000_000)


def _timed_events(torch, fn, reps):
    """Seconds per call of fn: reps back-to-back calls between two HIP events on
    the launch stream (torch's current stream), after one warm-up call and with
    a spin kernel queued ahead (_queue_ahead).  Calls queue behind each other,
    so the host's launch latency is hidden as it is in a pipeline of such
    calls; the median of three such runs."""
    stream = torch.cuda.current_stream()
    fn()
    torch.cuda.synchronize()
    runs = []
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        _queue_ahead(torch, stream)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        runs.append(e0.elapsed_time(e1) * 1e-3 / reps)
    return statistics.median(runs)


def bench_extension(reps: int = 10):
    """The quadratic extension ring R_q[Y]/(Y^2+1), q = 3221225473: 2^17
    elements of n = 64 (u32).  Algorithmic bytes per element: mul reads a, b
    and writes the product (3 x 512 B); norm reads 512 B and writes 256 B;
    inverse reads 512 B and writes 512 B + 1 flag byte."""
    import torch

    from lzx_amd import device as D
    from lzx_amd import extension as E
    from lzx_amd import ring as R

    count, n = 1 << 17, 64
    q = R.Modulus32.Q_32.q
    ring = R.RingConfig(R.Modulus32.Q_32, 6)
    rng = np.random.default_rng(0xE2)
    a = D.to_device(rng.integers(0, q, count * 2 * n, dtype=np.uint64).astype(np.uint32))
    b = D.to_device(rng.integers(0, q, count * 2 * n, dtype=np.uint64).astype(np.uint32))
    out = D.empty(count * 2 * n, np.uint32)
    nrm = D.empty(count * n, np.uint32)
    t_mul = _timed_events(torch, lambda: E.mul_batch(ring, a, b, out), reps)
    t_norm = _timed_events(torch, lambda: E.norm_batch(ring, a, nrm), reps)
    t_inv = _timed_events(torch, lambda: E.inverse_batch(ring, a, out), reps)
    gbs = lambda byts, t: round(byts * count / t / 1e9, 1)
    return {"workload": "R_q[Y]/(Y^2+1), q=3221225473, 2^17 elements x n=64, u32",
            "mul_us": round(t_mul * 1e6, 2), "mul_gbs": gbs(1536, t_mul), "mul_per_s": round(count / t_mul, 1),
            "mul_hbm_frac": round(1536 * count / t_mul / 1e9 / HBM_PEAK_GBS, 4),
            "norm_us": round(t_norm * 1e6, 2), "norm_gbs": gbs(768, t_norm),
            "inverse_us": round(t_inv * 1e6, 2), "inverse_gbs": gbs(1025, t_inv)}


def bench_packing(reps: int = 10):
    """Witness packing, HBM-bound maps: 2^24 Goldilocks values into 22-bit limbs
    (8 B read + 12 B written per value), balanced base-16 digits as i8 (17 per
    value: 8 B read + 17 B written), and 2^30 bits at 31 per coefficient
    (31 B read + 4 B written per coefficient); ring degree 64, q = 3221225473."""
    import torch

    from lzx_amd import device as D
    from lzx_amd import packing as PK
    from lzx_amd import ring as R

    cfg = R.RingConfig(R.Modulus32.Q_32, 6)
    rng = np.random.default_rng(0xB7)
    nv = 1 << 24
    vals = D.to_device(rng.integers(0, 0xFFFFFFFF00000001, nv, dtype=np.uint64))
    t_pack = _timed_events(torch, lambda: PK.pack_field_elements(cfg, vals), reps)
    g = PK.GadgetDecomposition(16, 17)
    t_dec = _timed_events(torch, lambda: g.decompose_batch(vals, digit_bytes=1), reps)
    nbits = 1 << 30
    bits = torch.randint(0, 2, (nbits,), dtype=torch.uint8, device="cuda")
    t_bits = _timed_events(torch, lambda: PK.pack_bits(cfg, bits), reps)
    coeffs = -(-nbits // 31)
    return {"workload": "witness packing, 2^24 values / 2^30 bits, n=64, q=3221225473",
            "pack_field_us": round(t_pack * 1e6, 2), "pack_field_gbs": round(20 * nv / t_pack / 1e9, 1),
            "decompose_i8_us": round(t_dec * 1e6, 2), "decompose_i8_gbs": round(25 * nv / t_dec / 1e9, 1),
            "pack_bits_us": round(t_bits * 1e6, 2), "pack_bits_gbs": round(35 * coeffs / t_bits / 1e9, 1)}


def bench_uneven_batch(reps: int = 5, autotune: bool = True):
    """The uneven commitment batch: 512 jobs (workloads.uneven_batch_jobs(0xC5)),
    n in {64, 128, 256}, q in {3221225473, 2013265921, 1125899437080577},
    k = 16, W = 1, m = floor(2^(8+8u)); one persistent launch per batch,
    replayed from a hipGraph.  Bytes = sum over jobs of key + digits + output;
    the key stream (77 GB) dominates, so the bound is HBM."""
    import torch

    from lzx_amd import commit_batch as CB

    t0 = time.perf_counter()
    wl = CB.uneven_batch_workload()
    setup_s = time.perf_counter() - t0
    b = wl["batch"]
    if autotune:
        b.autotune()
    info = b.info()
    b.run()
    b.check()
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(reps):
        a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        _queue_ahead(torch, stream)
        a.record(stream)
        b.run()
        e.record(stream)
        torch.cuda.synchronize()
        times.append(a.elapsed_time(e))
    b.check()
    ms = statistics.median(times)
    total = wl["key_bytes"] + wl["digit_bytes"] + wl["out_bytes"]
    res = {"workload": "uneven commitment batch: 512 jobs, n in {64,128,256}, q in {Q_32, 2013265921, "
                       "1125899437080577}, k=16, W=1, m=floor(2^(8+8u)); one persistent launch in a hipGraph",
           "ms_per_batch": round(ms, 3), "ring_macs": wl["ring_macs"],
           "ring_macs_per_s": wl["ring_macs"] / (ms * 1e-3), "bytes": total,
           "gbs": round(total / (ms * 1e-3) / 1e9, 1), "hbm_frac": round(total / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "work_items": info["items"], "scratch_bytes": info["scratch_bytes"],
           "chunks_by_class": info["chunks"], "autotuned": autotune, "setup_s": round(setup_s, 2)}
    del wl, b
    torch.cuda.empty_cache()
    return res


def bench_c1(reps: int = 50):
    """The plain Ajtai workload: Q_32, d = 64, k = 8, m = 1024, A from seed
    [0x42; 32], one witness from splitmix64(0xC1) in [-8, 8) (8,192
    ring-MACs).  Times the drop-in host call (lzx_ajtai_commit_host: the
    witness crosses PCIe in and the commitment back, one synchronisation) and
    the device-resident digit commit on each path."""
    import torch

    from lzx_amd import commitment as CM
    from lzx_amd import device as D
    from lzx_amd import ring as R
    from lzx_amd import workloads as W

    k, m = 8, 1024
    pk = CM.AjtaiPublicKey.from_seed(CM.AjtaiParams(R.RingConfig(R.Modulus32.Q_32, 6), k, m, 8), bytes([0x42]) * 32)
    digits = W.small_digits(0xC1, m * 64, zero_half=False).reshape(1, m, 64)
    s = W.digits_to_residues(digits[0], R.Modulus32.Q_32.q)
    t_host = _median_s(lambda: pk.commit(s), reps, warmup=3)
    want = pk.commit(s).rows_array
    dd = D.to_device(digits)
    res = {"workload": "plain Ajtai commit, q=3221225473, d=64, k=8, m=1024, W=1, seed [0x42;32], "
                       "witness splitmix64(0xC1) in [-8,8)",
           "ring_macs": k * m,
           "host_call": {"entry": "lzx_ajtai_commit_host (AjtaiPublicKey::commit signature, PCIe in and out)",
                         "us": round(t_host * 1e6, 2), "ring_macs_per_s": k * m / t_host}}
    prev = os.environ.get("LZX_AJTAI_MAC")
    for name, path in (("plain_ntt_path", CM.COMMIT_NTT), ("ntt_path", CM.COMMIT_NTT), ("int8_path", CM.COMMIT_INT8),
                       ("auto", CM.COMMIT_AUTO)):
        if name == "plain_ntt_path":
            os.environ["LZX_AJTAI_MAC"] = "valu"  # the NTT-domain commit with the VALU multiply-accumulate
        elif prev is None:
            os.environ.pop("LZX_AJTAI_MAC", None)
        else:
            os.environ["LZX_AJTAI_MAC"] = prev
        out = D.empty(k * 64, np.uint32)
        t = _timed_events(torch, lambda: pk.commit_digits_batch(dd, 1, out, path=path), 20)
        same = bool(np.array_equal(D.to_host(out, np.uint32).reshape(k, 64), want))
        res[name] = {"us": round(t * 1e6, 2), "ring_macs_per_s": k * m / t, "equals_host_call": same}
    names = {CM.COMMIT_NTT: "ntt_path", CM.COMMIT_INT8: "int8_path", CM.COMMIT_AUTO: "not measured"}
    res["auto"]["choice"] = names[pk.commit_path_choice(1)]  # measured on AUTO's first call of the shape
    res["auto"]["faster_path"] = min(("ntt_path", "int8_path"), key=lambda p: res[p]["us"])
    return res, (pk, s, want)


def cpu_baseline_secondary(c1=None, reps: int = 5):
    """One-core oracle baselines of the secondary workloads (the reference is
    single-threaded), each with a labelled all-cores column, the CPU model and
    thread count recorded:
      - the plain Ajtai workload at full size (median of reps);
      - the large Ajtai workload by its exact scaling rule: one witness of the
        full key (k = 16, m = 2^16, A^ derived on the GPU and checked equal to
        the oracle's at sampled entries) is timed; a W-witness commit is W
        such independent commits (ajtai.rs:118-153 has no cross-witness
        work), so 64 x the sample is the workload's time;
      - the 2^24 sum-check prove at full size: the reference's AVX-512F
        kernels (its default where the CPU has avx512f) and its scalar
        LZX_NO_SIMD=1 mode, one core each, plus all cores."""
    from lzx_amd import commitment as CM
    from lzx_amd import device as D
    from lzx_amd import ring as R
    from lzx_amd import sumcheck as S
    from lzx_amd import workloads as W
    from oracle import oracle as O

    cpu = host_cpu()
    th = cpu["threads"]
    res = {"cpu": cpu["model"], "nproc": cpu["nproc"], "all_cores_threads": th,
           "avx512f": O.cpu_has_avx512f(), "build": "gcc -O3 -march=x86-64-v2 (+ avx512f kernels per function, "
                                                    "selected at run time by CPUID)"}
    ot = O.Ntt(*O.Q_32, 6)
    # plain Ajtai (C1) at full size
    if c1 is not None:
        pk, s, want = c1
        Ah = pk.matrix_ntt()
        t1 = _median_s(lambda: O.ajtai_commit(ot, 8, 1024, Ah, s), reps)
        ta = _median_s(lambda: O.ajtai_commit_par(ot, 8, 1024, Ah, s, th), reps)
        ok = bool(np.array_equal(O.ajtai_commit(ot, 8, 1024, Ah, s)[1], want))
        res["c1_plain_ajtai"] = {"ring_macs_per_s": 8192 / t1, "us": round(t1 * 1e6, 1), "cores": 1, "kind": "port",
                                 "sample": f"full workload, oracle cached-NTT commit (ajtai.rs:118-153), median of {reps}",
                                 "equals_gpu": ok,
                                 "all_cores": {"ring_macs_per_s": 8192 / ta, "us": round(ta * 1e6, 1), "cores": th}}
    # large Ajtai (C3): one witness of the full key, times 64
    k, m = 16, 1 << 16
    pk = CM.AjtaiPublicKey.from_seed(CM.AjtaiParams(R.RingConfig(R.Modulus32.Q_32, 6), k, m, 8), bytes([0x43]) * 32)
    Ah = pk.matrix_ntt()
    A = pk.matrix()
    for idx in (0, 777777, k * m - 1):
        assert np.array_equal(ot.forward(A[idx])[1], Ah[idx]), "GPU-derived key differs from the oracle"
    del A, pk
    s0 = W.digits_to_residues(W.small_digits(0xC3, m * 64).reshape(m, 64), O.Q_32[0])
    t1 = _median_s(lambda: O.ajtai_commit(ot, k, m, Ah, s0), 3, warmup=0)
    ta = _median_s(lambda: O.ajtai_commit_par(ot, k, m, Ah, s0, th), 3, warmup=0)
    res["c3_ajtai"] = {"ring_macs_per_s": k * m / t1, "s_per_witness": round(t1, 3),
                       "s_per_c3_commit": round(64 * t1, 2), "cores": 1, "kind": "port",
                       "sample": "witness 0 of the full key (k=16, m=2^16), median of 3; C3 = 64 independent witnesses",
                       "all_cores": {"ring_macs_per_s": k * m / ta, "s_per_c3_commit": round(64 * ta, 2),
                                     "cores": th}}
    del Ah
    # sum-check (C4) at full size
    nv = 24
    r = W.uniform_gl(0xC4E, nv)
    facs = [D.to_host(S.eq_table(r), np.uint64)] + [W.uniform_gl(x, 1 << nv) for x in (0xC4A, 0xC4B, 0xC4C)]
    terms = [(1, [0, 1, 2]), (O.GL_P - 1, [0, 3])]
    vp = S.VirtualPolynomial(nv)
    for f in facs:
        vp.add_factor(D.to_device(f))
    for c, ids in terms:
        vp.add_term(c, ids)
    claim = vp.sum_over_hypercube()
    gpu = S.prove(vp, claim, S.Transcript(b"lzx-mi355x-sumcheck"))
    del vp
    out = {}

    def run(simd, threads):
        st, o = O.sumcheck_prove(nv, facs, terms, claim, O.Transcript(b"lzx-mi355x-sumcheck"), simd=simd,
                                 threads=threads)
        assert st == 0, st
        out["same"] = bool(np.array_equal(o["rounds"], gpu.rounds) and o["final_claim"] == gpu.final_claim)

    sc = {"workload": "2^24 tables, eq*(a*b-c), degree 3, transcript challenges (full size)", "kind": "port"}
    if O.cpu_has_avx512f():
        t = _median_s(lambda: run(1, 1), 3, warmup=0)
        sc["avx512_1core"] = {"ms_per_prove": round(t * 1e3, 1), "cores": 1, "same_proof_as_gpu": out["same"]}
        t = _median_s(lambda: run(1, th), 3, warmup=0)
        sc["avx512_all_cores"] = {"ms_per_prove": round(t * 1e3, 1), "cores": th, "same_proof_as_gpu": out["same"]}
    else:
        sc["avx512_1core"] = None
        sc["avx512_note"] = "this CPU reports no avx512f; the reference would run its scalar path"
    t = _median_s(lambda: run(0, 1), 1, warmup=0)
    sc["scalar_1core"] = {"ms_per_prove": round(t * 1e3, 1), "cores": 1, "same_proof_as_gpu": out["same"],
                          "note": "the reference's LZX_NO_SIMD=1 mode; one run"}
    res["c4_sumcheck"] = sc
    return res


def _copy_calibration(torch):
    """A plain 64 MiB device copy of known size, for scaling the HBM counters."""
    src = torch.zeros(1 << 24, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)  # 64 MiB read + 64 MiB write per dispatch
    torch.cuda.synchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only-ntt", action="store_true", help="skip the secondary workloads (profiling runs)")
    ap.add_argument("--only-batch", action="store_true", help="run only the uneven commitment batch leg")
    ap.add_argument("--only-sumcheck", action="store_true", help="run only the 2^24 sum-check prove leg")
    ap.add_argument("--only-c1", action="store_true", help="run only the plain Ajtai (k=8, m=1024) leg")
    ap.add_argument("--only-cpu", action="store_true", help="run only the CPU baseline legs")
    ap.add_argument("--only-ring50", action="store_true", help="run only the 50-bit quadratic-slot ring leg")
    ap.add_argument("--only-ntt-q32", action="store_true", help="run only the default-modulus (q > 2^31) NTT leg")
    ap.add_argument("--only-ajtai", action="store_true", help="run only the large Ajtai commit leg (both paths)")
    ap.add_argument("--only-ntt-large", action="store_true",
                    help="run only the 4 x 2^18-polynomial NTT leg (both kernel classes)")
    ap.add_argument("--only-extension", action="store_true", help="run only the R_q[Y]/(Y^2+1) leg")
    ap.add_argument("--only-packing", action="store_true", help="run only the witness-packing leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batch leg: use cached (LZX_AUTOTUNE_CACHE) or default chunks instead of tuning")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a plain 64 MiB device copy (counter calibration for --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the path would shard polynomials/primes with no exchange", file=sys.stderr)
        sys.exit(2)
    import torch

    from lzx_amd import device as D
    D.require_gpu()
    if args.only_ring50:
        print(json.dumps({"ring50": bench_ring50()}), flush=True)
        return
    if args.only_ntt_q32:
        print(json.dumps({"ntt_q32": bench_ntt_q32()}), flush=True)
        return
    if args.only_ajtai:
        print(json.dumps({"ajtai_commit": bench_ajtai()}), flush=True)
        return
    if args.only_ntt_large:
        print(json.dumps({"ntt_large": bench_ntt_large()}), flush=True)
        return
    if args.only_extension:
        print(json.dumps({"extension": bench_extension()}), flush=True)
        return
    if args.only_packing:
        print(json.dumps({"packing": bench_packing()}), flush=True)
        return
    if args.only_c1:
        print(json.dumps({"c1_plain_ajtai": bench_c1()[0]}), flush=True)
        return
    if args.only_cpu:
        print(json.dumps({"cpu_baseline": cpu_baseline_ntt(), "secondary": cpu_baseline_secondary(bench_c1()[1])}),
              flush=True)
        return
    if args.only_sumcheck:
        print(json.dumps({"sumcheck_prove": bench_sumcheck(reps=5)}), flush=True)
        return
    if args.only_batch:
        res = bench_uneven_batch(autotune=not args.no_autotune)
        if args.copy_calib:
            _copy_calibration(torch)
        print(json.dumps({"uneven_batch": res}), flush=True)
        return
    r = bench_ntt(args.steps, args.warmup)
    if args.copy_calib:
        _copy_calibration(torch)
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
        "roofline": {"bound": "hbm", "kernel": NTT_FWD_KERNEL[r["path"]],
                     "achieved": round(r["fwd_gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["fwd_gbs"] / HBM_PEAK_GBS, 4), "traffic": NTT_FWD_TRAFFIC_BYTES[r["path"]],
                     "alg_bytes": r["bytes_per_pass"],
                     "fwd_us": round(r["fwd_ms"] * 1e3, 3), "inv_us": round(r["inv_ms"] * 1e3, 3),
                     "inv_gbs": round(r["inv_gbs"], 2)},
        "cpu_baseline": None if args.no_cpu else cpu_baseline_ntt(),
    }
    copy_gbs = achievable_copy_gbs(torch)
    line["roofline"]["achievable_peak"] = {"value": round(copy_gbs, 1), "unit": "GB/s",
                                           "how": "device-to-device copy of 128 MiB (read + write)",
                                           "frac": round(r["fwd_gbs"] / copy_gbs, 4)}
    if not args.only_ntt:
        c1, c1_inputs = bench_c1()
        line["secondary"] = {"c1_plain_ajtai": c1, "ajtai_commit": bench_ajtai(), "sumcheck_prove": bench_sumcheck(),
                             "uneven_batch": bench_uneven_batch(), "ring50": bench_ring50(),
                             "extension": bench_extension(), "packing": bench_packing(),
                             "ntt_q32": bench_ntt_q32(), "ntt_large": bench_ntt_large()}
        if not args.no_cpu:
            line["secondary"]["cpu_baseline"] = cpu_baseline_secondary(c1_inputs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

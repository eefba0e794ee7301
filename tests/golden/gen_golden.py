"""Write tests/golden/*.json: the known answers the oracle is pinned against.

Two kinds of fixture, both data only:
  reference_kats.json  known-answer values transcribed from the reference's
                       own tests (file:line cited per entry);
  vectors.json         input/output vectors computed HERE from first
                       principles, independently of the C oracle: Python big
                       integers for the modular arithmetic and closed forms,
                       hashlib (FIPS 202) for SHAKE-256, and the reference's
                       counter-indexed derivation rule
                       (crates/lattice-ring/tests/uniform_check.rs:6-36).

Run:  python tests/golden/gen_golden.py
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

HERE = Path(__file__).resolve().parent
P = 0xFFFFFFFF00000001

MODULI = {  # q: (two_adicity, generator)
    3221225473: (30, 5), 12289: (12, 11), 2013265921: (27, 31),
    1811939329: (26, 13), 2113929217: (25, 5), 998244353: (23, 3),
}


def sm64(seed):
    s = seed & (2**64 - 1)
    while True:
        s = (s + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        yield z ^ (z >> 31)


def xof(domain: bytes, msg: bytes, n: int) -> bytes:
    """SHAKE256("LZXX" || u32le|d| || d || u32le|m| || m) (transcript.rs:229-237)."""
    buf = b"LZXX" + len(domain).to_bytes(4, "little") + domain + len(msg).to_bytes(4, "little") + msg
    return hashlib.shake_256(buf).digest(n)


def uniform_counter(q, log_n, domain, seed, index):
    """uniform_check.rs:6-36: coefficient windows at byte 8 + 4*counter."""
    n = 1 << log_n
    salt = domain + seed + index.to_bytes(8, "little")
    limit = 2**32 - (2**32 % q)
    stream = xof(b"uniform", salt, 8 + 4 * (4 * n + 64))
    out, c = [], 0
    while len(out) < n:
        cand = int.from_bytes(stream[8 + 4 * c: 12 + 4 * c], "little")
        if cand < limit:
            out.append(cand % q)
        c += 1
    return out


def small_secret(q, log_n, m, bound, seed):
    """ajtai.rs:307-348."""
    n = 1 << log_n
    span = 2 * bound + 1
    limit = ((2**32 - 1) // span) * span
    res = []
    for i in range(m):
        coeffs, ctr = [], 0
        while len(coeffs) < n:
            inp = seed + i.to_bytes(8, "little") + ctr.to_bytes(8, "little")
            by = xof(b"ajtai-secret", inp, max(n * 16, 64))
            for c in range(0, len(by) - 3, 4):
                if len(coeffs) == n:
                    break
                v = int.from_bytes(by[c:c + 4], "little")
                if v < limit:
                    coeffs.append(((v % span) - bound) % q)
            ctr += 1
        res.append(coeffs)
    return res


def closed_ntt(a, q, g, log_n):
    n = 1 << log_n
    psi = pow(g, (q - 1) >> (log_n + 1), q)
    return [sum(a[i] * pow(psi, i * (2 * k + 1), q) for i in range(n)) % q for k in range(n)]


def negacyclic(a, b, q):
    n = len(a)
    out = [0] * n
    for i in range(n):
        if a[i] == 0:
            continue
        for j in range(n):
            if i + j < n:
                out[i + j] += a[i] * b[j]
            else:
                out[i + j - n] -= a[i] * b[j]
    return [x % q for x in out]


def mle_eval(table, point):
    """Multilinear extension at `point`, variable 0 = most significant bit."""
    m = len(point)
    total = 0
    for b in range(1 << m):
        w = 1
        for i in range(m):
            bit = (b >> (m - 1 - i)) & 1
            w = w * (point[i] if bit else (1 - point[i])) % P
        total += w * table[b]
    return total % P


def sumcheck_brute(num_vars, factors, terms, challenges):
    """g_j(t) = sum_x P(r_0..r_{j-1}, t, x) by hypercube enumeration."""
    d = max(len(fs) for _, fs in terms)
    rounds = []
    for j in range(num_vars):
        row = []
        rest = num_vars - j - 1
        for t in range(d + 1):
            acc = 0
            for x in range(1 << rest):
                xb = [(x >> (rest - 1 - i)) & 1 for i in range(rest)]
                pt = list(challenges[:j]) + [t] + xb
                vals = [mle_eval(f, pt) for f in factors]
                for c, fs in terms:
                    prod = c
                    for fi in fs:
                        prod = prod * vals[fi] % P
                    acc += prod
            row.append(acc % P)
        rounds.append(row)
    claims = [mle_eval(f, challenges) for f in factors]
    return rounds, claims


def reference_kats():
    return {
        "goldilocks_sub": {"source": "lattice-core/src/field.rs:270-284", "cases": [
            [1, 18403641651327692187, 43102418086892135], [5, 18446744069414584320, 6],
            [0, 1, 18446744069414584320], [18446744069414584320, 18446744069414584320, 0],
            [1234567890123456789, 9876543210987654321, 9804768748550386789],
            [1, 877746185904530010, 17568997883510054312]]},
        "goldilocks_add": {"source": "lattice-core/src/field.rs:285-303", "cases": [
            [1, 18403641651327692187, 18403641651327692188], [5, 18446744069414584320, 4], [0, 1, 1],
            [18446744069414584320, 18446744069414584320, 18446744069414584319],
            [1234567890123456789, 9876543210987654321, 11111111101111111110],
            [1, 877746185904530010, 877746185904530011]]},
        "goldilocks_mul": {"source": "lattice-core/src/field.rs:304-312", "cases": [
            [1, 18403641651327692187, 18403641651327692187], [5, 18446744069414584320, 18446744069414584316],
            [0, 1, 0], [18446744069414584320, 18446744069414584320, 1]]},
        "sha3_256": {"source": "lattice-core/src/keccak.rs:416-434", "cases": [
            ["", "a7ffc6f8bf1ed76651c14756a061d662f580ff4de43b49fa82d80a4b80f8434a"],
            ["616263", "3a985da74fe225b2045c172d6bd390bd855f086e3e9d525b46bfe24511431532"],
            ["00" * 200, "2b43036c229ba512995f91fdb46fcd5327a4dc834d86d6e0f58a08053346dc2e"]]},
        "shake256_32": {"source": "lattice-core/src/keccak.rs:436-443", "cases": [
            ["", "46b9dd2b0ba88d13233b3feb743eeb243fcd52ea62b81b82b50c27646ed5762f"]]},
        "reduce_i64_q32": {"source": "lattice-ring/src/modulus.rs:262-264", "q": 3221225473, "cases": [
            [-1, 3221225472], [-3221225473, 0], [-3221225468, 5]]},
        "lagrange": {"source": "lattice-sumcheck/src/sumcheck.rs:503-509",
                     "evals": [1, 3, 7], "r": 5, "value": 31},
        "ntt_rejects": {"source": "lattice-ring/src/ntt.rs:470-479",
                        "q": 12289, "two_adicity": 12, "ok_log_n": 11, "bad_log_n": 12},
        "ajtai_shapes": {"source": "lattice-commitment/src/ajtai.rs:493-498",
                         "cases": [[4, 2, 3, 256, 255], [5, 2, 8, 8388608, 8], [6, 3, 12, 64, 3], [6, 1, 16, 1024, 1]]},
    }


def vectors():
    out = {"ntt": [], "uniform": [], "small_secret": [], "ajtai": [], "sumcheck": []}
    rng = sm64(0x601D)
    for q, (adic, g) in MODULI.items():
        for log_n in (1, 3, 6):
            if log_n + 1 > adic:
                continue
            a = [next(rng) % q for _ in range(1 << log_n)]
            out["ntt"].append({"q": q, "two_adicity": adic, "generator": g, "log_n": log_n, "input": a,
                               "psi": pow(g, (q - 1) >> (log_n + 1), q), "forward": closed_ntt(a, q, g, log_n)})
    # uniform_check.rs:62-81 grid (Q_32, "ajtai-A")
    for log_n in (3, 5, 6):
        for idx in (0, 1, 7, 255, 4096, 1 << 20):
            for tag in (0, 1, 42, 255):
                out["uniform"].append({"q": 3221225473, "log_n": log_n, "domain": "ajtai-A", "seed_byte": tag,
                                       "index": idx,
                                       "coeffs": uniform_counter(3221225473, log_n, b"ajtai-A", bytes([tag]) * 32, idx)})
    for bound, seed in ((8, b"secret"), (256, b"diff"), (1, b"x")):
        out["small_secret"].append({"q": 3221225473, "log_n": 4, "m": 3, "bound": bound, "seed": seed.decode(),
                                    "coeffs": small_secret(3221225473, 4, 3, bound, seed)})
    # cached == naive shapes of ajtai.rs:493-498 with key seed [9;32] and secret b"diff"
    q = 3221225473
    for log_n, k, m, bound, density in reference_kats()["ajtai_shapes"]["cases"]:
        A = [uniform_counter(q, log_n, b"ajtai-A", bytes([9]) * 32, idx) for idx in range(k * m)]
        s = small_secret(q, log_n, m, density, b"diff")
        t = []
        for i in range(k):
            acc = [0] * (1 << log_n)
            for j in range(m):
                prod = negacyclic(A[i * m + j], s[j], q)
                acc = [(x + y) % q for x, y in zip(acc, prod)]
            t.append(acc)
        out["ajtai"].append({"q": q, "log_n": log_n, "k": k, "m": m, "norm_bound": bound, "density": density,
                             "key_seed_byte": 9, "secret_seed": "diff", "commitment": t})
    # sum-check: a 3-factor polynomial shaped like sumcheck.rs:382-400, fixed challenges
    for num_vars in (1, 2, 4):
        rs = sm64(0x5C + num_vars)
        factors = [[next(rs) % P for _ in range(1 << num_vars)] for _ in range(3)]
        terms = [[3, [0, 1]], [5, [1, 2, 0]], [11, [2]]]
        chal = [next(rs) % P for _ in range(num_vars)]
        rounds, claims = sumcheck_brute(num_vars, factors, terms, chal)
        claim = sum(sum(c * __import__("math").prod(factors[f][x] for f in fs) for c, fs in terms)
                    for x in range(1 << num_vars)) % P
        out["sumcheck"].append({"num_vars": num_vars, "factors": factors, "terms": terms, "challenges": chal,
                                "claim": claim, "rounds": rounds, "factor_claims": claims})
    out["c1"] = c1_vector()
    return out


def c1_vector():
    """The plain Ajtai workload: Q_32, d = 64, k = 8, m = 1024, one witness.
    A[idx] = uniform_from_seed("ajtai-A", [0x42; 32], idx) by the counter
    rule; witness coefficient c (row-major [j][c]) = (x_c >> 60) - 8 for the
    splitmix64(0xC1) outputs x_0, x_1, ...; t_i = sum_j A_ij * s_j in
    Z_q[X]/(X^64 + 1) by exact negacyclic schoolbook in int64 (|A s| terms
    < 2^32 * 8, at most 2^16 of them per output: < 2^51)."""
    import numpy as np
    q, log_n, k, m = 3221225473, 6, 8, 1024
    n = 1 << log_n
    A = np.array([uniform_counter(q, log_n, b"ajtai-A", bytes([0x42]) * 32, idx) for idx in range(k * m)],
                 dtype=np.int64).reshape(k, m, n)
    rs = sm64(0xC1)
    s = np.array([(next(rs) >> 60) - 8 for _ in range(m * n)], dtype=np.int64).reshape(m, n)
    # negacyclic Toeplitz of s_j: T[j][r][c] = s_j[r - c] (c <= r), -s_j[r - c + n] (c > r)
    r_idx, c_idx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    T = np.where(c_idx <= r_idx, s[:, (r_idx - c_idx) % n], -s[:, (r_idx - c_idx) % n])
    t = np.einsum("jrc,ijc->ir", T, A)
    assert np.abs(t).max() < 2**62
    return {"q": q, "log_n": log_n, "k": k, "m": m, "key_seed_byte": 0x42, "witness_seed": 0xC1,
            "witness": s.reshape(-1).tolist(), "commitment": np.mod(t, q).tolist()}


def main():
    (HERE / "reference_kats.json").write_text(json.dumps(reference_kats(), indent=1) + "\n")
    (HERE / "vectors.json").write_text(json.dumps(vectors()) + "\n")
    print("wrote", HERE / "reference_kats.json", HERE / "vectors.json")


if __name__ == "__main__":
    main()

"""GPU parity of the uneven commitment batch (one persistent launch over jobs
of mixed ring degree n in {64, 128, 256}, modulus and length).

Every job must equal AjtaiPublicKey::commit semantics
(crates/lattice-commitment/src/ajtai.rs:118-153): t_i = sum_j A_ij s_j in
Z_q[X]/(X^n + 1), checked against the oracle's exact schoolbook for small
jobs and by evaluation at roots of X^n + 1 for large ones.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    from lzx_amd import commit_batch as CB
    from lzx_amd import device as D
    from lzx_amd import workloads as W
    D.require_gpu()
    return CB, D, W


def _rand_A(rng, q, k, m, n):
    dt = np.uint32 if q < (1 << 32) else np.uint64
    return (rng.integers(0, q, size=(k, m, n), dtype=np.uint64)).astype(dt)


def _roots(q, g, n, count=2):
    psi = O.powmod64(g, (q - 1) // (2 * n), q)
    return [O.powmod64(psi, 2 * i + 1, q) for i in range(count)]


def _run(CB, D, specs, rng, k=16):
    """specs: list of (q, g, log_n, key_m, job_ms).  Returns per job (A, digits, m, t, q, g, n)."""
    keys, jobs, meta = [], [], []
    for q, g, log_n, key_m, job_ms in specs:
        n = 1 << log_n
        A = _rand_A(rng, q, k, key_m, n)
        key = CB.WideKey.from_host(q, g, A)
        keys.append(key)
        for m in job_ms:
            d = rng.integers(-8, 8, size=(max(m, 1), n)).astype(np.int8)
            d[rng.random(d.shape[0]) < 0.3] = 0  # zero ring elements (ajtai.rs:130-137 zero-skip)
            dd = D.to_device(d)
            out = D.empty(k * n, key.dtype)
            jobs.append((key, dd, m, out))
            meta.append((A, d, m, out, q, g, n, key.dtype))
    batch = CB.CommitBatch(jobs)
    return keys, batch, jobs, meta


def _check_exact(meta):
    from lzx_amd import device as D
    for A, d, m, out, q, g, n, dt in meta:
        got = D.to_host(out, dt).reshape(A.shape[0], n).astype(np.uint64)
        want = O.commit_wide_schoolbook(q, A, d[:m], m=m)
        assert np.array_equal(got, want), (q, n, m)


ALL_CLASSES = [
    (2013265921, 31, 6, 40, [0, 1, 17, 40]),
    (2013265921, 31, 7, 33, [33, 5]),
    (2013265921, 31, 8, 20, [20, 3]),
    (3221225473, 5, 6, 37, [37, 2]),
    (3221225473, 5, 7, 25, [25]),
    (3221225473, 5, 8, 18, [18, 7]),
    (1125899437080577, 5, 6, 36, [36, 1]),
    (1125899437080577, 5, 7, 30, [30, 9]),
    (1125899437080577, 5, 8, 19, [19, 2]),
]


def test_every_class_matches_schoolbook(env):
    CB, D, W = env
    rng = np.random.default_rng(0xB0)
    keys, batch, jobs, meta = _run(CB, D, ALL_CLASSES, rng)
    info = batch.info()
    assert info["items"] > 0 and all(c > 0 for c in info["chunks"])
    batch.run()
    batch.check()
    _check_exact(meta)


@pytest.mark.parametrize("chunk", ["1", "3", "16", "4096"])
def test_chunk_sizes_and_ragged_items(env, monkeypatch, chunk):
    """Work items of any size, including ones that end mid sub-batch."""
    CB, D, W = env
    monkeypatch.setenv("LZX_BATCH_CHUNK", chunk)
    rng = np.random.default_rng(int(chunk))
    specs = [(2013265921, 31, 6, 150, [150, 149]), (1125899437080577, 5, 8, 70, [70, 33]),
             (3221225473, 5, 7, 90, [90])]
    keys, batch, jobs, meta = _run(CB, D, specs, rng, k=5)
    assert set(c for c in batch.info()["chunks"] if c) == {int(chunk)}
    batch.run()
    batch.check()
    _check_exact(meta)


def test_extreme_digits_and_unreduced_key(env):
    """|digit| up to 128 and key entries given as non-canonical representatives."""
    CB, D, W = env
    rng = np.random.default_rng(5)
    for q, g, log_n, dt in [(2013265921, 31, 6, np.uint32), (3221225473, 5, 7, np.uint32),
                            (1125899437080577, 5, 8, np.uint64)]:
        n, k, m = 1 << log_n, 3, 9
        hi = (1 << 32) if dt == np.uint32 else (1 << 62)
        A = rng.integers(0, hi, size=(k, m, n), dtype=np.uint64).astype(dt)
        key = CB.WideKey.from_host(q, g, A)
        d = rng.choice(np.array([-128, 127, -1, 0, 1], dtype=np.int8), size=(m, n))
        out = D.empty(k * n, key.dtype)
        b = CB.CommitBatch([(key, D.to_device(d), m, out)])
        b.run()
        b.check()
        got = D.to_host(out, key.dtype).reshape(k, n).astype(np.uint64)
        assert np.array_equal(got, O.commit_wide_schoolbook(q, A, d))


def test_graph_replays_with_eager_work_between(env):
    CB, D, W = env
    import torch
    rng = np.random.default_rng(7)
    keys, batch, jobs, meta = _run(CB, D, ALL_CLASSES[::2], rng, k=4)
    batch.run()
    batch.check()
    first = [D.to_host(j[3], m[7]).copy() for j, m in zip(jobs, meta)]
    for rep in range(3):
        for j in jobs:
            j[3].zero_()
        x = torch.randn(1 << 20, device="cuda")
        (x * 2).sum().item()  # unrelated eager work and a synchronization between replays
        batch.run()
        batch.check()
        for j, m, f in zip(jobs, meta, first):
            assert np.array_equal(D.to_host(j[3], m[7]), f), rep
    _check_exact(meta)


def test_autotune_cache_roundtrip(env, monkeypatch, tmp_path):
    CB, D, W = env
    path = tmp_path / "autotune.txt"
    monkeypatch.setenv("LZX_AUTOTUNE_CACHE", str(path))
    monkeypatch.delenv("LZX_BATCH_CHUNK", raising=False)
    rng = np.random.default_rng(9)
    specs = [(2013265921, 31, 6, 600, [600, 300]), (1125899437080577, 5, 8, 200, [200])]
    keys, batch, jobs, meta = _run(CB, D, specs, rng)
    batch.autotune()
    chunks = batch.info()["chunks"]
    text = path.read_text()
    assert "commit_batch kind=0 log_n=6 k=16" in text and "commit_batch kind=2 log_n=8 k=16" in text
    for line in text.splitlines():  # the cache file also keeps other modules' choices
        key, val = line.split("\t")
        if not key.startswith("commit_batch "):
            continue
        kind, log_n = int(key.split("kind=")[1].split()[0]), int(key.split("log_n=")[1].split()[0])
        assert chunks[3 * kind + log_n - 6] == int(val)
    batch.run()
    batch.check()
    _check_exact(meta)
    # a new batch of the same classes picks the cached choice up
    b2 = CB.CommitBatch(jobs)
    assert b2.info()["chunks"] == chunks


def test_large_jobs_by_evaluation(env):
    """m up to 2^16 - 1 (the batch's largest): t_i(z) = sum_j A_ij(z) s_j(z) at roots z of X^n + 1."""
    CB, D, W = env
    rng = np.random.default_rng(11)
    specs = [(2013265921, 31, 6, 65535, [65535, 40000]), (1125899437080577, 5, 8, 9000, [9000]),
             (3221225473, 5, 7, 20000, [20000])]
    keys, batch, jobs, meta = _run(CB, D, specs, rng, k=16)
    batch.run()
    batch.check()
    for A, d, m, out, q, g, n, dt in meta:
        t = D.to_host(out, dt).reshape(A.shape[0], n)
        for z in _roots(q, g, n):
            assert O.commit_wide_check_at(q, A, d, t, z, m=m) == 0, (q, n, m)


def test_error_paths(env):
    CB, D, W = env
    from lzx_amd._lib import LzxError
    rng = np.random.default_rng(3)
    A = _rand_A(rng, 2013265921, 2, 4, 64)
    key = CB.WideKey.from_host(2013265921, 31, A)
    d = D.to_device(np.zeros((8, 64), np.int8))
    out = D.empty(2 * 64, np.uint32)
    with pytest.raises(LzxError) as e:
        CB.CommitBatch([(key, d, 8, out)])  # m beyond the key's columns
    assert e.value.code == 4
    with pytest.raises(LzxError) as e:
        CB.WideKey.from_host(2013265921, 31, A[:, :, :32].copy())  # n = 32 unsupported
    assert e.value.code == 30
    with pytest.raises(LzxError) as e:  # 2n = 512 does not divide 3329 - 1 = 2^8 * 13
        CB.WideKey.from_host(3329, 3, _rand_A(rng, 3329, 1, 1, 256))
    assert e.value.code == 2
    with pytest.raises(LzxError) as e:  # generator 1: psi^n = 1, not -1
        CB.WideKey.from_host(2013265921, 1, A)
    assert e.value.code == 30
    with pytest.raises(LzxError) as e:  # beyond 2^50
        CB.WideKey.from_host((1 << 50) + 1, 3, _rand_A(rng, 1 << 40, 1, 1, 64))
    assert e.value.code == 30
    with pytest.raises(LzxError) as e:  # digits pointer not 16-byte aligned
        CB.CommitBatch([(key, d[0][1:], 1, out)])
    assert e.value.code == 30


def test_full_uneven_batch_all_jobs(env):
    """The 512-job batch at full size.  Every job is checked by evaluation at
    a root z of X^n + 1 against its class key's evaluations A_ij(z), computed
    once per class; sampled jobs of every modulus at a second root; the
    smallest job of each class exactly by schoolbook."""
    CB, D, W = env
    wl = CB.uneven_batch_workload()
    b = wl["batch"]
    b.run()
    b.check()
    spec = wl["spec"]
    by_class = {}
    for i, (m, log_n, (q, g)) in enumerate(spec):
        by_class.setdefault(CB.class_of(q, log_n), []).append(i)
    checked = 0
    for c, idx in sorted(by_class.items()):
        m0, log_n, (q, g) = spec[idx[0]]
        n = 1 << log_n
        key = wl["jobs"][idx[0]][0]
        A = D.to_host(wl["A"][c], key.dtype)  # (16, class m, n)
        z1, z2 = _roots(q, g, n)
        Az1 = O.eval_key_at(q, A, z1)
        idx.sort(key=lambda i: spec[i][0])
        for pos, i in enumerate(idx):
            m = spec[i][0]
            _, d, _, out = wl["jobs"][i]
            dig = d.cpu().numpy()
            t = D.to_host(out, key.dtype).reshape(16, n)
            assert O.commit_check_at_evals(q, Az1, dig, t, z1, m=m) == 0, (i, m, n, q)
            if pos in (0, len(idx) // 2, len(idx) - 1):
                assert O.commit_wide_check_at(q, A, dig, t, z2, m=m) == 0, (i, m, n, q)
            if pos == 0:
                want = O.commit_wide_schoolbook(q, np.ascontiguousarray(A[:, :m]), dig)
                assert np.array_equal(t.astype(np.uint64), want), i
            checked += 1
        del A
    assert checked == 512

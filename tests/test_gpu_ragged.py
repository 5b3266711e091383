"""Ragged batches: many transforms of uneven lengths and kinds in one
persistent launch.  Every job is checked against the oracle (and the full
4096-job workload against the reference engine), the persistent launch
against the per-transform calls, and graph replays against the eager result
(several replays with eager work and synchronisation in between).  Bar:
rel-L2 <= 1e-6 (complex64/float32), <= 1e-13 (complex128/float64)."""
import numpy as np
import pytest
import torch

import rocket_fft_amd as R
from helpers import REF, oracle_out, rel_l2, tol_of

pytestmark = pytest.mark.gpu

LENGTHS = [64, 100, 127, 256, 1000, 1536, 2187, 4096, 4099, 4913, 8192, 9973, 12000,
           20000, 65536, 65537, 3, 2, 1, 17]


def make_jobs(dtype_c, lengths, rs):
    dtype_r = np.float32 if dtype_c == np.complex64 else np.float64
    jobs, host = [], []
    for i, n in enumerate(lengths):
        kind = ("c2c", "r2c", "c2r")[i % 3]
        fwd = bool(i % 2)
        fct = 1.0 if i % 4 else 0.5
        if kind == "c2c":
            x = (rs.standard_normal(n) + 1j * rs.standard_normal(n)).astype(dtype_c)
            y = np.zeros(n, dtype_c)
        elif kind == "r2c":
            x = rs.standard_normal(n).astype(dtype_r)
            y = np.zeros(n // 2 + 1, dtype_c)
        else:
            h = n // 2 + 1
            x = (rs.standard_normal(h) + 1j * rs.standard_normal(h)).astype(dtype_c)
            y = np.zeros(n, dtype_r)
        tin = torch.from_numpy(x).cuda()
        tout = torch.from_numpy(y).cuda()
        jobs.append((kind, fwd, n, tin, tout, fct))
        host.append((kind, fwd, n, x, y, fct))
    return jobs, host


def expected(host):
    out = []
    for kind, fwd, n, x, y, fct in host:
        out.append(oracle_out(kind, x, y, np.array([0]), fwd, fct))
    return out


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("autotune", [False, True])
def test_ragged_matches_oracle(dtype, autotune):
    rs = np.random.RandomState(5)
    jobs, host = make_jobs(dtype, LENGTHS, rs)
    want = expected(host)
    tol = tol_of(dtype)
    print(f"ragged create dtype={dtype.__name__} autotune={autotune}", flush=True)
    b = R.RaggedBatch(jobs, double=(dtype == np.complex128), autotune=autotune)
    print("ragged info", b.info(), flush=True)
    torch.cuda.synchronize()
    # creating (and tuning) a batch transforms nothing: inputs unchanged,
    # outputs still the zeros they were made with
    for (kind, fwd, n, tin, tout, fct), (_, _, _, x, y, _) in zip(jobs, host):
        assert np.array_equal(tin.cpu().numpy(), x), f"batch creation changed input n={n}"
        assert not tout.any(), f"batch creation wrote output n={n}"
    b.execute()
    torch.cuda.synchronize()
    for (kind, fwd, n, tin, tout, fct), w in zip(jobs, want):
        e = rel_l2(tout.cpu().numpy(), w)
        assert e <= tol, f"ragged {kind} n={n} fwd={fwd}: {e:.3e}"
    got = [j[4].cpu().numpy().copy() for j in jobs]
    for j in jobs:
        j[4].zero_()
    b.execute_serial()
    torch.cuda.synchronize()
    for (kind, fwd, n, tin, tout, fct), g in zip(jobs, got):
        e = rel_l2(tout.cpu().numpy(), g)
        assert e <= tol, f"serial vs persistent {kind} n={n}: {e:.3e}"
    b.close()


def test_ragged_graph_replays():
    rs = np.random.RandomState(6)
    jobs, host = make_jobs(np.complex64, LENGTHS[:12], rs)
    b = R.RaggedBatch(jobs, double=False, autotune=False)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.execute()
    torch.cuda.synchronize()
    eager = [j[4].cpu().numpy().copy() for j in jobs]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        b.execute()
    for rep in range(3):
        for j in jobs:
            j[4].zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for (kind, fwd, n, tin, tout, fct), e in zip(jobs, eager):
            assert rel_l2(tout.cpu().numpy(), e) == 0.0, f"replay {rep} {kind} n={n}"
    b.close()


def test_ragged_workload_sample():
    """A 256-job slice of the ragged benchmark workload (log-uniform lengths,
    good sizes and raw lengths, c2c and r2c) through the persistent launch,
    checked against the per-transform calls."""
    spec = R.ragged_workload(njobs=256, seed=5)
    jobs = []
    for kind, n in spec:
        if kind == "c2c":
            x = torch.randn(n, dtype=torch.complex64, device="cuda")
            y = torch.empty(n, dtype=torch.complex64, device="cuda")
        else:
            x = torch.randn(n, dtype=torch.float32, device="cuda")
            y = torch.empty(n // 2 + 1, dtype=torch.complex64, device="cuda")
        jobs.append((kind, True, n, x, y, 1.0))
    b = R.RaggedBatch(jobs, double=False, autotune=True)
    b.execute()
    torch.cuda.synchronize()
    got = [j[4].clone() for j in jobs]
    b.execute_serial()
    torch.cuda.synchronize()
    worst = max(float(torch.linalg.norm(g - j[4]) / torch.linalg.norm(j[4]))
                for g, j in zip(got, jobs))
    assert worst <= 1e-6, worst
    # spot-check against torch.fft in FP64 (independent of this engine)
    for kind, fwd, n, x, y, fct in jobs[:16]:
        ref = (torch.fft.fft(x.to(torch.complex128)) if kind == "c2c"
               else torch.fft.rfft(x.to(torch.float64)))
        assert float(torch.linalg.norm(y.to(ref.dtype) - ref) / torch.linalg.norm(ref)) <= 1e-6
    b.close()


@pytest.mark.skipif(not REF.available(), reason="reference engine not built")
def test_ragged_full_workload_vs_reference():
    """The whole ragged configuration: 4096 transforms, lengths log-uniform in
    64..65536 (70% good sizes, 30% raw incl. primes), 60% c2c complex64 and
    40% r2c float32, inputs N(0,1), one persistent launch; every job against
    the reference engine's own output for the same input."""
    spec = R.ragged_workload(njobs=4096, seed=5)
    rs = np.random.RandomState(5)
    jobs, host = [], []
    for kind, n in spec:
        if kind == "c2c":
            x = (rs.standard_normal(n) + 1j * rs.standard_normal(n)).astype(np.complex64)
            y = np.zeros(n, np.complex64)
        else:
            x = rs.standard_normal(n).astype(np.float32)
            y = np.zeros(n // 2 + 1, np.complex64)
        host.append((kind, n, x, y))
        jobs.append((kind, True, n, torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), 1.0))
    print("case full ragged workload: create", flush=True)
    b = R.RaggedBatch(jobs, double=False, autotune=True)
    print("ragged info", b.info(), flush=True)
    b.execute()
    torch.cuda.synchronize()
    worst, where = 0.0, None
    for (kind, n, x, y), j in zip(host, jobs):
        want = np.zeros_like(y)
        getattr(REF, kind)(x, want, np.array([0], dtype=np.uint64), True, 1.0, 1)
        e = rel_l2(j[4].cpu().numpy(), want)
        if e > worst:
            worst, where = e, (kind, n)
    print(f"worst job vs reference: {worst:.3e} at {where}", flush=True)
    assert worst <= 1e-6, (worst, where)
    b.close()


def test_ragged_inplace_jobs_with_autotune():
    """In-place jobs (input buffer == output buffer) under the autotuner:
    tuning runs on scratch copies, so the first execute() transforms the
    caller's data exactly once."""
    rs = np.random.RandomState(31)
    lengths = [64, 1000, 4099, 4096, 65536, 257]
    jobs, want = [], []
    for i, n in enumerate(lengths):
        x = (rs.standard_normal(n) + 1j * rs.standard_normal(n)).astype(np.complex64)
        t = torch.from_numpy(x).cuda()
        fwd = bool(i % 2 == 0)
        jobs.append(("c2c", fwd, n, t, t, 1.0))
        want.append(oracle_out("c2c", x, x, np.array([0]), fwd, 1.0))
    print("case in-place ragged jobs with autotune", flush=True)
    b = R.RaggedBatch(jobs, double=False, autotune=True)
    b.execute()
    torch.cuda.synchronize()
    for (kind, fwd, n, t, _, _), w in zip(jobs, want):
        e = rel_l2(t.cpu().numpy(), w)
        assert e <= 1e-6, f"in-place c2c n={n}: {e:.3e}"
    b.close()


def test_ragged_rejects_bad_jobs():
    """Job tensors are validated before anything reaches the GPU."""
    x = torch.zeros(64, dtype=torch.complex64, device="cuda")
    y = torch.zeros(64, dtype=torch.complex64, device="cuda")
    with pytest.raises(TypeError, match="CUDA tensor"):
        R.RaggedBatch([("c2c", True, 64, x.cpu(), y, 1.0)])
    with pytest.raises(TypeError, match="complex128"):
        R.RaggedBatch([("c2c", True, 64, x, y, 1.0)], double=True)
    with pytest.raises(TypeError, match="float32"):
        R.RaggedBatch([("r2c", True, 64, x, y, 1.0)])
    with pytest.raises(ValueError, match="needs 128"):
        R.RaggedBatch([("c2c", True, 128, x, y, 1.0)])
    z = torch.zeros((64, 2), dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError, match="contiguous"):
        R.RaggedBatch([("c2c", True, 64, z[:, 0], y, 1.0)])
    with pytest.raises(ValueError, match="unknown kind"):
        R.RaggedBatch([("dct", True, 64, x, y, 1.0)])


def test_ragged_tune_cache_on_disk(tmp_path):
    """Autotune results persist in the cache directory: a process with an
    empty memory cache reads them back (autotuned == 3) and runs the same
    configuration with bit-identical output; unreadable lines are skipped;
    a new signature is timed and appended."""
    rs = np.random.RandomState(21)
    jobs, _ = make_jobs(np.complex64, [64, 1000, 4099, 257, 8192, 12], rs)
    R.set_tune_cache(tmp_path, clear_memory=True)
    try:
        b1 = R.RaggedBatch(jobs, autotune=True)
        i1 = b1.info()
        assert i1["autotuned"] == 1
        files = list(tmp_path.iterdir())
        assert len(files) == 1 and files[0].name.startswith("ragged_tune_gfx")
        assert [l.split()[0] for l in files[0].read_text().splitlines()] == ["rfh1"]
        b1.execute()
        torch.cuda.synchronize()
        first = [j[4].clone() for j in jobs]
        with open(files[0], "a") as f:
            f.write("garbage line\nrfh1 zz\n")
        R.set_tune_cache(tmp_path, clear_memory=True)  # as a fresh process
        b2 = R.RaggedBatch(jobs, autotune=True)
        i2 = b2.info()
        assert i2["autotuned"] == 3
        assert (i2["grid"], i2["lds_bytes"]) == (i1["grid"], i1["lds_bytes"])
        for j in jobs:
            j[4].zero_()
        b2.execute()
        torch.cuda.synchronize()
        for a, j in zip(first, jobs):
            assert torch.equal(a, j[4])
        assert R.RaggedBatch(jobs, autotune=True).info()["autotuned"] == 2
        jobs2, _ = make_jobs(np.complex64, [64, 1000, 4099, 257, 8192, 13], rs)
        assert R.RaggedBatch(jobs2, autotune=True).info()["autotuned"] == 1
        good = [l for l in files[0].read_text().splitlines()
                if l.startswith("rfh1 ") and len(l.split()) == 6]
        assert len(good) == 2
    finally:
        R.set_tune_cache(None)

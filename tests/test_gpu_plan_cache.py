"""On-disk plan cache on the GPU (rfh_plan_cache): transforms run with a
cache directory write one record per plan; after the in-memory plans are
dropped, the same transforms load the records (no rebuild) and give
bit-identical outputs; a corrupted record is ignored, rebuilt and rewritten,
and the outputs stay bit-identical.  Shapes cover a direct plan, a four-step
(two-pass) plan, Bluestein on one LDS pass and two-pass Bluestein (65537),
and a real-input plan."""
import os

import numpy as np
import pytest

import rocket_fft_amd as R
from helpers import crandn, on_gpu, rrandn

pytestmark = pytest.mark.gpu

SHAPES = [("c2c", (3, 1000), np.complex64), ("c2c", (2, 1 << 16), np.complex64),
          ("c2c", (3, 4099), np.complex128), ("c2c", (2, 65537), np.complex64),
          ("r2c", (3, 4096), np.float64)]


def run_all():
    outs = []
    for kind, shape, dt in SHAPES:
        rs = np.random.RandomState(shape[1])
        if kind == "c2c":
            x = crandn(rs, shape, dt)
            outs.append(on_gpu(R, "c2c", x, x, np.array([1]), True, 1.0))
        else:
            x = rrandn(rs, shape, dt)
            y = np.zeros((shape[0], shape[1] // 2 + 1), np.complex128)
            outs.append(on_gpu(R, "r2c", x, y, np.array([1]), True, 1.0))
    return outs


def test_plan_cache_on_disk(tmp_path):
    d = tmp_path / "plans"
    d.mkdir()
    R.plan_cache(str(d), clear_memory=True)
    try:
        h0, w0 = R.plan_cache_stats()
        first = run_all()
        h1, w1 = R.plan_cache_stats()
        files = sorted(os.listdir(d))
        print("plan files", files, flush=True)
        assert w1 - w0 == len(files) >= len(SHAPES) and h1 == h0
        assert all(f.startswith("plan_gfx") and f.endswith(".bin") for f in files)

        R.plan_cache(str(d), clear_memory=True)
        second = run_all()
        h2, w2 = R.plan_cache_stats()
        assert w2 == w1 and h2 - h1 == len(files)
        for a, b in zip(first, second):
            assert np.array_equal(a, b)

        victim = d / next(f for f in files if "_65537_" in f)
        raw = bytearray(victim.read_bytes())
        raw[len(raw) // 2] ^= 0x5A
        victim.write_bytes(bytes(raw))
        R.plan_cache(str(d), clear_memory=True)
        third = run_all()
        h3, w3 = R.plan_cache_stats()
        assert w3 - w2 == 1 and h3 - h2 == len(files) - 1
        for a, b in zip(first, third):
            assert np.array_equal(a, b)
    finally:
        R.plan_cache(None)

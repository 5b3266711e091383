"""Concurrent callers: 8 host threads call the entry points at once on one GPU
(mirrors rocket-fft's tests/test_multithreading.py:15-31, which maps
42 inputs of 2^20 over an 8-worker pool).  Here every result is checked
against the oracle: c2c on host arrays, r2c on device tensors, DCT-II on
both, half the threads on a stream of their own (separate workspaces), the
others sharing torch's default stream (one workspace, serialised).  The first
failing task fails the test."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

import rocket_fft_amd as R
from helpers import oracle_out, rel_l2

pytestmark = pytest.mark.gpu

N = 1 << 20


def task_inputs(i):
    rs = np.random.RandomState(100 + i)
    kind = ("c2c", "r2c", "dct")[i % 3]
    if kind == "c2c":
        x = rs.standard_normal(N) + 1j * rs.standard_normal(N)
        want = oracle_out("c2c", x, x, np.array([0]), True, 1.0)
    elif kind == "r2c":
        x = rs.standard_normal(N)
        want = oracle_out("r2c", x, np.zeros(N // 2 + 1, np.complex128), np.array([0]), True, 1.0)
    else:
        x = rs.uniform(-1, 1, N)
        want = oracle_out("dct", x, x, np.array([0]), 2, 1.0, False)
    return kind, x, want


def run_task(args):
    i, kind, x, want = args
    own_stream = i % 2 == 1
    on_device = (i // 3) % 2 == 0 or kind == "r2c"
    ctx = torch.cuda.stream(torch.cuda.Stream()) if own_stream else torch.cuda.stream(
        torch.cuda.current_stream())
    with ctx:
        if on_device:
            tx = torch.from_numpy(x).cuda()
            ty = torch.empty(want.shape, dtype=torch.from_numpy(want).dtype, device="cuda")
        else:
            tx, ty = x, np.zeros_like(want)
        if kind == "c2c":
            R.c2c(tx, ty, np.array([0]), True, 1.0, 1)
        elif kind == "r2c":
            R.r2c(tx, ty, np.array([0]), True, 1.0, 1)
        else:
            R.dct(tx, ty, np.array([0]), 2, 1.0, False, 1)
        if on_device:
            torch.cuda.current_stream().synchronize()
            ty = ty.cpu().numpy()
    e = rel_l2(ty, want)
    assert e <= 1e-13, f"task {i} ({kind}, device={on_device}, own stream={own_stream}): {e:.3e}"
    return e


def test_eight_threads():
    tasks = [(i, *task_inputs(i)) for i in range(24)]
    print("case 8 threads x 24 tasks of 2^20 (c2c host, r2c device, dct both)", flush=True)
    worst = 0.0
    with ThreadPoolExecutor(max_workers=8) as ex:
        for e in ex.map(run_task, tasks):  # re-raises the first failure
            worst = max(worst, e)
    print(f"worst rel-L2 vs oracle: {worst:.3e}", flush=True)

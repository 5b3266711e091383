"""Workspaces and HIP graph capture: a graph captured over a transform that
uses a device workspace keeps working after a larger eager call on the same
stream has grown that workspace, and a call that would have to grow the
workspace during capture fails with a clear error instead of allocating."""
import numpy as np
import pytest

import rocket_fft_amd as R
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def fresh_stream():
    """A new HIP stream (default priority) of its own, so that the
    library's per-stream workspace starts empty whatever ran before."""
    import ctypes
    import torch
    hip = ctypes.CDLL("libamdhip64.so")
    h = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(h)) == 0
    return torch.cuda.ExternalStream(h.value), hip, h


def test_graph_survives_workspace_growth():
    import torch
    n = 1 << 16  # four-step plan: the intermediate lives in the workspace
    g = torch.Generator(device="cuda")
    g.manual_seed(21)
    x = torch.randn((2, n), dtype=torch.complex64, device="cuda", generator=g)
    y = torch.empty_like(x)
    big = torch.randn((64, 1 << 18), dtype=torch.complex64, device="cuda", generator=g)
    big_out = torch.empty_like(big)
    s, hip, h = fresh_stream()
    ax = np.array([1])
    print("case graph + workspace growth", flush=True)
    with torch.cuda.stream(s):
        R.c2c(x, y, ax, True, 1.0, 1)  # plan and workspace before capture
    torch.cuda.synchronize()
    eager = y.clone()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        R.c2c(x, y, ax, True, 1.0, 1)
    torch.cuda.synchronize()
    for rep in range(3):
        y.zero_()
        # a larger eager call on the same stream grows the workspace the
        # graph's intermediate lives in
        with torch.cuda.stream(s):
            R.c2c(big, big_out, ax, True, 1.0, 1)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        e = rel_l2(y.cpu().numpy(), eager.cpu().numpy())
        assert e == 0.0, f"replay {rep} after workspace growth: rel-L2 {e:.3e}"
    want = np.fft.fft(big.cpu().numpy().astype(np.complex128)[:4], axis=1)
    assert rel_l2(big_out.cpu().numpy()[:4], want) <= 1e-6


def test_growth_during_capture_is_an_error():
    import torch
    n = 1 << 17
    x_small = torch.randn((1, n), dtype=torch.complex64, device="cuda")
    x_large = torch.randn((16, n), dtype=torch.complex64, device="cuda")
    y_small, y_large = torch.empty_like(x_small), torch.empty_like(x_large)
    s, hip, h = fresh_stream()
    ax = np.array([1])
    print("case growth during capture", flush=True)
    with torch.cuda.stream(s):
        R.c2c(x_small, y_small, ax, True, 1.0, 1)  # plan cached, small workspace
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with pytest.raises(RuntimeError, match="while the stream is capturing"):
        with torch.cuda.graph(graph, stream=s):
            R.c2c(x_large, y_large, ax, True, 1.0, 1)
    torch.cuda.synchronize()
    # the library still works afterwards
    with torch.cuda.stream(s):
        R.c2c(x_large, y_large, ax, True, 1.0, 1)
    torch.cuda.synchronize()
    want = np.fft.fft(x_large.cpu().numpy().astype(np.complex128), axis=1)
    assert rel_l2(y_large.cpu().numpy(), want) <= 1e-6


@pytest.mark.parametrize("mode", ["fp32", "mfma_fp16"])
def test_graph_replays_pipelined_kernels(mode):
    """The persistent pipelined kernels (lines of 4096 and the 2^20 passes,
    in the FP32 and the matrix-core form) captured in one graph: every
    replay, with eager transforms and torch work on the same stream in
    between, equals the eager result bit for bit."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(31)
    xa = torch.randn((2048, 4096), dtype=torch.complex64, device="cuda", generator=g)
    xb = torch.randn((3, 1 << 20), dtype=torch.complex64, device="cuda", generator=g)
    other = torch.randn((512, 4096), dtype=torch.complex64, device="cuda", generator=g)
    ya, yb, yo = torch.empty_like(xa), torch.empty_like(xb), torch.empty_like(other)
    s, hip, h = fresh_stream()
    ax = np.array([1])
    prev = R.set_precision(mode)
    try:
        print(f"case graph replays of the pipelined kernels, mode {mode}", flush=True)
        with torch.cuda.stream(s):
            R.c2c(xa, ya, ax, True, 1.0, 1)
            R.c2c(xb, yb, ax, False, 1.0 / (1 << 20), 1)
        torch.cuda.synchronize()
        ea, eb = ya.clone(), yb.clone()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            R.c2c(xa, ya, ax, True, 1.0, 1)
            R.c2c(xb, yb, ax, False, 1.0 / (1 << 20), 1)
        torch.cuda.synchronize()
        for rep in range(3):
            ya.zero_()
            yb.zero_()
            with torch.cuda.stream(s):
                R.c2c(other, yo, ax, True, 1.0, 1)  # eager work between replays
                yo.mul_(0.5)
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(ya, ea), f"replay {rep}: lines of 4096 differ"
            assert torch.equal(yb, eb), f"replay {rep}: 2^20 lines differ"
    finally:
        R.set_precision(prev)
    want = np.fft.fft(xa[:8].cpu().numpy().astype(np.complex128), axis=1)
    assert rel_l2(ea[:8].cpu().numpy(), want) <= 1e-6

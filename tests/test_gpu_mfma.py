"""Split-precision form: radix-16 stages on the matrix cores.

* The operand layout the kernels assume for v_mfma_f32_16x16x32_{f16,bf16}
  is checked exactly with small integers (exact in FP16/BF16/FP32).
* complex64 transforms of length 4096 in each precision mode against the
  long-double oracle on both input sets of the precision configuration:
  iid N(0,1), and magnitudes log-uniform over 1e-3..1e3 with uniform phase.
  Every matrix-core case asserts through the launch counter that the
  matrix-core kernel ran (no silent FP32 fallback), and the two-way BF16
  study split must be measurably less accurate than FP32.
"""
import numpy as np
import pytest
import torch

import rocket_fft_amd as R
from helpers import oracle_out, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bf16", [False, True])
def test_mfma_operand_layout(bf16):
    rs = np.random.RandomState(1)
    A = rs.randint(-4, 5, (16, 32)).astype(np.float32)
    B = rs.randint(-4, 5, (32, 16)).astype(np.float32)
    al = np.zeros((64, 8), np.float32)
    bl = np.zeros((64, 8), np.float32)
    for l in range(64):
        g = l // 16
        al[l] = A[l % 16, 8 * g:8 * g + 8]
        bl[l] = B[8 * g:8 * g + 8, l % 16]
    d = R.mfma_probe(torch.from_numpy(al).cuda(), torch.from_numpy(bl).cuda(), bf16).cpu().numpy()
    D = A @ B
    want = np.zeros((64, 4), np.float32)
    for l in range(64):
        g = l // 16
        want[l] = D[4 * g:4 * g + 4, l % 16]
    assert np.array_equal(d, want)


def inputs(kind, rows, rs):
    n = 4096
    if kind == "normal":
        x = rs.standard_normal((rows, n)) + 1j * rs.standard_normal((rows, n))
    else:
        mag = np.exp(rs.uniform(np.log(1e-3), np.log(1e3), (rows, n)))
        x = mag * np.exp(2j * np.pi * rs.uniform(size=(rows, n)))
    return x.astype(np.complex64)


@pytest.mark.parametrize("mode,tol", [("fp32", 1e-6), ("mfma_fp16", 1e-6), ("mfma_bf16", 1e-6),
                                      ("mfma_bf16_2split", 1e-4)])
@pytest.mark.parametrize("kind", ["normal", "loguniform"])
def test_mfma_4096_vs_fp64(mode, tol, kind):
    rs = np.random.RandomState(4)
    x = inputs(kind, 32, rs)
    errs = {}
    for m in ("fp32", mode):
        prev = R.set_precision(m)
        try:
            for fwd, fct in ((True, 1.0), (False, 1.0 / 4096)):
                t = torch.from_numpy(x).cuda()
                y = torch.empty_like(t)
                before = R.mfma_launches(None if m == "fp32" else m)
                R.c2c(t, y, np.array([1]), fwd, fct, 1)
                torch.cuda.synchronize()
                ran = R.mfma_launches(None if m == "fp32" else m) - before
                assert ran == (0 if m == "fp32" else 1), f"mode {m}: {ran} matrix-core launches"
                want = oracle_out("c2c", x.astype(np.complex128), x.astype(np.complex128),
                                  np.array([1]), fwd, fct)
                e = rel_l2(y.cpu().numpy(), want)
                errs[(m, fwd)] = e
                print(f"mode={m} input={kind} forward={fwd}: rel-L2 vs FP64 = {e:.3e}", flush=True)
                assert e <= (1e-6 if m == "fp32" else tol)
        finally:
            R.set_precision(prev)
    if mode == "mfma_bf16_2split":  # ~16-bit operands: clearly worse than FP32
        for fwd in (True, False):
            assert errs[(mode, fwd)] > 5 * errs[("fp32", fwd)]


@pytest.mark.parametrize("kind", ["normal", "loguniform"])
def test_mfma_4096_pipelined_vs_fp64(kind):
    """2051 lines of 4096 (the persistent pipelined matrix-core kernel: four
    lines per tile, at least two tiles per CU, a partial last tile) in every
    mode, both directions, against the long-double oracle; the per-launch
    timing labels name the kernel family."""
    rs = np.random.RandomState(6)
    x = inputs(kind, 2051, rs)
    want = {fwd: oracle_out("c2c", x.astype(np.complex128), x.astype(np.complex128),
                            np.array([1]), fwd, fct)
            for fwd, fct in ((True, 1.0), (False, 1.0 / 4096))}
    for mode, tol in (("mfma_fp16", 1e-6), ("mfma_bf16", 1e-6), ("mfma_bf16_2split", 1e-4)):
        prev = R.set_precision(mode)
        try:
            for fwd, fct in ((True, 1.0), (False, 1.0 / 4096)):
                t = torch.from_numpy(x).cuda()
                y = torch.empty_like(t)
                R.kernel_timing_report()
                R.kernel_timing(True)
                try:
                    R.c2c(t, y, np.array([1]), fwd, fct, 1)
                    torch.cuda.synchronize()
                finally:
                    R.kernel_timing(False)
                labels = R.kernel_timing_report()
                assert any(k.startswith("k_pipe4096_mfma c2c n=4096") for k in labels), labels
                e = rel_l2(y.cpu().numpy(), want[fwd])
                print(f"mode={mode} pipelined input={kind} forward={fwd}: rel-L2 vs FP64 = "
                      f"{e:.3e}", flush=True)
                assert e <= tol
        finally:
            R.set_precision(prev)


@pytest.mark.parametrize("mode", ["mfma_fp16", "mfma_bf16", "mfma_bf16_2split"])
@pytest.mark.parametrize("kind", ["normal", "loguniform"])
def test_mfma_2p20_vs_fp64(mode, kind):
    """complex64 lines of 2^20 = 1024 x 1024: both four-step passes run their
    radix-16 stages on the matrix cores (one pass-A and one pass-B launch per
    chunk of lines), both directions, against the long-double oracle."""
    n = 1 << 20
    rs = np.random.RandomState(5)
    if kind == "normal":
        x = rs.standard_normal((2, n)) + 1j * rs.standard_normal((2, n))
    else:
        x = np.exp(rs.uniform(np.log(1e-3), np.log(1e3), (2, n))) * \
            np.exp(2j * np.pi * rs.uniform(size=(2, n)))
    x = x.astype(np.complex64)
    tol = 1e-4 if mode == "mfma_bf16_2split" else 1e-6
    prev = R.set_precision(mode)
    try:
        for fwd, fct in ((True, 1.0), (False, 1.0 / n)):
            t = torch.from_numpy(x).cuda()
            y = torch.empty_like(t)
            before = R.mfma_launches(mode)
            R.c2c(t, y, np.array([1]), fwd, fct, 1)
            torch.cuda.synchronize()
            assert R.mfma_launches(mode) - before == 2, "both passes on the matrix cores"
            want = oracle_out("c2c", x.astype(np.complex128), x.astype(np.complex128),
                              np.array([1]), fwd, fct)
            e = rel_l2(y.cpu().numpy(), want)
            print(f"mode={mode} 2^20 input={kind} forward={fwd}: rel-L2 vs FP64 = {e:.3e}",
                  flush=True)
            assert e <= tol
    finally:
        R.set_precision(prev)


def test_mfma_mode_only_touches_4096():
    """Other lengths keep the plain path whatever the mode (the contiguous
    length-1024 lines here are not four-step passes)."""
    rs = np.random.RandomState(3)
    x = (rs.standard_normal((4, 1024)) + 1j * rs.standard_normal((4, 1024))).astype(np.complex64)
    prev = R.set_precision("mfma_bf16_2split")
    try:
        t = torch.from_numpy(x).cuda()
        y = torch.empty_like(t)
        before = R.mfma_launches()
        R.c2c(t, y, np.array([1]), True, 1.0, 1)
        torch.cuda.synchronize()
        assert R.mfma_launches() == before
        assert rel_l2(y.cpu().numpy(), np.fft.fft(x.astype(np.complex128), axis=1)) <= 1e-6
    finally:
        R.set_precision(prev)

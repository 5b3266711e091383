"""Parity at the stated sizes of the benchmark configurations, through the C
ABI, against the reference engine itself (compiled from its sources into
oracle/_ref) on the same seeded inputs, and for the precision configuration
against an FP64 transform.  Bar: rel-L2 <= 1e-13 (complex128/float64), <= 1e-6
(complex64/float32).  Inputs follow the configurations' seeds (numpy legacy
RandomState, seed = configuration number).  The ragged configuration (all
4096 jobs) is in test_gpu_ragged.py."""
import os

import numpy as np
import pytest
import torch

import rocket_fft_amd as R
from helpers import REF, oracle_out, rel_l2, tol_of

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not REF.available(), reason="reference engine not built")]

THREADS = max(1, min(16, os.cpu_count() or 1))


def ref(fname, ain, aout_like, *args):
    out = np.zeros_like(aout_like)
    getattr(REF, fname)(ain, out, *args, THREADS)
    return out


def gpu(fname, x, out_like, *args):
    tx = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    ty = torch.from_numpy(np.zeros_like(out_like)).cuda()
    getattr(R, fname)(tx, ty, *args, 1)
    torch.cuda.synchronize()
    return ty.cpu().numpy()


def test_config1_c128_1024x4096_axis1():
    rs = np.random.RandomState(1)
    x = rs.standard_normal((1024, 4096)) + 1j * rs.standard_normal((1024, 4096))
    ax = np.array([1], dtype=np.uint64)
    print("case config 1: c2c complex128 (1024, 4096) axis 1", flush=True)
    got = gpu("c2c", x, x, ax, True, 1.0)
    e = rel_l2(got, ref("c2c", x, x, ax, True, 1.0))
    print(f"  vs reference {e:.3e}", flush=True)
    assert e <= 1e-13
    # and a slice of rows against the long-double oracle
    e2 = rel_l2(got[:32], oracle_out("c2c", x[:32], x[:32], ax, True, 1.0))
    assert e2 <= 1e-13


def test_config2a_r2c_c2r_256cubed():
    rs = np.random.RandomState(2)
    x = rs.standard_normal((256, 256, 256)).astype(np.float32)
    ax = np.array([0, 1, 2], dtype=np.uint64)
    spec_like = np.zeros((256, 256, 129), np.complex64)
    print("case config 2a: r2c float32 256^3 axes (0,1,2)", flush=True)
    got = gpu("r2c", x, spec_like, ax, True, 1.0)
    want = ref("r2c", x, spec_like, ax, True, 1.0)
    e = rel_l2(got, want)
    print(f"  r2c vs reference {e:.3e}", flush=True)
    assert e <= 1e-6
    print("case config 2a: c2r back, fct = 1/256^3", flush=True)
    fct = 1.0 / 256 ** 3
    back = gpu("c2r", want, x, ax, False, fct)  # the same spectrum into both engines
    e = rel_l2(back, ref("c2r", want, x, ax, False, fct))
    print(f"  c2r vs reference {e:.3e}", flush=True)
    assert e <= 1e-6
    rt = gpu("c2r", got, x, ax, False, fct)
    assert rel_l2(rt, x) <= 1e-6


def test_config2b_dct2_ortho_512():
    rs = np.random.RandomState(2)
    rs.standard_normal((256, 256, 256))  # the first draw of seed 2 is the 3-D field
    d = rs.uniform(-1.0, 1.0, (512, 512))
    ax = np.array([0, 1], dtype=np.uint64)
    print("case config 2b: dct-II ortho float64 512^2", flush=True)
    got = gpu("dct", d, d, ax, 2, 1.0 / 1024, True)
    e = rel_l2(got, ref("dct", d, d, ax, 2, 1.0 / 1024, True))
    assert e <= 1e-13


@pytest.mark.parametrize("n", [1000, 1536, 2187, 4913, 4099, 65537])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_config3_batch256(n, dtype):
    rs = np.random.RandomState(3)
    x = (rs.standard_normal((256, n)) + 1j * rs.standard_normal((256, n))).astype(dtype)
    ax = np.array([1], dtype=np.uint64)
    tol = tol_of(dtype)
    print(f"case config 3: c2c {np.dtype(dtype).name} (256, {n})", flush=True)
    fwd = gpu("c2c", x, x, ax, True, 1.0)
    want = ref("c2c", x, x, ax, True, 1.0)
    e = rel_l2(fwd, want)
    print(f"  forward vs reference {e:.3e}", flush=True)
    assert e <= tol
    inv = gpu("c2c", want, x, ax, False, 1.0 / n)
    e = rel_l2(inv, ref("c2c", want, x, ax, False, 1.0 / n))
    print(f"  inverse vs reference {e:.3e}", flush=True)
    assert e <= tol
    # forward error against an FP64 transform (numpy) and the round trip
    exact = np.fft.fft(x.astype(np.complex128), axis=1)
    assert rel_l2(fwd, exact) <= tol
    assert rel_l2(gpu("c2c", fwd, x, ax, False, 1.0 / n), x) <= 2 * tol


def precision_inputs(kind, shape, rs):
    if kind == "normal":
        x = rs.standard_normal(shape) + 1j * rs.standard_normal(shape)
    else:
        mag = np.exp(rs.uniform(np.log(1e-3), np.log(1e3), shape))
        x = mag * np.exp(2j * np.pi * rs.uniform(size=shape))
    return x.astype(np.complex64)


@pytest.mark.parametrize("shape", [(64, 1 << 20), (16384, 4096)])
@pytest.mark.parametrize("kind", ["normal", "loguniform"])
def test_config4_precision_forms(shape, kind):
    """Both shapes, both input sets: the plain FP32 path and every
    matrix-core mode that claims rel-L2 <= 1e-6 against FP64 (the reference
    engine's complex128 transform of the promoted input), plus the plain
    path against the reference's own complex64 output."""
    rs = np.random.RandomState(4)
    x = precision_inputs(kind, shape, rs)
    ax = np.array([1], dtype=np.uint64)
    print(f"case config 4: {shape} {kind}: FP64 reference", flush=True)
    x64 = x.astype(np.complex128)
    exact = ref("c2c", x64, x64, ax, True, 1.0)
    del x64
    tx = torch.from_numpy(x).cuda()
    ty = torch.empty_like(tx)
    errs = {}
    for mode in ("fp32", "mfma_fp16", "mfma_bf16"):
        prev = R.set_precision(mode)
        try:
            before = R.mfma_launches(None if mode == "fp32" else mode)
            R.c2c(tx, ty, np.array([1]), True, 1.0, 1)
            torch.cuda.synchronize()
            ran = R.mfma_launches(None if mode == "fp32" else mode) - before
        finally:
            R.set_precision(prev)
        got = ty.cpu().numpy()
        errs[mode] = rel_l2(got, exact)
        print(f"  {mode}: rel-L2 vs FP64 {errs[mode]:.3e} ({ran} matrix-core launches)",
              flush=True)
        assert errs[mode] <= 1e-6, (mode, errs[mode])
        if mode == "fp32":
            assert ran == 0
            e = rel_l2(got, ref("c2c", x, x, ax, True, 1.0))
            print(f"  fp32 vs reference complex64 {e:.3e}", flush=True)
            assert e <= 1e-6
        else:
            assert ran >= 1, f"{mode} did not run on the matrix cores"

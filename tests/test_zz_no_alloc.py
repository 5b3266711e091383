"""The allocation-free c2c entry point (rfh_c2c_noalloc) against the ordinary
numba_c2c on the same inputs: two calls in a row, the second on a smaller
input, plus its refusals (a size beyond the static scratch, a length without
a cached plan, host memory).  Same kernels, so the outputs must agree to the
parity bar (in practice bit for bit)."""
import numpy as np
import pytest
import torch

import rocket_fft_amd as R
from helpers import rel_l2, tol_of

pytestmark = pytest.mark.gpu


def pair(shape, dtype, axes, fwd, fct, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randn(shape, dtype=dtype, device="cuda", generator=g)
    want = torch.empty_like(x)
    R.c2c(x, want, np.array(axes), fwd, fct, 1)  # builds the plans
    got = torch.full_like(x, float("nan"))
    R.c2c_noalloc(x, got, np.array(axes), fwd, fct)
    torch.cuda.synchronize()
    return got.cpu().numpy(), want.cpu().numpy()


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_noalloc_matches_c2c_twice(dtype):
    npdt = np.complex64 if dtype == torch.complex64 else np.complex128
    # first call: four-step lines of 2^20, more lines than one scratch chunk
    big = (12, 1 << 20) if dtype == torch.complex64 else (6, 1 << 20)
    print(f"case noalloc {dtype} first {big}", flush=True)
    got, want = pair(big, dtype, [1], True, 1.0, 1)
    e = rel_l2(got, want)
    assert e <= tol_of(npdt), e
    print(f"  rel-L2 {e:.3e} bit-equal {np.array_equal(got, want)}", flush=True)
    # second call: a smaller input, several axes, inverse with a factor
    print(f"case noalloc {dtype} second (64, 4099)", flush=True)
    got, want = pair((64, 4099), dtype, [1, 0], False, 0.25, 2)
    e = rel_l2(got, want)
    assert e <= tol_of(npdt), e
    print(f"  rel-L2 {e:.3e} bit-equal {np.array_equal(got, want)}", flush=True)


def test_noalloc_refusals():
    x = torch.randn((2, 1 << 21), dtype=torch.complex128, device="cuda")
    y = torch.empty_like(x)
    R.c2c(x, y, np.array([1]), True, 1.0, 1)  # plan cached; needs > 64 MiB scratch
    with pytest.raises(RuntimeError, match="B of scratch"):
        R.c2c_noalloc(x, y, np.array([1]), True, 1.0)
    z = torch.randn((2, 7919 * 3), dtype=torch.complex64, device="cuda")
    with pytest.raises(RuntimeError, match="no plan on this GPU yet"):
        R.c2c_noalloc(z, torch.empty_like(z), np.array([1]), True, 1.0)
    with pytest.raises(TypeError, match="CUDA tensors only"):
        R.c2c_noalloc(np.zeros((2, 8), np.complex64), np.zeros((2, 8), np.complex64),
                      np.array([1]), True, 1.0)
    torch.cuda.synchronize()

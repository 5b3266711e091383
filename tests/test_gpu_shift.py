"""Shifted transforms on the GPU (rfh_c2c_shifted): numpy.roll / fftshift /
ifftshift fused into the loads and stores of the transform passes.

Bar: the fused result equals np.roll of the same library's unfused c2c output
bit for bit (the rolls only move where elements are read and written), and
the oracle's c2c of the rolled input, rolled, to the rel-L2 bar (1e-6
complex64, 1e-13 complex128).  Covers one-pass lengths on the specialised
power-of-two kernel (64, 1024, 4096) and the generic kernel (1000, odd 243,
whose fftshift is not a half turn), four-step lengths (2^16, 2^20 complex64),
strided and transposed arrays, in place, several axes with mixed signs,
Bluestein lengths whose transform is one LDS pass (4099 complex64, 1031
complex128), three passes (65537 complex64, 4099 complex128) or around a
multi-pass transform (prime 1000003), multi-pass lengths (2^21), a roll the
plan cannot fuse (one element on 2^20: nothing written, front-end falls
back), the real-transform front-ends, and the matrix-core precision mode
(shifted lines run on the FP32 kernels)."""
import numpy as np
import pytest

import rocket_fft_amd as R
import rocket_fft_amd.fft as F
from helpers import crandn, oracle_out, rel_l2, to_gpu, tol_of

pytestmark = pytest.mark.gpu


def np_roll(x, shifts, axes):
    return np.roll(x, [int(v) for v in shifts], axis=tuple(axes)) if shifts is not None else x


def run(x, axes, forward, fct, si, so, inplace=False):
    import torch
    t = to_gpu(x)
    out = t if inplace else torch.zeros_like(t)
    fused = R.c2c_shifted(t, out, np.asarray(axes, np.int64), forward, fct, si, so)
    torch.cuda.synchronize()
    return fused, out.cpu().numpy()


def unfused(x, axes, forward, fct):
    import torch
    t = to_gpu(x)
    out = torch.zeros_like(t)
    R.c2c(t, out, np.asarray(axes, np.int64), forward, fct, 1)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def check_case(x, axes, forward, fct, si, so, inplace=False):
    print(f"case shifted c2c shape={x.shape} {x.dtype} axes={axes} si={si} so={so} "
          f"fwd={forward} inplace={inplace}", flush=True)
    fused, got = run(x, axes, forward, fct, si, so, inplace)
    assert fused
    xr = np_roll(x, si, axes)
    want_bits = np_roll(unfused(xr, axes, forward, fct), so, axes)
    assert np.array_equal(got, want_bits), "fused differs from roll(unfused)"
    want = np_roll(oracle_out("c2c", xr, xr, np.asarray(axes), forward, fct), so, axes)
    e = rel_l2(got, want)
    assert e <= tol_of(x.dtype), f"rel-L2 {e:.3e} vs oracle"


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("n", [64, 1000, 243, 1024, 4096, 1 << 16])
def test_fftshift_one_axis(dtype, n):
    rs = np.random.RandomState(n)
    x = crandn(rs, (3, n), dtype)
    check_case(x, [1], True, 1.0, [-(n // 2)], [n // 2])      # centred forward
    check_case(x, [1], False, 1.0 / n, [n // 2], [-(n // 2)])  # its inverse pairing
    check_case(x, [1], True, 1.0, None, [n // 2])              # fftshift(fft(x))


def test_fftshift_4096_pipelined_complex64():
    """2051 lines of 4096: the persistent pipelined kernel with rotated loads
    and stores (partial last tile), bit for bit against roll(unfused)."""
    n = 4096
    rs = np.random.RandomState(5)
    x = crandn(rs, (2051, n), np.complex64)
    check_case(x, [1], True, 1.0, [-(n // 2)], [n // 2])
    check_case(x, [1], False, 1.0 / n, [n // 2 + 5], [-(n // 2) + 3])


def test_fftshift_2p20_complex64():
    n = 1 << 20
    rs = np.random.RandomState(20)
    x = crandn(rs, (2, n), np.complex64)
    check_case(x, [1], True, 1.0, [-(n // 2)], [n // 2])


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_shift_several_axes_strided_inplace(dtype):
    rs = np.random.RandomState(3)
    x = crandn(rs, (16, 24, 10), dtype)
    check_case(x, [0, 1, 2], True, 1.0, [3, -5, 7], [-8, 12, 1])
    check_case(x, [2, 0], False, 0.25, [4, -1], None)
    check_case(x, [1], True, 1.0, [7], [-7], inplace=True)
    xt = np.ascontiguousarray(crandn(rs, (40, 64), dtype)).T  # transposed view
    import torch
    t = torch.from_numpy(np.ascontiguousarray(xt.T)).cuda().T  # strided device view
    out = torch.zeros(xt.shape, dtype=t.dtype, device="cuda")
    assert R.c2c_shifted(t, out, np.array([0, 1]), True, 1.0, [32, -20], [-32, 20])
    torch.cuda.synchronize()
    want = np_roll(unfused(np_roll(xt, [32, -20], [0, 1]), [0, 1], True, 1.0), [-32, 20], [0, 1])
    assert np.array_equal(out.cpu().numpy(), want)


@pytest.mark.parametrize("dtype,n", [(np.complex64, 4099), (np.complex128, 1031)])
def test_fftshift_bluestein_one_pass(dtype, n):
    """Bluestein lengths whose whole transform is one LDS pass (chirp, FFT,
    kernel product, FFT, chirp): the rolls rotate its loads and stores."""
    kind = R.gpu_schedule(n, dtype == np.complex128)[0]
    print(f"case bluestein one-pass shift n={n} {dtype.__name__} kind={kind}", flush=True)
    assert kind == 1
    rs = np.random.RandomState(n)
    x = crandn(rs, (3, n), dtype)
    check_case(x, [1], True, 1.0, [-(n // 2)], [n // 2])
    check_case(x, [1], False, 1.0 / n, [n // 2], [-(n // 2)])
    check_case(x, [1], True, 1.0, [7], None)


@pytest.mark.parametrize("dtype,n", [(np.complex64, 65537), (np.complex128, 4099)])
def test_fftshift_bluestein_three_pass(dtype, n):
    """Bluestein over three passes (chirped columns, rows, final columns):
    the rolls rotate pass A's chirped loads and pass C's final stores over
    the whole line."""
    kind = R.gpu_schedule(n, dtype == np.complex128)[0]
    print(f"case bluestein three-pass shift n={n} {dtype.__name__} kind={kind}", flush=True)
    assert kind == 3
    rs = np.random.RandomState(n)
    x = crandn(rs, (3, n), dtype)
    check_case(x, [1], True, 1.0, [-(n // 2)], [n // 2])
    check_case(x, [1], False, 1.0 / n, [n // 2], [-(n // 2)])
    check_case(x, [1], True, 1.0, [1], [n - 1])


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_fftshift_multi_pass(dtype):
    """2^21: a column pass of n1 and the row sub-plan of M = 2^21 / n1; the
    input roll (a multiple of M) rotates the columns, the output roll (a
    multiple of n1) the sub-plan's row outputs."""
    n = 1 << 21
    kind = R.gpu_schedule(n, dtype == np.complex128)[0]
    print(f"case multi-pass shift n={n} {dtype.__name__} kind={kind}", flush=True)
    assert kind == 4
    rs = np.random.RandomState(21)
    x = crandn(rs, (2, n), dtype)
    check_case(x, [1], True, 1.0, [-(n // 2)], [n // 2])
    check_case(x, [1], False, 1.0 / n, [n // 2], None)


def test_fftshift_bluestein_multi_pass():
    """Prime 1000003: Bluestein around a multi-pass transform; the rolls
    rotate the elementwise chirp passes' loads (input) and stores (output)."""
    n = 1000003
    kind = R.gpu_schedule(n, False)[0]
    print(f"case bluestein multi-pass shift n={n} kind={kind}", flush=True)
    assert kind == 5
    rs = np.random.RandomState(7)
    x = crandn(rs, (2, n), np.complex64)
    check_case(x, [1], True, 1.0, [-(n // 2)], [n // 2])
    check_case(x, [1], False, 1.0 / n, [12345], None)


def test_not_fusable_writes_nothing_and_frontend_falls_back():
    """2^20 complex64 with an input roll of one element: not a whole column
    of the four-step plan, so nothing is fused."""
    import torch
    n = 1 << 20
    rs = np.random.RandomState(4)
    x = crandn(rs, (2, n), np.complex64)
    t = to_gpu(x)
    out = torch.full_like(t, 7.0)
    assert not R.c2c_shifted(t, out, np.array([1]), True, 1.0, [1], None)
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())
    got = F.fftn_shifted(t, axes=[1], shift_in=[1], shift_out="fftshift").cpu().numpy()
    want = np.fft.fftshift(np.fft.fft(np.roll(x.astype(np.complex128), 1, axis=1), axis=1),
                           axes=1)
    assert rel_l2(got, want) <= 1e-6


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_centered_frontend_matches_numpy(dtype):
    rs = np.random.RandomState(5)
    x = crandn(rs, (32, 48, 20), dtype)
    x64 = x.astype(np.complex128)
    tol = tol_of(dtype)
    got = F.centered_fftn(to_gpu(x)).cpu().numpy()
    want = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x64)))
    assert rel_l2(got, want) <= tol
    back = F.centered_ifftn(to_gpu(got)).cpu().numpy()
    assert rel_l2(back, x64) <= 2 * tol
    got = F.fftn_shifted(x, axes=(0, 2), norm="ortho", shift_out="fftshift")  # host arrays
    want = np.fft.fftshift(np.fft.fftn(x64, axes=(0, 2), norm="ortho"), axes=(0, 2))
    assert isinstance(got, np.ndarray) and rel_l2(got, want) <= tol


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rfftn_shifted_matches_numpy(dtype):
    """rfftn with rolls along the leading axes: the real pass of the last
    axis, then the complex passes with rotated loads and stores; bit for bit
    against roll(rfftn(roll(x))) of the same library, and numpy's FP64
    result to the rel-L2 bar."""
    import torch
    rs = np.random.RandomState(9)
    x = rs.standard_normal((32, 48, 20)).astype(dtype)
    tol = tol_of(np.complex64 if dtype == np.float32 else np.complex128)
    lead = (0, 1)
    got = F.rfftn_shifted(to_gpu(x), shift_in="ifftshift", shift_out="fftshift").cpu().numpy()
    xr = np.fft.ifftshift(x, axes=lead)
    bits = np.fft.fftshift(F.rfftn(to_gpu(xr)).cpu().numpy(), axes=lead)
    assert np.array_equal(got, bits), "fused differs from roll(rfftn(roll))"
    want = np.fft.fftshift(np.fft.rfftn(xr.astype(np.float64)), axes=lead)
    assert rel_l2(got, want) <= tol
    got = F.rfftn_shifted(x, axes=(2, 0), norm="ortho", shift_out=[5])  # host array
    want = np.roll(np.fft.rfftn(x.astype(np.float64), axes=(2, 0), norm="ortho"), 5, axis=2)
    assert isinstance(got, np.ndarray) and rel_l2(got, want) <= tol
    torch.cuda.synchronize()


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_irfftn_shifted_matches_numpy(dtype):
    """irfftn with rolls along the leading axes (the complex passes take
    them, then the real pass), and the round trip with rfftn_shifted."""
    rs = np.random.RandomState(10)
    tol = tol_of(dtype)
    x = rs.standard_normal((24, 30, 16))
    spec = np.fft.rfftn(x).astype(dtype)
    lead = (0, 1)
    got = F.irfftn_shifted(to_gpu(spec), shift_in=[3, -4], shift_out="fftshift").cpu().numpy()
    want = np.fft.fftshift(np.fft.irfftn(np.roll(spec.astype(np.complex128), [3, -4], lead)),
                           axes=lead)
    assert rel_l2(got, want) <= tol
    xs = x.astype(np.float32 if dtype == np.complex64 else np.float64)
    fwd = F.rfftn_shifted(to_gpu(xs), shift_in="ifftshift", shift_out="fftshift")
    back = F.irfftn_shifted(fwd, shift_in="ifftshift", shift_out="fftshift").cpu().numpy()
    assert rel_l2(back, xs) <= 2 * tol
    got = F.irfftn_shifted(spec, s=(24, 30), axes=(0, 2), shift_out=[-7])  # host array
    want = np.roll(np.fft.irfftn(spec.astype(np.complex128), s=(24, 30), axes=(0, 2)), -7,
                   axis=0)
    assert isinstance(got, np.ndarray) and rel_l2(got, want) <= tol


def test_shifted_runs_fp32_kernels_in_matrix_core_mode():
    import torch
    rs = np.random.RandomState(6)
    x = crandn(rs, (8, 4096), np.complex64)
    ref = np_roll(unfused(np_roll(x, [-2048], [1]), [1], True, 1.0), [2048], [1])
    prev = R.set_precision("mfma_fp16")
    try:
        before = R.mfma_launches()
        fused, got = run(x, [1], True, 1.0, [-2048], [2048])
        assert fused and R.mfma_launches() == before
    finally:
        R.set_precision(prev)
    assert np.array_equal(got, ref)

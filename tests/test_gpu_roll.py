"""roll / fftshift / ifftshift (rocket_fft_amd.fft) on the GPU gather kernel
against numpy, bit-exact: the reference overloads numpy.roll
(rocket_fft/overloads.py:762-855) and fftshift / ifftshift (:1253-1310)."""
import numpy as np
import pytest
import torch

from rocket_fft_amd import fft as F

pytestmark = pytest.mark.gpu


def same(got, want):
    got = got.cpu().numpy() if hasattr(got, "cpu") else got
    assert got.shape == want.shape and got.dtype == want.dtype, (got.shape, got.dtype)
    assert np.array_equal(got, want)


CASES = [((7,), 3, None), ((7,), -10, None), ((4, 5), 2, None), ((4, 5), (1, 2), None),
         ((4, 5), 1, 0), ((4, 5), -3, 1), ((4, 5), (1, 2), (0, 1)), ((3, 4, 5), 7, -1),
         ((3, 4, 5), (2, -1, 9), (0, 1, 2)), ((3, 4, 5), (1, 1), (1, 1)), ((3, 4, 5), 4, (0, 2)),
         ((2, 3, 4, 5), (3, -2), (-1, 0)), ((1, 1, 6), 5, 2), ((6, 0, 3), 2, 0)]


@pytest.mark.parametrize("shape,shift,axis", CASES)
@pytest.mark.parametrize("dtype", [np.float32, np.complex128, np.int8, np.int16, np.int64,
                                   np.complex64])
def test_roll_numpy(shape, shift, axis, dtype):
    rs = np.random.RandomState(sum(shape))
    x = (rs.standard_normal(shape) * 50).astype(dtype)
    same(F.roll(x, shift, axis), np.roll(x, shift, axis))


@pytest.mark.parametrize("shape,shift,axis", CASES)
def test_roll_tensor_strided(shape, shift, axis):
    """CUDA tensors, including transposed and sliced (non-contiguous) views."""
    rs = np.random.RandomState(3)
    x = rs.standard_normal(shape).astype(np.float64)
    t = torch.from_numpy(x).cuda()
    same(F.roll(t, shift, axis), np.roll(x, shift, axis))
    if len(shape) >= 2:
        tt = t.transpose(0, -1)
        same(F.roll(tt, shift, axis), np.roll(np.swapaxes(x, 0, -1), shift, axis))
    big = torch.from_numpy(rs.standard_normal((9, 11))).cuda()
    view = big[1::2, ::3]
    same(F.roll(view, 2, 1), np.roll(view.cpu().numpy(), 2, 1))


def test_roll_edge_cases():
    assert F.roll(3.5, 1) == np.roll(3.5, 1)
    x = np.arange(12.0).reshape(3, 4)
    with pytest.raises(ValueError):
        F.roll(x, 1, axis=2)
    with pytest.raises(ValueError):
        F.roll(x, np.ones((2, 2), dtype=int))
    b = np.array([True, False, True, True])
    same(F.roll(b, 1), np.roll(b, 1))


@pytest.mark.parametrize("shape", [(5,), (6,), (4, 7), (5, 6, 3)])
@pytest.mark.parametrize("axes", [None, 0, -1, (0, -1)])
def test_fftshift_ifftshift(shape, axes):
    rs = np.random.RandomState(len(shape))
    x = rs.standard_normal(shape) + 1j * rs.standard_normal(shape)
    same(F.fftshift(x, axes), np.fft.fftshift(x, axes))
    same(F.ifftshift(x, axes), np.fft.ifftshift(x, axes))
    t = torch.from_numpy(x).cuda()
    same(F.fftshift(t, axes), np.fft.fftshift(x, axes))
    same(F.ifftshift(F.fftshift(t, axes), axes), x)

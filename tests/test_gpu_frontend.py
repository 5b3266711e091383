"""NumPy/SciPy-style front-end (rocket_fft_amd.fft) against numpy.fft and
scipy.fft, as the reference's differential suites do
(tests/test_numpy_compare.py, tests/test_scipy_compare.py): every function x
norm x axes / s, on host arrays and on CUDA tensors.  DST with a non-default
norm is compared only at the default, as there (README.md:61-65)."""
import numpy as np
import pytest
import scipy.fft as sfft
import torch

from rocket_fft_amd import fft as F

pytestmark = pytest.mark.gpu

NORMS = [None, "backward", "ortho", "forward"]


def close(got, want, tol):
    got = got.cpu().numpy() if hasattr(got, "cpu") else got
    assert got.shape == want.shape, (got.shape, want.shape)
    den = max(np.linalg.norm(want), 1e-300)
    e = np.linalg.norm(got - want) / den
    assert e <= tol, e


@pytest.fixture(params=["numpy", "torch"])
def kind(request):
    return request.param


def arr(x, kind):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda() if kind == "torch" else x


@pytest.mark.parametrize("norm", NORMS)
@pytest.mark.parametrize("fn", ["fft", "ifft"])
def test_fft_1d(fn, norm, kind):
    rs = np.random.RandomState(0)
    x = rs.standard_normal((6, 30)) + 1j * rs.standard_normal((6, 30))
    for n, axis in ((None, -1), (40, -1), (20, 0), (None, 0)):
        close(getattr(F, fn)(arr(x, kind), n=n, axis=axis, norm=norm),
              getattr(np.fft, fn)(x, n=n, axis=axis, norm=norm), 1e-13)
    xr = rs.standard_normal((6, 30))  # real input -> c2c_sym
    close(getattr(F, fn)(arr(xr, kind), norm=norm), getattr(np.fft, fn)(xr, norm=norm), 1e-13)


@pytest.mark.parametrize("norm", NORMS)
@pytest.mark.parametrize("fn", ["fft2", "ifft2", "fftn", "ifftn"])
def test_fft_nd(fn, norm, kind):
    rs = np.random.RandomState(1)
    x = rs.standard_normal((8, 6, 10)) + 1j * rs.standard_normal((8, 6, 10))
    close(getattr(F, fn)(arr(x, kind), norm=norm), getattr(np.fft, fn)(x, norm=norm), 1e-13)
    close(getattr(F, fn)(arr(x, kind), s=(5, 12), axes=(0, 2), norm=norm),
          getattr(np.fft, fn)(x, s=(5, 12), axes=(0, 2), norm=norm), 1e-13)
    x32 = x.astype(np.complex64)
    got = getattr(F, fn)(arr(x32, kind), norm=norm)
    assert str(got.dtype).endswith("complex64")
    close(got, getattr(np.fft, fn)(x32.astype(np.complex128), norm=norm), 1e-6)


@pytest.mark.parametrize("norm", NORMS)
@pytest.mark.parametrize("fn,inv", [("rfft", "irfft"), ("rfft2", "irfft2"), ("rfftn", "irfftn")])
def test_real(fn, inv, norm, kind):
    rs = np.random.RandomState(2)
    x = rs.standard_normal((8, 6, 11))
    y = getattr(F, fn)(arr(x, kind), norm=norm)
    close(y, getattr(np.fft, fn)(x, norm=norm), 1e-13)
    yh = getattr(np.fft, fn)(x, norm=norm)
    close(getattr(F, inv)(arr(yh, kind), norm=norm), getattr(np.fft, inv)(yh, norm=norm), 1e-13)
    if fn == "rfft":
        close(getattr(F, inv)(arr(yh, kind), n=15, norm=norm),
              getattr(np.fft, inv)(yh, n=15, norm=norm), 1e-13)


@pytest.mark.parametrize("norm", NORMS)
def test_hermitian(norm, kind):
    rs = np.random.RandomState(3)
    x = rs.standard_normal((5, 9)) + 1j * rs.standard_normal((5, 9))
    close(F.hfft(arr(x, kind), norm=norm), np.fft.hfft(x, norm=norm), 1e-13)
    close(F.hfft(arr(x, kind), n=12, norm=norm), np.fft.hfft(x, n=12, norm=norm), 1e-13)
    xr = rs.standard_normal((5, 16))
    close(F.ihfft(arr(xr, kind), norm=norm), np.fft.ihfft(xr, norm=norm), 1e-13)


def test_helpers():
    for n in (1, 2, 7, 8):
        assert np.allclose(F.fftfreq(n, 0.3), np.fft.fftfreq(n, 0.3))
        assert np.allclose(F.rfftfreq(n, 0.3), np.fft.rfftfreq(n, 0.3))
    x = np.arange(30).reshape(5, 6)
    assert np.array_equal(F.fftshift(x), np.fft.fftshift(x))
    assert np.array_equal(F.ifftshift(x, axes=1), np.fft.ifftshift(x, axes=1))
    t = torch.arange(30).reshape(5, 6).cuda()
    assert np.array_equal(F.fftshift(t).cpu().numpy(), np.fft.fftshift(x))
    for n in (1, 7, 13, 97, 4099, 65537):
        assert F.next_fast_len(n) == sfft.next_fast_len(n)
        assert F.next_fast_len(n, True) == sfft.next_fast_len(n, True)


@pytest.mark.parametrize("norm", NORMS)
@pytest.mark.parametrize("type", [1, 2, 3, 4])
@pytest.mark.parametrize("fn", ["dct", "idct", "dctn", "idctn"])
def test_dct(fn, type, norm, kind):
    rs = np.random.RandomState(4)
    x = rs.standard_normal((7, 9))
    close(getattr(F, fn)(arr(x, kind), type=type, norm=norm),
          getattr(sfft, fn)(x, type=type, norm=norm), 1e-13)
    xc = x + 1j * rs.standard_normal((7, 9))
    close(getattr(F, fn)(arr(xc, kind), type=type, norm=norm),
          getattr(sfft, fn)(xc, type=type, norm=norm), 1e-13)


@pytest.mark.parametrize("type", [1, 2, 3, 4])
@pytest.mark.parametrize("fn", ["dst", "idst", "dstn", "idstn"])
def test_dst_default_norm(fn, type, kind):
    rs = np.random.RandomState(5)
    x = rs.standard_normal((7, 9))
    close(getattr(F, fn)(arr(x, kind), type=type), getattr(sfft, fn)(x, type=type), 1e-13)


@pytest.mark.parametrize("norm", NORMS)
@pytest.mark.parametrize("fn", ["fft", "ifft", "rfft", "fftn", "rfftn", "irfftn"])
def test_scipy_c2c_and_real(fn, norm, kind):
    rs = np.random.RandomState(6)
    x = rs.standard_normal((6, 8))
    if fn == "irfftn":
        x = np.fft.rfftn(x)
    elif fn in ("fft", "ifft", "fftn"):
        x = x + 1j * rs.standard_normal((6, 8))
    close(getattr(F, fn)(arr(x, kind), norm=norm), getattr(sfft, fn)(x, norm=norm), 1e-13)


def test_errors():
    x = np.zeros((4, 4))
    with pytest.raises(ValueError):
        F.fft(x, norm="bogus")
    with pytest.raises(ValueError):
        F.fftn(x, axes=(0, 0))
    with pytest.raises(ValueError):
        F.fftn(x, s=(4, 4), axes=(0,))
    with pytest.raises(ValueError):
        F.dct(x, type=5)
    with pytest.raises(TypeError):
        F.rfft(x + 0j)

"""FFTLog fast Hankel transforms (rocket_fft_amd.fft.fht / ifht / fhtoffset)
against the checks of the reference's tests/test_fftlog.py: the Fortran
FFTLog tables (golden fixture), the fht -> ifht identity, the singular
cases, the exact transform of a power law (closed form), and scipy.fft.

The reference forces the last coefficient real for every n
(rocket_fft/overloads.py:877); SciPy does so only for even n, so odd lengths
are compared with a host restatement of the reference's formula instead."""
import json
import os
import warnings

import numpy as np
import pytest
import scipy.fft as sfft
import torch
from scipy.special import poch

from rocket_fft_amd import fft as F

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fftlog_fortran.json")


def host_fht(a, dln, mu, offset, bias, inverse):
    """The reference's fht / ifht (overloads.py:858-902, 1755-1809) on the
    host in FP64: weights, rfft, times u (or over conj(u)), irfft, reversed."""
    n = a.shape[-1]
    j = np.arange(n)
    jc = (n - 1) / 2
    if bias:
        a = a * (np.exp(bias * ((j - jc) * dln + offset)) if inverse
                 else np.exp(-bias * (j - jc) * dln))
    u = F._fht_coeff(n, dln, mu, offset, bias)
    A = np.fft.rfft(a, axis=-1)
    if inverse:
        with np.errstate(divide="ignore", invalid="ignore"):
            A = np.where(u == 0, 0, A / np.conj(np.where(u == 0, 1, u)))
    else:
        A = A * np.where(np.isinf(u), 0, u)
    out = np.fft.irfft(A, n, axis=-1)[..., ::-1]
    if bias:
        out = out * (np.exp(bias * (j - jc) * dln) if inverse
                     else np.exp(-bias * ((j - jc) * dln + offset)))
    return out


def test_fhtoffset_vs_scipy():
    rs = np.random.RandomState(1)
    for _ in range(20):
        dln, mu, ini, bias = rs.uniform(-1, 1), rs.uniform(-2, 2), rs.uniform(-1, 1), rs.uniform(-1, 1)
        assert abs(F.fhtoffset(dln, mu, ini, bias) - sfft.fhtoffset(dln, mu, ini, bias)) < 1e-13


def test_fht_coeff_vs_scipy_even():
    from scipy.fft._fftlog_backend import fhtcoeff
    for n, dln, mu, off, bias in ((16, 1.2, 0.3, 0.0, 0.0), (64, -0.4, 1.7, 0.5, 0.3),
                                  (42, 1.0, 1.0, 1.0, 0.0)):
        u = F._fht_coeff(n, dln, mu, off, bias)
        assert np.max(np.abs(u - fhtcoeff(n, dln, mu, off, bias))) <= 1e-13 * np.max(np.abs(u))


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(4))
def test_fht_agrees_with_fortran_fftlog(case):
    doc = json.load(open(GOLDEN))
    c = doc["cases"][case]
    mu = doc["mu"]
    r = np.logspace(-4, 4, 16)
    dln = np.log(r[1] / r[0])
    a = r ** (mu + 1) * np.exp(-r ** 2 / 2)
    offset = F.fhtoffset(dln, mu, bias=c["bias"]) if c["optimal_offset"] else 0.0
    got = F.fht(a, dln, mu, offset=offset, bias=c["bias"])
    np.testing.assert_allclose(got, c["expected"], rtol=1e-7)


@pytest.mark.gpu
@pytest.mark.parametrize("optimal", [True, False])
@pytest.mark.parametrize("offset", [0.0, 1.0, -1.0])
@pytest.mark.parametrize("bias", [0, 0.1, -0.1])
@pytest.mark.parametrize("n", [64, 63])
def test_fht_identity(n, bias, offset, optimal):
    rng = np.random.RandomState(3491349965)
    a = rng.standard_normal(n)
    dln = rng.uniform(-1, 1)
    mu = rng.uniform(-2, 2)
    if optimal:
        offset = F.fhtoffset(dln, mu, initial=offset, bias=bias)
    A = F.fht(a, dln, mu, offset=offset, bias=bias)
    np.testing.assert_allclose(A, host_fht(a, dln, mu, offset, bias, False), rtol=1e-12,
                               atol=1e-12 * np.max(np.abs(A)))
    a_ = F.ifht(A, dln, mu, offset=offset, bias=bias)
    np.testing.assert_allclose(a, a_, rtol=1e-10, atol=1e-12)


@pytest.mark.gpu
def test_fht_special_cases():
    rng = np.random.RandomState(3491349965)
    a = rng.standard_normal(64)
    dln = rng.uniform(-1, 1)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        F.fht(a, dln, -4.0, bias=1.0)   # xp, xm both non-positive integers: well defined
        F.fht(a, dln, -2.5, bias=0.5)   # only xm: well defined
    with pytest.warns(RuntimeWarning):
        F.fht(a, dln, -3.5, bias=0.5)   # only xp: singular transform
    with pytest.warns(RuntimeWarning):
        F.ifht(a, dln, -2.5, bias=0.5)  # only xm: singular inverse


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 63])
def test_fht_exact(n):
    """A power law r^gamma biased with q = gamma transforms exactly:
    A(k) = (2 / k)^gamma * poch((mu + 1 - gamma) / 2, gamma)."""
    rng = np.random.RandomState(3491349965)
    mu = rng.uniform(0, 3)
    gamma = rng.uniform(-1 - mu, 1 / 2)
    r = np.logspace(-2, 2, n)
    a = r ** gamma
    dln = np.log(r[1] / r[0])
    offset = F.fhtoffset(dln, mu, initial=0.0, bias=gamma)
    A = F.fht(a, dln, mu, offset=offset, bias=gamma)
    k = np.exp(offset) / r[::-1]
    At = (2 / k) ** gamma * poch((mu + 1 - gamma) / 2, gamma)
    np.testing.assert_allclose(A, At, rtol=1e-7)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", (np.float32, np.float64, np.uint8, np.int32, np.int64))
def test_compare_with_scipy(dtype):
    a = np.arange(42).astype(dtype)
    args = (dtype(1.0), dtype(1.0), dtype(1.0), dtype(0.0))
    for ours, theirs in ((F.fht, sfft.fht), (F.ifht, sfft.ifht)):
        got, want = ours(a, *args), theirs(a, *args)
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5 * np.max(np.abs(want)))
        assert got.dtype == want.dtype


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 1e-5)])
@pytest.mark.parametrize("n", [128, 127, 4096])
def test_fht_batched_tensor(dtype, tol, n):
    """Batched lines in HBM (last axis), even and odd n, both precisions."""
    rs = np.random.RandomState(n)
    a = rs.standard_normal((3, 5, n))
    dln, mu, off, bias = 4.0 / n, 0.7, 0.2, 0.3  # bias weights within e^{+-0.6}
    t = torch.from_numpy(a).to(dtype).cuda()
    got = F.fht(t, dln, mu, offset=off, bias=bias)
    assert got.is_cuda and got.dtype == dtype and got.shape == t.shape
    cast = np.float32 if dtype == torch.float32 else np.float64
    want = host_fht(t.cpu().numpy().astype(np.float64), *(float(cast(v)) for v in (dln, mu, off, bias)), False)
    e = np.linalg.norm(got.cpu().numpy() - want) / np.linalg.norm(want)
    assert e <= tol, e
    back = F.ifht(got, dln, mu, offset=off, bias=bias)
    e2 = float(torch.linalg.norm((back - t).double()) / torch.linalg.norm(t.double()))
    assert e2 <= 10 * tol, e2


@pytest.mark.parametrize("dtype", (np.complex64, np.complex128))
def test_raises_complex(dtype):
    a = np.arange(42).astype(dtype)
    with pytest.raises(TypeError):
        F.fht(a, 1.0, 2.0, 3.0, 4.0)
    with pytest.raises(TypeError):
        F.ifht(a, 1.0, 2.0, 3.0, 4.0)

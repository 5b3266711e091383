"""The GPU Bluestein plans pad to any length N >= 2n - 2 (rather than the
reference's good_size(2n - 1), _pocketfft_hdronly.h:2765-2767), so that
n = 2^k + 1 (65537) runs on N = 2^(k+1).  This pins the mathematics the plan
relies on, in FP64 numpy against numpy's own FFT: Bluestein's algorithm
(chirp e^{i pi m^2/n}, circular convolution of length N) gives the DFT for
N = 2n - 2 and every larger N, and fails for N = 2n - 3 (so the bound is
tight and a test of it can fail)."""
import numpy as np
import pytest


def bluestein(x, N, forward=True):
    n = len(x)
    m = np.arange(n)
    sgn = 1 if forward else -1
    # chirp with the exponent m^2 mod 2n exact in integers
    bk = np.exp(1j * np.pi * sgn * ((m * m) % (2 * n)) / n)
    a = np.zeros(N, complex)
    a[:n] = x * np.conj(bk)
    b = np.zeros(N, complex)
    b[:n] = bk
    b[N - n + 1:] = bk[1:][::-1]  # b[N - m] = b[m]; with N = 2n - 2, slot n-1 twice
    if N == 2 * n - 2:
        assert np.isclose(b[n - 1], bk[n - 1])
    y = np.fft.ifft(np.fft.fft(a) * np.fft.fft(b))[:n]
    return y * np.conj(bk)


@pytest.mark.parametrize("n", [2, 3, 5, 17, 97, 257, 1031, 4099])
def test_bluestein_any_length_from_2n_minus_2(n):
    rs = np.random.RandomState(n)
    x = rs.standard_normal(n) + 1j * rs.standard_normal(n)
    want = np.fft.fft(x)
    for N in (2 * n - 2, 2 * n - 1, 2 * n + 7, 4 * n):
        if N < n:
            continue
        got = bluestein(x, N)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), N
    inv = bluestein(x, 2 * n - 2, forward=False)
    assert np.allclose(inv, np.fft.ifft(x) * n, rtol=0, atol=1e-10 * np.abs(x).sum())


@pytest.mark.parametrize("n", [5, 17, 97, 1031])
def test_bluestein_length_bound_is_tight(n):
    rs = np.random.RandomState(n + 1)
    x = rs.standard_normal(n) + 1j * rs.standard_normal(n)
    want = np.fft.fft(x)
    N = 2 * n - 3
    a = np.zeros(N, complex)
    m = np.arange(n)
    bk = np.exp(1j * np.pi * ((m * m) % (2 * n)) / n)
    a[:n] = x * np.conj(bk)
    b = np.zeros(N, complex)
    b[:n - 1] = bk[:n - 1]
    b[N - n + 1:] = bk[1:][::-1]  # overwrites b[n - 2] with b[n - 1]'s value
    y = np.fft.ifft(np.fft.fft(a) * np.fft.fft(b))[:n] * np.conj(bk)
    # one wrong kernel value: far above FP64 rounding (1e-15), if small for
    # large n (the chirp angles of m = n-1 and n-2 differ by 3 pi / n)
    assert np.linalg.norm(y - want) > 1e-9 * np.linalg.norm(want)

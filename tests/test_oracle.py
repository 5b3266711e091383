"""Pins the oracle (oracle/pfft_oracle.c + oracle/oracle.py) before anything
is checked against it (CPU only):

* the reference engine's own outputs on seeded inputs, committed as
  tests/golden/reference_vectors.npz by tests/golden/make_golden.py;
* the reference engine itself (oracle/_ref) on fresh inputs, where built;
* SciPy's FFTW known-answer vectors for DCT/DST I-IV on linspace(0, n-1, n),
  the same data the reference's test suite uses
  (tests/test_scipy_testsuite.py:1208-1293, decimals of :1386-1423);
* closed forms and the defining O(n^2) sums in long double (brute force),
  independent of the oracle's FFT code;
* planted mistakes: a dropped term, a wrong sign and a transposed index in
  the oracle's output must fail these checks.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import ref as REF
from helpers import rel_l2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "reference_vectors.npz")
FFTW_DIR = "/usr/local/lib/python3.10/dist-packages/scipy/fftpack/tests"


def golden():
    return np.load(GOLDEN, allow_pickle=False)


def golden_cases():
    g = golden()
    for i, m in enumerate(g["meta"]):
        yield i, json.loads(str(m)), g[f"in_{i:03d}"], g[f"out_{i:03d}"]


def _tol(dtype):
    return 2e-6 if np.dtype(dtype) in (np.float32, np.complex64) else 5e-15


def test_golden_vectors():
    worst = 0.0
    for i, m, x, want in golden_cases():
        got = np.zeros_like(want)
        getattr(O, m["fname"])(x, got, np.array(m["axes"]), *m["args"])
        e = rel_l2(got, want)
        worst = max(worst, e if want.dtype in (np.float64, np.complex128) else 0.0)
        assert e <= _tol(want.dtype), f"case {i} {m}: {e:.3e}"
    assert worst < 5e-15


def test_golden_integer_facts():
    g = golden()
    for n, c, r in zip(g["good_size_n"], g["good_size_cmplx"], g["good_size_real"]):
        assert O.good_size(int(n), False) == int(c)
        assert O.good_size(int(n), True) == int(r)
    for n, c, r in zip(g["big_n"], g["big_good_cmplx"], g["big_good_real"]):
        assert O.good_size(int(n), False) == int(c)
        assert O.good_size(int(n), True) == int(r)
    for n, cf, rf, bc, br in zip(g["fac_n"], g["cfftp"], g["rfftp"], g["blue_c"], g["blue_r"]):
        assert ",".join(map(str, O.cfftp_factors(int(n)))) == str(cf)
        assert ",".join(map(str, O.rfftp_factors(int(n)))) == str(rf)
        assert O.plan_c(int(n))[1] == int(bc)
        assert O.plan_r(int(n))[1] == int(br)


@pytest.mark.skipif(not REF.available(), reason="reference engine not built")
@pytest.mark.parametrize("n", [4096, 1536, 2187, 4913, 8232, 65537])
def test_against_reference_engine(n):
    rs = np.random.RandomState(n)
    x = rs.standard_normal((2, n)) + 1j * rs.standard_normal((2, n))
    for fwd in (True, False):
        a, b = np.zeros_like(x), np.zeros_like(x)
        O.c2c(x, a, [1], fwd, 1.0)
        REF.c2c(x, b, np.array([1], np.uint64), fwd, 1.0, 1)
        assert rel_l2(a, b) < 5e-15


@pytest.mark.skipif(not REF.available(), reason="reference engine not built")
def test_plan_facts_against_reference_engine():
    for n in list(range(2, 3000)) + [65537, 131220, 1 << 20, 4099, 999983]:
        assert O.cfftp_factors(n) == REF.cfftp_factors(n)
        assert O.rfftp_factors(n) == REF.rfftp_factors(n)
        assert O.plan_c(n) == REF.plan_c(n)
        assert O.plan_r(n) == REF.plan_r(n)


def _fftw(type_, kind, n, dt):
    f = "fftw_double_ref.npz" if dt == np.float64 else "fftw_single_ref.npz"
    z = np.load(os.path.join(FFTW_DIR, f), allow_pickle=False)
    return np.linspace(0, n - 1, n).astype(dt), z[f"{kind}_{type_}_{n}"]


FFTW_SIZES = [2, 3, 4, 12, 15, 16, 17, 32, 64, 128, 256, 512, 1024]
# decimals used by the reference's suite (test_scipy_testsuite.py:1386-1423)
DEC = {("dct", 1, np.float64): 10, ("dct", 1, np.float32): 4,
       ("dct", 2, np.float64): 14, ("dct", 2, np.float32): 5,
       ("dct", 3, np.float64): 14, ("dct", 3, np.float32): 5,
       ("dct", 4, np.float64): 13, ("dct", 4, np.float32): 6,
       ("dst", 1, np.float64): 13, ("dst", 1, np.float32): 4,
       ("dst", 2, np.float64): 14, ("dst", 2, np.float32): 6,
       ("dst", 3, np.float64): 14, ("dst", 3, np.float32): 7,
       ("dst", 4, np.float64): 13, ("dst", 4, np.float32): 6}


@pytest.mark.skipif(not os.path.isdir(FFTW_DIR), reason="SciPy FFTW vectors absent")
@pytest.mark.parametrize("kind", ["dct", "dst"])
@pytest.mark.parametrize("type_", [1, 2, 3, 4])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_fftw_known_answers(kind, type_, dt):
    for n in FFTW_SIZES:
        if kind == "dct" and type_ == 1 and n < 2:
            continue
        x, want = _fftw(type_, kind, n, dt)
        got = np.zeros(n, dt)
        getattr(O, kind)(x, got, [0], type_, 1.0, False)
        # the reference suite's comparison: decimals relative to max |y|
        dec = DEC[(kind, type_, dt)]
        np.testing.assert_array_almost_equal(got / np.max(np.abs(want)),
                                             want / np.max(np.abs(want)), decimal=dec)


def _naive_dcst(x, type_, cosine):
    """Defining sums evaluated directly in long double."""
    n = len(x)
    xl = x.astype(np.longdouble)
    k = np.arange(n, dtype=np.longdouble)[:, None]
    j = np.arange(n, dtype=np.longdouble)[None, :]
    pi = np.longdouble(np.pi)
    if type_ == 1 and cosine:
        m = np.cos(pi * k * j / (n - 1)) * 2
        m[:, 0] = 1
        m[:, -1] = np.cos(pi * k[:, 0] * (n - 1) / (n - 1))
        return m @ xl
    if type_ == 1:
        return 2 * np.sin(pi * (k + 1) * (j + 1) / (n + 1)) @ xl
    if type_ == 2:
        f = np.cos if cosine else np.sin
        kk = k if cosine else k + 1
        return 2 * f(pi * kk * (2 * j + 1) / (2 * n)) @ xl
    if type_ == 3:
        if cosine:
            m = 2 * np.cos(pi * j * (2 * k + 1) / (2 * n))
            m[:, 0] = 1
        else:
            m = 2 * np.sin(pi * (j + 1) * (2 * k + 1) / (2 * n))
            m[:, -1] = (-1.0) ** k[:, 0]
        return m @ xl
    f = np.cos if cosine else np.sin
    return 2 * f(pi * (2 * k + 1) * (2 * j + 1) / (4 * n)) @ xl


@pytest.mark.parametrize("type_", [1, 2, 3, 4])
@pytest.mark.parametrize("cosine", [True, False])
@pytest.mark.parametrize("n", [2, 3, 5, 8, 13, 64])
def test_dcst_definitions(type_, cosine, n):
    rs = np.random.RandomState(n)
    x = rs.standard_normal(n)
    got = np.zeros(n)
    (O.dct if cosine else O.dst)(x, got, [0], type_, 1.0, False)
    want = _naive_dcst(x, type_, cosine).astype(np.float64)
    assert rel_l2(got, want) < 1e-14


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 49, 97, 243, 1000, 4099])
def test_c2c_brute_force(n):
    rs = np.random.RandomState(n)
    x = (rs.standard_normal((2, n)) + 1j * rs.standard_normal((2, n)))
    for fwd in (True, False):
        got = np.zeros_like(x)
        O.c2c(x, got, [1], fwd, 1.0)
        want = O.brute_dft_lines(x, fwd)
        assert rel_l2(got, want) < 2e-15


def test_closed_forms():
    n = 360
    imp = np.zeros(n, complex)
    imp[0] = 1
    got = np.zeros_like(imp)
    O.c2c(imp, got, [0], True, 1.0)
    assert np.max(np.abs(got - 1)) < 1e-15          # impulse -> constant
    const = np.ones(n, complex)
    O.c2c(const, got, [0], True, 1.0)
    want = np.zeros(n, complex)
    want[0] = n
    assert np.max(np.abs(got - want)) < 1e-12       # constant -> impulse
    k0 = 7
    tone = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    O.c2c(tone, got, [0], True, 1.0)
    want = np.zeros(n, complex)
    want[k0] = n
    assert np.max(np.abs(got - want)) < 1e-11       # pure tone -> one bin
    # linearity, Parseval, shift theorem on a Bluestein length
    rs = np.random.RandomState(0)
    m = 4099
    a = rs.standard_normal(m) + 1j * rs.standard_normal(m)
    b = rs.standard_normal(m) + 1j * rs.standard_normal(m)
    fa, fb, fab = (np.zeros(m, complex) for _ in range(3))
    O.c2c(a, fa, [0], True, 1.0)
    O.c2c(b, fb, [0], True, 1.0)
    O.c2c(2 * a - 3j * b, fab, [0], True, 1.0)
    assert rel_l2(fab, 2 * fa - 3j * fb) < 1e-14
    assert abs(np.sum(np.abs(fa) ** 2) / m - np.sum(np.abs(a) ** 2)) / np.sum(np.abs(a) ** 2) < 1e-14
    fs = np.zeros(m, complex)
    O.c2c(np.roll(a, 5), fs, [0], True, 1.0)
    assert rel_l2(fs, fa * np.exp(-2j * np.pi * 5 * np.arange(m) / m)) < 1e-13


def test_hermitian_and_hartley_identities():
    """T/test_low_level_interface.py:76-144: genuine Hartley = Re+Im of fftn,
    Hartley applied twice = N * identity."""
    rs = np.random.RandomState(3)
    a = rs.rand(32, 17, 9) - 0.5
    out = np.zeros_like(a)
    O.r2r_genuine_hartley(a, out, [0, 1, 2], 1.0)
    f = np.fft.fftn(a)
    assert rel_l2(out, f.real + f.imag) < 1e-15
    twice = np.zeros_like(a)
    O.r2r_separable_hartley(a, out, [0, 1, 2], 1.0)
    O.r2r_separable_hartley(out, twice, [0, 1, 2], 1.0 / a.size)
    assert rel_l2(twice, a) < 1e-15
    full = np.zeros(a.shape, complex)
    O.c2c_sym(a, full, [0, 1, 2], True, 1.0)
    assert rel_l2(full, f) < 1e-15


# ------------------------------------------------- planted mistakes (CPU)

def _planted(monkeypatch, fn):
    """Replaces oracle.c2c_lines with a faulty version for one check."""
    real = O.c2c_lines
    monkeypatch.setattr(O, "c2c_lines", lambda x2, fwd, fct: fn(real(x2, fwd, fct)))


@pytest.mark.parametrize("mistake", ["drop_term", "wrong_sign", "transpose"])
def test_planted_mistakes_are_caught(monkeypatch, mistake):
    if mistake == "drop_term":
        _planted(monkeypatch, lambda y: y - y[:, :1] * 0 - np.where(
            np.arange(y.shape[1]) == 1, y, 0) * 1e-3)
    elif mistake == "wrong_sign":
        _planted(monkeypatch, lambda y: np.conj(y))
    else:
        _planted(monkeypatch, lambda y: y[:, ::-1].copy())
    caught = False
    for i, m, x, want in golden_cases():
        if m["fname"] != "c2c" or x.shape[-1] < 8:
            continue
        got = np.zeros_like(want)
        O.c2c(x, got, np.array(m["axes"]), *m["args"])
        if rel_l2(got, want) > _tol(want.dtype):
            caught = True
            break
    assert caught, f"planted '{mistake}' not detected by the golden vectors"

"""CPU checks of the product library (no GPU compute):

* librocketfft_hip.so loads and exports every function include/rocket_fft_hip.h
  declares (the ten numba_* symbols of rocket_fft/_pocketfft_numba.cpp plus the
  rfh_* additions);
* the host plan layer is bit-exact against the oracle and the reference's
  golden facts: good_size (H:588-650), FFTPACK factor orders (H:1750-1778,
  H:2629-2652), Bluestein decision and n2 (H:2843-2861, H:2885-2903);
* argument typing mirrors rocket_fft/pocketfft.py:147-178 and
  tests/test_low_level_interface.py:200-325 (TypeError where the reference
  raises TypingError), before anything reaches the device.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
import ref as REF
import rocket_fft_amd as R

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "rocket_fft_hip.h")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "reference_vectors.npz")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:numba|rfh)_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_declared_symbols():
    lib = ctypes.CDLL(R.library_path())
    syms = declared_symbols()
    assert len([s for s in syms if s.startswith("numba_")]) == 10
    for s in syms:
        assert hasattr(lib, s), f"{s} not exported"


def test_good_size_bit_exact():
    for n in range(0, 30000):
        assert R.good_size(n) == O.good_size(n, False)
        assert R.good_size(n, True) == O.good_size(n, True)
    rs = np.random.RandomState(0)
    for n in rs.randint(1, 1 << 40, size=300):
        assert R.good_size(int(n)) == O.good_size(int(n), False)
        assert R.good_size(int(n), True) == O.good_size(int(n), True)


def test_good_size_golden():
    g = np.load(GOLDEN, allow_pickle=False)
    for n, c, r in zip(g["good_size_n"], g["good_size_cmplx"], g["good_size_real"]):
        assert R.good_size(int(n)) == int(c) and R.good_size(int(n), True) == int(r)
    for n, c, r in zip(g["big_n"], g["big_good_cmplx"], g["big_good_real"]):
        assert R.good_size(int(n)) == int(c) and R.good_size(int(n), True) == int(r)


def test_plan_facts_bit_exact():
    g = np.load(GOLDEN, allow_pickle=False)
    for n, cf, rf, bc, br in zip(g["fac_n"], g["cfftp"], g["rfftp"], g["blue_c"], g["blue_r"]):
        f = R.plan_facts(int(n))
        assert ",".join(map(str, f["cfftp"])) == str(cf)
        assert ",".join(map(str, f["rfftp"])) == str(rf)
        assert f["blue_c"][1] == int(bc) and f["blue_r"][1] == int(br)
    for n in list(range(2049, 12000)) + [65537, 131220, 1 << 20, 4099, 999983, 2 ** 31 - 1]:
        f = R.plan_facts(n)
        assert f["cfftp"] == O.cfftp_factors(n)
        assert f["rfftp"] == O.rfftp_factors(n)
        assert f["blue_c"] == O.plan_c(n) and f["blue_r"] == O.plan_r(n)


@pytest.mark.skipif(not REF.available(), reason="reference engine not built")
def test_plan_facts_against_reference_engine():
    for n in [4096, 1000, 1536, 2187, 4913, 4099, 65537, 1 << 20, 256, 512, 64, 65536, 2011]:
        f = R.plan_facts(n)
        assert f["cfftp"] == REF.cfftp_factors(n)
        assert f["blue_c"] == REF.plan_c(n)
        assert f["blue_r"] == REF.plan_r(n)


# ---------------------------------------------------------------- typing

def _arrays():
    a = np.zeros((4, 4), np.complex128)
    return a, np.zeros_like(a), np.array([0, 1])


@pytest.mark.parametrize("axes_shape", ((1, 1), (1, 1, -1), ()))
def test_typing_raise_axes_shape(axes_shape):
    a = np.zeros(42, np.complex128)
    with pytest.raises(TypeError):
        R.c2c(a, a, np.array(0, dtype=np.uint64).reshape(axes_shape), True, 1.0, 1)


@pytest.mark.parametrize("forward", (1, 1.0, np.int8(1), np.uint8(1)))
def test_typing_raise_forward(forward):
    a = np.zeros(42, np.complex128)
    with pytest.raises(TypeError):
        R.c2c(a, a, np.array([0], np.uint64), forward, 1.0, 1)


@pytest.mark.parametrize("fct", (np.int64(1), np.int32(1), 1, True, 1j))
def test_typing_raise_fct(fct):
    a = np.zeros(42, np.complex128)
    with pytest.raises(TypeError):
        R.c2c(a, a, np.array([0], np.uint64), True, fct, 1)


@pytest.mark.parametrize("nthreads", (True, 1.0))
def test_typing_raise_nthreads(nthreads):
    a = np.zeros(42, np.complex128)
    with pytest.raises(TypeError):
        R.c2c(a, a, np.array([0], np.uint64), True, 1.0, nthreads)


def test_typing_raise_ndim():
    a = np.zeros(42, np.complex128)
    ax = np.array([0], np.uint64)
    with pytest.raises(TypeError):
        R.c2c(a.reshape(1, 42), a, ax, True, 1.0, 1)
    with pytest.raises(TypeError):
        R.c2c(a, a.reshape(1, 42), ax, True, 1.0, 1)
    with pytest.raises(ValueError):
        R.c2c(a, a, np.array([1]), True, 1.0, 1)


@pytest.mark.parametrize("n", (1.0, 1j))
@pytest.mark.parametrize("real", (1.0, 1j))
def test_good_size_raise(n, real):
    with pytest.raises(TypeError):
        R.good_size(n, real)


@pytest.mark.parametrize("v", (True, np.int8(42), np.uint8(42), np.int16(42), np.uint16(42),
                               np.int32(42), np.uint32(42), np.int64(42), np.uint64(42)))
def test_good_size_noraise(v):
    assert R.good_size(v, v) == O.good_size(int(v), bool(v))


def test_dct_type_range():
    a = np.zeros((4, 4))
    with pytest.raises(ValueError):
        R.dct(a, a, np.array([0]), 5, 1.0, False, 1)
    with pytest.raises(TypeError):
        R.dct(a, a, np.array([0]), 2.0, 1.0, False, 1)


def test_aliases():
    assert R.separable_hartley is R.r2r_separable_hartley
    assert R.genuine_hartley is R.r2r_genuine_hartley
    assert R.fftpack is R.r2r_fftpack

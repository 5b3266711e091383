"""ctypes front of the reference engine built by ``oracle/Makefile`` into
``oracle/_ref/libref_pocketfft.so``.

TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).  The library is the
reference's own ``rocket_fft/_pocketfft_hdronly.h`` compiled with the
reference's flags; ``oracle/ref_driver.cpp`` restates the dispatch of
``rocket_fft/_pocketfft_numba.cpp:25-223`` with plain pointers.  Every function
here takes NumPy arrays with the reference's low-level signature
(``rocket_fft/__init__.pyi:6-105``) and writes ``aout`` in place.

``available()`` is False wherever the library was not built (a checkout
without the reference sources); callers skip then.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "_ref", "libref_pocketfft.so")
_lib = None

_u64 = ctypes.c_uint64
_ptr = ctypes.c_void_p
_dbl = ctypes.c_double
_int = ctypes.c_int
_ARGS_CMPLX = [_int, _u64, _ptr, _ptr, _ptr, _ptr, _u64, _int, _dbl, _u64, _ptr, _ptr]
_ARGS_REAL = [_int, _u64, _ptr, _ptr, _ptr, _ptr, _u64, _u64, _dbl, _int, _u64, _ptr, _ptr]


def available():
    return os.path.exists(_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"reference engine not built: {_PATH}")
        L = ctypes.CDLL(_PATH)
        for name in ("ref_c2c", "ref_r2c", "ref_c2r", "ref_c2c_sym"):
            getattr(L, name).argtypes = _ARGS_CMPLX
            getattr(L, name).restype = _int
        for name in ("ref_dct", "ref_dst"):
            getattr(L, name).argtypes = _ARGS_REAL
            getattr(L, name).restype = _int
        L.ref_r2r_fftpack.argtypes = [_int, _u64, _ptr, _ptr, _ptr, _ptr, _u64, _int, _int,
                                      _dbl, _u64, _ptr, _ptr]
        L.ref_r2r_fftpack.restype = _int
        L.ref_hartley.argtypes = [_int, _int, _u64, _ptr, _ptr, _ptr, _ptr, _u64, _dbl, _u64,
                                  _ptr, _ptr]
        L.ref_hartley.restype = _int
        L.ref_good_size.argtypes = [_u64, _int]
        L.ref_good_size.restype = _u64
        L.ref_largest_prime_factor.argtypes = [_u64]
        L.ref_largest_prime_factor.restype = _u64
        L.ref_cost_guess.argtypes = [_u64]
        L.ref_cost_guess.restype = _dbl
        for name in ("ref_cfftp_factors", "ref_rfftp_factors"):
            getattr(L, name).argtypes = [_u64, _ptr, _int]
            getattr(L, name).restype = _int
        for name in ("ref_plan_c", "ref_plan_r"):
            getattr(L, name).argtypes = [_u64, ctypes.POINTER(_u64)]
            getattr(L, name).restype = _int
        L.ref_twiddle.argtypes = [_u64, _u64, _int, _ptr]
        L.ref_twiddle.restype = None
        _lib = L
    return _lib


def _geom(shape_arr, ain, aout, axes):
    nd = shape_arr.ndim
    shape = np.array(shape_arr.shape, dtype=np.intp)
    sin = np.array(ain.strides, dtype=np.intp)
    sout = np.array(aout.strides, dtype=np.intp)
    ax = np.ascontiguousarray(np.asarray(axes), dtype=np.uint64)
    keep = (shape, sin, sout, ax)
    return nd, keep


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _is_dbl_c(a):
    return a.dtype.itemsize == 16


def _is_dbl_r(a):
    return a.dtype.itemsize == 8


def _call(fn, dbl, shape_arr, ain, aout, axes, *mid):
    nd, (shape, sin, sout, ax) = _geom(shape_arr, ain, aout, axes)
    rc = fn(int(dbl), nd, _p(shape), _p(sin), _p(sout), _p(ax), ax.size, *mid, _p(ain), _p(aout))
    if rc != 0:
        raise ValueError("reference engine raised")
    return aout


def c2c(ain, aout, axes, forward, fct, nthreads=1):
    return _call(lib().ref_c2c, _is_dbl_c(ain), ain, ain, aout, axes, int(forward), float(fct),
                 int(nthreads))


def r2c(ain, aout, axes, forward, fct, nthreads=1):
    return _call(lib().ref_r2c, _is_dbl_r(ain), ain, ain, aout, axes, int(forward), float(fct),
                 int(nthreads))


def c2r(ain, aout, axes, forward, fct, nthreads=1):
    return _call(lib().ref_c2r, _is_dbl_c(ain), aout, ain, aout, axes, int(forward), float(fct),
                 int(nthreads))


def c2c_sym(ain, aout, axes, forward, fct, nthreads=1):
    return _call(lib().ref_c2c_sym, _is_dbl_r(ain), ain, ain, aout, axes, int(forward),
                 float(fct), int(nthreads))


def dct(ain, aout, axes, type, fct, ortho, nthreads=1):
    return _call(lib().ref_dct, _is_dbl_r(ain), ain, ain, aout, axes, int(type), float(fct),
                 int(ortho), int(nthreads))


def dst(ain, aout, axes, type, fct, ortho, nthreads=1):
    return _call(lib().ref_dst, _is_dbl_r(ain), ain, ain, aout, axes, int(type), float(fct),
                 int(ortho), int(nthreads))


def r2r_fftpack(ain, aout, axes, real2hermitian, forward, fct, nthreads=1):
    return _call(lib().ref_r2r_fftpack, _is_dbl_r(ain), ain, ain, aout, axes,
                 int(real2hermitian), int(forward), float(fct), int(nthreads))


def r2r_separable_hartley(ain, aout, axes, fct, nthreads=1):
    nd, (shape, sin, sout, ax) = _geom(ain, ain, aout, axes)
    rc = lib().ref_hartley(int(_is_dbl_r(ain)), 0, nd, _p(shape), _p(sin), _p(sout), _p(ax),
                           ax.size, float(fct), int(nthreads), _p(ain), _p(aout))
    assert rc == 0
    return aout


def r2r_genuine_hartley(ain, aout, axes, fct, nthreads=1):
    nd, (shape, sin, sout, ax) = _geom(ain, ain, aout, axes)
    rc = lib().ref_hartley(int(_is_dbl_r(ain)), 1, nd, _p(shape), _p(sin), _p(sout), _p(ax),
                           ax.size, float(fct), int(nthreads), _p(ain), _p(aout))
    assert rc == 0
    return aout


def good_size(n, real):
    return int(lib().ref_good_size(int(n), int(bool(real))))


def largest_prime_factor(n):
    return int(lib().ref_largest_prime_factor(int(n)))


def cost_guess(n):
    return float(lib().ref_cost_guess(int(n)))


def cfftp_factors(n):
    buf = np.zeros(64, dtype=np.uint64)
    k = lib().ref_cfftp_factors(int(n), _p(buf), 64)
    return [int(x) for x in buf[:k]]


def rfftp_factors(n):
    buf = np.zeros(64, dtype=np.uint64)
    k = lib().ref_rfftp_factors(int(n), _p(buf), 64)
    return [int(x) for x in buf[:k]]


def plan_c(n):
    """(uses_bluestein, n2) of the complex plan for length n."""
    n2 = _u64(0)
    b = lib().ref_plan_c(int(n), ctypes.byref(n2))
    return bool(b), int(n2.value)


def plan_r(n):
    n2 = _u64(0)
    b = lib().ref_plan_r(int(n), ctypes.byref(n2))
    return bool(b), int(n2.value)


def twiddle(n, count=None, dbl=True):
    """sincos_2pibyn<T>(n)[k] = e^{+2 pi i k/n} for k < count, as complex128."""
    count = n if count is None else count
    out = np.zeros(2 * count, dtype=np.float64)
    lib().ref_twiddle(int(n), int(count), int(bool(dbl)), _p(out))
    return out[0::2] + 1j * out[1::2]

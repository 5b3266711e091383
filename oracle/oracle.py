"""N-D driver of the C restatement oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg may use it, as the checker.  The product never imports it.

The functions mirror rocket-fft's low-level API (``rocket_fft/__init__.pyi:6-105``)
and write into ``aout``.  They restate the N-D semantics of the reference
engine (``rocket_fft/_pocketfft_hdronly.h``, "H") on top of the 1-D routines of
``pfft_oracle.c``:

* c2c / dct / dst / r2r: one axis after another, the first reading ``ain``
  and later ones the running result; ``fct`` applied on the first axis only
  (H:3568-3607).  An empty axes list leaves ``aout`` untouched.
* r2c: real transform along ``axes[-1]`` (conjugated when not forward), then
  c2c over the remaining axes (H:3955-3975, H:3723-3779).
* c2r: c2c over the leading axes, then the halfcomplex back-transform along
  ``axes[-1]``, which reads only the real part of bins 0 and n/2
  (H:3995-4023, H:3780-3852).
* c2c_sym: r2c, then the Hermitian completion of numba_c2c_sym
  (rocket_fft/_pocketfft_numba.cpp:111-143, rev_iter H:3383-3444) with the
  exact overwrite order of its C-order loop.
* Hartley: H:3661-3705 (separable), H:4061-4091 (genuine).
* r2r_fftpack: sign flips of ExecR2R (H:3854-3873).

Arithmetic is long double inside the C library; results are rounded to
``aout``'s dtype at the end.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u64 = ctypes.c_uint64
_p = ctypes.c_void_p
_d = ctypes.c_double
_i = ctypes.c_int


def build():
    """Compile liboracle.so (gcc, seconds)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            build()
        L = ctypes.CDLL(_PATH)
        L.or_c2c.argtypes = [_u64, _u64, _p, _p, _i, _d]
        L.or_c2c.restype = None
        L.or_r2hc.argtypes = [_u64, _u64, _p, _p, _d]
        L.or_r2hc.restype = None
        L.or_hc2r.argtypes = [_u64, _u64, _p, _p, _d]
        L.or_hc2r.restype = None
        L.or_dcst.argtypes = [_u64, _u64, _p, _p, _i, _i, _i, _d]
        L.or_dcst.restype = _i
        L.or_brute_dft.argtypes = [_u64, _u64, _p, _p, _i]
        L.or_brute_dft.restype = None
        for name in ("or_good_size_cmplx", "or_good_size_real", "or_largest_prime_factor"):
            getattr(L, name).argtypes = [_u64]
            getattr(L, name).restype = _u64
        L.or_cost_guess.argtypes = [_u64]
        L.or_cost_guess.restype = _d
        for name in ("or_cfftp_factors", "or_rfftp_factors"):
            getattr(L, name).argtypes = [_u64, _p, _i]
            getattr(L, name).restype = _i
        for name in ("or_plan_c", "or_plan_r"):
            getattr(L, name).argtypes = [_u64, ctypes.POINTER(_u64)]
            getattr(L, name).restype = _i
        L.or_twiddle.argtypes = [_u64, _u64, _p]
        L.or_twiddle.restype = None
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


# ------------------------------------------------------------ integer facts

def good_size(n, real):
    n = int(n)
    return int(lib().or_good_size_real(n) if real else lib().or_good_size_cmplx(n))


def largest_prime_factor(n):
    return int(lib().or_largest_prime_factor(int(n)))


def cost_guess(n):
    return float(lib().or_cost_guess(int(n)))


def cfftp_factors(n):
    buf = np.zeros(64, dtype=np.uint64)
    k = lib().or_cfftp_factors(int(n), _ptr(buf), 64)
    return [int(x) for x in buf[:k]]


def rfftp_factors(n):
    buf = np.zeros(64, dtype=np.uint64)
    k = lib().or_rfftp_factors(int(n), _ptr(buf), 64)
    return [int(x) for x in buf[:k]]


def plan_c(n):
    n2 = _u64(0)
    b = lib().or_plan_c(int(n), ctypes.byref(n2))
    return bool(b), int(n2.value)


def plan_r(n):
    n2 = _u64(0)
    b = lib().or_plan_r(int(n), ctypes.byref(n2))
    return bool(b), int(n2.value)


def twiddle(n, count=None):
    count = n if count is None else count
    out = np.zeros(2 * count)
    lib().or_twiddle(int(n), int(count), _ptr(out))
    return out[0::2] + 1j * out[1::2]


# ------------------------------------------------------------- 1-D kernels

def _lines(x, ax):
    """Move axis ``ax`` last and flatten the rest: (contiguous 2-D, restore)."""
    xm = np.moveaxis(x, ax, -1)
    lead = xm.shape[:-1]
    flat = np.ascontiguousarray(xm).reshape(-1, xm.shape[-1])

    def restore(y2):
        return np.moveaxis(y2.reshape(lead + (y2.shape[-1],)), -1, ax)

    return flat, restore


def c2c_lines(x2, forward, fct):
    x2 = np.ascontiguousarray(x2, dtype=np.complex128)
    out = np.empty_like(x2)
    if x2.size:
        lib().or_c2c(x2.shape[0], x2.shape[1], _ptr(x2), _ptr(out), int(bool(forward)),
                     float(fct))
    return out


def r2hc_lines(x2, fct):
    x2 = np.ascontiguousarray(x2, dtype=np.float64)
    out = np.empty_like(x2)
    if x2.size:
        lib().or_r2hc(x2.shape[0], x2.shape[1], _ptr(x2), _ptr(out), float(fct))
    return out


def hc2r_lines(x2, fct):
    x2 = np.ascontiguousarray(x2, dtype=np.float64)
    out = np.empty_like(x2)
    if x2.size:
        lib().or_hc2r(x2.shape[0], x2.shape[1], _ptr(x2), _ptr(out), float(fct))
    return out


def dcst_lines(x2, type, cosine, ortho, fct):
    x2 = np.ascontiguousarray(x2, dtype=np.float64)
    out = np.empty_like(x2)
    if x2.size:
        rc = lib().or_dcst(x2.shape[0], x2.shape[1], _ptr(x2), _ptr(out), int(type),
                           int(bool(cosine)), int(bool(ortho)), float(fct))
        if rc != 0:
            raise ValueError("invalid DCT/DST type or length")
    return out


def brute_dft_lines(x2, forward):
    x2 = np.ascontiguousarray(x2, dtype=np.complex128)
    out = np.empty_like(x2)
    lib().or_brute_dft(x2.shape[0], x2.shape[1], _ptr(x2), _ptr(out), int(bool(forward)))
    return out


def hc_to_complex(hc, n, conj=False):
    """FFTPACK halfcomplex rows -> n//2+1 complex bins (H:3748-3755)."""
    h = n // 2 + 1
    out = np.zeros(hc.shape[:-1] + (h,), dtype=np.complex128)
    out[..., 0] = hc[..., 0]
    m = (n - 1) // 2
    sgn = -1.0 if conj else 1.0
    out[..., 1:m + 1] = hc[..., 1:2 * m:2] + 1j * sgn * hc[..., 2:2 * m + 1:2]
    if n % 2 == 0 and n > 1:
        out[..., n // 2] = hc[..., n - 1]
    return out


def complex_to_hc(c, n, conj=False):
    """n//2+1 complex bins -> halfcomplex; bins 0 and n/2 keep only their real
    part (H:3826-3841)."""
    hc = np.zeros(c.shape[:-1] + (n,), dtype=np.float64)
    hc[..., 0] = c[..., 0].real
    m = (n - 1) // 2
    sgn = -1.0 if conj else 1.0
    hc[..., 1:2 * m:2] = c[..., 1:m + 1].real
    hc[..., 2:2 * m + 1:2] = sgn * c[..., 1:m + 1].imag
    if n % 2 == 0 and n > 1:
        hc[..., n - 1] = c[..., n // 2].real
    return hc


# -------------------------------------------------------------- N-D drivers

def _axes(axes):
    return [int(a) for a in np.asarray(axes).reshape(-1)]


def _store(aout, val):
    if np.iscomplexobj(aout):
        aout[...] = val
    else:
        aout[...] = np.real(val)
    return aout


def _c2c_nd(x, axes, forward, fct):
    for i, ax in enumerate(axes):
        flat, restore = _lines(x, ax)
        x = restore(c2c_lines(flat, forward, fct if i == 0 else 1.0))
    return x


def c2c(ain, aout, axes, forward, fct, nthreads=1):
    axes = _axes(axes)
    if 0 in ain.shape or not axes:
        return aout
    return _store(aout, _c2c_nd(ain.astype(np.complex128), axes, forward, fct))


def _r2c_nd(ain, axes, forward, fct):
    last = axes[-1]
    n = ain.shape[last]
    flat, restore = _lines(ain.astype(np.float64), last)
    y = restore(hc_to_complex(r2hc_lines(flat, fct), n, conj=not forward))
    return _c2c_nd(y, axes[:-1], forward, 1.0)


def r2c(ain, aout, axes, forward, fct, nthreads=1):
    axes = _axes(axes)
    if 0 in ain.shape:
        return aout
    return _store(aout, _r2c_nd(ain, axes, forward, fct))


def c2r(ain, aout, axes, forward, fct, nthreads=1):
    axes = _axes(axes)
    if 0 in aout.shape:
        return aout
    last = axes[-1]
    n = aout.shape[last]
    x = np.take(ain.astype(np.complex128), np.arange(n // 2 + 1), axis=last)
    x = _c2c_nd(x, axes[:-1], forward, 1.0)
    flat, restore = _lines(x, last)
    y = hc2r_lines(complex_to_hc(flat, n, conj=forward), fct)
    return _store(aout, restore(y))


def _plane_order(shape, axes, last, k):
    """For the plane ``index[last] == k``: C-order rank of each position and
    of its mirror (every axis in ``axes`` reflected, (n - i) mod n)."""
    plane_shape = tuple(s for d, s in enumerate(shape) if d != last)
    dims = [d for d in range(len(shape)) if d != last]
    grids = np.indices(plane_shape)
    rev = [((-g) % shape[d]) if d in axes else g for g, d in zip(grids, dims)]
    lin = np.ravel_multi_index(tuple(grids), plane_shape) if plane_shape else np.array(0)
    rlin = np.ravel_multi_index(tuple(rev), plane_shape) if plane_shape else np.array(0)
    return lin, rlin


def _mirror(a, axes, skip_last=True):
    """a[rev(idx)] over the transformed axes (all but the last)."""
    for d in set(axes[:-1]) if skip_last else set(axes):
        a = np.take(a, (-np.arange(a.shape[d])) % a.shape[d], axis=d)
    return a


def _sym_fill(R, shape, axes, combine_upper, combine_plane):
    last = axes[-1]
    n = shape[last]
    h = n // 2 + 1
    out = np.empty(shape, dtype=R.dtype)
    idx_last = [slice(None)] * len(shape)
    idx_last[last] = slice(0, h)
    out[tuple(idx_last)] = R
    if h < n:
        Rm = _mirror(R, axes)
        src = np.take(Rm, n - np.arange(h, n), axis=last)
        idx_last[last] = slice(h, n)
        out[tuple(idx_last)] = combine_upper(src)
    planes = [0] + ([n // 2] if n % 2 == 0 and n > 1 else [])
    for k in planes:
        P = np.take(R, k, axis=last)
        lin, rlin = _plane_order(shape, axes, last, k)
        flatP = P.reshape(-1)
        Pr = flatP[rlin.reshape(-1)].reshape(P.shape)
        val = combine_plane(P, Pr, lin > rlin, lin == rlin)
        sl = [slice(None)] * len(shape)
        sl[last] = k
        out[tuple(sl)] = val
    return out


def c2c_sym(ain, aout, axes, forward, fct, nthreads=1):
    axes = _axes(axes)
    if 0 in ain.shape:
        return aout
    R = _r2c_nd(ain, axes, forward, fct)

    def plane(P, Pr, larger, fixed):
        return np.where(fixed, np.conj(P), np.where(larger, np.conj(Pr), P))

    return _store(aout, _sym_fill(R, ain.shape, axes, np.conj, plane))


def _dcst_nd(ain, axes, type, ortho, fct, cosine):
    x = ain.astype(np.float64)
    for i, ax in enumerate(axes):
        flat, restore = _lines(x, ax)
        x = restore(dcst_lines(flat, type, cosine, ortho, fct if i == 0 else 1.0))
    return x


def dct(ain, aout, axes, type, fct, ortho, nthreads=1):
    axes = _axes(axes)
    if type < 1 or type > 4:
        raise ValueError("invalid DCT type")
    if 0 in ain.shape or not axes:
        return aout
    return _store(aout, _dcst_nd(ain, axes, type, ortho, fct, True))


def dst(ain, aout, axes, type, fct, ortho, nthreads=1):
    axes = _axes(axes)
    if type < 1 or type > 4:
        raise ValueError("invalid DST type")
    if 0 in ain.shape or not axes:
        return aout
    return _store(aout, _dcst_nd(ain, axes, type, ortho, fct, False))


def r2r_fftpack(ain, aout, axes, real2hermitian, forward, fct, nthreads=1):
    axes = _axes(axes)
    if 0 in ain.shape or not axes:
        return aout
    x = ain.astype(np.float64)
    for i, ax in enumerate(axes):
        flat, restore = _lines(x, ax)
        f = fct if i == 0 else 1.0
        buf = flat.copy()
        if (not real2hermitian) and forward:
            buf[:, 2::2] *= -1
        buf = r2hc_lines(buf, f) if forward else hc2r_lines(buf, f)
        if real2hermitian and not forward:
            buf[:, 2::2] *= -1
        x = restore(buf)
    return _store(aout, x)


def _hartley_lines(flat, fct):
    n = flat.shape[1]
    hc = r2hc_lines(flat, fct)
    out = np.empty_like(hc)
    out[:, 0] = hc[:, 0]
    m = (n - 1) // 2
    k = np.arange(1, m + 1)
    out[:, k] = hc[:, 2 * k - 1] + hc[:, 2 * k]
    out[:, n - k] = hc[:, 2 * k - 1] - hc[:, 2 * k]
    if n % 2 == 0 and n > 1:
        out[:, n // 2] = hc[:, n - 1]
    return out


def r2r_separable_hartley(ain, aout, axes, fct, nthreads=1):
    axes = _axes(axes)
    if 0 in ain.shape or not axes:
        return aout
    x = ain.astype(np.float64)
    for i, ax in enumerate(axes):
        flat, restore = _lines(x, ax)
        x = restore(_hartley_lines(flat, fct if i == 0 else 1.0))
    return _store(aout, x)


def r2r_genuine_hartley(ain, aout, axes, fct, nthreads=1):
    axes = _axes(axes)
    if 0 in ain.shape:
        return aout
    if len(axes) == 1:
        return r2r_separable_hartley(ain, aout, axes, fct, nthreads)
    T = _r2c_nd(ain, axes, True, fct)

    def upper(src):
        return src.real - src.imag

    def plane(P, Pr, larger, fixed):
        return np.where(fixed, P.real - P.imag,
                        np.where(larger, P.real + P.imag, Pr.real - Pr.imag))

    full = _sym_fill(T, ain.shape, axes, upper, plane)
    # half-region positions outside the planes: re + im of T
    last = axes[-1]
    n = ain.shape[last]
    res = np.real(full).copy()
    mid = [slice(None)] * ain.ndim
    hi = n // 2 if n % 2 == 0 else n // 2 + 1
    mid[last] = slice(1, hi)
    Tm = T[tuple(mid)]
    res[tuple(mid)] = Tm.real + Tm.imag
    return _store(aout, res)

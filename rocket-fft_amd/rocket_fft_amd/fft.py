"""NumPy- and SciPy-style FFT functions over the MI355X engine.

The interface of ``numpy.fft`` and ``scipy.fft`` as rocket-fft overloads it
(``rocket_fft/overloads.py``), built on this package's low-level calls, with
the same rules:

* axes and ``s`` (``overloads.py:421-489``): a single ``axis``; ``axes``
  wrapped and checked unique; ``s`` without ``axes`` selects the last
  ``len(s)`` axes;
* ``norm`` -> ``fct`` (``:513-552``): backward / ortho / forward, counted on
  the input shape (c2c, r2c) or the output shape (c2r); DCT/DST count
  ``2 (n + delta)`` per axis with delta = -1 (DCT-I), +1 (DST-I), 0 otherwise;
* zero padding / cropping to ``s`` and dtype promotion (``:48-77``,
  ``:575-609``): float32/complex64 stay single precision, everything else
  becomes float64/complex128;
* real input to a complex transform goes through ``c2c_sym``; ``hfft`` /
  ``ihfft`` are c2r / r2c with the direction flipped (``:1202-1213``);
* DCT/DST: inverse types map 2 <-> 3 (``:680-712``), ``ortho`` comes from
  ``orthogonalize`` or ``norm == "ortho"`` (``:715-732``), complex input is
  transformed as real and imaginary parts.

Arrays may be NumPy arrays (host memory, staged through the GPU) or torch
CUDA tensors (computed in HBM); results are of the same kind.
"""
import numbers

import numpy as np

from . import c2c, c2c_sym, c2r, dct as _dct_ll, dst as _dst_ll, good_size, r2c

__all__ = [
    "fft", "ifft", "fft2", "ifft2", "fftn", "ifftn", "rfft", "irfft", "rfft2", "irfft2",
    "rfftn", "irfftn", "hfft", "ihfft", "fftfreq", "rfftfreq", "fftshift", "ifftshift",
    "dct", "idct", "dst", "idst", "dctn", "idctn", "dstn", "idstn", "next_fast_len",
]


# ---------------------------------------------------------------- arrays

def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _dtype_pair(a):
    """(real dtype, complex dtype) names for the transform of a."""
    name = str(a.dtype).replace("torch.", "")
    single = name in ("float32", "complex64")
    return ("float32", "complex64") if single else ("float64", "complex128")


def _is_complex(a):
    return a.is_complex() if _is_torch(a) else np.iscomplexobj(a)


def _empty(like, shape, dtype):
    if _is_torch(like):
        import torch
        return torch.empty(tuple(shape), dtype=getattr(torch, dtype), device=like.device)
    return np.empty(tuple(shape), dtype=dtype)


def _zeros(like, shape, dtype):
    if _is_torch(like):
        import torch
        return torch.zeros(tuple(shape), dtype=getattr(torch, dtype), device=like.device)
    return np.zeros(tuple(shape), dtype=dtype)


def _cast(a, dtype):
    if _is_torch(a):
        import torch
        t = getattr(torch, dtype)
        return a if a.dtype == t else a.to(t)
    return a if a.dtype == np.dtype(dtype) else a.astype(dtype)


def _asarray(a):
    if _is_torch(a) or isinstance(a, np.ndarray):
        return a
    return np.asarray(a)


# ----------------------------------------------------------- axes and s

def _wrap(ax, ndim):
    ax = int(ax)
    if ax < -ndim or ax >= ndim:
        raise ValueError("axis exceeds dimensionality of input")
    return ax % ndim


def _shape_and_axes(a, s, axes):
    nd = a.ndim
    if s is not None and isinstance(s, numbers.Integral):
        s = [s]
    if axes is not None and isinstance(axes, numbers.Integral):
        axes = [axes]
    if s is not None:
        s = [int(v) for v in s]
        if any(v < 1 for v in s):
            raise ValueError("Invalid number of data points specified")
    if axes is None:
        if s is None:
            axes = list(range(nd))
        else:
            if len(s) > nd:
                raise ValueError("Shape requires more axes than are present.")
            axes = list(range(nd - len(s), nd))
    else:
        axes = [_wrap(ax, nd) for ax in axes]
        if len(set(axes)) != len(axes):
            raise ValueError("All axes must be unique.")
        if s is not None and len(s) != len(axes):
            raise ValueError("When given, axes and shape arguments have to be of the same length.")
    return s, axes


def _pad_crop(a, s, axes, dtype):
    """Zero-pad / crop a to s on axes and cast to dtype (overloads.py:575-609)."""
    if s is None:
        return _cast(a, dtype)
    shape = list(a.shape)
    for n, ax in zip(s, axes):
        shape[ax] = n
    if list(a.shape) == shape:
        return _cast(a, dtype)
    out = _zeros(a, shape, dtype)
    idx = tuple(slice(0, min(p, q)) for p, q in zip(a.shape, shape))
    out[idx] = _cast(a, dtype)[idx]
    return out


def _count(shape, axes, delta=None):
    n = 1.0
    for ax in axes:
        n *= shape[ax] if delta is None else 2.0 * (shape[ax] + delta)
    return n


def _fct(shape, axes, norm, forward, delta=None):
    if norm is None or norm == "backward":
        return 1.0 if forward else 1.0 / _count(shape, axes, delta)
    if norm == "ortho":
        return 1.0 / np.sqrt(_count(shape, axes, delta))
    if norm == "forward":
        return 1.0 / _count(shape, axes, delta) if forward else 1.0
    raise ValueError("Invalid norm value; should be 'backward', 'ortho' or 'forward'.")


def _ax(axes):
    return np.asarray(axes, dtype=np.int64)


# ------------------------------------------------------------- drivers

def _c2cn(a, s, axes, norm, forward, overwrite_x=False):
    a = _asarray(a)
    s, axes = _shape_and_axes(a, s, axes)
    rdt, cdt = _dtype_pair(a)
    if _is_complex(a):
        x = _pad_crop(a, s, axes, cdt)
        out = x if (overwrite_x and x is a) else _empty(x, x.shape, cdt)
        c2c(x, out, _ax(axes), forward, _fct(x.shape, axes, norm, forward), 1)
        return out
    x = _pad_crop(a, s, axes, rdt)
    out = _empty(x, x.shape, cdt)
    c2c_sym(x, out, _ax(axes), forward, _fct(x.shape, axes, norm, forward), 1)
    return out


def _r2cn(a, s, axes, norm, forward):
    a = _asarray(a)
    if _is_complex(a):
        raise TypeError(f"unsupported dtype {a.dtype}")
    s, axes = _shape_and_axes(a, s, axes)
    rdt, cdt = _dtype_pair(a)
    x = _pad_crop(a, s, axes, rdt)
    shape = list(x.shape)
    shape[axes[-1]] = shape[axes[-1]] // 2 + 1
    out = _empty(x, shape, cdt)
    r2c(x, out, _ax(axes), forward, _fct(x.shape, axes, norm, forward), 1)
    return out


def _c2rn(a, s, axes, norm, forward):
    a = _asarray(a)
    s, axes = _shape_and_axes(a, s, axes)
    rdt, cdt = _dtype_pair(a)
    xin = _pad_crop(a, s, axes, cdt)
    shape = list(a.shape)
    shape[axes[-1]] = (shape[axes[-1]] - 1) * 2
    if s is not None:  # overloads.py:655-666: s sets the last array axis
        last = a.ndim - 1
        for i, ax in enumerate(axes):
            if ax == last:
                shape[last] = s[i]
                break
    if shape[axes[-1]] < 1:
        raise ValueError(f"Invalid number of data points ({shape[axes[-1]]}) specified")
    out = _empty(xin, shape, rdt)
    c2r(xin, out, _ax(axes), forward, _fct(shape, axes, norm, forward), 1)
    return out


def _one(n):
    return None if n is None else [n]


# ------------------------------------------------------------ numpy.fft

def fft(a, n=None, axis=-1, norm=None, overwrite_x=False, workers=None, plan=None, out=None):
    return _c2cn(a, _one(n), [axis], norm, True, overwrite_x)


def ifft(a, n=None, axis=-1, norm=None, overwrite_x=False, workers=None, plan=None, out=None):
    return _c2cn(a, _one(n), [axis], norm, False, overwrite_x)


def fft2(a, s=None, axes=(-2, -1), norm=None, overwrite_x=False, workers=None, plan=None,
         out=None):
    return _c2cn(a, s, axes, norm, True, overwrite_x)


def ifft2(a, s=None, axes=(-2, -1), norm=None, overwrite_x=False, workers=None, plan=None,
          out=None):
    return _c2cn(a, s, axes, norm, False, overwrite_x)


def fftn(a, s=None, axes=None, norm=None, overwrite_x=False, workers=None, plan=None, out=None):
    return _c2cn(a, s, axes, norm, True, overwrite_x)


def ifftn(a, s=None, axes=None, norm=None, overwrite_x=False, workers=None, plan=None, out=None):
    return _c2cn(a, s, axes, norm, False, overwrite_x)


def rfft(a, n=None, axis=-1, norm=None, overwrite_x=False, workers=None, plan=None, out=None):
    return _r2cn(a, _one(n), [axis], norm, True)


def irfft(a, n=None, axis=-1, norm=None, overwrite_x=False, workers=None, plan=None, out=None):
    return _c2rn(a, _one(n), [axis], norm, False)


def rfft2(a, s=None, axes=(-2, -1), norm=None, overwrite_x=False, workers=None, plan=None,
          out=None):
    return _r2cn(a, s, axes, norm, True)


def irfft2(a, s=None, axes=(-2, -1), norm=None, overwrite_x=False, workers=None, plan=None,
           out=None):
    return _c2rn(a, s, axes, norm, False)


def rfftn(a, s=None, axes=None, norm=None, overwrite_x=False, workers=None, plan=None,
          out=None):
    return _r2cn(a, s, axes, norm, True)


def irfftn(a, s=None, axes=None, norm=None, overwrite_x=False, workers=None, plan=None,
           out=None):
    return _c2rn(a, s, axes, norm, False)


def hfft(a, n=None, axis=-1, norm=None, overwrite_x=False, workers=None, plan=None, out=None):
    return _c2rn(a, _one(n), [axis], norm, True)


def ihfft(a, n=None, axis=-1, norm=None, overwrite_x=False, workers=None, plan=None, out=None):
    return _r2cn(a, _one(n), [axis], norm, False)


def fftfreq(n, d=1.0, device=None):
    """Sample frequencies (overloads.py:1232-1243)."""
    val = 1.0 / (n * d)
    res = np.empty(n, dtype=np.int64)
    N = (n - 1) // 2 + 1
    res[:N] = np.arange(N)
    res[N:] = np.arange(-(n // 2), 0)
    return res * val


def rfftfreq(n, d=1.0, device=None):
    val = 1.0 / (n * d)
    return np.arange(n // 2 + 1) * val


def _shift(x, axes, sign):
    if axes is None:
        axes = list(range(x.ndim))
    elif isinstance(axes, numbers.Integral):
        axes = [axes]
    shifts = [sign * (x.shape[ax] // 2) for ax in axes]
    if _is_torch(x):
        import torch
        return torch.roll(x, shifts, dims=list(axes))
    return np.roll(x, shifts, axis=tuple(axes))


def fftshift(x, axes=None):
    return _shift(_asarray(x), axes, +1)


def ifftshift(x, axes=None):
    return _shift(_asarray(x), axes, -1)


# ------------------------------------------------------------ scipy.fft

def _type(type, forward):
    if type not in (1, 2, 3, 4):
        raise ValueError("Invalid type; must be one of (1, 2, 3, 4).")
    if forward:
        return type
    return {2: 3, 3: 2}.get(type, type)


def _r2rn(x, type, s, axes, norm, overwrite_x, orthogonalize, forward, cosine):
    x = _asarray(x)
    s, axes = _shape_and_axes(x, s, axes)
    rdt, cdt = _dtype_pair(x)
    cplx = _is_complex(x)
    dt = cdt if cplx else rdt
    xin = _pad_crop(x, s, axes, dt)
    out = xin if (overwrite_x and xin is x) else _empty(xin, xin.shape, dt)
    t = _type(type, forward)
    delta = (-1.0 if cosine else 1.0) if t == 1 else 0.0
    fct = _fct(xin.shape, axes, norm, forward, delta)
    ortho = bool(orthogonalize) if orthogonalize is not None else norm == "ortho"
    f = _dct_ll if cosine else _dst_ll
    if cplx:  # real and imaginary parts separately (overloads.py:993-1000)
        if _is_torch(xin):
            ri, ro = xin.real.contiguous(), _empty(xin, xin.shape, rdt)
            ii, io = xin.imag.contiguous(), _empty(xin, xin.shape, rdt)
        else:
            ri, ro = np.ascontiguousarray(xin.real), np.empty(xin.shape, rdt)
            ii, io = np.ascontiguousarray(xin.imag), np.empty(xin.shape, rdt)
        f(ri, ro, _ax(axes), t, fct, ortho, 1)
        f(ii, io, _ax(axes), t, fct, ortho, 1)
        out[...] = ro + 1j * io
        return out
    f(xin, out, _ax(axes), t, fct, ortho, 1)
    return out


def dct(x, type=2, n=None, axis=-1, norm=None, overwrite_x=False, workers=None,
        orthogonalize=None):
    return _r2rn(x, type, _one(n), [axis], norm, overwrite_x, orthogonalize, True, True)


def idct(x, type=2, n=None, axis=-1, norm=None, overwrite_x=False, workers=None,
         orthogonalize=None):
    return _r2rn(x, type, _one(n), [axis], norm, overwrite_x, orthogonalize, False, True)


def dst(x, type=2, n=None, axis=-1, norm=None, overwrite_x=False, workers=None,
        orthogonalize=None):
    return _r2rn(x, type, _one(n), [axis], norm, overwrite_x, orthogonalize, True, False)


def idst(x, type=2, n=None, axis=-1, norm=None, overwrite_x=False, workers=None,
         orthogonalize=None):
    return _r2rn(x, type, _one(n), [axis], norm, overwrite_x, orthogonalize, False, False)


def dctn(x, type=2, s=None, axes=None, norm=None, overwrite_x=False, workers=None,
         orthogonalize=None):
    return _r2rn(x, type, s, axes, norm, overwrite_x, orthogonalize, True, True)


def idctn(x, type=2, s=None, axes=None, norm=None, overwrite_x=False, workers=None,
          orthogonalize=None):
    return _r2rn(x, type, s, axes, norm, overwrite_x, orthogonalize, False, True)


def dstn(x, type=2, s=None, axes=None, norm=None, overwrite_x=False, workers=None,
         orthogonalize=None):
    return _r2rn(x, type, s, axes, norm, overwrite_x, orthogonalize, True, False)


def idstn(x, type=2, s=None, axes=None, norm=None, overwrite_x=False, workers=None,
          orthogonalize=None):
    return _r2rn(x, type, s, axes, norm, overwrite_x, orthogonalize, False, False)


def next_fast_len(target, real=False):
    """scipy.fft.next_fast_len through good_size (overloads.py:1861-1870)."""
    return good_size(target, real)

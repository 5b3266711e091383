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
import warnings

import numpy as np

from . import (c2c, c2c_shifted, c2c_sym, c2r, dct as _dct_ll, dst as _dst_ll, good_size,
               line_scale, r2c, roll_device)

__all__ = [
    "fft", "ifft", "fft2", "ifft2", "fftn", "ifftn", "rfft", "irfft", "rfft2", "irfft2",
    "rfftn", "irfftn", "hfft", "ihfft", "fftfreq", "rfftfreq", "fftshift", "ifftshift",
    "dct", "idct", "dst", "idst", "dctn", "idctn", "dstn", "idstn", "next_fast_len",
    "fht", "ifht", "fhtoffset", "roll", "fftn_shifted", "centered_fftn", "centered_ifftn",
    "rfftn_shifted", "irfftn_shifted",
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


def _roll_gpu(x, shifts):
    """Cyclic shift of every axis by shifts[k] in one gather kernel (rfh_roll).
    Host arrays are copied to the current CUDA device and back."""
    import torch
    host = not _is_torch(x)
    if host:
        try:
            t = torch.from_numpy(np.ascontiguousarray(x)).cuda()
        except TypeError:
            raise TypeError(f"roll: unsupported dtype {x.dtype}") from None
    else:
        t = x
    out = torch.empty(tuple(t.shape), dtype=t.dtype, device=t.device)
    if t.numel():
        roll_device(t, out, shifts)
    return out.cpu().numpy() if host else out


def roll(a, shift, axis=None):
    """numpy.roll (overloads.py:762-855): with ``axis`` None the flattened
    array is shifted by the sum of ``shift`` (a scalar or 1-D sequence);
    otherwise ``axis`` and ``shift`` broadcast against each other, shifts
    on one axis add up, and every axis moves in the same gather kernel."""
    if np.isscalar(a):
        return np.asarray(a)
    x = _asarray(a)
    if axis is None:
        sh = np.asarray(shift)
        if sh.ndim > 1:
            raise ValueError("'shift' must be a scalar or 1D sequence")
        flat = x.reshape(-1) if _is_torch(x) else np.ravel(x)
        return _roll_gpu(flat, [int(sh.sum())]).reshape(tuple(x.shape))
    ax, sh = np.broadcast_arrays(np.asarray(axis), np.asarray(shift))
    shifts = [0] * x.ndim
    for k, v in zip(ax.ravel(), sh.ravel()):
        k = int(k)
        if k >= x.ndim or k < -x.ndim:
            raise ValueError("axis is out of bounds")
        shifts[k % x.ndim] += int(v)
    return _roll_gpu(x, shifts)


def _shift(x, axes, sign):
    if axes is None:
        axes = list(range(x.ndim))
    elif isinstance(axes, numbers.Integral):
        axes = [axes]
    shifts = [0] * x.ndim
    for ax in axes:
        ax = _wrap(ax, x.ndim)
        shifts[ax] += sign * (x.shape[ax] // 2)
    return _roll_gpu(x, shifts)


def fftshift(x, axes=None):
    return _shift(_asarray(x), axes, +1)


def ifftshift(x, axes=None):
    return _shift(_asarray(x), axes, -1)


# ------------------------------------------- transforms with fused shifts

def _shift_amounts(spec, shape, axes):
    """Roll amounts along axes for a shift spec: None, "fftshift" (n // 2),
    "ifftshift" (-(n // 2)) or integers (one, or one per axis)."""
    if spec is None:
        return None
    if isinstance(spec, str):
        if spec == "fftshift":
            return [shape[ax] // 2 for ax in axes]
        if spec == "ifftshift":
            return [-(shape[ax] // 2) for ax in axes]
        raise ValueError("shift must be None, 'fftshift', 'ifftshift' or integers")
    v = np.broadcast_to(np.asarray(spec, dtype=np.int64), (len(axes),))
    return [int(x) for x in v]


def fftn_shifted(a, s=None, axes=None, norm=None, forward=True, shift_in=None,
                 shift_out=None):
    """roll(fftn(roll(a, shift_in), s, axes, norm), shift_out) (ifftn when
    forward is False), the rolls along the transformed axes.  A shift is None,
    "fftshift", "ifftshift" or integer roll amounts.  numpy composes these as
    separate calls (overloads.py:1253-1310 for the shifts); here the rolls
    become index rotations in the loads of each axis's first pass and the
    stores of its last pass (rfh_c2c_shifted), with a bit-identical result.
    Rolls a plan cannot fuse (four-step or multi-pass rolls that are not
    whole columns / rows) fall back to one separate gather pass on each
    side."""
    a = _asarray(a)
    s, axes = _shape_and_axes(a, s, axes)
    rdt, cdt = _dtype_pair(a)
    x = _pad_crop(a, s, axes, cdt)
    si = _shift_amounts(shift_in, x.shape, axes)
    so = _shift_amounts(shift_out, x.shape, axes)
    fct = _fct(x.shape, axes, norm, forward)
    out = _empty(x, x.shape, cdt)
    if x.ndim == 0 or min(x.shape) == 0:
        return out
    if c2c_shifted(x, out, _ax(axes), forward, fct, si, so, 1):
        return out
    if si is not None:
        shifts = [0] * x.ndim
        for ax, v in zip(axes, si):
            shifts[ax] += v
        x = _roll_gpu(x, shifts)
    c2c(x, out, _ax(axes), forward, fct, 1)
    if so is not None:
        shifts = [0] * x.ndim
        for ax, v in zip(axes, so):
            shifts[ax] += v
        out = _roll_gpu(out, shifts)
    return out


def rfftn_shifted(a, s=None, axes=None, norm=None, shift_in=None, shift_out=None):
    """roll(rfftn(roll(a, shift_in), s, axes, norm), shift_out), the rolls
    along the leading transformed axes (axes[:-1]; the half-spectrum axis is
    not rolled).  The real transform of the last axis runs first (it commutes
    with rolls along the other axes), then the complex passes of the leading
    axes in place, with the rolls as index rotations of their loads and
    stores (rfh_c2c_shifted): no extra pass, and a result bit-identical to
    the composition.  Plans that cannot fuse a roll fall back to a separate
    gather pass."""
    a = _asarray(a)
    if _is_complex(a):
        raise TypeError(f"unsupported dtype {a.dtype}")
    s, axes = _shape_and_axes(a, s, axes)
    rdt, cdt = _dtype_pair(a)
    x = _pad_crop(a, s, axes, rdt)
    shape = list(x.shape)
    shape[axes[-1]] = shape[axes[-1]] // 2 + 1
    out = _empty(x, shape, cdt)
    if x.ndim == 0 or min(x.shape) == 0:
        return out
    lead = list(axes[:-1])
    r2c(x, out, _ax([axes[-1]]), True, _fct(x.shape, axes, norm, True), 1)
    if not lead:
        return out
    si = _shift_amounts(shift_in, x.shape, lead)
    so = _shift_amounts(shift_out, x.shape, lead)
    if c2c_shifted(out, out, _ax(lead), True, 1.0, si, so, 1):
        return out
    y = out
    if si is not None:
        shifts = [0] * x.ndim
        for ax, v in zip(lead, si):
            shifts[ax] += v
        y = _roll_gpu(out, shifts)
    c2c(y, out, _ax(lead), True, 1.0, 1)
    if so is not None:
        shifts = [0] * x.ndim
        for ax, v in zip(lead, so):
            shifts[ax] += v
        out = _roll_gpu(out, shifts)
    return out


def irfftn_shifted(a, s=None, axes=None, norm=None, shift_in=None, shift_out=None):
    """roll(irfftn(roll(a, shift_in), s, axes, norm), shift_out), the rolls
    along the leading transformed axes (axes[:-1]).  The complex passes of
    the leading axes run first, into a complex temporary, with both rolls as
    index rotations of their loads and stores (the real pass of the last
    axis commutes with rolls along the other axes), then the real pass:
    the same passes as irfftn.  Plans that cannot fuse a roll fall back to
    a separate gather pass."""
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
    fct = _fct(shape, axes, norm, False)
    lead = list(axes[:-1])
    if not lead or min(xin.shape) == 0:
        c2r(xin, out, _ax(axes), False, fct, 1)
        return out
    si = _shift_amounts(shift_in, xin.shape, lead)
    so = _shift_amounts(shift_out, xin.shape, lead)
    tmp = _empty(xin, xin.shape, cdt)
    if not c2c_shifted(xin, tmp, _ax(lead), False, 1.0, si, so, 1):
        y = xin
        if si is not None:
            shifts = [0] * xin.ndim
            for ax, v in zip(lead, si):
                shifts[ax] += v
            y = _roll_gpu(xin, shifts)
        c2c(y, tmp, _ax(lead), False, 1.0, 1)
        if so is not None:
            shifts = [0] * xin.ndim
            for ax, v in zip(lead, so):
                shifts[ax] += v
            tmp = _roll_gpu(tmp, shifts)
    c2r(tmp, out, _ax([axes[-1]]), False, fct, 1)
    return out


def centered_fftn(a, s=None, axes=None, norm=None):
    """fftshift(fftn(ifftshift(a))) over axes, the DFT of data centred on
    index n // 2, in one fused transform."""
    return fftn_shifted(a, s, axes, norm, True, "ifftshift", "fftshift")


def centered_ifftn(a, s=None, axes=None, norm=None):
    """fftshift(ifftn(ifftshift(a))) over axes, the inverse of centered_fftn."""
    return fftn_shifted(a, s, axes, norm, False, "ifftshift", "fftshift")


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


# --------------------------------------------------------- FFTLog (fht)

def _loggamma_pair(xp, xm, y):
    from scipy.special import loggamma
    return loggamma(xp + 1j * y), loggamma(xm + 1j * y)


def _fht_coeff(n, dln, mu, offset, bias):
    """FFTLog coefficients u_m, m = 0..n/2 (overloads.py:858-883), from
    u_m = 2^q (2 / k0r0)^{2 i y_m} Gamma(xp + i y_m) / Gamma(xm - i y_m) with
    xp, xm = (mu + 1 +- q) / 2, y_m = pi m / (n dln), ln(k0r0) = offset,
    q = bias; Gamma(xm - iy) = conj(Gamma(xm + iy)) gives the real part
    Re lnG(xp+iy) - Re lnG(xm+iy) + q ln2 and the phase
    Im lnG(xp+iy) + Im lnG(xm+iy) + 2 y (ln2 - offset)."""
    xp = (mu + 1 + bias) / 2
    xm = (mu + 1 - bias) / 2
    y = np.linspace(0, np.pi * (n // 2) / (n * dln), n // 2 + 1)
    gp, gm = _loggamma_pair(xp, xm, y)
    ln2 = np.log(2)
    u = np.exp((gp.real - gm.real + ln2 * bias) + 1j * (gp.imag + gm.imag + 2 * y * (ln2 - offset)))
    u.imag[-1] = 0  # the last coefficient is real (overloads.py:877)
    if not np.isfinite(u[0]):
        # Gamma(xp) / Gamma(xm) with both at poles: its finite limit
        from scipy.special import poch
        u[0] = 2 ** bias * poch(xm, xp - xm)
    return u


def _fht_real_dtype(a):
    if _is_torch(a):
        if a.is_complex():
            raise TypeError("The 1st argument must be a real array.")
        return "float32" if str(a.dtype) == "torch.float32" else "float64"
    if np.iscomplexobj(a):
        raise TypeError("The 1st argument must be a real array.")
    return "float32" if a.dtype == np.float32 else "float64"


def _fht(a, dln, mu, offset, bias, inverse):
    """fht / ifht on the GPU (overloads.py:1755-1809): bias weights, r2c,
    the coefficient step and the reversal in one pass (rfh_line_scale), c2r,
    output weights.  Host arrays are copied to the current CUDA device and
    back."""
    import torch
    a = _asarray(a)
    rt = _fht_real_dtype(a)
    host = not _is_torch(a)
    cast = np.float32 if rt == "float32" else np.float64
    # the scalars are cast to the array's precision, as in the reference
    dln, mu, offset, bias = (float(cast(v)) for v in (dln, mu, offset, bias))
    t = torch.as_tensor(np.asarray(a, dtype=rt)) if host else a.to(getattr(torch, rt))
    x = t.to("cuda" if host else t.device).contiguous().clone()
    n = x.shape[-1]
    dbl = rt == "float64"
    lines = x.numel() // n if n else 0
    j = np.arange(n)
    jc = (n - 1) / 2
    if bias != 0:
        w_in = np.exp(bias * ((j - jc) * dln + offset)) if inverse else np.exp(-bias * (j - jc) * dln)
        line_scale(x, lines, n, n, w_in, False, dbl)
    u = _fht_coeff(n, dln, mu, offset, bias)
    h = n // 2 + 1
    ph = np.exp(2j * np.pi * np.arange(h) / n)
    if not inverse:
        if np.isinf(u[0]):
            warnings.warn("singular transform; consider changing the bias", RuntimeWarning)
            u = u.copy()
            u[0] = 0
        c = np.conj(u) * ph
    else:
        if u[0] == 0:
            warnings.warn("singular inverse transform; consider changing the bias",
                          RuntimeWarning)
        with np.errstate(divide="ignore", invalid="ignore"):
            c = np.where(u == 0, 0, ph / np.where(u == 0, 1, u))
    ct = "complex128" if dbl else "complex64"
    X = torch.empty(tuple(x.shape[:-1]) + (h,), dtype=getattr(torch, ct), device=x.device)
    ax = np.array([x.ndim - 1])
    r2c(x, X, ax, True, 1.0)
    line_scale(X, lines, h, h, c, True, dbl)
    out = torch.empty_like(x)
    c2r(X, out, ax, False, 1.0 / n)
    if bias != 0:
        w_out = (1 / np.exp(-bias * (j - jc) * dln) if inverse
                 else np.exp(-bias * ((j - jc) * dln + offset)))
        line_scale(out, lines, n, n, w_out, False, dbl)
    return out.cpu().numpy() if host else out


def fht(a, dln, mu, offset=0.0, bias=0.0):
    """scipy.fft.fht: fast Hankel transform (FFTLog) over the last axis."""
    return _fht(a, dln, mu, offset, bias, False)


def ifht(A, dln, mu, offset=0.0, bias=0.0):
    """scipy.fft.ifht: inverse fast Hankel transform over the last axis."""
    return _fht(A, dln, mu, offset, bias, True)


def fhtoffset(dln, mu, initial=0.0, bias=0.0):
    """scipy.fft.fhtoffset: the low-ringing offset nearest to ``initial``
    (overloads.py:1845-1858): ln(kr) moved by the fraction of dln that makes
    the phase of the coefficient at the Nyquist frequency a multiple of pi."""
    xp = (mu + 1 + bias) / 2
    xm = (mu + 1 - bias) / 2
    y = np.pi / (2 * dln)
    gp, gm = _loggamma_pair(xp, xm, y)
    arg = (np.log(2) - initial) / dln + (gp.imag + gm.imag) / np.pi
    return initial + (arg - np.round(arg)) * dln

"""rocket_fft_amd -- rocket-fft's low-level transform API on MI355X (gfx950).

Same functions, argument order and meaning as rocket-fft's low-level
interface (``rocket_fft/__init__.py:18-34``, typed in
``rocket_fft/__init__.pyi:6-105``)::

    c2c(ain, aout, axes, forward, fct, nthreads)
    r2c(ain, aout, axes, forward, fct, nthreads)
    c2r(ain, aout, axes, forward, fct, nthreads)
    c2c_sym(ain, aout, axes, forward, fct, nthreads)
    dct(ain, aout, axes, type, fct, ortho, nthreads)
    dst(ain, aout, axes, type, fct, ortho, nthreads)
    r2r_separable_hartley(ain, aout, axes, fct, nthreads)
    r2r_genuine_hartley(ain, aout, axes, fct, nthreads)
    r2r_fftpack(ain, aout, axes, real2hermitian, forward, fct, nthreads)
    good_size(n, real)

Each call goes through the C ABI of ``librocketfft_hip.so`` (the same ten
``numba_*`` symbols rocket-fft's Numba intrinsics bind, see
``rocket_fft/pocketfft.py:32-128``), with the Numba array struct built here
instead of by Numba.  Arrays may be NumPy arrays (host memory, staged through
the GPU) or torch CUDA tensors (computed in place in HBM, on torch's current
stream, asynchronously like any torch op).  There is no CPU fallback: the
functions raise if the HIP library is missing.
"""
import ctypes
import numbers
import os

import numpy as np

__all__ = [
    "c2c", "r2c", "c2r", "c2c_sym", "dct", "dst", "r2r_separable_hartley",
    "r2r_genuine_hartley", "r2r_fftpack", "good_size", "separable_hartley",
    "genuine_hartley", "fftpack", "library_path", "last_error", "plan_facts",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librocketfft_hip.so")
_lib = None


def library_path():
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(
            f"rocket_fft_amd: {_LIB_PATH} is not built; run __graft_entry__.build() "
            "or python rocket-fft_amd/rocket_fft_amd/_build.py")
    L = ctypes.CDLL(_LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    u64, vp, b, d = ctypes.c_uint64, ctypes.c_void_p, ctypes.c_bool, ctypes.c_double
    L.numba_good_size.argtypes = [u64, b]
    L.numba_good_size.restype = u64
    for name in ("numba_c2c", "numba_r2c", "numba_c2r", "numba_c2c_sym"):
        getattr(L, name).argtypes = [u64, vp, vp, vp, b, d, u64]
        getattr(L, name).restype = None
    for name in ("numba_dct", "numba_dst"):
        getattr(L, name).argtypes = [u64, vp, vp, vp, u64, d, b, u64]
        getattr(L, name).restype = None
    L.numba_r2r_fftpack.argtypes = [u64, vp, vp, vp, b, b, d, u64]
    L.numba_r2r_fftpack.restype = None
    for name in ("numba_r2r_separable_hartley", "numba_r2r_genuine_hartley"):
        getattr(L, name).argtypes = [u64, vp, vp, vp, d, u64]
        getattr(L, name).restype = None
    L.rfh_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.rfh_last_error.restype = ctypes.c_int
    L.rfh_clear_error.argtypes = []
    L.rfh_clear_error.restype = None
    L.rfh_set_stream.argtypes = [vp, b]
    L.rfh_set_stream.restype = ctypes.c_int
    for name in ("rfh_cfftp_factors", "rfh_rfftp_factors"):
        getattr(L, name).argtypes = [u64, vp, ctypes.c_int]
        getattr(L, name).restype = ctypes.c_int
    L.rfh_plan_choice.argtypes = [u64, b, ctypes.POINTER(u64)]
    L.rfh_plan_choice.restype = ctypes.c_int
    L.rfh_gpu_schedule.argtypes = [u64, b, ctypes.POINTER(ctypes.c_int), vp, ctypes.c_int,
                                   ctypes.POINTER(u64), ctypes.POINTER(u64)]
    L.rfh_gpu_schedule.restype = ctypes.c_int
    L.rfh_set_fast_path.argtypes = [b]
    L.rfh_set_fast_path.restype = ctypes.c_int
    _lib = L
    return L


def last_error():
    """(code, message) of the calling thread's last failed call."""
    buf = ctypes.create_string_buffer(512)
    code = _load().rfh_last_error(buf, 512)
    return code, buf.value.decode()


# ------------------------------------------------------------ array structs

_struct_cache = {}


def _arystruct_type(ndim):
    t = _struct_cache.get(ndim)
    if t is None:
        t = type(f"arystruct_{ndim}", (ctypes.Structure,), {"_fields_": [
            ("meminfo", ctypes.c_void_p), ("parent", ctypes.c_void_p),
            ("nitems", ctypes.c_int64), ("itemsize", ctypes.c_int64),
            ("data", ctypes.c_void_p),
            ("shape_and_strides", ctypes.c_int64 * (2 * max(ndim, 1)))]})
        _struct_cache[ndim] = t
    return t


def _is_torch(a):
    return type(a).__module__.startswith("torch")


_TORCH_ITEMSIZE = {"torch.float32": 4, "torch.float64": 8, "torch.complex64": 8,
                   "torch.complex128": 16}


def _describe(a):
    """(ndim, shape, byte strides, itemsize, data pointer, kind, is_cuda)."""
    if _is_torch(a):
        key = str(a.dtype)
        if key not in _TORCH_ITEMSIZE:
            raise TypeError(f"unsupported tensor dtype {a.dtype}")
        isz = _TORCH_ITEMSIZE[key]
        kind = "c" if a.is_complex() else "f"
        return (a.dim(), tuple(a.shape), tuple(s * isz for s in a.stride()), isz,
                a.data_ptr(), kind, a.is_cuda)
    if not isinstance(a, np.ndarray):
        raise TypeError("arrays must be numpy arrays or torch tensors")
    if a.dtype.kind not in "fc" or a.dtype.itemsize not in (4, 8, 16):
        raise TypeError(f"unsupported array dtype {a.dtype}")
    return (a.ndim, a.shape, a.strides, a.dtype.itemsize, a.ctypes.data, a.dtype.kind, False)


def _struct(a):
    nd, shape, strides, isz, ptr, kind, cuda = _describe(a)
    S = _arystruct_type(nd)()
    S.nitems = int(np.prod(shape)) if nd else 1
    S.itemsize = isz
    S.data = ptr
    for i in range(nd):
        S.shape_and_strides[i] = shape[i]
        S.shape_and_strides[nd + i] = strides[i]
    return S, nd, kind, cuda


def _axes_struct(axes, ndim):
    ax = np.asarray(axes)
    if ax.ndim != 1:
        raise TypeError("Axes must be a one-dimensional array")
    # any integer or float axes array is copied to uint64 (pocketfft.py:155-167)
    ax = np.ascontiguousarray(ax, dtype=np.int64)
    if ax.size and (ax.min() < 0 or ax.max() >= ndim):
        raise ValueError("axis out of range")
    ax = ax.astype(np.uint64)
    S = _arystruct_type(1)()
    S.nitems = ax.size
    S.itemsize = 8
    S.data = ax.ctypes.data
    S.shape_and_strides[0] = ax.size
    S.shape_and_strides[1] = 8
    return S, ax


def _check_bool(name, v):
    if not isinstance(v, (bool, np.bool_)):
        raise TypeError(f"'{name}' must be a boolean")


def _check_int(name, v):
    if isinstance(v, (bool, np.bool_)) or not isinstance(v, (numbers.Integral, np.integer)):
        raise TypeError(f"'{name}' must be an integer")


def _check_float(name, v):
    # rocket_fft/pocketfft.py widens float arguments only; integer fct fails
    # typing there (tests/test_low_level_interface.py:218-225)
    if isinstance(v, (bool, np.bool_, numbers.Integral, np.integer)) or not isinstance(
            v, (numbers.Real, np.floating)):
        raise TypeError(f"'{name}' must be a floating-point number")


def _call(fname, ain, aout, axes, *mid):
    S_in, nd, kin, cin = _struct(ain)
    S_out, nd2, kout, cout = _struct(aout)
    if nd != nd2:
        raise TypeError("Input and output array must have the same number of dimensions")
    S_ax, keep = _axes_struct(axes, nd)
    L = _load()
    stream = None
    if cin or cout:
        import torch
        stream = torch.cuda.current_stream(aout.device if cout else ain.device).cuda_stream
        L.rfh_set_stream(ctypes.c_void_p(stream), False)
    try:
        L.rfh_clear_error()
        getattr(L, fname)(nd, ctypes.byref(S_in), ctypes.byref(S_out), ctypes.byref(S_ax), *mid)
        code, msg = last_error()
    finally:
        if stream is not None:
            L.rfh_set_stream(None, True)
    del keep
    if code:
        raise RuntimeError(f"{fname}: {msg}")
    return aout


def _cmplx_args(forward, fct, nthreads):
    _check_bool("forward", forward)
    _check_float("fct", fct)
    _check_int("nthreads", nthreads)
    return bool(forward), float(fct), int(nthreads)


def c2c(ain, aout, axes, forward, fct, nthreads=1):
    """Complex-to-complex DFT over ``axes`` (_pocketfft_numba.cpp:31-49)."""
    f, s, t = _cmplx_args(forward, fct, nthreads)
    return _call("numba_c2c", ain, aout, axes, f, s, t)


NOT_FUSABLE = 5


def c2c_shifted(ain, aout, axes, forward, fct, shift_in=None, shift_out=None, nthreads=1):
    """c2c with numpy.roll shifts fused into the transform (rfh_c2c_shifted):
    aout = roll(c2c(roll(ain, shift_in, axes)), shift_out, axes), shift_in[i]
    and shift_out[i] (None: no roll) along axes[i], each axis once.  Returns
    True when the rolls were fused; False, with nothing written, when a
    length's plan cannot fuse them (the caller then rolls separately)."""
    f, s, t = _cmplx_args(forward, fct, nthreads)
    S_in, nd, kin, cin = _struct(ain)
    S_out, nd2, kout, cout = _struct(aout)
    if nd != nd2:
        raise TypeError("Input and output array must have the same number of dimensions")
    S_ax, keep = _axes_struct(axes, nd)
    nax = len(np.atleast_1d(np.asarray(axes)))

    def shifts(v):
        if v is None:
            return None
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.int64), (nax,)))
        return a

    si, so = shifts(shift_in), shifts(shift_out)
    L = _load()
    i64p = ctypes.POINTER(ctypes.c_int64)
    L.rfh_c2c_shifted.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_bool, ctypes.c_double,
                                  ctypes.c_uint64, i64p, i64p]
    L.rfh_c2c_shifted.restype = ctypes.c_int
    stream = None
    if cin or cout:
        import torch
        stream = torch.cuda.current_stream(aout.device if cout else ain.device).cuda_stream
        L.rfh_set_stream(ctypes.c_void_p(stream), False)
    try:
        L.rfh_clear_error()
        rc = L.rfh_c2c_shifted(nd, ctypes.byref(S_in), ctypes.byref(S_out), ctypes.byref(S_ax),
                               f, s, t,
                               None if si is None else si.ctypes.data_as(i64p),
                               None if so is None else so.ctypes.data_as(i64p))
        code, msg = last_error()
    finally:
        if stream is not None:
            L.rfh_set_stream(None, True)
    del keep
    if rc == NOT_FUSABLE:
        return False
    if rc or code:
        raise RuntimeError(f"rfh_c2c_shifted: {msg}")
    return True


def c2c_noalloc(ain, aout, axes, forward, fct):
    """c2c that allocates no device memory (rfh_c2c_noalloc): device tensors
    only, enqueued on torch's current stream without synchronising; its
    four-step intermediates live in the library's static 64 MiB scratch array
    on the GPU.  Plans come from the plan cache: run c2c once with the same
    shape first.  Calls that overlap on one GPU share the scratch and race."""
    S_in, nd, kin, cin = _struct(ain)
    S_out, nd2, kout, cout = _struct(aout)
    if nd != nd2:
        raise TypeError("Input and output array must have the same number of dimensions")
    if not (cin and cout):
        raise TypeError("c2c_noalloc takes CUDA tensors only")
    _check_bool("forward", forward)
    _check_float("fct", fct)
    S_ax, keep = _axes_struct(axes, nd)
    L = _load()
    import torch
    L.rfh_c2c_noalloc.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_bool, ctypes.c_double,
                                  ctypes.c_void_p]
    L.rfh_c2c_noalloc.restype = ctypes.c_int
    stream = torch.cuda.current_stream(aout.device).cuda_stream
    L.rfh_clear_error()
    rc = L.rfh_c2c_noalloc(nd, ctypes.byref(S_in), ctypes.byref(S_out), ctypes.byref(S_ax),
                           bool(forward), float(fct), ctypes.c_void_p(stream))
    del keep
    if rc:
        raise RuntimeError(f"rfh_c2c_noalloc: {last_error()[1]}")
    return aout


def r2c(ain, aout, axes, forward, fct, nthreads=1):
    """Real-to-half-complex DFT (_pocketfft_numba.cpp:91-109)."""
    f, s, t = _cmplx_args(forward, fct, nthreads)
    return _call("numba_r2c", ain, aout, axes, f, s, t)


def c2r(ain, aout, axes, forward, fct, nthreads=1):
    """Half-complex-to-real DFT; shape from ``aout`` (_pocketfft_numba.cpp:145-163)."""
    f, s, t = _cmplx_args(forward, fct, nthreads)
    return _call("numba_c2r", ain, aout, axes, f, s, t)


def c2c_sym(ain, aout, axes, forward, fct, nthreads=1):
    """Real-to-full-complex DFT (_pocketfft_numba.cpp:111-143)."""
    f, s, t = _cmplx_args(forward, fct, nthreads)
    return _call("numba_c2c_sym", ain, aout, axes, f, s, t)


def _real_args(type, fct, ortho, nthreads):
    _check_int("type", type)
    _check_float("fct", fct)
    _check_bool("ortho", ortho)
    _check_int("nthreads", nthreads)
    if type < 1 or type > 4:
        raise ValueError("invalid DCT/DST type")
    return int(type), float(fct), bool(ortho), int(nthreads)


def dct(ain, aout, axes, type, fct, ortho, nthreads=1):
    """DCT type 1-4 (_pocketfft_numba.cpp:51-69)."""
    return _call("numba_dct", ain, aout, axes, *_real_args(type, fct, ortho, nthreads))


def dst(ain, aout, axes, type, fct, ortho, nthreads=1):
    """DST type 1-4 (_pocketfft_numba.cpp:71-89)."""
    return _call("numba_dst", ain, aout, axes, *_real_args(type, fct, ortho, nthreads))


def r2r_separable_hartley(ain, aout, axes, fct, nthreads=1):
    """Hartley transform axis by axis (_pocketfft_numba.cpp:185-203)."""
    _check_float("fct", fct)
    _check_int("nthreads", nthreads)
    return _call("numba_r2r_separable_hartley", ain, aout, axes, float(fct), int(nthreads))


def r2r_genuine_hartley(ain, aout, axes, fct, nthreads=1):
    """N-D Hartley transform (_pocketfft_numba.cpp:205-223)."""
    _check_float("fct", fct)
    _check_int("nthreads", nthreads)
    return _call("numba_r2r_genuine_hartley", ain, aout, axes, float(fct), int(nthreads))


def r2r_fftpack(ain, aout, axes, real2hermitian, forward, fct, nthreads=1):
    """FFTPACK halfcomplex transforms (_pocketfft_numba.cpp:165-183)."""
    _check_bool("real2hermitian", real2hermitian)
    f, s, t = _cmplx_args(forward, fct, nthreads)
    return _call("numba_r2r_fftpack", ain, aout, axes, bool(real2hermitian), f, s, t)


def good_size(n, real=False):
    """Smallest fast length >= n (_pocketfft_numba.cpp:25-29); bit-exact."""
    if not isinstance(n, (numbers.Integral, np.integer, bool, np.bool_)):
        raise TypeError("The first argument 'n' must be an integer")
    if not isinstance(real, (numbers.Integral, np.integer, bool, np.bool_)):
        raise TypeError("The second argument 'real' must be a boolean")
    return int(_load().numba_good_size(int(n), bool(real)))


separable_hartley = r2r_separable_hartley
genuine_hartley = r2r_genuine_hartley
fftpack = r2r_fftpack


def plan_facts(n):
    """Reference plan facts for length n: complex and real FFTPACK factor lists
    and the Bluestein decisions (with n2)."""
    L = _load()
    buf = np.zeros(64, dtype=np.uint64)
    k = L.rfh_cfftp_factors(int(n), buf.ctypes.data, 64)
    cf = [int(x) for x in buf[:k]]
    k = L.rfh_rfftp_factors(int(n), buf.ctypes.data, 64)
    rf = [int(x) for x in buf[:k]]
    n2c, n2r = ctypes.c_uint64(0), ctypes.c_uint64(0)
    bc = L.rfh_plan_choice(int(n), False, ctypes.byref(n2c))
    br = L.rfh_plan_choice(int(n), True, ctypes.byref(n2r))
    return {"cfftp": cf, "rfftp": rf, "blue_c": (bool(bc), int(n2c.value)),
            "blue_r": (bool(br), int(n2r.value))}


PRECISION_MODES = {"fp32": 0, "mfma_fp16": 1, "mfma_bf16": 2, "mfma_bf16_2split": 3}


def set_precision(mode):
    """Arithmetic of complex64 length-4096 transforms: "fp32" (plain
    butterflies), "mfma_fp16" (FP16 matrix cores, hi/lo splits), "mfma_bf16"
    (BF16, hi/mid/lo), "mfma_bf16_2split" (BF16 hi/lo, study).  Returns the
    previous mode's name."""
    L = _load()
    L.rfh_set_precision.argtypes = [ctypes.c_int]
    L.rfh_set_precision.restype = ctypes.c_int
    prev = L.rfh_set_precision(PRECISION_MODES[mode])
    return {v: k for k, v in PRECISION_MODES.items()}[prev]


def plan_cache(directory, clear_memory=False):
    """Keep plan records (twiddle and Bluestein tables) on disk under
    `directory` (None: off), keyed by kind, length, precision, LDS budget,
    stride class and GPU architecture (rfh_plan_cache); `clear_memory` also
    drops the plans built so far in this process."""
    L = _load()
    L.rfh_plan_cache.argtypes = [ctypes.c_char_p, ctypes.c_bool]
    L.rfh_plan_cache.restype = ctypes.c_int
    d = b"" if directory is None else os.fsencode(directory)
    if L.rfh_plan_cache(d, bool(clear_memory)) != 0:
        raise RuntimeError(last_error()[1])


def plan_cache_stats():
    """(plans loaded from disk, plans written to disk) so far."""
    L = _load()
    L.rfh_plan_cache_stats.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.rfh_plan_cache_stats.restype = ctypes.c_int
    hits, writes = ctypes.c_uint64(), ctypes.c_uint64()
    L.rfh_plan_cache_stats(ctypes.byref(hits), ctypes.byref(writes))
    return int(hits.value), int(writes.value)


def plan_record_roundtrip(n, double_precision, lds_budget, interleaved=False, real=False,
                          path=None):
    """Host-only check of the plan file format: 0 when the record of the
    plan survives a write to `path` and a read back (rfh_plan_record_roundtrip)."""
    L = _load()
    L.rfh_plan_record_roundtrip.argtypes = [ctypes.c_uint64, ctypes.c_bool, ctypes.c_uint64,
                                            ctypes.c_bool, ctypes.c_bool, ctypes.c_char_p]
    L.rfh_plan_record_roundtrip.restype = ctypes.c_int
    rc = L.rfh_plan_record_roundtrip(int(n), bool(double_precision), int(lds_budget),
                                     bool(interleaved), bool(real), os.fsencode(path))
    if rc < 0:
        raise RuntimeError(last_error()[1])
    return rc


def kernel_timing(enabled):
    """Per-launch HIP-event timing of the line kernels on their own streams
    (rfh_kernel_timing); off by default, must be off while capturing graphs."""
    L = _load()
    L.rfh_kernel_timing.argtypes = [ctypes.c_bool]
    L.rfh_kernel_timing.restype = ctypes.c_int
    L.rfh_kernel_timing(bool(enabled))


def kernel_timing_report():
    """{label: (launches, total_ms, algorithmic_bytes)} of the launches timed
    since the last report (waits for them)."""
    L = _load()
    L.rfh_kernel_timing_report.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.rfh_kernel_timing_report.restype = ctypes.c_int
    buf = ctypes.create_string_buffer(1 << 20)
    if L.rfh_kernel_timing_report(buf, len(buf)) >= len(buf):
        raise RuntimeError("kernel timing report truncated")
    out = {}
    for line in buf.value.decode().splitlines():
        label, cnt, ms, nbytes = line.split("\t")
        out[label] = (int(cnt), float(ms), float(nbytes))
    return out


def mfma_launches(mode=None):
    """Matrix-core kernel launches so far in `mode` ("mfma_fp16", "mfma_bf16",
    "mfma_bf16_2split"), or in all modes (None) (rfh_mfma_launches)."""
    L = _load()
    L.rfh_mfma_launches.argtypes = [ctypes.c_int]
    L.rfh_mfma_launches.restype = ctypes.c_uint64
    return int(L.rfh_mfma_launches(0 if mode is None else PRECISION_MODES[mode]))


def mfma_probe(a_lanes, b_lanes, bf16=False):
    """d (64 x 4) = one 16x16x32 matrix-core product of device operands laid
    out per lane (64 x 8 each), float32 torch CUDA tensors."""
    import torch
    L = _load()
    d = torch.zeros((64, 4), dtype=torch.float32, device=a_lanes.device)
    L.rfh_mfma_probe.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_void_p]
    rc = L.rfh_mfma_probe(int(bf16), a_lanes.data_ptr(), b_lanes.data_ptr(), d.data_ptr(),
                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    if rc:
        raise RuntimeError(last_error()[1])
    return d


def line_scale(t, nlines, length, stride, w, complex_lines, double):
    """In-place per-line scaling of a torch CUDA tensor (rfh_line_scale):
    real lines times w (complex_lines False), or complex lines X -> conj(X) w
    (w complex, complex_lines True).  Strides in elements."""
    L = _load()
    L.rfh_line_scale.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                 ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_bool,
                                 ctypes.c_void_p]
    L.rfh_line_scale.restype = ctypes.c_int
    import torch
    if complex_lines:
        wv = np.ascontiguousarray(np.asarray(w, dtype=np.complex128).view(np.float64))
    else:
        wv = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
    stream = torch.cuda.current_stream(t.device).cuda_stream
    rc = L.rfh_line_scale(ctypes.c_void_p(t.data_ptr()), int(nlines), int(length), int(stride),
                          wv.ctypes.data_as(ctypes.c_void_p), 1 if complex_lines else 0,
                          bool(double), ctypes.c_void_p(stream))
    if rc:
        raise RuntimeError(f"rfh_line_scale: {last_error()[1]}")
    return t


def roll_device(t, out, shifts):
    """out[i] = t[(i - shifts) mod shape] over every axis (rfh_roll): t a
    torch CUDA tensor of any strides, out a C-contiguous tensor of the same
    shape and dtype on the same device, enqueued on the current stream."""
    L = _load()
    L.rfh_roll.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    L.rfh_roll.restype = ctypes.c_int
    import torch
    if tuple(out.shape) != tuple(t.shape) or out.dtype != t.dtype or not out.is_contiguous():
        raise ValueError("roll_device: out must be a contiguous tensor like t")
    nd = t.dim()
    shape = np.array(t.shape, dtype=np.int64)
    strides = np.array(t.stride(), dtype=np.int64)
    sh = np.array(shifts, dtype=np.int64).reshape(nd)
    stream = torch.cuda.current_stream(t.device).cuda_stream
    rc = L.rfh_roll(ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(out.data_ptr()), nd,
                    shape.ctypes.data_as(ctypes.c_void_p), strides.ctypes.data_as(ctypes.c_void_p),
                    sh.ctypes.data_as(ctypes.c_void_p), t.element_size(), ctypes.c_void_p(stream))
    if rc:
        raise RuntimeError(f"rfh_roll: {last_error()[1]}")
    return out


def set_fast_path(enabled):
    """Process-wide switch of the specialised power-of-two kernels (tests)."""
    _load().rfh_set_fast_path(bool(enabled))


def gpu_schedule(n, double=False):
    """(kind, radices, n1, n2) of the GPU plan for a complex transform of n
    (needs a GPU: plans live in device memory)."""
    L = _load()
    kind = ctypes.c_int(0)
    rad = (ctypes.c_int * 32)()
    n1, n2 = ctypes.c_uint64(0), ctypes.c_uint64(0)
    k = L.rfh_gpu_schedule(int(n), bool(double), ctypes.byref(kind), ctypes.addressof(rad), 32,
                           ctypes.byref(n1), ctypes.byref(n2))
    if k < 0:
        raise RuntimeError(last_error()[1])
    return kind.value, list(rad[:k]), n1.value, n2.value


# ----------------------------------------------------------- ragged batches

class _Job(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("forward", ctypes.c_int32), ("n", ctypes.c_uint64),
                ("inp", ctypes.c_void_p), ("out", ctypes.c_void_p), ("fct", ctypes.c_double)]


JOB_KINDS = {"c2c": 0, "r2c": 1, "c2r": 2}


class RaggedBatch:
    """Many independent transforms of uneven lengths and kinds, planned once
    and run in one persistent launch (``rfh_batch_*`` in the C ABI).

    jobs: iterable of (kind, forward, n, inp, out, fct) with kind in
    {"c2c", "r2c", "c2r"}, n the transform length (the real length for
    r2c/c2r) and inp/out contiguous CUDA tensors holding one line each:
    c2c n complex -> n complex; r2c n real -> n//2+1 complex; c2r n//2+1
    complex -> n real.  All jobs share one precision (double=True for
    complex128/float64).  The tensors must stay alive while the batch is used.
    """

    def __init__(self, jobs, double=False, autotune=True):
        import torch
        L = _load()
        self._lib = L
        self._keep = []
        arr = (_Job * max(1, len(jobs)))()
        device = None
        for i, (kind, forward, n, inp, out, fct) in enumerate(jobs):
            _check_bool("forward", forward)
            if kind not in JOB_KINDS:
                raise ValueError(f"job {i}: unknown kind {kind!r}")
            _check_int("n", n)
            if n < 0:
                raise ValueError(f"job {i}: negative length")
            rdt = torch.float64 if double else torch.float32
            cdt = torch.complex128 if double else torch.complex64
            h = n // 2 + 1
            want_in = {"c2c": (cdt, n), "r2c": (rdt, n), "c2r": (cdt, h)}[kind]
            want_out = {"c2c": (cdt, n), "r2c": (cdt, h), "c2r": (rdt, n)}[kind]
            for name, t, (dt, count) in (("input", inp, want_in), ("output", out, want_out)):
                if not (_is_torch(t) and t.is_cuda):
                    raise TypeError(f"job {i}: {name} must be a CUDA tensor")
                if device is None:
                    device = t.device
                elif t.device != device:
                    raise ValueError(f"job {i}: {name} is on {t.device}, the batch on {device}")
                if t.dtype != dt:
                    raise TypeError(f"job {i} ({kind}, double={bool(double)}): {name} must be "
                                    f"{dt}, got {t.dtype}")
                if not t.is_contiguous():
                    raise ValueError(f"job {i}: {name} must be contiguous")
                if t.numel() < count:
                    raise ValueError(f"job {i}: {name} holds {t.numel()} elements, "
                                     f"the transform needs {count}")
            arr[i].kind = JOB_KINDS[kind]
            arr[i].forward = int(bool(forward))
            arr[i].n = int(n)
            arr[i].inp = inp.data_ptr()
            arr[i].out = out.data_ptr()
            arr[i].fct = float(fct)
            self._keep += [inp, out]
        self.njobs = len(jobs)
        h = ctypes.c_void_p()
        stream = torch.cuda.current_stream().cuda_stream
        L.rfh_batch_create.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_bool,
                                       ctypes.c_bool, ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_void_p)]
        rc = L.rfh_batch_create(ctypes.addressof(arr), len(jobs), bool(double), bool(autotune),
                                ctypes.c_void_p(stream), ctypes.byref(h))
        if rc:
            raise RuntimeError("rfh_batch_create: " + last_error()[1])
        self._h = h

    def _run(self, fname):
        import torch
        fn = getattr(self._lib, fname)
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        rc = fn(self._h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        if rc:
            raise RuntimeError(f"{fname}: " + last_error()[1])

    def execute(self):
        """One persistent launch over all jobs, on torch's current stream."""
        self._run("rfh_batch_execute")

    def execute_serial(self):
        """The same jobs, one drop-in call per transform (comparison point)."""
        self._run("rfh_batch_execute_serial")

    def info(self):
        g, l, a, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._lib.rfh_batch_info.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4
        self._lib.rfh_batch_info(self._h, ctypes.byref(g), ctypes.byref(l), ctypes.byref(a),
                                 ctypes.byref(ms))
        return {"grid": g.value, "lds_bytes": l.value, "autotuned": a.value,
                "tuned_ms": ms.value}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.rfh_batch_destroy.argtypes = [ctypes.c_void_p]
            self._lib.rfh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_tune_cache(directory=None, clear_memory=False):
    """Directory of the on-disk ragged-batch autotune cache (None: none),
    read and appended by RaggedBatch(autotune=True); clear_memory forgets the
    results remembered by this process (rfh_tune_cache)."""
    L = _load()
    L.rfh_tune_cache.argtypes = [ctypes.c_char_p, ctypes.c_bool]
    L.rfh_tune_cache.restype = ctypes.c_int
    d = None if directory is None else os.fsencode(str(directory))
    L.rfh_tune_cache(d, bool(clear_memory))


def ragged_workload(njobs=4096, seed=5, lo=64, hi=65536, good_frac=0.7, c2c_frac=0.6):
    """The ragged-batch workload: lengths log-uniform in [lo, hi] from
    RandomState(seed); good_frac of them rounded up with good_size (complex
    rule for c2c jobs, real rule for r2c jobs), the rest kept (primes
    included); c2c_frac complex c2c jobs, the rest real r2c jobs.
    Returns a list of (kind, n)."""
    rs = np.random.RandomState(seed)
    raw = np.exp(rs.uniform(np.log(lo), np.log(hi), njobs)).astype(np.int64)
    good = rs.uniform(size=njobs) < good_frac
    kind_c2c = rs.uniform(size=njobs) < c2c_frac
    out = []
    for n, g, c in zip(raw, good, kind_c2c):
        n = int(max(2, n))
        if g:
            n = good_size(n, not c)
        out.append(("c2c" if c else "r2c", n))
    return out

"""Shared helpers of the test suite: seeded inputs, error norms, and running
one low-level call on the GPU (torch CUDA tensors) or through the oracle."""
import numpy as np

import oracle as O
import ref as REF

TOL = {np.float32: 1e-6, np.float64: 1e-13, np.complex64: 1e-6, np.complex128: 1e-13}
CTYPE = {np.float32: np.complex64, np.float64: np.complex128}
RTYPE = {np.complex64: np.float32, np.complex128: np.float64}


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    den = np.sqrt(np.sum(np.abs(b) ** 2))
    num = np.sqrt(np.sum(np.abs(a - b) ** 2))
    return float(num / den) if den > 0 else float(num)


def tol_of(dtype):
    return TOL[np.dtype(dtype).type]


def crandn(rs, shape, dtype):
    x = rs.standard_normal(shape) + 1j * rs.standard_normal(shape)
    return x.astype(dtype)


def rrandn(rs, shape, dtype):
    return rs.standard_normal(shape).astype(dtype)


def to_gpu(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def on_gpu(lib, fname, ain, aout_like, *args, inplace=False):
    """Run lib.fname on CUDA copies; returns the output as numpy."""
    import torch
    tin = to_gpu(ain)
    tout = tin if inplace else torch.from_numpy(np.zeros_like(aout_like)).cuda()
    getattr(lib, fname)(tin, tout, *args)
    torch.cuda.synchronize()
    return tout.cpu().numpy()


def oracle_out(fname, ain, aout_like, *args):
    out = np.zeros_like(aout_like)
    getattr(O, fname)(ain, out, *args)
    return out


def ref_out(fname, ain, aout_like, *args):
    out = np.zeros_like(aout_like)
    getattr(REF, fname)(ain, out, *args)
    return out

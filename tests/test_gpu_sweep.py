"""Seeded randomized parity sweep over shapes, strides, axes and kinds.

Each case draws a shape from lengths that exercise every plan form (direct,
specialised power-of-two, four-step, multi-level strided, Bluestein, the
half-length real path), takes a strided or reversed view of a device
tensor, transforms over a random set of axes and compares with scipy.fft in
FP64 on the host.  Tolerances: rel-L2 <= 1e-6 (complex64 / float32),
<= 1e-13 (complex128 / float64)."""
import numpy as np
import pytest
import scipy.fft as sfft
import torch

import rocket_fft_amd as R
from helpers import rel_l2

pytestmark = pytest.mark.gpu

LENGTHS = [1, 2, 3, 5, 16, 17, 60, 100, 128, 243, 256, 343, 1000, 1217, 2048, 3000, 4099,
           8192]
CAP = 1 << 21  # elements per case


def draw(rs):
    while True:
        nd = int(rs.randint(1, 4))
        shape = [int(rs.choice(LENGTHS)) for _ in range(nd)]
        if int(np.prod(shape)) <= CAP:
            return shape


def view_of(t, rs):
    """A strided (step 2 on one dim), reversed, transposed or plain view."""
    kind = int(rs.randint(4))
    if kind == 0 and t.dim() >= 2:
        return t.transpose(0, t.dim() - 1)
    if kind == 1:
        d = int(rs.randint(t.dim()))
        if t.shape[d] >= 4:
            idx = [slice(None)] * t.dim()
            idx[d] = slice(None, None, 2)
            return t[tuple(idx)]
    if kind == 2:
        return t.flip(0)  # torch flips by copy; keeps the case contiguous
    return t


@pytest.mark.parametrize("seed", range(48))
def test_sweep(seed):
    rs = np.random.RandomState(1000 + seed)
    shape = draw(rs)
    dbl = bool(rs.randint(2))
    op = ["c2c", "c2c", "r2c", "c2r", "dct2", "dct3"][int(rs.randint(6))]
    nax = int(rs.randint(1, len(shape) + 1))
    axes = sorted(rs.choice(len(shape), nax, replace=False).tolist())
    rdt = torch.float64 if dbl else torch.float32
    cdt = torch.complex128 if dbl else torch.complex64
    tol = 1e-13 if dbl else 1e-6
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    print(f"case seed={seed} op={op} shape={shape} axes={axes} dbl={dbl}", flush=True)
    if op in ("c2c", "r2c", "dct2", "dct3"):
        base = torch.randn(shape, dtype=cdt if op == "c2c" else rdt, device="cuda", generator=g)
        x = view_of(base, rs)
    if op == "c2c":
        fwd = bool(rs.randint(2))
        y = torch.empty(x.shape, dtype=cdt, device="cuda")
        R.c2c(x, y, np.array(axes), fwd, 0.5, 1)
        xh = x.cpu().numpy().astype(np.complex128)
        want = 0.5 * (sfft.fftn(xh, axes=axes) if fwd else sfft.ifftn(xh, axes=axes, norm="forward"))
    elif op == "r2c":
        hs = list(x.shape)
        hs[axes[-1]] = hs[axes[-1]] // 2 + 1
        y = torch.empty(hs, dtype=cdt, device="cuda")
        R.r2c(x, y, np.array(axes), True, 1.0, 1)
        want = sfft.rfftn(x.cpu().numpy().astype(np.float64), axes=axes)
    elif op == "c2r":
        n = list(shape)
        hs = list(shape)
        hs[axes[-1]] = hs[axes[-1]] // 2 + 1
        spec = torch.randn(hs, dtype=cdt, device="cuda", generator=g)
        x = view_of(spec, rs)
        out_shape = list(x.shape)
        last = axes[-1]
        n_last = 2 * (x.shape[last] - 1) + int(rs.randint(2)) if x.shape[last] > 1 else 1
        out_shape[last] = n_last
        y = torch.empty(out_shape, dtype=rdt, device="cuda")
        R.c2r(x, y, np.array(axes), False, 1.0, 1)
        xh = x.cpu().numpy().astype(np.complex128)
        # scipy's irfftn, like the reference, runs the leading axes first and
        # ignores the imaginary parts of bins 0 and n/2 of the last axis
        s = [out_shape[a] for a in axes]
        want = sfft.irfftn(xh, s=s, axes=axes, norm="forward")
    else:
        typ = 2 if op == "dct2" else 3
        y = torch.empty(x.shape, dtype=rdt, device="cuda")
        R.dct(x, y, np.array(axes), typ, 1.0, False, 1)
        want = sfft.dctn(x.cpu().numpy().astype(np.float64), type=typ, axes=axes)
    torch.cuda.synchronize()
    e = rel_l2(y.cpu().numpy(), want)
    assert e <= tol, f"rel-L2 {e:.3e}"

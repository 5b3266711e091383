"""GPU parity: every low-level entry point through the C ABI against the
oracle (long-double restatement) and, where it was built, the reference
engine itself.  Bar: rel-L2 <= 1e-13 (complex128/float64), <= 1e-6
(complex64/float32).  Inputs are seeded N(0,1)."""
import numpy as np
import pytest

import rocket_fft_amd as R
from helpers import (CTYPE, REF, crandn, on_gpu, oracle_out, ref_out, rel_l2, rrandn,
                     tol_of)

pytestmark = pytest.mark.gpu


def check(fname, ain, aout_like, *args, inplace=False, tol=None):
    print(f"case {fname} in={ain.shape}{ain.dtype} axes/args={args}", flush=True)
    got = on_gpu(R, fname, ain, aout_like, *args, inplace=inplace)
    tol = tol if tol is not None else tol_of(aout_like.dtype)
    want = oracle_out(fname, ain, aout_like, *args)
    e = rel_l2(got, want)
    assert e <= tol, f"{fname} vs oracle: rel-L2 {e:.3e} > {tol:.0e}"
    if REF.available():
        e2 = rel_l2(got, ref_out(fname, ain, aout_like, *args))
        assert e2 <= tol, f"{fname} vs reference: rel-L2 {e2:.3e} > {tol:.0e}"
    return e


C2C_LENGTHS = [1, 2, 3, 4, 5, 7, 8, 11, 13, 16, 17, 30, 32, 47, 64, 97, 100, 128, 243, 256,
               343, 512, 1000, 1024, 1536, 2187, 4096, 4913, 8192, 4099, 8232, 16384]


@pytest.mark.parametrize("n", C2C_LENGTHS)
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_c2c_lengths(n, dtype):
    rs = np.random.RandomState(n)
    x = crandn(rs, (3, n), dtype)
    check("c2c", x, x, np.array([1]), True, 1.0)
    check("c2c", x, x, np.array([1]), False, 1.0 / n)


@pytest.mark.parametrize("n,dtype", [(2187, np.complex64), (4099, np.complex64),
                                     (8232, np.complex64), (2187, np.complex128),
                                     (4116, np.complex128), (4913, np.complex128)])
def test_wide_and_narrow_line_groups(n, dtype):
    """A launch of at most one line group per CU runs the line kernel with
    1024 (complex64, from 2048 elements per group) / 512 (complex128, from
    2048) threads per group, a larger batch with 256: 7 lines take the
    first, 300 lines (more groups than the MI355X's 256 CUs) the second.
    Both against the oracle, and the 7 lines of both launches against each
    other."""
    rs = np.random.RandomState(n + 17)
    x = crandn(rs, (300, n), dtype)
    check("c2c", x, x, np.array([1]), True, 1.0)
    check("c2c", x[:7].copy(), x[:7], np.array([1]), True, 1.0)
    narrow = on_gpu(R, "c2c", x, x, np.array([1]), True, 1.0)
    wide = on_gpu(R, "c2c", x[:7].copy(), x[:7], np.array([1]), True, 1.0)
    e = rel_l2(wide, narrow[:7])
    assert e <= tol_of(dtype), f"wide vs narrow groups: rel-L2 {e:.3e}"


@pytest.mark.parametrize("n,dtype", [(4913, np.complex128), (4928, np.complex128),
                                     (4929, np.complex128), (9856, np.complex64),
                                     (9857, np.complex64)])
@pytest.mark.parametrize("lines", [3, 300])
def test_one_pass_limit(n, dtype, lines):
    """Lengths at the one-LDS-pass limit (two line buffers beside the line
    kernel's 5 KB of static LDS in 160 KB: 4928 complex128, 9856 complex64
    elements) and one past it (four-step passes), few and many lines."""
    rs = np.random.RandomState(n + lines)
    x = crandn(rs, (lines, n), dtype)
    check("c2c", x, x, np.array([1]), True, 1.0)
    check("c2c", x, x, np.array([1]), False, 1.0 / n)


@pytest.mark.parametrize("shape,dtype", [((1000, 24), np.complex64), ((243, 40), np.complex64),
                                         ((600, 12), np.complex128), ((1200, 6), np.complex128)])
def test_wide_line_groups_interleaved(shape, dtype):
    """Lines along axis 0 (interleaved in memory, several adjacent lines per
    group): a few groups of at least 2048 elements, so the wide-group
    launch with the line-fastest loads and stores."""
    rs = np.random.RandomState(shape[0] + shape[1])
    x = crandn(rs, shape, dtype)
    check("c2c", x, x, np.array([0]), True, 1.0)
    check("c2c", x, x, np.array([0]), False, 1.0 / shape[0])


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_c2c_two_pass_long(dtype):
    rs = np.random.RandomState(7)
    x = crandn(rs, (2, 65536), dtype)
    check("c2c", x, x, np.array([1]), True, 1.0)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_c2c_bluestein_two_pass(dtype):
    rs = np.random.RandomState(8)
    x = crandn(rs, (2, 65537), dtype)
    check("c2c", x, x, np.array([1]), True, 1.0)
    check("c2c", x, x, np.array([1]), False, 1.0 / 65537)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_bluestein_two_pass_pow2_columns(dtype):
    """n = 50174 = 2 * 25087: Bluestein length 100352 = 256 * 392, so the
    column pass is a power-of-two length and runs on the specialised line
    kernel, which must apply the chirp (and the zero padding from m = n on)
    while it loads.  c2c both ways; r2c / c2r of the same n run the half
    length 25087 (a Bluestein prime) between the pair pre/post kernels."""
    n = 50174
    dbl = dtype == np.complex128
    kind, _, n1, n2 = R.gpu_schedule(n, dbl)
    print(f"case bluestein pow2 columns n={n} {dtype.__name__} schedule={kind, n1, n2}", flush=True)
    assert kind == 3 and n1 * n2 == 100352 and n1 & (n1 - 1) == 0 and n1 <= 4096
    rs = np.random.RandomState(9)
    x = crandn(rs, (2, n), dtype)
    check("c2c", x, x, np.array([1]), True, 1.0)
    check("c2c", x, x, np.array([1]), False, 1.0 / n)
    xr = rrandn(rs, (2, n), np.float64 if dbl else np.float32)
    yr = np.zeros((2, n // 2 + 1), dtype)
    check("r2c", xr, yr, np.array([1]), True, 1.0)
    xc = crandn(rs, (2, n // 2 + 1), dtype)
    check("c2r", xc, xr, np.array([1]), False, 1.0 / n)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("n,lines", [(32, 1000), (64, 777), (32, 3), (64, 1)])
def test_short_lines_partial_tiles(dtype, n, lines):
    """Contiguous lines of 32 / 64 (2 / 4 lanes per line on the specialised
    kernel): line counts that leave partial tiles, both directions, in
    place."""
    rs = np.random.RandomState(n + lines)
    x = crandn(rs, (lines, n), dtype)
    check("c2c", x, x, np.array([1]), True, 1.0)
    check("c2c", x, x, np.array([1]), False, 1.0 / n, inplace=True)


@pytest.mark.parametrize("lines", [3, 17])
def test_two_pass_2p20_pipelined(lines):
    """complex64 2^20 = 1024 x 1024: both four-step passes run on the
    persistent pipelined kernel (checked through the per-launch timing
    labels); 3 lines are 192 tiles, fewer than the CUs.  Both directions
    against the oracle."""
    import torch
    n = 1 << 20
    rs = np.random.RandomState(lines)
    x = crandn(rs, (lines, n), np.complex64)
    R.kernel_timing_report()
    R.kernel_timing(True)
    try:
        check("c2c", x, x, np.array([1]), True, 1.0)
    finally:
        torch.cuda.synchronize()
        R.kernel_timing(False)
    labels = R.kernel_timing_report()
    print("labels", sorted(labels), flush=True)
    assert any(k.startswith("k_pass1024 passA n=1024") for k in labels)
    assert any(k.startswith("k_pass1024 passB n=1024") for k in labels)
    check("c2c", x, x, np.array([1]), False, 1.0 / n)


@pytest.mark.parametrize("lines", [2048, 2051])
def test_c2c_4096_pipelined(lines):
    """complex64 lines of 4096, enough of them for the persistent pipelined
    kernel (4-line tiles, at least two per CU; checked through the per-launch
    timing labels); 2051 lines leave a partial last tile.  Both directions
    against the oracle."""
    import torch
    n = 4096
    rs = np.random.RandomState(lines)
    x = crandn(rs, (lines, n), np.complex64)
    R.kernel_timing_report()
    R.kernel_timing(True)
    try:
        check("c2c", x, x, np.array([1]), True, 1.0)
    finally:
        torch.cuda.synchronize()
        R.kernel_timing(False)
    labels = R.kernel_timing_report()
    print("labels", sorted(labels), flush=True)
    assert any(k.startswith("k_pipe4096 c2c n=4096 c64") for k in labels)
    check("c2c", x, x, np.array([1]), False, 1.0 / n)


def _labels_of(fn):
    import torch
    R.kernel_timing_report()
    R.kernel_timing(True)
    try:
        fn()
        torch.cuda.synchronize()
    finally:
        R.kernel_timing(False)
    return R.kernel_timing_report()


@pytest.mark.parametrize("case", ["step2", "batch2d"])
def test_pipelined_kernels_on_views(case):
    """The persistent pipelined kernels on non-contiguous device views:
    lines of 4096 and of 2^20 with an element step of 2 (every other element
    of longer rows), and lines of 4096 under two unmergeable batch
    dimensions (a transposed batch); against the oracle."""
    import torch
    rs = np.random.RandomState(11)
    if case == "step2":
        shapes = [((2048, 8192), "k_pipe4096 c2c n=4096"), ((3, 1 << 21), "k_pass1024 passA")]
    else:
        shapes = [((1024, 2, 4096), "k_pipe4096 c2c n=4096")]
    for shape, family in shapes:
        base = crandn(rs, shape, np.complex64)
        tb = torch.from_numpy(base).cuda()
        if case == "step2":
            view, hv = tb[:, ::2], base[:, ::2]
        else:
            view, hv = tb.transpose(0, 1), base.transpose(1, 0, 2)
        ax = np.array([view.dim() - 1])
        print(f"case pipelined view {case} base={shape} view={tuple(view.shape)} "
              f"strides={view.stride()}", flush=True)
        out = torch.zeros(view.shape, dtype=torch.complex64, device="cuda")
        labels = _labels_of(lambda: R.c2c(view, out, ax, True, 1.0))
        assert any(k.startswith(family) for k in labels), sorted(labels)
        want = oracle_out("c2c", np.ascontiguousarray(hv), np.ascontiguousarray(hv), ax, True, 1.0)
        e = rel_l2(out.cpu().numpy(), want)
        assert e <= 1e-6, e
        assert torch.equal(tb, torch.from_numpy(base).cuda()), "input modified"


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("n", [8193, 32749, 65537])
def test_bluestein_pow2_rows(dtype, n):
    """Bluestein lengths whose padding from 2n - 2 is a power of two (8193 and
    65537 = 2^k + 1 pad to exactly 2^(k+1); prime 32749 to 65536) split into
    power-of-two columns and rows, so both the column passes and the row pass
    (forward FFT, kernel product, inverse FFT, inverse twiddle, in place) run
    on the specialised kernel.  Contiguous and strided (axis 0) lines, both
    directions."""
    dbl = dtype == np.complex128
    kind, _, n1, n2 = R.gpu_schedule(n, dbl)
    print(f"case bluestein pow2 rows n={n} {dtype.__name__} schedule={kind, n1, n2}", flush=True)
    assert kind == 3 and n1 * n2 >= 2 * n - 2 and n2 & (n2 - 1) == 0
    rs = np.random.RandomState(n)
    x = crandn(rs, (3, n), dtype)
    check("c2c", x, x, np.array([1]), True, 1.0)
    check("c2c", x, x, np.array([1]), False, 1.0 / n)
    xt = crandn(rs, (n, 5), dtype)
    check("c2c", xt, xt, np.array([0]), True, 1.0)
    check("c2c", xt, xt, np.array([0]), False, 0.5, inplace=True)


@pytest.mark.parametrize("shape,axes", [((16, 24, 10), (0, 1, 2)), ((16, 24, 10), (2, 0)),
                                        ((33, 17), (0,)), ((33, 17), (1, 0)),
                                        ((5, 6, 7, 8), (3, 1)), ((64, 3), (0,))])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_c2c_axes(shape, axes, dtype):
    rs = np.random.RandomState(1)
    x = crandn(rs, shape, dtype)
    check("c2c", x, x, np.array(axes), True, 0.5)
    check("c2c", x, x, np.array(axes), False, 1.0, inplace=True)


# 3086 = 2 * 1543, 18134 = 2 * 9067 (half length a Bluestein prime: pairs ->
# c2c of n/2 -> split twiddle), 40000 and 65536 (half length beyond one LDS line)
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 15, 16, 30, 64, 101, 128, 255, 256,
                               512, 1000, 4097, 4099, 8192, 3086, 18134, 40000, 65536])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_r2c_c2r_lengths(n, dtype):
    rs = np.random.RandomState(n + 3)
    x = rrandn(rs, (4, n), dtype)
    h = np.zeros((4, n // 2 + 1), CTYPE[dtype])
    check("r2c", x, h, np.array([1]), True, 1.0)
    check("r2c", x, h, np.array([1]), False, 2.0)
    X = crandn(rs, (4, n // 2 + 1), CTYPE[dtype])
    check("c2r", X, x, np.array([1]), False, 1.0 / n)
    check("c2r", X, x, np.array([1]), True, 1.0)


@pytest.mark.parametrize("shape,axes", [((8, 6, 10), (0, 1, 2)), ((8, 6, 10), (2, 0)),
                                        ((9, 7), (0, 1)), ((12, 5), (0,)), ((4, 4, 4), (1,)),
                                        ((3086, 3), (0,)), ((3, 3086), (1, 0))])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_r2c_c2r_axes(shape, axes, dtype):
    rs = np.random.RandomState(2)
    x = rrandn(rs, shape, dtype)
    hs = list(shape)
    hs[axes[-1]] = shape[axes[-1]] // 2 + 1
    h = np.zeros(hs, CTYPE[dtype])
    check("r2c", x, h, np.array(axes), True, 1.0)
    X = crandn(rs, hs, CTYPE[dtype])
    check("c2r", X, x, np.array(axes), False, 1.0)
    full = np.zeros(shape, CTYPE[dtype])
    check("c2c_sym", x, full, np.array(axes), True, 1.0)
    check("c2c_sym", x, full, np.array(axes), False, 0.5)


@pytest.mark.parametrize("kind", ["dct", "dst"])
@pytest.mark.parametrize("type", [1, 2, 3, 4])
@pytest.mark.parametrize("ortho", [False, True])
@pytest.mark.parametrize("n", [2, 3, 8, 15, 64, 100])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dct_dst(kind, type, ortho, n, dtype):
    rs = np.random.RandomState(n)
    x = rrandn(rs, (5, n), dtype)
    check(kind, x, x, np.array([1]), type, 1.0, ortho)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dct2_ortho_2d(dtype):
    rs = np.random.RandomState(2)
    x = rs.uniform(-1, 1, (64, 48)).astype(dtype)
    check("dct", x, x, np.array([0, 1]), 2, 1.0 / np.sqrt(4 * 64 * 48), True)
    check("dst", x, x, np.array([1, 0]), 3, 1.0, True, inplace=True)
    x = rs.uniform(-1, 1, (512, 512)).astype(dtype)  # the bench image
    check("dct", x, x, np.array([0, 1]), 2, 1.0 / 1024, True)
    check("dct", x, x, np.array([1, 0]), 3, 1.0 / 1024, True)
    check("dst", x, x, np.array([0, 1]), 2, 1.0, True)


@pytest.mark.parametrize("r2h", [False, True])
@pytest.mark.parametrize("fwd", [False, True])
@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 10, 64])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fftpack(r2h, fwd, n, dtype):
    rs = np.random.RandomState(n)
    x = rrandn(rs, (3, n), dtype)
    check("r2r_fftpack", x, x, np.array([1]), r2h, fwd, 1.0)


@pytest.mark.parametrize("shape,axes", [((10,), (0,)), ((127,), (0,)), ((16, 17), (0, 1)),
                                        ((1, 9), (0, 1)), ((9, 1), (0, 1)),
                                        ((6, 5, 7), (0, 1, 2)), ((6, 5, 7), (2, 1))])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_hartley(shape, axes, dtype):
    rs = np.random.RandomState(4)
    x = rrandn(rs, shape, dtype)
    check("r2r_separable_hartley", x, x, np.array(axes), 1.0)
    check("r2r_genuine_hartley", x, x, np.array(axes), 1.0)


def test_host_arrays_roundtrip():
    """NumPy (host) arrays go through the staging path."""
    rs = np.random.RandomState(5)
    x = crandn(rs, (8, 100), np.complex128)
    y = np.zeros_like(x)
    R.c2c(x, y, np.array([1]), True, 1.0, 1)
    assert rel_l2(y, oracle_out("c2c", x, x, np.array([1]), True, 1.0)) < 1e-13
    z = np.zeros_like(x)
    R.c2c(y, z, np.array([1]), False, 1.0 / 100, 1)
    assert rel_l2(z, x) < 1e-13


def test_strided_views():
    """Non-contiguous and negative-stride views (T/test_numpy_testsuite.py:323-351):
    device tensors with a step, and host arrays with negative strides."""
    import torch
    rs = np.random.RandomState(6)
    base = crandn(rs, (20, 40), np.complex128)
    tb = torch.from_numpy(base).cuda()
    view = tb[::2, 1::3]
    out = torch.zeros(view.shape, dtype=torch.complex128, device="cuda")
    R.c2c(view, out, np.array([0, 1]), True, 1.0)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), np.fft.fftn(base[::2, 1::3])) < 1e-13
    hv = base[::2, ::-3]
    ho = np.zeros(hv.shape, np.complex128)[::-1]
    R.c2c(hv, ho, np.array([0, 1]), True, 1.0)
    assert rel_l2(ho, np.fft.fftn(hv)) < 1e-13


def test_zero_size_noop():
    import torch
    x = torch.zeros((0, 8), dtype=torch.complex64, device="cuda")
    R.c2c(x, x, np.array([1]), True, 1.0)


def test_schedule_report():
    """Prints the GPU schedule of the configuration lengths (informational)."""
    for n in (256, 512, 1000, 1536, 2187, 4096, 4913, 4099, 8232, 65537, 1 << 20):
        for dbl in (False, True):
            print(n, "c128" if dbl else "c64", R.gpu_schedule(n, dbl), flush=True)


@pytest.mark.parametrize("fast", [True, False])
@pytest.mark.parametrize("n", [2, 8, 32, 64, 128, 256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_fast_and_generic_paths(fast, n, dtype):
    """Power-of-two lengths through the specialised kernels and, switched
    off, through the generic ones: contiguous lines, interleaved lines
    (axis 0), r2c/c2r on the last axis and on axis 0."""
    R.set_fast_path(fast)
    try:
        rs = np.random.RandomState(n)
        x = crandn(rs, (n, 24), dtype)
        check("c2c", x, x, np.array([0]), True, 1.0)
        check("c2c", x.T.copy(), x.T, np.array([1]), False, 0.5)
        rt = np.float32 if dtype == np.complex64 else np.float64
        xr = rrandn(rs, (6, 2 * n), rt)
        check("r2c", xr, np.zeros((6, n + 1), dtype), np.array([1]), True, 1.0)
        check("c2r", crandn(rs, (6, n + 1), dtype), xr, np.array([1]), True, 1.0)
        xr0 = rrandn(rs, (2 * n, 10), rt)
        check("r2c", xr0, np.zeros((n + 1, 10), dtype), np.array([0]), False, 1.0)
        check("c2r", crandn(rs, (n + 1, 10), dtype), xr0, np.array([0]), False, 1.0)
    finally:
        R.set_fast_path(True)


@pytest.mark.parametrize("n", [1 << 18, 1 << 20])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_two_pass_pow2_fast(dtype, n):
    """Four-step power-of-two lengths on the specialised kernels (2^20 in
    complex128: interleaved column tiles of 1024 with the split exchange)."""
    rs = np.random.RandomState(11)
    x = crandn(rs, (2, n), dtype)
    check("c2c", x, x, np.array([1]), True, 1.0)
    check("c2c", x, x, np.array([1]), False, 1.0 / n)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_two_pass_2p20_generic(dtype):
    """The 2^20 four-step passes with the specialised and pipelined kernels
    switched off: the generic line kernel runs both passes."""
    R.set_fast_path(False)
    try:
        rs = np.random.RandomState(12)
        x = crandn(rs, (2, 1 << 20), dtype)
        check("c2c", x, x, np.array([1]), True, 1.0)
        check("c2c", x, x, np.array([1]), False, 1.0 / (1 << 20))
    finally:
        R.set_fast_path(True)


@pytest.mark.parametrize("kind,type,n", [("dct", 2, 4099), ("dst", 3, 4099), ("dct", 4, 8192),
                                         ("dct", 1, 6000), ("dst", 1, 5000), ("dct", 3, 12000),
                                         ("dst", 2, 12000)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dct_long_unfused(kind, type, n, dtype):
    """Work lengths that need Bluestein or two passes run pre -> c2c -> post."""
    rs = np.random.RandomState(n)
    x = rrandn(rs, (3, n), dtype)
    check(kind, x, x, np.array([1]), type, 1.0, True)


def test_separable_hartley_strided_axis0():
    rs = np.random.RandomState(12)
    x = rrandn(rs, (300, 7), np.float64)
    check("r2r_separable_hartley", x, x, np.array([0]), 1.0)
    check("r2r_fftpack", x, x, np.array([0]), True, True, 1.0)


@pytest.mark.parametrize("rows,n,dtype", [(80, 1 << 18, np.complex64), (40, 1 << 18, np.complex128),
                                          (160, 65537, np.complex64), (65, 1 << 20, np.complex64)])
def test_two_pass_chunked(rows, n, dtype):
    """Batches whose two-pass intermediate exceeds one chunk (128 MiB; 512 MiB
    for the complex64 1024 x 1024 plans) run chunk by chunk, the last chunk
    short; checked against an FP64 transform (torch.fft, complex128)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(rows)
    tdt = torch.complex64 if dtype == np.complex64 else torch.complex128
    x = torch.randn((rows, n), dtype=tdt, device="cuda", generator=g)
    y = torch.empty_like(x)
    print(f"case chunked c2c rows={rows} n={n} {dtype.__name__}", flush=True)
    R.c2c(x, y, np.array([1]), True, 1.0, 1)
    ref = torch.fft.fft(x.to(torch.complex128), dim=1)
    torch.cuda.synchronize()
    e = float(torch.linalg.norm(y.to(torch.complex128) - ref) / torch.linalg.norm(ref))
    assert e <= tol_of(dtype), e


@pytest.mark.parametrize("shape,dtype", [((8192, 40), np.complex64), ((5000, 33), np.complex64),
                                         ((3000, 17), np.complex128), ((4099, 20), np.complex64),
                                         ((4099, 12), np.complex128), ((2048, 24), np.complex128)])
def test_strided_long_axis(shape, dtype):
    """Long lines along a strided axis (lines interleaved in memory): lengths
    whose direct form cannot keep enough adjacent lines in one LDS group run
    as four-step passes with a line-fastest intermediate (4099: Bluestein)."""
    rs = np.random.RandomState(shape[0] % 1000)
    x = crandn(rs, shape, dtype)
    check("c2c", x, x, np.array([0]), True, 1.0)
    check("c2c", x, x, np.array([0]), False, 1.0 / shape[0])


@pytest.mark.parametrize("n,rows,dtype,kind", [
    (1 << 21, 2, np.complex64, 4), (3 << 20, 1, np.complex128, 4), (1 << 22, 1, np.complex64, 4),
    (1000003, 1, np.complex64, 5), (1000003, 1, np.complex128, 5)])
def test_multi_pass_lengths(n, rows, dtype, kind):
    """Lengths beyond one two-pass split: a column pass and the rows by
    their own plan (kind 4), and primes whose Bluestein length needs it
    (kind 5); against numpy's FP64 transform, both directions, and along a
    strided axis."""
    k, _, n1, n2 = R.gpu_schedule(n, dtype == np.complex128)
    print(f"case multi-pass n={n} {dtype.__name__} schedule={k, n1, n2}", flush=True)
    assert k == kind
    rs = np.random.RandomState(n % 997)
    x = crandn(rs, (rows, n), dtype)
    tol = tol_of(dtype)
    want = np.fft.fft(x.astype(np.complex128), axis=1)
    got = on_gpu(R, "c2c", x, x, np.array([1]), True, 1.0)
    assert rel_l2(got, want) <= tol
    got = on_gpu(R, "c2c", x, x, np.array([1]), False, 1.0 / n)
    assert rel_l2(got, np.fft.ifft(x.astype(np.complex128), axis=1)) <= tol
    if rows == 2:
        xt = np.ascontiguousarray(x.T)
        got = on_gpu(R, "c2c", xt, xt, np.array([0]), True, 1.0)
        assert rel_l2(got, want.T) <= tol

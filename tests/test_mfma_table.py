"""Host checks of the DFT-16 operand table the matrix-core kernels load
(rfh_mfma_operands; no GPU needed).

For each precision mode and direction the table must (a) hold values exact in
the operand format (FP16: numpy float16 round trip; BF16: the low 16 bits of
the float32 pattern zero), (b) sum over its split parts to the FP64 root to
the split's precision, and (c) laid out per lane as the hardware A operand
(lane l, g = l // 16: row l % 16, columns 8 g .. 8 g + 7), form
A_re = [Re F, -Im F] and A_im = [Im F, Re F] with F[k][r] = e^{dir 2 pi i k r/16},
so that [A_re; A_im] [Re v; Im v] is numpy's DFT of v (forward) or 16 times
its inverse DFT (backward)."""
import ctypes

import numpy as np
import pytest

import rocket_fft_amd as R

TERMS = {1: 2, 2: 3, 3: 2}
# bound on |sum of parts - root| for roots of modulus <= 1
PREC = {1: 2.0 ** -22, 2: 2.0 ** -24, 3: 2.0 ** -16}


def table(mode, direction):
    L = R._load()
    L.rfh_mfma_operands.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    L.rfh_mfma_operands.restype = ctypes.c_int
    n = L.rfh_mfma_operands(mode, direction, None, 0)
    assert n == TERMS[mode] * 2 * 64 * 8
    out = np.zeros(n, np.float32)
    assert L.rfh_mfma_operands(mode, direction, out.ctypes.data, n) == n
    return out.reshape(TERMS[mode], 2, 64, 8)


def matrices(tab):
    """(A_re, A_im) 16 x 32 as sums of the parts, and the per-part arrays."""
    parts = np.zeros((tab.shape[0], 2, 16, 32))
    for l in range(64):
        g = l // 16
        parts[:, :, l % 16, 8 * g:8 * g + 8] = tab[:, :, l, :]
    full = parts.sum(axis=0)
    return full[0], full[1], parts


@pytest.mark.parametrize("mode", [1, 2, 3])
@pytest.mark.parametrize("direction", [-1, 1])
def test_operand_table(mode, direction):
    tab = table(mode, direction).astype(np.float64)
    if mode == 1:
        assert np.array_equal(tab.astype(np.float16).astype(np.float64), tab)
    else:
        bits = tab.astype(np.float32).view(np.uint32)
        assert not np.any(bits & 0xFFFF)
    a_re, a_im, _ = matrices(tab)
    k = np.arange(16)[:, None]
    r = np.arange(16)[None, :]
    F = np.exp(direction * 2j * np.pi * k * r / 16)
    want_re = np.hstack([F.real, -F.imag])
    want_im = np.hstack([F.imag, F.real])
    assert np.max(np.abs(a_re - want_re)) <= PREC[mode]
    assert np.max(np.abs(a_im - want_im)) <= PREC[mode]
    rs = np.random.RandomState(mode)
    v = rs.standard_normal((16, 5)) + 1j * rs.standard_normal((16, 5))
    B = np.vstack([v.real, v.imag])
    X = a_re @ B + 1j * (a_im @ B)
    ref = np.fft.fft(v, axis=0) if direction < 0 else 16 * np.fft.ifft(v, axis=0)
    assert np.linalg.norm(X - ref) / np.linalg.norm(ref) <= 4 * PREC[mode]


def test_bad_mode_rejected():
    L = R._load()
    L.rfh_mfma_operands.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    L.rfh_mfma_operands.restype = ctypes.c_int
    assert L.rfh_mfma_operands(0, -1, None, 0) == -1
    assert L.rfh_mfma_operands(1, 0, None, 0) == -1

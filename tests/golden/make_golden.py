"""Writes tests/golden/reference_vectors.npz: inputs and outputs of the
reference engine (rocket_fft/_pocketfft_hdronly.h compiled in place by
oracle/Makefile into oracle/_ref) on small seeded cases, plus its integer
plan facts.  Run in a checkout where oracle/_ref is built:

    make -C oracle ref && python tests/golden/make_golden.py

The file holds data only (numeric arrays and a unicode case table; no
pickled objects), so the CPU tests can pin the oracle against the
reference's own output where the reference is not available.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import ref as REF  # noqa: E402

OUT = os.path.join(HERE, "reference_vectors.npz")


def cases():
    rs = np.random.RandomState(2024)

    def cx(shape, dt):
        return (rs.standard_normal(shape) + 1j * rs.standard_normal(shape)).astype(dt)

    def rx(shape, dt):
        return rs.standard_normal(shape).astype(dt)

    out = []
    for dt in (np.complex64, np.complex128):
        for n in (1, 2, 3, 5, 7, 8, 11, 12, 13, 16, 17, 30, 49, 64, 97, 100, 120, 1000, 4099):
            rows = 2 if n > 1000 else 3
            out.append(("c2c", cx((rows, n), dt), None, ([1], True, 1.0)))
            out.append(("c2c", cx((rows, n), dt), None, ([1], False, 1.0 / n)))
        out.append(("c2c", cx((6, 10, 4), dt), None, ([0, 2, 1], True, 0.5)))
    for dt in (np.float32, np.float64):
        ct = np.complex64 if dt == np.float32 else np.complex128
        for n in (1, 2, 3, 6, 9, 16, 31, 64, 101, 256):
            x = rx((3, n), dt)
            out.append(("r2c", x, np.zeros((3, n // 2 + 1), ct), ([1], True, 1.0)))
            out.append(("r2c", x, np.zeros((3, n // 2 + 1), ct), ([1], False, 1.0)))
            X = cx((3, n // 2 + 1), ct)
            out.append(("c2r", X, np.zeros((3, n), dt), ([1], False, 1.0 / n)))
            out.append(("c2r", X, np.zeros((3, n), dt), ([1], True, 1.0)))
        for shape, axes in (((6, 8, 5), [0, 1, 2]), ((7, 9), [1, 0]), ((4, 6), [0])):
            x = rx(shape, dt)
            hs = list(shape)
            hs[axes[-1]] = shape[axes[-1]] // 2 + 1
            out.append(("r2c", x, np.zeros(hs, ct), (axes, True, 1.0)))
            out.append(("c2r", cx(hs, ct), np.zeros(shape, dt), (axes, False, 1.0)))
            out.append(("c2c_sym", x, np.zeros(shape, ct), (axes, True, 1.0)))
            out.append(("r2r_genuine_hartley", x, None, (axes, 1.0)))
            out.append(("r2r_separable_hartley", x, None, (axes, 1.0)))
        for kind in ("dct", "dst"):
            for t in (1, 2, 3, 4):
                for n in (2, 3, 8, 15, 64):
                    for ortho in (False, True):
                        out.append((kind, rx((2, n), dt), None, ([1], t, 1.0, ortho)))
            out.append((kind, rx((8, 12), dt), None, ([0, 1], 2, 1.0 / 24, True)))
        for r2h in (False, True):
            for fwd in (False, True):
                for n in (1, 2, 3, 4, 7, 10):
                    out.append(("r2r_fftpack", rx((2, n), dt), None, ([1], r2h, fwd, 1.0)))
    return out


def main():
    if not REF.available():
        raise SystemExit("build oracle/_ref first (make -C oracle ref)")
    arrays, meta = {}, []
    for i, (fname, x, aout, args) in enumerate(cases()):
        y = np.zeros_like(x) if aout is None else aout
        getattr(REF, fname)(x, y, np.array(args[0], dtype=np.uint64), *args[1:])
        arrays[f"in_{i:03d}"] = x
        arrays[f"out_{i:03d}"] = y
        meta.append(json.dumps({"fname": fname, "axes": list(args[0]), "args": list(args[1:])}))
    ns = np.arange(1, 4097)
    arrays["good_size_n"] = ns
    arrays["good_size_cmplx"] = np.array([REF.good_size(n, False) for n in ns], dtype=np.int64)
    arrays["good_size_real"] = np.array([REF.good_size(n, True) for n in ns], dtype=np.int64)
    big = np.array([4097, 4099, 4913, 8191, 65537, 100003, 131071, 1 << 20, 999983, 1234567],
                   dtype=np.int64)
    arrays["big_n"] = big
    arrays["big_good_cmplx"] = np.array([REF.good_size(n, False) for n in big], dtype=np.int64)
    arrays["big_good_real"] = np.array([REF.good_size(n, True) for n in big], dtype=np.int64)
    fac_n = np.arange(2, 2049)
    arrays["fac_n"] = fac_n
    arrays["cfftp"] = np.array([",".join(map(str, REF.cfftp_factors(n))) for n in fac_n])
    arrays["rfftp"] = np.array([",".join(map(str, REF.rfftp_factors(n))) for n in fac_n])
    arrays["blue_c"] = np.array([REF.plan_c(n)[1] for n in fac_n], dtype=np.int64)
    arrays["blue_r"] = np.array([REF.plan_r(n)[1] for n in fac_n], dtype=np.int64)
    arrays["meta"] = np.array(meta)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT}: {len(meta)} cases, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()

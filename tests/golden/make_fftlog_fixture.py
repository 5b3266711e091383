"""Writes tests/golden/fftlog_fortran.json: the four Fortran FFTLog result
tables that rocket-fft's tests/test_fftlog.py:25-125 compares fht against
(numbers only, read from that file as text).  Run in a checkout next to the
reference:

    python tests/golden/make_fftlog_fixture.py /root/reference
"""
import json
import os
import re
import sys


def main(ref_root):
    src = open(os.path.join(ref_root, "tests", "test_fftlog.py")).read()
    blocks = re.findall(r"theirs = \[(.*?)\]", src, re.S)
    vals = [[float(v) for v in re.findall(r"[-+]?\d+\.\d+e[-+]?\d+", b)] for b in blocks[:4]]
    assert len(vals) == 4 and all(len(v) == 16 for v in vals), [len(v) for v in vals]
    doc = {
        "about": "Fortran FFTLog outputs for a(r) = r^(mu+1) exp(-r^2/2), r = logspace(-4, 4, "
                 "16), mu = 0.3 (data of rocket-fft tests/test_fftlog.py:25-125, itself from "
                 "SciPy's test_fftlog.py). Cases: offset 0 / bias 0; optimal offset "
                 "(fhtoffset) with bias 0, 0.8 and -0.8.",
        "mu": 0.3,
        "cases": [
            {"bias": 0.0, "optimal_offset": False, "expected": vals[0]},
            {"bias": 0.0, "optimal_offset": True, "expected": vals[1]},
            {"bias": 0.8, "optimal_offset": True, "expected": vals[2]},
            {"bias": -0.8, "optimal_offset": True, "expected": vals[3]},
        ],
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fftlog_fortran.json")
    with open(out, "w") as f:
        json.dump(doc, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

"""Host checks of the plan record file format (rfh_plan_record_roundtrip; no
GPU needed).  Every plan kind -- direct, two-pass, multi-pass, Bluestein in
one LDS pass, Bluestein two-pass and Bluestein around a multi-pass sub-plan --
and the real-input plans must survive a write and a read back field for
field.  The file layout is checked independently of the library: the magic,
the key fields (kind, length, precision, LDS budget, stride class), the plan
kind, and an FNV-1a-64 checksum of everything before it, recomputed here."""
import struct

import pytest

import rocket_fft_amd as R

# (n, double, lds budget, interleaved lines, real, expected plan kind)
# kinds: 0 direct, 1 Bluestein one pass, 2 two-pass, 3 Bluestein two-pass,
#        4 multi-pass, 5 Bluestein multi-pass; None: real plan
CASES = [
    (4096, False, 131072, False, False, 0),
    (1 << 16, False, 131072, False, False, 2),
    (1 << 20, True, 131072, True, False, 2),
    (1 << 22, False, 16384, False, False, 4),
    (1009, False, 131072, False, False, 1),
    (65537, False, 131072, False, False, 3),
    (65537, True, 131072, True, False, 3),
    (100003, False, 8192, False, False, 5),
    (1, False, 131072, False, False, 0),
    (4096, False, 131072, False, True, None),
    (65537, True, 131072, False, True, None),
    (7, True, 131072, False, True, None),
]


def fnv1a64(b):
    h = 1469598103934665603
    for c in b:
        h = ((h ^ c) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.mark.parametrize("n,dbl,budget,inter,real,kind", CASES)
def test_plan_record_roundtrip(tmp_path, n, dbl, budget, inter, real, kind):
    path = tmp_path / "plan.bin"
    assert R.plan_record_roundtrip(n, dbl, budget, inter, real, path=str(path)) == 0
    f = path.read_bytes()
    assert f[:8] == b"RFHPLAN1"
    assert struct.unpack("<Q", f[-8:])[0] == fnv1a64(f[:-8])
    k, nn, d, bb, it = struct.unpack("<BQBQB", f[8:27])
    assert (k, nn, d, bb, it) == (1 if real else 0, n, int(dbl), budget,
                                  0 if real else int(inter))
    if not real:
        assert struct.unpack("<i", f[27:31])[0] == kind


def test_plan_record_rejects_zero_budget(tmp_path):
    with pytest.raises(RuntimeError):
        R.plan_record_roundtrip(4096, False, 0, path=str(tmp_path / "x.bin"))

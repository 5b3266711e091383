#!/usr/bin/env python3
"""Throughput of the transform engine on one MI355X.

Headline workload (one "step"): rocket-fft's precision configuration --
forward c2c of complex64 lines of 2^20 (batch 64) and of complex64 lines of
4096 (batch 16384), 1 GiB of input each, inputs iid N(0,1) from a seeded
generator, resident in HBM before timing.  The step runs in the fastest
arithmetic that keeps rel-L2 <= 1e-6 against FP64 on both input sets of the
configuration (iid N(0,1), and magnitudes log-uniform over 1e-3..1e3 with
uniform phase): the plain FP32 butterflies or a split-precision matrix-core
form; every form's time and error are reported beside it.
value = algorithmic bytes per step (each transform reads its input once and
writes its output once: 2 x 1,073,741,824 B) / time per step.  The step is
captured once in a HIP graph and replayed.

roofline: the kernel with the largest share of the step (HIP events on the
stream it runs on, over graph replays of its own launches), algorithmic bytes
per launch / mean launch time against 8 TB/s; traffic from rocprofv3 counters
(see TRAFFIC below).
cpu_baseline: the reference engine compiled from its sources (oracle/_ref),
the same step on all host cores, bounded to about ten seconds; the other
configurations are timed on the host beside their GPU figures.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "rocket-fft_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = ("batched FFT GB/s (%HBM peak): MFMA & ragged vs plain GPU path at stated rel-L2")
HBM_PEAK_GBS = 8000.0
SHAPES = ((64, 1 << 20), (16384, 4096))
BYTES_PER_SHAPE = 2 * 64 * (1 << 20) * 8
BYTES_STEP = 2 * BYTES_PER_SHAPE
MODES = ("fp32", "mfma_fp16", "mfma_bf16")
# HBM bytes per launch of the step's kernels, measured with
#   rocprofv3 --pmc FETCH_SIZE -- python3 bench.py --probe 64x1048576:fp32
#   (then WRITE_SIZE; and the same for 16384x4096:fp32)
# each counter scaled by what it reads for the probe's device copies of the
# same array in the same run (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000);
# None where not measured for a kernel.
TRAFFIC = {
    "k_pass1024 passA n=1024 c64 strided": 1074434252,  # 1.0006 x the 1,073,741,824 algorithmic
    "k_pass1024 passB n=1024 c64": 1074200323,          # 1.0004 x
    "k_pipe4096 c2c n=4096 c64": 1074480235,            # 1.0007 x 1,073,741,824
}


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def precision_sets(torch, shape, seed=4):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    sets = {"normal": torch.randn(shape, dtype=torch.complex64, device="cuda", generator=g)}
    mag = torch.exp(torch.empty(shape, device="cuda").uniform_(np.log(1e-3), np.log(1e3),
                                                               generator=g))
    ph = torch.empty(shape, device="cuda").uniform_(0, 2 * np.pi, generator=g)
    sets["loguniform"] = torch.polar(mag, ph).to(torch.complex64)
    return sets


def err_vs_fp64(torch, x, y, rows=None):
    """rel-L2 of y against torch's complex128 FFT of x along axis 1 (in
    slices of rows, to bound the FP64 temporaries)."""
    step = rows or max(1, (1 << 24) // x.shape[1])
    num = den = 0.0
    for lo in range(0, x.shape[0], step):
        ref = torch.fft.fft(x[lo:lo + step].to(torch.complex128), dim=1)
        num += float(torch.sum(torch.abs(y[lo:lo + step].to(torch.complex128) - ref) ** 2))
        den += float(torch.sum(torch.abs(ref) ** 2))
        del ref
    return float(np.sqrt(num / den))


def graph_time(torch, fn, stream, reps, batches=5, inner=1):
    """Capture fn `inner` times in one graph and replay it: `batches` timed
    batches of `reps` replays (events on `stream`); median over batches of
    the mean ms per call of fn (inner > 1 keeps short calls from being
    timed at the host's graph-launch rate)."""
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        for _ in range(inner):
            fn()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(batches):
        with torch.cuda.stream(stream):
            e0.record(stream)
            for _ in range(reps):
                g.replay()
            e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / (reps * inner))
    return float(np.median(times)), g


def with_mode(R, mode, fn):
    prev = R.set_precision(mode)
    try:
        return fn()
    finally:
        R.set_precision(prev)


def precision_forms(torch, R, stream, data, reps):
    """Per arithmetic form: ms of each shape (graph replays), GB/s, rel-L2
    against FP64 on both input sets, matrix-core launches."""
    out = {}
    ax = np.array([1])
    for mode in MODES:
        entry = {}
        for shape, sets in data.items():
            name = f"{shape[0]}x{shape[1]}"
            x = sets["normal"]
            y = torch.empty_like(x)
            before = R.mfma_launches(None if mode == "fp32" else mode)
            with torch.cuda.stream(stream):
                with_mode(R, mode, lambda: R.c2c(x, y, ax, True, 1.0, 1))
            torch.cuda.synchronize()
            launched = R.mfma_launches(None if mode == "fp32" else mode) - before
            ms, _ = with_mode(R, mode, lambda: graph_time(
                torch, lambda: R.c2c(x, y, ax, True, 1.0, 1), stream, reps))
            errs = {}
            for set_name, xs in sets.items():
                ys = torch.empty_like(xs)
                with torch.cuda.stream(stream):
                    with_mode(R, mode, lambda: R.c2c(xs, ys, ax, True, 1.0, 1))
                torch.cuda.synchronize()
                errs[set_name] = err_vs_fp64(torch, xs, ys)
                del ys
            entry[name] = {"ms": round(ms, 4), "GB/s": round(BYTES_PER_SHAPE / ms / 1e6, 1),
                           "rel_l2_vs_fp64": errs, "matrix_core_launches": launched}
            del y
        entry["step_ms"] = round(sum(v["ms"] for v in entry.values() if isinstance(v, dict)), 4)
        entry["meets_1e-6"] = all(e <= 1e-6 for v in entry.values() if isinstance(v, dict)
                                  for e in v["rel_l2_vs_fp64"].values())
        out[mode] = entry
    base = out["fp32"]["step_ms"]
    for mode in out:
        out[mode]["speedup_vs_fp32"] = round(base / out[mode]["step_ms"], 3)
        for shape_name in [k for k in out[mode] if "x" in k and isinstance(out[mode][k], dict)]:
            out[mode][shape_name]["speedup_vs_fp32"] = round(
                out["fp32"][shape_name]["ms"] / out[mode][shape_name]["ms"], 3)
    return out


def roofline(torch, R, stream, mode, xs, ys, steps):
    """Per-launch HIP-event timing of every line kernel of the step (events
    recorded by the library on the stream each kernel is launched on), over
    `steps` eager steps; the dominant kernel is the one with the largest
    total time.  achieved = its algorithmic bytes per launch (the launch's
    lines read once and written once) / its mean launch time."""
    ax = np.array([1])
    prev = R.set_precision(mode)
    try:
        with torch.cuda.stream(stream):
            for x, y in zip(xs, ys):
                R.c2c(x, y, ax, True, 1.0, 1)
        torch.cuda.synchronize()
        R.kernel_timing_report()  # drop anything recorded before
        R.kernel_timing(True)
        try:
            with torch.cuda.stream(stream):
                for _ in range(steps):
                    for x, y in zip(xs, ys):
                        R.c2c(x, y, ax, True, 1.0, 1)
            torch.cuda.synchronize()
        finally:
            R.kernel_timing(False)
        rep = R.kernel_timing_report()
    finally:
        R.set_precision(prev)
    label, (n, total_ms, nbytes) = max(rep.items(), key=lambda kv: kv[1][1])
    avg_ms = total_ms / n
    per_launch = nbytes / n
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    kernels = {k: {"launches_per_step": v[0] / steps, "ms_per_step": round(v[1] / steps, 4),
                   "avg_launch_ms": round(v[1] / v[0], 5),
                   "GB/s": round(v[2] / v[0] / (v[1] / v[0] * 1e-3) / 1e9, 1)}
               for k, v in rep.items()}
    key = label
    info = {"bound": "hbm", "kernel": label, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "bytes_per_launch": int(per_launch), "avg_launch_ms": round(avg_ms, 5),
            "traffic": TRAFFIC.get(key), "step_kernels": kernels}
    if info["traffic"]:
        info["traffic_over_algorithmic"] = round(info["traffic"] / per_launch, 3)
    return info


def cpu_threads():
    return max(1, os.cpu_count() or 1)


def cpu_baseline(seconds):
    """The reference engine on the headline step (both shapes, forward c2c
    complex64), nthreads = all host cores, repeated to about `seconds`."""
    import ref as REF
    if not REF.available():
        return None
    threads = cpu_threads()
    rs = np.random.RandomState(4)
    arrs = []
    for shape in SHAPES:
        x = (rs.standard_normal(shape) + 1j * rs.standard_normal(shape)).astype(np.complex64)
        arrs.append((x, np.empty_like(x)))
    ax = np.array([1], dtype=np.uint64)

    def step():
        for x, y in arrs:
            REF.c2c(x, y, ax, True, 1.0, threads)

    step()  # plans
    n, t0 = 0, time.perf_counter()
    while True:
        step()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 400:
            break
    return {"value": round(BYTES_STEP * n / el / 1e9, 3), "unit": "GB/s", "cores": threads,
            "nproc": os.cpu_count(), "kind": "reference",
            "sample": f"{n} full steps (forward c2c complex64 64x2^20 and 16384x4096) of the "
                      f"reference engine built from its sources with its own flags, "
                      f"nthreads={threads}, {el:.1f} s"}


def cpu_time(fn, seconds):
    fn()
    n, t0 = 0, time.perf_counter()
    while True:
        fn()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            return el / n * 1e3


def cpu_configs(seconds):
    """Reference engine times of the other configurations (host cores)."""
    import ref as REF
    import rocket_fft_amd as R
    if not REF.available():
        return None
    th = cpu_threads()
    out = {"cores": th}
    rs = np.random.RandomState(1)
    x1 = rs.standard_normal((1024, 4096)) + 1j * rs.standard_normal((1024, 4096))
    y1 = np.empty_like(x1)
    ms = cpu_time(lambda: REF.c2c(x1, y1, np.array([1], np.uint64), True, 1.0, th), seconds)
    out["c2c_c128_1024x4096_axis1"] = {"ms": round(ms, 3), "GB/s": round(2 * x1.nbytes / ms / 1e6, 2)}
    del x1, y1
    rs = np.random.RandomState(2)
    x = rs.standard_normal((256, 256, 256)).astype(np.float32)
    d = rs.uniform(-1.0, 1.0, (512, 512))
    X = np.empty((256, 256, 129), np.complex64)
    x2 = np.empty_like(x)
    yd = np.empty_like(d)
    a3, a2 = np.array([0, 1, 2], np.uint64), np.array([0, 1], np.uint64)

    def step2():
        REF.r2c(x, X, a3, True, 1.0, th)
        REF.c2r(X, x2, a3, False, 1.0 / 256 ** 3, th)
        REF.dct(d, yd, a2, 2, 1.0 / 1024, True, th)

    ms = cpu_time(step2, seconds)
    b2 = 2 * (x.nbytes + X.nbytes) + 2 * d.nbytes
    out["rfftn_irfftn_256cubed_dct2_512sq"] = {"ms": round(ms, 3), "GB/s": round(b2 / ms / 1e6, 2)}
    rs = np.random.RandomState(3)
    for n in (1000, 4099, 65537):
        for dt, name in ((np.complex64, "c64"), (np.complex128, "c128")):
            xc = (rs.standard_normal((256, n)) + 1j * rs.standard_normal((256, n))).astype(dt)
            yc = np.empty_like(xc)
            ms = cpu_time(lambda: REF.c2c(xc, yc, np.array([1], np.uint64), True, 1.0, th),
                          seconds / 4)
            out[f"{name}_n{n}_batch256"] = {"ms": round(ms, 3),
                                            "GB/s": round(2 * xc.nbytes / ms / 1e6, 2)}
    spec = R.ragged_workload(njobs=4096, seed=5)
    rs = np.random.RandomState(5)
    jobs, nbytes = [], 0
    for kind, n in spec:
        if kind == "c2c":
            xx = (rs.standard_normal(n) + 1j * rs.standard_normal(n)).astype(np.complex64)
            jobs.append((REF.c2c, xx, np.empty_like(xx)))
            nbytes += 16 * n
        else:
            xx = rs.standard_normal(n).astype(np.float32)
            jobs.append((REF.r2c, xx, np.empty(n // 2 + 1, np.complex64)))
            nbytes += 4 * n + 8 * (n // 2 + 1)
    a0 = np.array([0], np.uint64)

    def ragged():
        for f, xx, yy in jobs:
            f(xx, yy, a0, True, 1.0, 1)

    ms = cpu_time(ragged, seconds)
    out["ragged_4096_jobs_one_call_each_nthreads1"] = {"ms": round(ms, 3),
                                                        "GB/s": round(nbytes / ms / 1e6, 2)}
    return out


def extra_configs(torch, R, stream, reps):
    """Plain c2c throughput of the batched complex128 configuration."""
    out = {}
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    a = torch.randn((1024, 4096), dtype=torch.complex128, device="cuda", generator=g)
    b = torch.empty_like(a)
    ax = np.array([1])
    with torch.cuda.stream(stream):
        R.c2c(a, b, ax, True, 1.0, 1)
    torch.cuda.synchronize()
    ms, _ = graph_time(torch, lambda: R.c2c(a, b, ax, True, 1.0, 1), stream, reps)
    nbytes = 2 * a.numel() * a.element_size()
    out["c2c_c128_1024x4096_axis1"] = {"ms": round(ms, 4), "GB/s": round(nbytes / ms / 1e6, 1),
                                       "hbm_frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 3)}
    return out


def real3d_config(torch, R, stream, reps):
    """r2c of a float32 256^3 field over axes (0,1,2), c2r back with
    fct = 1/256^3, and DCT-II ortho of a float64 512^2 image (one step)."""
    rs = np.random.RandomState(2)
    x_h = rs.standard_normal((256, 256, 256)).astype(np.float32)
    d_h = rs.uniform(-1.0, 1.0, (512, 512))
    x = torch.from_numpy(x_h).cuda()
    d = torch.from_numpy(d_h).cuda()
    X = torch.empty((256, 256, 129), dtype=torch.complex64, device="cuda")
    x2 = torch.empty_like(x)
    y = torch.empty_like(d)
    a3, a2 = np.array([0, 1, 2]), np.array([0, 1])

    def step():
        R.r2c(x, X, a3, True, 1.0, 1)
        R.c2r(X, x2, a3, False, 1.0 / 256 ** 3, 1)
        R.dct(d, y, a2, 2, 1.0 / 1024, True, 1)

    with torch.cuda.stream(stream):
        step()
    torch.cuda.synchronize()
    ms, _ = graph_time(torch, step, stream, reps)
    nbytes = 2 * (x_h.nbytes + 256 * 256 * 129 * 8) + 2 * d_h.nbytes
    rt = rel_l2(x2.cpu().numpy(), x_h)
    spec = np.fft.rfftn(x_h[:, :, :].astype(np.float64), axes=(0, 1, 2))
    slab = rel_l2(X.cpu().numpy()[3].astype(np.complex128), spec[3])
    return {"ms": round(ms, 4), "GB/s": round(nbytes / ms / 1e6, 1),
            "hbm_frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 3),
            "roundtrip_rel_l2": rt, "r2c_slab_vs_numpy_fp64_rel_l2": slab}


def lengths_config(torch, R, stream, reps):
    """Mixed-radix and prime lengths, batch 256, complex64 and complex128:
    forward c2c timed (ten calls per captured graph, replayed), forward error
    against numpy's FP64 transform on the host, and the forward -> inverse
    (fct = 1/N) round trip."""
    out = {}
    rs = np.random.RandomState(3)
    ax = np.array([1])
    for n in (1000, 1536, 2187, 4913, 4099, 65537):
        x64 = rs.standard_normal((256, n)) + 1j * rs.standard_normal((256, n))
        want = np.fft.fft(x64, axis=1)
        for dt, npdt, name in ((torch.complex64, np.complex64, "c64"),
                               (torch.complex128, np.complex128, "c128")):
            xh = x64.astype(npdt)
            a = torch.from_numpy(xh).cuda()
            b = torch.empty_like(a)
            c = torch.empty_like(a)
            with torch.cuda.stream(stream):
                R.c2c(a, b, ax, True, 1.0, 1)
                R.c2c(b, c, ax, False, 1.0 / n, 1)
            torch.cuda.synchronize()
            ref = np.fft.fft(xh.astype(np.complex128), axis=1) if npdt == np.complex64 else want
            fwd_err = float(np.linalg.norm(b.cpu().numpy() - ref) / np.linalg.norm(ref))
            rt_err = float(np.linalg.norm(c.cpu().numpy() - xh) / np.linalg.norm(xh))
            ms, _ = graph_time(torch, lambda: R.c2c(a, b, ax, True, 1.0, 1), stream,
                               max(1, reps // 10), inner=10)
            nbytes = 2 * a.numel() * a.element_size()
            out[f"{name}_n{n}"] = {"ms": round(ms, 4), "transforms_per_s": round(256 / ms * 1e3),
                                   "GB/s": round(nbytes / ms / 1e6, 1),
                                   "hbm_frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 3),
                                   "rel_l2_fwd_vs_fp64": fwd_err, "rel_l2_roundtrip": rt_err}
            del a, b, c
    return out


def ragged_config(torch, R, stream):
    """4096 transforms of log-uniform lengths 64..65536 (70% good sizes, 30%
    raw incl. primes; 60% c2c complex64, 40% r2c float32) in one persistent
    launch, against one drop-in call per transform."""
    spec = R.ragged_workload(njobs=4096, seed=5)
    jobs, nbytes = [], 0
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    for kind, n in spec:
        if kind == "c2c":
            x = torch.randn(n, dtype=torch.complex64, device="cuda", generator=g)
            y = torch.empty(n, dtype=torch.complex64, device="cuda")
            nbytes += 16 * n
        else:
            x = torch.randn(n, dtype=torch.float32, device="cuda", generator=g)
            y = torch.empty(n // 2 + 1, dtype=torch.complex64, device="cuda")
            nbytes += 4 * n + 8 * (n // 2 + 1)
        jobs.append((kind, True, n, x, y, 1.0))
    with torch.cuda.stream(stream):
        b = R.RaggedBatch(jobs, double=False, autotune=True)
        b.execute_serial()
        b.execute_serial()
    torch.cuda.synchronize()
    ref = [j[4].clone() for j in jobs]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        e0.record(stream)
        for _ in range(3):
            b.execute_serial()
        e1.record(stream)
    torch.cuda.synchronize()
    serial_ms = e0.elapsed_time(e1) / 3
    with torch.cuda.stream(stream):
        b.execute()
    torch.cuda.synchronize()
    ms, _ = graph_time(torch, b.execute, stream, 10)
    worst = max(float(torch.linalg.norm(r - j[4]) / torch.linalg.norm(r)) for r, j in zip(ref, jobs))
    info = b.info()
    b.close()
    # second comparison point: the jobs grouped by (kind, length), each group
    # stacked into one array and transformed by one batched call
    groups = {}
    for kind, fwd, n, x, y, fct in jobs:
        groups.setdefault((kind, n), []).append(x)
    batched = []
    for (kind, n), xs in groups.items():
        X = torch.stack(xs)
        Y = torch.empty((len(xs), n if kind == "c2c" else n // 2 + 1), dtype=torch.complex64,
                        device="cuda")
        batched.append((kind, X, Y))
    ax = np.array([1])

    def grouped():
        for kind, X, Y in batched:
            (R.c2c if kind == "c2c" else R.r2c)(X, Y, ax, True, 1.0, 1)

    with torch.cuda.stream(stream):
        grouped()
    torch.cuda.synchronize()
    grouped_ms, _ = graph_time(torch, grouped, stream, 3, batches=3)
    return {"jobs": len(jobs), "bytes": nbytes, "persistent_ms": round(ms, 4),
            "GB/s": round(nbytes / ms / 1e6, 1), "hbm_frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 3),
            "per_transform_ms": round(serial_ms, 3), "speedup_vs_per_transform": round(serial_ms / ms, 2),
            "grouped_batched_ms": round(grouped_ms, 3), "distinct_kind_length_groups": len(batched),
            "speedup_vs_grouped_batched": round(grouped_ms / ms, 2),
            "max_rel_l2_persistent_vs_per_transform": worst, "autotune": info}


def probe_ragged(torch, R):
    """The 4096-job ragged workload: planned (no autotune) and executed three
    times, no graph (trace and counter runs)."""
    spec = R.ragged_workload(njobs=4096, seed=5)
    jobs = []
    for kind, n in spec:
        if kind == "c2c":
            x = torch.randn(n, dtype=torch.complex64, device="cuda")
            y = torch.empty(n, dtype=torch.complex64, device="cuda")
        else:
            x = torch.randn(n, dtype=torch.float32, device="cuda")
            y = torch.empty(n // 2 + 1, dtype=torch.complex64, device="cuda")
        jobs.append((kind, True, n, x, y, 1.0))
    b = R.RaggedBatch(jobs, double=False, autotune=False)
    torch.cuda.synchronize()
    for _ in range(3):
        b.execute()
    torch.cuda.synchronize()
    b.close()
    print("probe done", flush=True)


def probe(torch, R, shape, mode):
    """Calibration copy (reads and writes the array once each, three times),
    then three transforms of `shape` in `mode`, no graph (counter runs)."""
    if shape == "ragged":
        return probe_ragged(torch, R)
    shape = tuple(int(v) for v in shape.split("x"))
    x = torch.randn(shape, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    for _ in range(3):
        with_mode(R, mode, lambda: R.c2c(x, y, np.array([1]), True, 1.0, 1))
    torch.cuda.synchronize()
    print("probe done", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--mode", choices=MODES, default=None,
                    help="arithmetic of the step (default: the fastest form within 1e-6)")
    ap.add_argument("--probe", default=None, metavar="SHAPE:MODE",
                    help="only a device copy and three transforms of one shape "
                         "('64x1048576:fp32', '16384x4096:mfma_fp16', 'ragged:fp32', ...), "
                         "for counter runs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: every transform of this workload fits one "
              "MI355X and the build runs no multi-GPU path (run with --gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import rocket_fft_amd as R

    if args.probe:
        shp, mode = args.probe.split(":")
        probe(torch, R, shp, mode)
        return
    stream = torch.cuda.Stream()
    data = {shape: precision_sets(torch, shape) for shape in SHAPES}
    forms = precision_forms(torch, R, stream, data, 10)
    mode = args.mode or min((m for m in MODES if forms[m]["meets_1e-6"]),
                            key=lambda m: forms[m]["step_ms"])
    xs = [data[s]["normal"] for s in SHAPES]
    ys = [torch.empty_like(x) for x in xs]
    ax = np.array([1])

    def step():
        for x, y in zip(xs, ys):
            R.c2c(x, y, ax, True, 1.0, 1)

    prev = R.set_precision(mode)
    try:
        with torch.cuda.stream(stream):
            for _ in range(max(1, args.warmup)):
                step()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            step()
        torch.cuda.synchronize()
        for _ in range(args.warmup):
            g.replay()
        torch.cuda.synchronize()
        # timed region: K replays of the captured step, synchronised both sides
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
    finally:
        R.set_precision(prev)
    value = BYTES_STEP / (ms * 1e-3) / 1e9
    parity = {f"{s[0]}x{s[1]}": forms[mode][f"{s[0]}x{s[1]}"]["rel_l2_vs_fp64"] for s in SHAPES}
    res = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (torch generator seed 4: iid N(0,1) complex; the log-uniform "
                "magnitude set for the error figures)",
        "config": {"workload": "precision configuration: forward c2c complex64 64 x 2^20 and "
                               "16384 x 4096 (axis 1), 1 GiB each",
                   "arithmetic": mode, "bytes_per_step": BYTES_STEP, "hip_graph": True},
        "hbm_frac": round(value / HBM_PEAK_GBS, 4),
        "parity": {"rel_l2_vs_fp64": parity, "bar": 1e-6},
        "precision_forms": forms,
        "roofline": roofline(torch, R, stream, mode, xs, ys, 10),
    }
    del data, xs, ys
    torch.cuda.empty_cache()
    if not args.no_extra:
        res["configs"] = extra_configs(torch, R, stream, 20)
        res["configs"]["real3d_r2c_c2r_256cubed_dct2_512sq"] = real3d_config(torch, R, stream, 20)
        res["configs"]["ragged_4096_c64_r2c"] = ragged_config(torch, R, stream)
        res["configs"]["lengths_batch256"] = lengths_config(torch, R, stream, 50)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        if not args.no_extra:
            res["cpu_configs"] = cpu_configs(2.0)
    print(json.dumps(res))


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Throughput of the transform engine on one MI355X.

Headline workload (one "step"): the real 3-D round trip and the 2-D DCT of
rocket-fft's r2c/c2r/DCT configuration --
  * r2c of a float32 256x256x256 field (iid N(0,1)) over axes (0,1,2) into
    complex64 256x256x129, then c2r back with fct = 1/256^3;
  * DCT-II with ortho=True over both axes of a float64 512x512 field
    (iid U(-1,1)), fct = 1/1024.
Inputs come from numpy RandomState(2) and are resident in HBM before timing.
value = algorithmic bytes per step (each transform reads its input once and
writes its output once: 2 x 134,742,016 + 4,194,304 B) / time per step.
The step is captured once in a HIP graph and replayed (eager per-call times
are reported beside it).

roofline: the dominant kernel of the step (the c2c pass over the strided
axes 0 and 1 of the 256x256x129 half spectrum; four launches per step, each
reading and writing 135,266,304 B) timed with HIP events on its own stream.
cpu_baseline: the reference engine compiled from its sources (oracle/_ref),
same step on the host cores, bounded to about ten seconds.
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
N3 = 256
N2D = 512
BYTES_R2C = N3 ** 3 * 4 + N3 * N3 * (N3 // 2 + 1) * 8
BYTES_DCT = 2 * N2D * N2D * 8
BYTES_STEP = 2 * BYTES_R2C + BYTES_DCT
BYTES_C2C_PASS = 2 * N3 * N3 * (N3 // 2 + 1) * 8
# HBM bytes per launch of the dominant kernel, measured with
#   rocprofv3 --pmc FETCH_SIZE -- python3 bench.py --probe   (then WRITE_SIZE)
# each counter scaled by what it reads for the device copy of known size
# (67,633,152 B each way) in the same run: FETCH_SIZE x 1.999, WRITE_SIZE
# x 1.000 on MI355X.  Mean of six launches: 67,786,089 B fetched +
# 68,978,544 B written = 1.011 x the algorithmic bytes.
TRAFFIC_C2C_PASS = 136764633


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def inputs():
    rs = np.random.RandomState(2)
    x = rs.standard_normal((N3, N3, N3)).astype(np.float32)
    d = rs.uniform(-1.0, 1.0, (N2D, N2D))
    return x, d


def graph_time(torch, fn, stream, reps, batches=5):
    """Capture fn once and replay it: `batches` timed batches of `reps`
    replays each (events on `stream`); the median over batches of the mean
    ms per replay, so that one slow batch does not move a short kernel's
    figure."""
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
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
        times.append(e0.elapsed_time(e1) / reps)
    return float(np.median(times)), g


def cpu_baseline(x, d, seconds):
    import ref as REF
    if not REF.available():
        return None
    threads = min(16, os.cpu_count() or 1)
    X = np.empty((N3, N3, N3 // 2 + 1), np.complex64)
    x2 = np.empty_like(x)
    y = np.empty_like(d)
    ax3 = np.array([0, 1, 2], dtype=np.uint64)
    ax2 = np.array([0, 1], dtype=np.uint64)

    def step():
        REF.r2c(x, X, ax3, True, 1.0, threads)
        REF.c2r(X, x2, ax3, False, 1.0 / N3 ** 3, threads)
        REF.dct(d, y, ax2, 2, 1.0 / 1024, True, threads)

    step()  # plans
    n, t0 = 0, time.perf_counter()
    while True:
        step()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 200:
            break
    return {"value": round(BYTES_STEP * n / el / 1e9, 3), "unit": "GB/s", "cores": threads,
            "kind": "reference",
            "sample": f"{n} full steps (r2c+c2r 256^3 f32, dctn-II 512^2 f64) of the reference "
                      f"engine built from its sources with its own flags, nthreads={threads}"}


def extra_configs(torch, R, stream, reps):
    """Plain complex c2c throughput of the other batched configurations."""
    out = {}
    rs = np.random.RandomState(4)
    cases = [("c2c_c128_1024x4096_axis1", (1024, 4096), torch.complex128),
             ("c2c_c64_16384x4096_axis1", (16384, 4096), torch.complex64),
             ("c2c_c64_64x1048576_axis1", (64, 1 << 20), torch.complex64)]
    for name, shape, dt in cases:
        a = torch.randn(shape, dtype=dt, device="cuda")
        b = torch.empty_like(a)
        ax = np.array([1])
        with torch.cuda.stream(stream):
            R.c2c(a, b, ax, True, 1.0, 1)
        torch.cuda.synchronize()
        ms, _ = graph_time(torch, lambda: R.c2c(a, b, ax, True, 1.0, 1), stream, reps)
        nbytes = 2 * a.numel() * a.element_size()
        out[name] = {"ms": round(ms, 4), "GB/s": round(nbytes / ms / 1e6, 1),
                     "hbm_frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 3)}
        del a, b
    return out


def lengths_config(torch, R, stream, reps):
    """Mixed-radix and prime lengths, batch 256, complex64 and complex128:
    forward c2c timed (graph replays), forward error against numpy's FP64
    transform on the host, and the forward -> inverse (fct = 1/N) round trip."""
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
            ms, _ = graph_time(torch, lambda: R.c2c(a, b, ax, True, 1.0, 1), stream, reps)
            nbytes = 2 * a.numel() * a.element_size()
            out[f"{name}_n{n}"] = {"ms": round(ms, 4), "transforms_per_s": round(256 / ms * 1e3),
                                   "GB/s": round(nbytes / ms / 1e6, 1),
                                   "hbm_frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 3),
                                   "rel_l2_fwd_vs_fp64": fwd_err, "rel_l2_roundtrip": rt_err}
            del a, b, c
    return out


def probe(torch, R):
    """Calibration copy (reads and writes BYTES_C2C_PASS/2 each) then the
    dominant kernel, three launches per strided axis, no graph."""
    X = torch.randn((N3, N3, N3 // 2 + 1), dtype=torch.complex64, device="cuda")
    Y = torch.empty_like(X)
    torch.cuda.synchronize()
    for _ in range(3):
        Y.copy_(X)
    torch.cuda.synchronize()
    for _ in range(3):
        R.c2c(X, X, np.array([1]), True, 1.0, 1)
        R.c2c(X, X, np.array([0]), True, 1.0, 1)
    torch.cuda.synchronize()
    print("probe done", flush=True)


def precision_config(torch, R, stream):
    """complex64 c2c of 16384 lines of 4096: plain FP32 butterflies against the
    matrix-core forms, GB/s and rel-L2 against an FP64 transform of the same
    input (torch.fft in complex128), on both input sets: iid N(0,1) and
    magnitudes log-uniform over 1e-3..1e3 with uniform phase."""
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    shape = (16384, 4096)
    sets = {"normal": torch.randn(shape, dtype=torch.complex64, device="cuda", generator=g)}
    mag = torch.exp(torch.empty(shape, device="cuda").uniform_(np.log(1e-3), np.log(1e3),
                                                               generator=g))
    ph = torch.empty(shape, device="cuda").uniform_(0, 2 * np.pi, generator=g)
    sets["loguniform"] = torch.polar(mag, ph).to(torch.complex64)
    del mag, ph
    out = {}
    nbytes = 2 * shape[0] * shape[1] * 8
    ax = np.array([1])
    for mode in ("fp32", "mfma_fp16", "mfma_bf16", "mfma_bf16_2split"):
        prev = R.set_precision(mode)
        try:
            entry = {}
            for name, x in sets.items():
                y = torch.empty_like(x)
                with torch.cuda.stream(stream):
                    R.c2c(x, y, ax, True, 1.0, 1)
                torch.cuda.synchronize()
                ms, _ = graph_time(torch, lambda: R.c2c(x, y, ax, True, 1.0, 1), stream, 10)
                err = 0.0
                for lo in range(0, shape[0], 2048):  # FP64 reference in slices
                    ref = torch.fft.fft(x[lo:lo + 2048].to(torch.complex128), dim=1)
                    err += float(torch.sum(torch.abs(y[lo:lo + 2048].to(torch.complex128) - ref) ** 2))
                    den = float(torch.sum(torch.abs(ref) ** 2))
                    entry.setdefault("_den_" + name, 0.0)
                    entry["_den_" + name] += den
                entry[name] = {"ms": round(ms, 4), "GB/s": round(nbytes / ms / 1e6, 1),
                               "rel_l2_vs_fp64": float(np.sqrt(err / entry.pop("_den_" + name)))}
                del y
            out[mode] = entry
        finally:
            R.set_precision(prev)
    base = out["fp32"]["normal"]["ms"]
    for mode in out:
        out[mode]["speedup_vs_fp32"] = round(base / out[mode]["normal"]["ms"], 3)
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
    return {"jobs": len(jobs), "bytes": nbytes, "persistent_ms": round(ms, 4),
            "GB/s": round(nbytes / ms / 1e6, 1), "hbm_frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 3),
            "per_transform_ms": round(serial_ms, 3), "speedup_vs_per_transform": round(serial_ms / ms, 2),
            "max_rel_l2_persistent_vs_per_transform": worst, "autotune": info}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--probe", action="store_true",
                    help="only launch the dominant kernel (3x per strided axis) after a "
                         "3x device copy of the same 135 MB array, for counter runs")
    ap.add_argument("--probe-two-pass", action="store_true",
                    help="only run the complex64 64 x 2^20 transform (axis 1) twice, for "
                         "kernel-trace and counter runs of the two-pass form")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: every transform of this workload fits one "
              "MI355X and the build runs no multi-GPU path (run with --gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import rocket_fft_amd as R

    if args.probe:
        probe(torch, R)
        return
    if args.probe_two_pass:
        a = torch.randn((64, 1 << 20), dtype=torch.complex64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(2):
            R.c2c(a, b, np.array([1]), True, 1.0, 1)
        torch.cuda.synchronize()
        print("probe done", flush=True)
        return
    x_h, d_h = inputs()
    dev = torch.device("cuda:0")
    x = torch.from_numpy(x_h).to(dev)
    d = torch.from_numpy(d_h).to(dev)
    X = torch.empty((N3, N3, N3 // 2 + 1), dtype=torch.complex64, device=dev)
    x2 = torch.empty_like(x)
    y = torch.empty_like(d)
    ax3 = np.array([0, 1, 2])
    ax2 = np.array([0, 1])
    stream = torch.cuda.Stream()

    def step():
        R.r2c(x, X, ax3, True, 1.0, 1)
        R.c2r(X, x2, ax3, False, 1.0 / N3 ** 3, 1)
        R.dct(d, y, ax2, 2, 1.0 / 1024, True, 1)

    with torch.cuda.stream(stream):
        for _ in range(max(1, args.warmup)):
            step()
    torch.cuda.synchronize()

    # eager per-call timing (host launch overhead included)
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for _ in range(args.steps):
            step()
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) / args.steps * 1e3

    # timed region: K replays of the captured step, barrier + sync both sides
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        step()
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.replay()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    value = BYTES_STEP / (ms * 1e-3) / 1e9

    # parity of this step's outputs (untimed): round trip, and the DCT against
    # the oracle
    import oracle as O
    rt = rel_l2(x2.cpu().numpy(), x_h)
    dct_ref = np.empty_like(d_h)
    O.dct(d_h, dct_ref, [0, 1], 2, 1.0 / 1024, True)
    dct_err = rel_l2(y.cpu().numpy(), dct_ref)
    # the spectrum against numpy's FP64 rfftn of the same field (one slab)
    spec = np.fft.rfftn(x_h.astype(np.float64))
    slab_err = rel_l2(X.cpu().numpy()[3].astype(np.complex128), spec[3])
    del spec

    # dominant kernel: c2c over the strided axes of the half spectrum
    def c2c_passes():
        R.c2c(X, X, np.array([1]), True, 1.0, 1)
        R.c2c(X, X, np.array([0]), True, 1.0, 1)

    with torch.cuda.stream(stream):
        c2c_passes()
    torch.cuda.synchronize()
    kms, _ = graph_time(torch, c2c_passes, stream, 20)
    kms /= 2.0
    achieved = BYTES_C2C_PASS / (kms * 1e-3) / 1e9

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
        "dtype": "f32+f64",
        "data": "synthetic (numpy RandomState(2): N(0,1) field, U(-1,1) image)",
        "config": {"workload": "rfftn/irfftn f32 256^3 axes (0,1,2) + dctn-II ortho f64 "
                               "512^2 axes (0,1)",
                   "bytes_per_step": BYTES_STEP, "hip_graph": True},
        "hbm_frac": round(value / HBM_PEAK_GBS, 4),
        "ms_per_step_eager": round(eager_ms, 4),
        "parity": {"roundtrip_rel_l2": rt, "dct_vs_oracle_rel_l2": dct_err,
                   "r2c_slab_vs_numpy_fp64_rel_l2": slab_err},
        "roofline": {"bound": "hbm",
                     "kernel": "k_fast<float, 256, 16 threads/line, 16 lines/block, interleaved> "
                               "(c2c over strided axis 0 / 1 of the 256x256x129 half spectrum)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "avg_launch_ms": round(kms, 5),
                     "bytes_per_launch": BYTES_C2C_PASS, "traffic": TRAFFIC_C2C_PASS,
                     "traffic_over_algorithmic": round(TRAFFIC_C2C_PASS / BYTES_C2C_PASS, 3)},
    }
    if not args.no_extra:
        res["configs"] = extra_configs(torch, R, stream, 20)
        res["configs"]["ragged_4096_c64_r2c"] = ragged_config(torch, R, stream)
        res["configs"]["precision_c64_16384x4096"] = precision_config(torch, R, stream)
        res["configs"]["lengths_batch256"] = lengths_config(torch, R, stream, 50)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(x_h, d_h, args.cpu_seconds)
    print(json.dumps(res))


if __name__ == "__main__":
    main()

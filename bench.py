#!/usr/bin/env python3
"""Benchmark: boost-factor frequency points/s on one MI355X, batched mixed
precision vs plain FP64.

Workload (one "step"): a 20-disk booster (eps=24, 1 mm disks, spacings
8 mm + U(-0.3, 0.3) mm from SplitMix64 seed 1000, disk radius 0.15 m) in front
of a mirror, recursive-Fourier `dancer` with the FFT propagator on a 256 x 256
grid spanning 1 m, 50 frequencies on 19.5-20.5 GHz, nmax = 2000 (2001 bounces),
all frequencies as one batched, device-resident sweep.  `value` is the mixed
precision sweep's points/s; the same sweep in FP64 is timed beside it and the
relative beta^2 error of every point against FP64 is reported.

Kernel timing: HIP events recorded on the engine's stream around every launch
of the two bounce kernels during the timed steps.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "boostfractor.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
NDISK, N, NFREQ, NMAX, DISKR = 20, 256, 50, 2000, 0.15


def workload():
    import BoostFractor as B  # noqa: F401
    from BoostFractor import workloads as W
    coords = W.grid_coords(N)
    sb = W.booster(coords, ndisk=NDISK, seed=1000)
    freqs = np.linspace(19.5e9, 20.5e9, NFREQ)
    return coords, sb, freqs


def timed_steps(plan, ctx, steps, warmup, timing):
    for _ in range(warmup):
        plan.run()
    ctx.synchronize()
    if timing:
        plan.timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    kt = plan.kernel_times() if timing else None
    if timing:
        plan.timing(False)
    return dt, kt


def cpu_baseline(coords, sb, freqs, budget_s=20.0):
    """FP64 CPU restatement (oracle, one frequency point per thread) on a bounded
    sample: a few bounces of as many points as threads, extrapolated linearly in
    bounces to the full nmax+1."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    oc = O.Coords(coords.X, coords.Y, coords.kX, coords.kY)
    os_ = O.Setup(sb.distance, sb.eps, sb.r)
    # calibrate: one bounce of one point
    t = time.perf_counter()
    O.dancer_sweep(os_, oc, 0, 0, freqs[:1], nthreads=1, diskR=DISKR)
    one = time.perf_counter() - t
    nb = max(1, min(NMAX + 1, int(budget_s / max(one, 1e-6))))
    sample_f = freqs[:threads]
    t = time.perf_counter()
    O.dancer_sweep(os_, oc, 0, nb - 1, sample_f, nthreads=threads, diskR=DISKR)
    dt = time.perf_counter() - t
    per_point = dt / len(sample_f) * threads * (NMAX + 1) / nb  # seconds per point per thread
    return {"value": threads / per_point, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": f"{len(sample_f)} frequency points x {nb} of {NMAX + 1} bounces on {threads} threads "
                      f"(FP64 C restatement of the reference dancer), extrapolated linearly in bounces"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--freq-batch", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the workload is a single-GPU sweep and this build has no multi-GPU "
              "path (frequency points would shard across GPUs with no exchange); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import BoostFractor as B
    from BoostFractor import workloads as W
    coords, sb, freqs = workload()
    ctx = B.context()
    dx = coords.X[1] - coords.X[0]
    K = N * N
    F = 2 * (2 * NDISK + 1)  # fields per bounce (dancer regions x 2 directions)

    print(f"[bench] workload: {NDISK}-disk booster, {N}x{N}, {NFREQ} freqs, nmax={NMAX}", file=sys.stderr, flush=True)
    plan32 = B.DancerPlan(0, NMAX, sb, coords, freqs, diskR=DISKR, precision="mixed", freq_batch=args.freq_batch)
    dt32, kt32 = timed_steps(plan32, ctx, args.steps, args.warmup, timing=True)
    _, p32 = plan32.fetch(fields=False)
    plan32.close()
    print(f"[bench] mixed: {dt32 / args.steps:.3f} s/step", file=sys.stderr, flush=True)

    plan64 = B.DancerPlan(0, NMAX, sb, coords, freqs, diskR=DISKR, precision="fp64", freq_batch=args.freq_batch)
    dt64, kt64 = timed_steps(plan64, ctx, args.fp64_steps, 1, timing=True)
    _, p64 = plan64.fetch(fields=False)
    plan64.close()
    print(f"[bench] fp64: {dt64 / args.fp64_steps:.3f} s/step", file=sys.stderr, flush=True)

    b32 = p32 * dx * dx / (math.pi * DISKR ** 2)
    b64 = p64 * dx * dx / (math.pi * DISKR ** 2)
    err = np.abs(b32 / b64 - 1.0)

    pts32 = NFREQ * args.steps / dt32
    pts64 = NFREQ * args.fp64_steps / dt64

    # roofline of the dominant kernel of the mixed path
    names = {"row_pass": "k_dancer_rows (row IFFT, tilt/surface, boundary update, mask, row FFT)",
             "col_pass": "k_cols (column FFT, exp(-i kz dz)/K, column IFFT)"}
    dom = max(kt32, key=lambda k: kt32[k][0])
    ms_tot, launches = kt32[dom]
    avg_ms = ms_tot / launches
    nb_per_launch = NFREQ * args.steps * (NMAX + 1) / launches  # frequency-bounces per launch
    alg_bytes = nb_per_launch * 2 * F * K * 8  # every field read once and written once, complex64
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    bw_ms = ctx.stream_copy_probe(1 << 30, reps=10)
    stream_gbs = 2 * (1 << 30) / (bw_ms * 1e-3) / 1e9

    out = {
        "metric": "boost-factor freq points/s on 1 MI355X: mixed-prec batched vs plain FP64",
        "value": round(pts32, 4),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt32 / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields/FFT, f64 phase arguments and Eout",
        "data": "synthetic (seeded SplitMix64 booster geometry, axion source from dance_intro)",
        "config": {"workload": "dancer, 20-disk booster eps=24, 256x256 grid over 1 m, 50 freqs 19.5-20.5 GHz, "
                               "nmax=2000, FFT propagator, one batched device-resident sweep",
                   "disks": NDISK, "grid": N, "frequencies": NFREQ, "nmax": NMAX},
        "fp64": {"value": round(pts64, 4), "unit": "points/s", "ms_per_step": round(dt64 / args.fp64_steps * 1e3, 3),
                 "steps": args.fp64_steps},
        "mixed_over_fp64": round(pts32 / pts64, 3),
        "beta2_rel_err_vs_fp64": {"max": float(np.max(err)), "median": float(np.median(err)),
                                  "p90": float(np.percentile(err, 90))},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": names[dom], "avg_launch_ms": round(avg_ms, 4),
                     "alg_bytes_per_launch": int(alg_bytes),
                     "measured_stream_copy_GBs": round(stream_gbs, 1)},
        "kernel_ms": {"mixed": {k: [round(v[0], 3), v[1]] for k, v in kt32.items()},
                      "fp64": {k: [round(v[0], 3), v[1]] for k, v in kt64.items()}},
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle, bounded sample)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(coords, sb, freqs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

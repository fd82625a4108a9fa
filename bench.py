#!/usr/bin/env python3
"""Benchmark: boost-factor frequency points/s on one MI355X, batched mixed
precision vs plain FP64.

Default workload (`--workload sweep`): the disk-position sweep, 4096 seeded
boosters (disk counts uniform in 10..80, spacings 8 mm + U(-0.5, 0.5) mm,
eps 24, 1 mm disks, radius 0.15 m, flat) in front of a mirror, `transformer`
mode matching with 9, 15 or 25 Bessel modes per geometry on a 256 x 256 grid
over 1 m, 100 frequencies on 19-21 GHz: 409,600 points.  The K timed steps
split the geometries into K disjoint contiguous slices, so K steps cover the
whole sweep once; `value` is the whole sweep's points over the timed time.
The sweep has no reduced-precision path (its products run on FP64 MFMA), so
`mixed_over_fp64` is 1.0.  The dancer workloads (`--workload dancer`,
`dancer512`) time a complex64-storage path against FP64 and report the ratio
and the beta^2 error of every point.

Kernel timing: the timed steps run without instrumentation; one more step
then records HIP events on the engine's stream around every launch of the
kernel classes, which gives the kernels' mean launch times.

Roofline: the dominant kernel's executed work per launch over its mean launch
time (sweep and transformer: FP64 MFMA flops of the grouped-Parseval
product; dancer: the bounce's algorithmic bytes, every field read once and
written once per bounce and frequency, 2 F K B_c with F = 82 fields, over
the two bounce kernels' combined time per frequency-bounce).  `traffic` is
HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE, each scaled by
the same counter's reading for a copy kernel of known size.

  python bench.py --summarize-pmc FETCH_DB WRITE_DB
  python bench.py --summarize-trace TRACE_DB
print the per-kernel counter summary and the kernel-trace statistics (the
profiles/ file format) from rocprofv3's rocpd databases.
"""
import argparse
import copy
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "boostfractor.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFS = 78.6
NDISK, N, NFREQ, NMAX, DISKR = 20, 256, 50, 2000, 0.15
# mode-matching workload: the same booster with per-disk tilts and roughness,
# 45 Laguerre-Gauss modes on 512^2, 200 frequencies
N3, NF3 = 512, 200
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_cols_reg_dancer_mixed_256x256_20disk.md")
TRAFFIC_PROFILE_TR = os.path.join(ROOT, "profiles", "transformer_512x512_45modes.md")
TRAFFIC_PROFILE_SW = os.path.join(ROOT, "profiles", "transformer_sweep_256x256_bessel.md")


def pmc_per_kernel(db_path, counter):
    """mean counter value per kernel name from a rocprofv3 (rocpd) database"""
    import sqlite3
    db = sqlite3.connect(db_path)
    acc = {}
    for name, val in db.execute("select kernel_name, value from counters_collection where counter_name = ?",
                                (counter,)):
        a = acc.setdefault(name, [0.0, 0])
        a[0] += val
        a[1] += 1
    return {k: (v[0] / v[1], v[1]) for k, v in acc.items()}


# kernel families of the mode-matching classes: a class's traffic is the mean over every dispatch of the
# kernels whose names start with the family prefix (the chain class runs k_chain_mfma and k_chain_tile)
PMC_FAMILIES = ("k_chain", "k_parseval_gemm", "k_series_eval_flat", "k_series_inverse", "k_batched_inverse_wave")


def summarize_pmc(fetch_db, write_db):
    f = pmc_per_kernel(fetch_db, "FETCH_SIZE")
    w = pmc_per_kernel(write_db, "WRITE_SIZE")
    lines = ["| kernel | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch |", "|---|---|---|---|"]
    for k in sorted(f, key=lambda k: -f[k][0]):
        lines.append(f"| `{k}` | {f[k][1]} | {f[k][0]:.1f} | {w.get(k, (float('nan'), 0))[0]:.1f} |")
    lines.append("")
    for k in f:
        if "k_stream_copy" in k:  # the 1 GiB copy of stream_copy_probe: the byte counters' calibration
            lines.append(f"calibration: copy_bytes={1 << 30} fetch_kib={f[k][0]:.1f} write_kib={w.get(k, (0.0, 0))[0]:.1f}")
    for fam in PMC_FAMILIES:
        names = [k for k in f if k.split("(")[0].split("<")[0].split("::")[-1].startswith(fam)]
        nd = sum(f[k][1] for k in names)
        if not nd:
            continue
        fk = sum(f[k][0] * f[k][1] for k in names) / nd
        wk = sum(w.get(k, (0.0, 0))[0] * f[k][1] for k in names) / nd
        lines.append(f"traffic: kernel={fam} freq_batch=1 fetch_kib={fk:.1f} write_kib={wk:.1f}")
    return "\n".join(lines)


def summarize_trace(db_path, top=12):
    import sqlite3
    db = sqlite3.connect(db_path)
    lines = ["| kernel | calls | total us | mean us | % |", "|---|---|---|---|---|"]
    for name, calls, total, mean, pct in db.execute(
            "select name, total_calls, total_duration, average, percentage from top_kernels limit ?", (top,)):
        lines.append(f"| `{name}` | {calls} | {total:.1f} | {mean:.2f} | {pct:.1f} |")
    return "\n".join(lines)


def traffic_from_profile(kernel_key, fb, path=TRAFFIC_PROFILE):
    """HBM bytes per launch of the dominant kernel from the committed counter
    summary, calibrated by the copy kernel's readings and scaled to the
    frequency batch used here (a launch moves bytes in proportion to it)."""
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"calibration: copy_bytes=(\d+) fetch_kib=([\d.]+) write_kib=([\d.]+)", txt)
    k = re.search(r"traffic: kernel=%s freq_batch=(\d+) fetch_kib=([\d.]+) write_kib=([\d.]+)" % kernel_key, txt)
    if not (m and k):
        return None
    copy_kib = int(m.group(1)) / 1024.0
    fetch = float(k.group(2)) * copy_kib / float(m.group(2))
    write = float(k.group(3)) * copy_kib / float(m.group(3))
    return (fetch + write) * 1024.0 * fb / int(k.group(1))


def workload():
    import BoostFractor as B  # noqa: F401
    from BoostFractor import workloads as W
    coords = W.grid_coords(N)
    sb = W.booster(coords, ndisk=NDISK, seed=1000)
    freqs = np.linspace(19.5e9, 20.5e9, NFREQ)
    return coords, sb, freqs


def timed_steps(plan, ctx, steps, warmup, timing):
    """wall time of `steps` uninstrumented runs; with `timing`, per-kernel HIP
    event times of one extra instrumented run"""
    for _ in range(warmup):
        plan.run()
    ctx.synchronize()
    print(f"[bench]   {warmup} warmup run(s) done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    print(f"[bench]   {steps} timed run(s) done", file=sys.stderr, flush=True)
    kt = None
    if timing:
        plan.timing(True)
        plan.run()
        ctx.synchronize()
        kt = plan.kernel_times()
        plan.timing(False)
    return dt, kt


def cpu_baseline(coords, sb, freqs, budget_s=20.0):
    """FP64 CPU restatement (oracle, one frequency point per thread) on a bounded
    sample: a few bounces of as many points as threads, extrapolated linearly in
    bounces to the full nmax+1."""
    from oracle import oracle as O
    model, host_cpus, threads = host_cpu_info()
    oc = O.Coords(coords.X, coords.Y, coords.kX, coords.kY)
    os_ = O.Setup(sb.distance, sb.eps, sb.r)
    # calibrate: one bounce of one point
    t = time.perf_counter()
    O.dancer_sweep(os_, oc, 0, 0, freqs[:1], nthreads=1, diskR=DISKR)
    one = time.perf_counter() - t
    nb = max(1, min(NMAX + 1, int(budget_s / max(one, 1e-6))))
    sample_f = np.resize(freqs, threads)  # one point per thread
    t = time.perf_counter()
    O.dancer_sweep(os_, oc, 0, nb - 1, sample_f, nthreads=threads, diskR=DISKR)
    dt = time.perf_counter() - t
    per_point = dt / len(sample_f) * threads * (NMAX + 1) / nb  # seconds per point per thread
    return {"value": threads / per_point, "unit": "points/s", "cores": threads, "kind": "port",
            "cpu_model": model, "host_logical_cpus": host_cpus,
            "sample": f"{len(sample_f)} frequency points x {nb} of {NMAX + 1} bounces on {threads} threads, the "
                      f"process's CPU share (FP64 C restatement of the reference dancer), extrapolated linearly in "
                      f"bounces"}


def cheerleader_cpu_baseline(coords, sb, freqs, nmax, budget_s=15.0):
    """FP64 CPU restatement of the cheerleader (oracle, one thread) on a bounded
    sample: one frequency point with nmax = 0 (set-up, prologue and one
    iteration) and with a few iterations more, extrapolated linearly in
    iterations to nmax + 1."""
    from oracle import oracle as O
    oc = O.Coords(coords.X, coords.Y, coords.kX, coords.kY)
    os_ = O.Setup(sb.distance, sb.eps, sb.r)
    t = time.perf_counter()
    O.cheerleader(os_, oc, 0, 0, freqs[0], diskR=DISKR)
    t0 = time.perf_counter() - t
    n1 = max(1, min(nmax, int(budget_s / max(t0, 1e-6))))
    t = time.perf_counter()
    O.cheerleader(os_, oc, 0, n1, freqs[0], diskR=DISKR)
    t1 = time.perf_counter() - t
    per_iter = max(t1 - t0, 0.0) / n1
    full = t0 + per_iter * nmax
    return {"value": 1.0 / full, "unit": "points/s", "cores": 1, "kind": "port",
            "sample": f"1 frequency point, {n1 + 1} of {nmax + 1} iterations (and 1 for the prologue time) on one "
                      f"thread (FP64 C restatement of the reference cheerleader), extrapolated linearly in "
                      f"iterations"}


def transformer_cpu_baseline(coords, sb, pat, kt):
    """FP64 CPU restatement on a bounded sample: the propagation matrix of one
    air region and one disk region (single thread), scaled to the 21 air and
    20 disk regions of one frequency point (the chain is negligible)."""
    from oracle import oracle as O
    oc = O.Coords(coords.X, coords.Y, coords.kX, coords.kY)
    P = pat.reshape(-1, coords.ny, coords.nx)
    K = kt.reshape(-1)
    lam = 299792458.0 / 20e9
    times = {}
    for name, r in (("air", 1), ("disk", 2)):
        t = time.perf_counter()
        O.propagation_matrix(sb.distance[r], DISKR, sb.eps[r], sb.relative_tilt_x[r], sb.relative_tilt_y[r],
                             sb.relative_surfaces[r], lam, oc, P, K)
        times[name] = time.perf_counter() - t
    n_air = sum(1 for r in range(1, len(sb.eps)) if sb.eps[r].real == 1.0)
    n_disk = len(sb.eps) - 1 - n_air
    per_point = n_air * times["air"] + n_disk * times["disk"]
    return {"value": 1.0 / per_point, "unit": "points/s", "cores": 1, "kind": "port",
            "sample": f"propagation matrices of one air ({times['air']:.1f} s) and one disk region "
                      f"({times['disk']:.1f} s) at 512^2 with 45 modes, one thread, scaled to {n_air} air + "
                      f"{n_disk} disk regions per point (FP64 C restatement of the reference transformer)"}


MM_CLASSES = ("wgram", "pgemm", "pseries", "sinverse", "inverse", "chain")
MM_KERNELS = {
    "pgemm": ("k_parseval_gemm", "k_parseval_gemm (v_mfma_f64_16x16x4f64, weights exp(-i kz dz)/K formed in "
              "registers or read from the spacing series' table): flat air regions, P = sum over k-bin groups of "
              "W_g H_g; over the packed upper triangles when the group Grams are Hermitian"),
    "wgram": ("k_wgram", "k_wgram (weighted Gram products A^H diag(w) B on v_mfma_f64_16x16x4f64)"),
    "inverse": ("k_batched_inverse_wave", "k_batched_inverse_wave (FP64 Gauss-Jordan with partial pivoting in place in "
                "LDS, one wave per matrix; FP64 VALU, whose MI355X peak equals the FP64 MFMA one)"),
    "sinverse": ("k_series_inverse", "k_series_inverse_mfma / k_series_inverse (coefficients of the gaps' inverse "
                 "series, E_0 = inv(D_0), E_p = -E_0 sum_j D_j E_(p-j): on v_mfma_f64_16x16x4f64 for 8 < n <= 32, "
                 "else FP64 VALU; blocks of columns per key)"),
    "chain": ("k_chain", "k_chain_mfma (n = 25: v_mfma_f64_16x16x4f64 on operands padded to 32, three MFMAs per "
              "complex product), k_chain_mfma2 (n = 15: padded to 16, the two products in separate waves) and "
              "k_chain_tile (n = 9: register-tiled FP64 VALU): bottom block row of the transfer-matrix chain, "
              "16 n^3 algorithmic flop per region; flat disks formed in the kernel from G, inv(G) and their "
              "phases"),
    "pseries": ("k_series_eval", "k_series_coef / k_series_eval (spacing-series weights and Horner sums of the "
                "gaps' matrices and inverses)"),
}


def mm_roofline(kt_, profile, alg_flops_per_s):
    """roofline of the dominant mode-matching kernel class (by HIP-event time):
    executed flops per launch over its mean launch time, against the FP64 peak;
    the survey's algorithmic figure (8 n^2 K per region and point, which the work
    reductions undercut) is reported beside it, not as `achieved`"""
    name = max(MM_CLASSES, key=lambda k: kt_[k][0])
    ms, launches, flops = kt_[name]
    achieved = flops / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    key, desc = MM_KERNELS[name]
    t = traffic_from_profile(key, 1, profile)
    return {"bound": "mfma", "achieved": round(achieved, 2), "peak": FP64_MFMA_PEAK_TFS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_MFMA_PEAK_TFS, 4), "traffic": None if t is None else int(t),
            "kernel": desc, "kernel_class": name, "avg_launch_ms": round(ms / max(launches, 1), 3),
            "executed_flops_per_launch": flops / max(launches, 1),
            "survey_algorithmic_TFs": round(alg_flops_per_s / 1e12, 1)}


def transformer_main(args):
    import BoostFractor as B
    from BoostFractor import workloads as W
    coords = W.grid_coords(N3)
    sb0 = W.rough_booster(coords, ndisk=NDISK, seed=1002)
    pat, kt = W.lg_modes(coords)
    modes = B.SeedModes(coords, ThreeDim=True, Mmax=5, Lmax=4, pattern_input=pat, kT_arr=kt)
    freqs = np.linspace(19.5e9, 20.5e9, NF3)
    n = modes.n
    ctx = B.context()
    print(f"[bench] transformer: {NDISK}-disk rough booster, {N3}x{N3}, {n} LG modes, {NF3} freqs",
          file=sys.stderr, flush=True)
    if not args.no_autotune:
        ctx.set_option("autotune", 1)  # the product's column blocks are timed in the (untimed) warmup
    # inputs resident: the plan uploads setup and modes once; each step recomputes the sweep
    plan = B.TransformerPlan(sb0, coords, modes, freqs, diskR=DISKR)
    for _ in range(args.warmup):
        plan.run()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.run()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ctx.timing(True)  # one more, instrumented step for the kernel means
    plan.run()
    ctx.synchronize()
    kt_ = {k: ctx.kernel_time(k) for k in MM_CLASSES}
    ctx.timing(False)
    boost = plan.fetch()
    plan.close()
    # the same sweep through the host-buffer API (setup and modes copied in, boost out)
    t1 = time.perf_counter()
    B.transformer(copy.deepcopy(sb0), coords, modes, f=freqs, diskR=DISKR)
    ctx.synchronize()
    host_dt = time.perf_counter() - t1
    bw_ms = ctx.stream_copy_probe(1 << 30, reps=3)  # also the byte-counter calibration of profiler runs
    alg = (len(sb0.distance) - 1) * 8.0 * n * n * N3 * N3  # survey: sum over regions 2..Nreg of 8 n^2 K
    roof = mm_roofline(kt_, TRAFFIC_PROFILE_TR, alg * NF3 * args.steps / dt)
    roof["kernel_ms_from"] = "one instrumented step after the timed ones"
    roof["measured_stream_copy_GBs"] = round(2 * (1 << 30) / (bw_ms * 1e-3) / 1e9, 1)
    beta2 = np.sum(np.abs(boost) ** 2, axis=-1)
    out = {
        "metric": "transformer boost-factor freq points/s on 1 MI355X (FP64)",
        "value": round(NF3 * args.steps / dt, 3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 booster, tilts N(0,50 urad), 10 um roughness; LG modes w0=0.1 m)",
        "note": "inputs resident (TransformerPlan); the host-buffer API (transformer) on the same sweep: "
                f"{NF3 / host_dt:.1f} points/s",
        "config": {"workload": f"transformer, 20-disk booster with per-disk tilt and roughness, {N3}x{N3} grid over "
                               f"1 m, {n} Laguerre-Gauss modes (p<=4, |l|<=4), {NF3} freqs 19.5-20.5 GHz",
                   "grid": N3, "modes": n, "frequencies": NF3},
        "roofline": roof,
        "kernel_ms": {k: [round(v[0], 2), v[1]] for k, v in kt_.items()},
        # useful flops (8 n^2 per complex multiply-add of the product) over each class's HIP-event time
        "kernel_tflops": {k: round(v[2] / (v[0] * 1e-3) / 1e12, 2) if v[0] > 0 else 0.0 for k, v in kt_.items()},
        "work_reductions": "air regions by k-bin groups (8 n^2 G per region, G = distinct (kx^2, ky^2) pairs); "
                           "tilted/rough disks by a frequency series of moment Grams",
        "tuning": dict(zip(("chosen", "ms"), ctx.tuning_report())),
        "beta2": {"min": float(beta2.min()), "max": float(beta2.max())},
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle, bounded sample)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = transformer_cpu_baseline(coords, sb0, pat, kt)
    print(json.dumps(out), flush=True)


# disk-position sweep: seeded boosters with 10..80 disks (spacings 8 mm +- 0.5 mm),
# 9/15/25 Bessel modes, 256^2, 100 frequencies on 19-21 GHz per geometry
N5, NF5, NCONF5 = 256, 100, 4096


def host_cpu_info():
    """CPU model, logical CPUs of the host, and the CPUs this process may run on
    (its share of the box; a one-GPU box gives a process 16 of its cores)"""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        share = min(share, int(omp))
    return model, os.cpu_count() or 1, max(1, share)


def sweep_cpu_baseline(coords, geoms, set_of, sets, freqs, budget_s=20.0):
    """FP64 CPU restatement of the reference transformer (the oracle) on a bounded
    sample: per mode set, the propagation matrix of one air and one disk region
    (the per-point cost is the sum over the geometry's regions; the chain is
    negligible), every matrix computed `threads` times concurrently on the
    process's CPU share; the whole sweep's CPU time then follows from the region
    counts of all its geometries."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    model, host_cpus, threads = host_cpu_info()
    oc = O.Coords(coords.X, coords.Y, coords.kX, coords.kY)
    lam = 299792458.0 / float(freqs[len(freqs) // 2])
    per = {}
    for k, md in enumerate(sets):
        pat, kt = md.patterns_flat()[:, 0], md.mode_kt.reshape(-1)
        for kind, eps, dz in (("air", complex(1.0, -0.0), 8e-3), ("disk", complex(24.0, -0.0), 1e-3)):
            def one(_):
                return O.propagation_matrix(dz, DISKR, eps, 0.0, 0.0, None, lam, oc, pat, kt)
            t = time.perf_counter()
            with ThreadPoolExecutor(max_workers=threads) as ex:
                list(ex.map(one, range(threads)))
            per[(k, kind)] = (time.perf_counter() - t) / threads  # seconds per matrix with every thread busy
    total = 0.0
    for g, k in zip(geoms, set_of):
        nd = (len(g.distance) - 2) // 2
        total += len(freqs) * ((nd + 1) * per[(int(k), "air")] + nd * per[(int(k), "disk")])
    npts = len(geoms) * len(freqs)
    return {"value": npts / total, "unit": "points/s", "cores": threads, "kind": "port",
            "cpu_model": model, "host_logical_cpus": host_cpus,
            "sample": f"propagation matrices of one air and one disk region per mode set (n = 9, 15, 25) at 256^2, "
                      f"each computed by {threads} threads at once on the process's {threads}-CPU share, scaled by "
                      f"the region counts of all {len(geoms)} geometries x {len(freqs)} frequencies (FP64 C "
                      f"restatement of the reference transformer; the chain is negligible)"}


def _eps_copies(geoms):
    """setups whose eps the transformer may mutate (NaN -> 1e30), sharing everything else"""
    out = []
    for g in geoms:
        c = copy.copy(g)
        c.eps = g.eps.copy()
        out.append(c)
    return out


SWEEP_WORKLOAD = ("disk-position sweep (transformer): {nconf} seeded boosters with 10-80 disks (spacings 8 mm "
                  "+- U(0.5 mm), eps 24, 1 mm, flat), 9/15/25 Bessel modes, 256x256 grid over 1 m, 100 freqs "
                  "19-21 GHz")


def sweep_main(args):
    """The disk-position sweep: 4096 seeded geometries x 100 frequencies.  The
    timed steps split the geometries into `steps` disjoint contiguous slices, so
    `steps` steps cover the whole sweep once; warmup steps run the first slices.
    Inputs resident: the grid, mode sets, their axion sources and the
    frequencies are uploaded once (TransformerBatchPlan); each step hands its
    slice's geometries over and computes every boost on the device (pattern
    transforms, group Grams, all propagation matrices, inverses, chains)."""
    import BoostFractor as B
    from BoostFractor import workloads as W
    coords = W.grid_coords(N5)
    nconf = args.nconf
    geoms, set_of = W.sweep_geometries(coords, nconf)
    sets = [B.SeedModes(coords, ThreeDim=True, Mmax=M, Lmax=L, diskR=DISKR) for M, L in W.SWEEP_MODESETS]
    freqs = np.linspace(19e9, 21e9, NF5)
    ctx = B.context()
    ndisks = [(len(g.distance) - 2) // 2 for g in geoms]
    K = max(1, args.steps)
    slices = [sl for sl in np.array_split(np.arange(nconf), K)]
    print(f"[bench] sweep: {nconf} geometries (disks {min(ndisks)}..{max(ndisks)}), {N5}x{N5}, {NF5} freqs, "
          f"{K} steps of ~{nconf // K} geometries", file=sys.stderr, flush=True)
    emits = [B.axion_induced_modes(coords, md, diskR=DISKR) for md in sets]
    if not args.no_autotune:
        ctx.set_option("autotune", 1)  # shapes timed during the (untimed) warmup, reported as `tuning`
    if args.batch_gib > 0:
        ctx.set_option("batch_bytes", int(args.batch_gib * (1 << 30)))
    plan = B.TransformerBatchPlan(coords, sets, freqs, diskR=DISKR, emits=emits)
    work = [(_eps_copies([geoms[i] for i in sl]), set_of[sl]) for sl in slices]

    def fresh(k):
        # the candidate geometries an optimiser hands over for step k: new objects every step (the run
        # replaces NaN permittivities in place, as the reference does), made before the timer starts
        gs, so = work[k % K]
        return _eps_copies(gs), so

    for w in range(args.warmup):
        gs, so = fresh(w)
        plan.run(gs, mode_set_of=so)
    ctx.synchronize()
    print(f"[bench]   {args.warmup} warmup step(s) done", file=sys.stderr, flush=True)
    inputs = [fresh(k) for k in range(K)]
    t0 = time.perf_counter()
    boosts = []
    for k in range(K):
        gs, so = inputs[k]
        plan.run(gs, mode_set_of=so)
        boosts += plan.fetch() if args.keep_boosts else []
    ctx.synchronize()
    dt = time.perf_counter() - t0
    print(f"[bench]   {K} timed steps done: {dt:.3f} s", file=sys.stderr, flush=True)
    pts = nconf * NF5 / dt
    ctx.timing(True)  # one more, instrumented step (slice 0) for the kernel means
    gs, so = fresh(0)
    plan.run(gs, mode_set_of=so)
    ctx.synchronize()
    kt_ = {k: ctx.kernel_time(k) for k in MM_CLASSES}
    host_prep, host_wait = ctx.kernel_time("host_prep"), ctx.kernel_time("host_wait")
    host_geo = ctx.kernel_time("host_geometries")
    ctx.timing(False)
    b0 = plan.fetch()
    plan.close()
    nmodes = [sets[s_].n for s_ in set_of]
    alg = sum((len(g.distance) - 1) * 8.0 * m * m * N5 * N5 for g, m in zip(geoms, nmodes)) * NF5
    roof = mm_roofline(kt_, TRAFFIC_PROFILE_SW, alg / dt)
    roof["kernel_ms_from"] = "one instrumented step (slice 0) after the timed ones"
    bw_ms = ctx.stream_copy_probe(1 << 30, reps=3)
    roof["measured_stream_copy_GBs"] = round(2 * (1 << 30) / (bw_ms * 1e-3) / 1e9, 1)
    b2 = np.concatenate([np.sum(np.abs(b) ** 2, axis=-1) for b in b0])
    out = {
        "metric": "boost-factor freq points/s on 1 MI355X: mixed-prec batched vs plain FP64",
        "value": round(pts, 2),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(dt / K * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 geometries; Bessel modes from SeedModes)",
        "config": {"workload": SWEEP_WORKLOAD.format(nconf=nconf), "geometries": nconf, "grid": N5,
                   "frequencies": NF5, "points": nconf * NF5,
                   "steps_cover": f"{K} disjoint slices of the {nconf} geometries, one per timed step"},
        "fp64": {"value": round(pts, 2), "unit": "points/s", "mixed_over_fp64": 1.0,
                 "note": "the sweep has no reduced-precision path: every product, inverse and chain is FP64 (a "
                         "complex64 Parseval product measured beta^2 errors up to 7.9e-2 against FP64, outside the "
                         "1e-4 bar), so the batched path timed here is the FP64 one and the ratio is 1.0"},
        "roofline": roof,
        "kernel_ms": {k: [round(v[0], 2), v[1]] for k, v in kt_.items()},
        # useful flops (8 n^2 per complex multiply-add of the product) over each class's HIP-event time
        "kernel_tflops": {k: round(v[2] / (v[0] * 1e-3) / 1e12, 2) if v[0] > 0 else 0.0 for k, v in kt_.items()},
        "beta2_slice0": {"min": float(b2.min()), "max": float(b2.max())},
        "host": {"prep_ms": round(host_prep[0], 2), "of_which_geometries_ms": round(host_geo[0], 2),
                 "wait_ms": round(host_wait[0], 2), "chunks": host_prep[1],
                 "note": "host work of the instrumented step (region lists, descriptors, launches) and its waits "
                         "for a free staging slot; the prep of a chunk overlaps the previous chunk's kernels"},
    }
    tuned, tune_ms = ctx.tuning_report()
    out["tuning"] = {"enabled": not args.no_autotune, "ms": round(tune_ms, 1), "chosen": tuned,
                     "note": "Parseval product column-block width (2: 128, 4: 256 matrix entries per block) per "
                             "matrix size, timed on the first flush of at least 256 blocks in the warmup; not in "
                             "the timed steps"}
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle, bounded sample)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = sweep_cpu_baseline(coords, geoms, set_of, sets, freqs)
    print(json.dumps(out), flush=True)


def cheerleader_main(args):
    """cheerleader (Gauss-Seidel recursive Fourier solver) on the 20-disk booster,
    256x256, 50 frequencies, 100 iterations (the reference test's nmax), mixed
    precision with the FP64 sweep beside it"""
    import BoostFractor as B
    coords, sb, freqs = workload()
    ctx = B.context()
    nmax = 100
    dx = coords.X[1] - coords.X[0]
    print(f"[bench] cheerleader: {NDISK}-disk booster, {N}x{N}, {NFREQ} freqs, nmax={nmax}", file=sys.stderr,
          flush=True)

    def timed(prec, steps, warmup):
        for _ in range(warmup):
            B.cheerleader(0, nmax, sb, coords, f=freqs, diskR=DISKR, precision=prec)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            E = B.cheerleader(0, nmax, sb, coords, f=freqs, diskR=DISKR, precision=prec)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        print(f"[bench] cheerleader {prec}: {dt / steps:.3f} s/step", file=sys.stderr, flush=True)
        return dt, np.sum(np.abs(E) ** 2, axis=(1, 2)) * dx * dx / (math.pi * DISKR ** 2)

    dt, b32 = timed("mixed", args.steps, args.warmup)
    dt64, b64 = timed("fp64", max(1, args.fp64_steps), 1)
    err = np.abs(b32 / b64 - 1.0)
    pts, pts64 = NFREQ * args.steps / dt, NFREQ * max(1, args.fp64_steps) / dt64
    out = {
        "metric": "cheerleader boost-factor freq points/s on 1 MI355X: mixed-prec batched vs plain FP64",
        "value": round(pts, 3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "storage": "complex64 fields (mixed precision: FFT, phase and update arithmetic in f64)",
        "data": "synthetic (seeded SplitMix64 booster geometry, axion source from dance_intro)",
        "config": {"workload": f"cheerleader, 20-disk booster eps=24, {N}x{N} grid over 1 m, {NFREQ} freqs "
                               f"19.5-20.5 GHz, nmax={nmax}, FFT propagator, host buffers in and out",
                   "disks": NDISK, "grid": N, "frequencies": NFREQ, "nmax": nmax},
        "fp64": {"value": round(pts64, 3), "unit": "points/s", "mixed_over_fp64": round(pts / pts64, 3),
                 "beta2_rel_err_vs_fp64": {"max": float(np.max(err)), "median": float(np.median(err))}},
        "roofline": None,
        "note": "the Gauss-Seidel sweep is a sequence of small launches (one region pair per step); "
                "per-kernel rooflines are reported for the dancer",
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle cheerleader, bounded sample)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cheerleader_cpu_baseline(coords, sb, freqs, nmax)
    print(json.dumps(out), flush=True)


def model1d_main(args):
    """1D analytical model sweep: 4096 spacing candidates of a 20-disk booster x
    1000 frequencies in one launch (disk_system_batch)"""
    import BoostFractor as B
    from BoostFractor import workloads as W
    ncand, nfreq, nd = 4096, 1000, NDISK
    rng = W.SplitMix64(7000)
    sp = 8e-3 + (2.0 * rng.uniforms(ncand * (nd + 1)) - 1.0).reshape(ncand, nd + 1) * 0.3e-3
    freqs = np.linspace(19.5e9, 20.5e9, nfreq)
    ctx = B.context()
    A1 = B.Analytical1D
    for _ in range(args.warmup):
        A1.disk_system_batch(freqs, sp[:64], disk_epsilon=24)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g, a = A1.disk_system_batch(freqs, sp, disk_epsilon=24)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    pts = ncand * nfreq * args.steps / dt
    out = {
        "metric": "1D analytical model boost-factor points/s on 1 MI355X (FP64)",
        "value": round(pts, 1), "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded SplitMix64 spacings)",
        "note": "host buffers in and out (65 MB of results per step)",
        "config": {"workload": f"disk_system (src/Analytical1D.py), {nd} disks eps=24, {ncand} spacing candidates "
                               f"8 mm +- 0.3 mm x {nfreq} freqs 19.5-20.5 GHz", "candidates": ncand,
                   "frequencies": nfreq},
        "boost_max": float(np.max(np.abs(a) ** 2)),
    }
    if not args.no_cpu_baseline:
        from oracle import analytic1d as OA
        t1 = time.perf_counter()
        nc = 8
        for c_ in range(nc):
            OA.disk_system(freqs, 24 + 0j, 1 + 0j, sp[c_], nd)
        cdt = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": round(nc * nfreq / cdt, 1), "unit": "points/s", "cores": 1, "kind": "port",
                               "sample": f"{nc} candidates x {nfreq} frequencies with the numpy restatement of "
                                         "src/Analytical1D.py (vectorised over frequencies, one thread)"}
    print(json.dumps(out), flush=True)


def main():
    global N, NFREQ, NMAX  # the dancer512 workload resizes the dancer sweep
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-autotune", action="store_true", help="use the launch-shape heuristics only")
    ap.add_argument("--no-graph", action="store_true",
                    help="dancer plans launch every bounce instead of replaying HIP graphs (profiler runs)")
    ap.add_argument("--fp32-steps", type=int, default=0,
                    help="dancer workloads: also time the FP32-arithmetic path (compared with FP64)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--freq-batch", type=int, default=0)
    ap.add_argument("--nmax", type=int, default=NMAX,
                    help="bounces - 1 (default: the workload's 2000; smaller only for profiler runs)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DB", "WRITE_DB"))
    ap.add_argument("--summarize-trace", metavar="TRACE_DB")
    ap.add_argument("--workload", choices=("dancer", "dancer512", "cheerleader", "transformer", "sweep", "model1d"),
                    default="sweep",
                    help="sweep (default): the disk-position sweep, 4096 geometries x 100 frequencies, the steps "
                         "splitting the geometries into disjoint slices; dancer: the 20-disk booster at 256x256, "
                         "50 frequencies, 2000 bounces; dancer512: the same booster at 512x512, 100 frequencies, "
                         "3000 bounces (run it with --steps 1 --warmup 0: ~1 min per sweep); cheerleader: the "
                         "Gauss-Seidel solver on the 20-disk booster; transformer: the mode-matching sweep of the "
                         "tilted, rough booster at 512x512 with 45 Laguerre-Gauss modes; model1d: the 1D "
                         "analytical model")
    ap.add_argument("--nconf", type=int, default=NCONF5, help="geometries of the sweep workload (default: all 4096)")
    ap.add_argument("--batch-gib", type=float, default=0.0,
                    help="sweep: matrix budget of a chunk in GiB (engine option batch_bytes; 0: the default)")
    ap.add_argument("--keep-boosts", action="store_true", help="fetch every step's boosts (sweep workload)")
    args = ap.parse_args()
    if args.summarize_pmc or args.summarize_trace:
        if args.summarize_trace:
            print(summarize_trace(args.summarize_trace))
        if args.summarize_pmc:
            print(summarize_pmc(*args.summarize_pmc))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the workload is a single-GPU sweep and this build has no multi-GPU "
              "path (frequency points would shard across GPUs with no exchange); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "transformer":
        transformer_main(args)
        return
    if args.workload == "sweep":
        sweep_main(args)
        return
    if args.workload == "model1d":
        model1d_main(args)
        return
    if args.workload == "cheerleader":
        cheerleader_main(args)
        return
    if args.workload == "dancer512":
        N, NFREQ = 512, 100
        if args.nmax == NMAX:
            args.nmax = NMAX = 3000

    import BoostFractor as B
    from BoostFractor import workloads as W
    coords, sb, freqs = workload()
    ctx = B.context()
    dx = coords.X[1] - coords.X[0]
    K = N * N
    F = 2 * (2 * NDISK + 1)  # fields per bounce (dancer regions x 2 directions)

    nmax = args.nmax
    print(f"[bench] workload: {NDISK}-disk booster, {N}x{N}, {NFREQ} freqs, nmax={nmax}", file=sys.stderr, flush=True)
    if not args.no_autotune:
        ctx.set_option("autotune", 1)  # plans time their launch shapes when created (untimed), see `tuning`
    if args.no_graph:
        ctx.set_option("graph", 0)
    plan32 = B.DancerPlan(0, nmax, sb, coords, freqs, diskR=DISKR, precision="mixed", freq_batch=args.freq_batch)
    info32 = plan32.info()
    fb32 = info32["freq_batch"]
    dt32, kt32 = timed_steps(plan32, ctx, args.steps, args.warmup, timing=True)
    _, p32 = plan32.fetch(fields=False)
    plan32.close()
    print(f"[bench] mixed: {dt32 / args.steps:.3f} s/step", file=sys.stderr, flush=True)

    pts32 = NFREQ * args.steps / dt32
    b32 = p32 * dx * dx / (math.pi * DISKR ** 2)
    fp64 = None
    if args.fp64_steps > 0:  # 0 only for profiler runs of the mixed sweep
        plan64 = B.DancerPlan(0, nmax, sb, coords, freqs, diskR=DISKR, precision="fp64", freq_batch=args.freq_batch)
        dt64, kt64 = timed_steps(plan64, ctx, args.fp64_steps, 1, timing=True)
        _, p64 = plan64.fetch(fields=False)
        fb64 = plan64.freq_batch
        plan64.close()
        print(f"[bench] fp64: {dt64 / args.fp64_steps:.3f} s/step", file=sys.stderr, flush=True)
        b64 = p64 * dx * dx / (math.pi * DISKR ** 2)
        err = np.abs(b32 / b64 - 1.0)
        pts64 = NFREQ * args.fp64_steps / dt64
        fp64 = {"value": round(pts64, 4), "unit": "points/s", "ms_per_step": round(dt64 / args.fp64_steps * 1e3, 3),
                "steps": args.fp64_steps, "freq_batch": fb64,
                "mixed_over_fp64": round(pts32 / pts64, 3),
                "beta2_rel_err_vs_fp64": {"max": float(np.max(err)), "median": float(np.median(err)),
                                          "p90": float(np.percentile(err, 90))},
                "kernel_ms": {k: [round(v[0], 3), v[1]] for k, v in kt64.items()}}
        if args.fp32_steps > 0:  # FP32 storage and FP32 arithmetic, timed and compared with FP64 beside
            planf = B.DancerPlan(0, nmax, sb, coords, freqs, diskR=DISKR, precision="fp32",
                                 freq_batch=args.freq_batch)
            dtf, _ = timed_steps(planf, ctx, args.fp32_steps, 1, timing=False)
            _, pf = planf.fetch(fields=False)
            fbf = planf.freq_batch
            planf.close()
            print(f"[bench] fp32: {dtf / args.fp32_steps:.3f} s/step", file=sys.stderr, flush=True)
            errf = np.abs(pf * dx * dx / (math.pi * DISKR ** 2) / b64 - 1.0)
            ptsf = NFREQ * args.fp32_steps / dtf
            fp64["fp32_arithmetic"] = {
                "value": round(ptsf, 4), "unit": "points/s", "freq_batch": fbf, "over_fp64": round(ptsf / pts64, 3),
                "beta2_rel_err_vs_fp64": {"max": float(np.max(errf)), "median": float(np.median(errf)),
                                          "p90": float(np.percentile(errf, 90))},
                "note": "complex64 fields with FP32 FFT, phase and update arithmetic (phase argument in FP64)"}

    # roofline of the mixed path's bounce: a frequency batch's bounce is two launches (column pass, fused
    # row pass) that together read and write every field once; the bounce's algorithmic bytes are divided
    # by the two launches' combined mean time (each pass alone moves the same fields, so neither gets the
    # whole bounce's bytes on its own time)
    col_ms, launches = kt32["col_pass"]
    row_ms, row_launches = kt32["row_pass"]
    avg_ms = (col_ms + row_ms) / launches  # one column + one row launch per frequency batch and bounce
    nb_per_launch = NFREQ * (nmax + 1) / launches  # frequency-bounces per launch pair (one instrumented run)
    alg_bytes = nb_per_launch * 2 * F * K * 8  # every field read once and written once, complex64
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    # HBM bytes of both passes per launch pair, from the committed counter summary (256 x 256 only)
    tc = traffic_from_profile("k_cols_reg", fb32) if N == 256 else None
    tr = traffic_from_profile("k_dancer_rows_reg4", fb32) if N == 256 else None
    traffic = (tc + tr) if (tc is not None and tr is not None) else None
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
        "dtype": "f64",
        "storage": "complex64 fields in HBM (mixed precision: every FFT, phase and boundary update in f64)",
        "data": "synthetic (seeded SplitMix64 booster geometry, axion source from dance_intro)",
        "config": {"workload": f"dancer, 20-disk booster eps=24, {N}x{N} grid over 1 m, {NFREQ} freqs "
                               f"19.5-20.5 GHz, nmax={nmax}, FFT propagator, one batched device-resident sweep",
                   "disks": NDISK, "grid": N, "frequencies": NFREQ, "nmax": nmax, "freq_batch": fb32},
        "fp64": fp64,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic if traffic is None else int(traffic),
                     "kernel": "dancer bounce: k_cols_reg (column FFT, exp(-i kz dz)/K, column IFFT) + "
                               "k_dancer_rows_reg (row IFFT, tilt/surface, boundary update, mask, row FFT)",
                     "avg_launch_ms": round(avg_ms, 4),
                     "col_pass_ms": round(col_ms / launches, 4), "row_pass_ms": round(row_ms / row_launches, 4),
                     "alg_bytes_per_launch": int(alg_bytes),
                     # the same algorithmic bytes over the whole step (both bounce kernels, the gaps between
                     # launches and the per-batch setup): the rate a user of the sweep sees
                     "whole_step_GBs": round(NFREQ * (nmax + 1) * 2 * F * K * 8 / (dt32 / args.steps) / 1e9, 1),
                     "measured_stream_copy_GBs": round(stream_gbs, 1)},
        "work_reductions": {"row_band": list(info32["band"]), "rows": N,
                            "note": "rows outside the aperture band skipped by both passes (exit field kept whole); "
                                    "band-pruned register FFTs (zero elements and unstored outputs skipped); "
                                    "identical k-space tables shared, quarter tables for complex64 fields; "
                                    "not folded into roofline.achieved"},
        "kernel_ms": {k: [round(v[0], 3), v[1]] for k, v in kt32.items()},
    }
    tuned, tune_ms = ctx.tuning_report()
    out["tuning"] = {"enabled": not args.no_autotune, "ms": round(tune_ms, 1), "chosen": tuned,
                     "note": "dancer plans: frequency batch, column-pass tile, row-pass rows (0: kernel default), "
                             "timed on short runs of the plan's inputs at plan creation; not in the timed steps"}
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle, bounded sample)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(coords, sb, freqs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

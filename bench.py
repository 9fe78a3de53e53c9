"""Benchmark: the search hot path on one MI355X.

One step = subband dedispersion of a resident 1024-channel x 2^22-sample
uint8 filterbank over 1024 DM trials (0-500, smearing-tolerance steps), then
in-place uint32->float conversion + 2 s running-median detrend + normalise of
the DM x time plane, then the FFA search (0.1-10 s, boxcar widths 1-32) of
every DM trial with peak collection.  The filterbank is generated on the GPU
before timing (seeded, 8 injected pulsars).

Prints one JSON line: DM-trial*samples/s over the whole step (wall clock with
device synchronisation on both sides); the roofline of the dominant kernel
(largest share of the step): its algorithmic bytes per launch over its
average launch time, from HIP events recorded between the launch stages on
the pipeline stream; every hot kernel's figures under "kernels"; and a CPU
baseline (gaffa's own CPU code, compiled from its sources, on a bounded
sample).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "dedisp+FFA DM-trial·samples/s, % HBM roofline, vs plain GPU at equal recall"
UNIT = "DM-trial·samples/s"
HBM_PEAK_GBS = 8000.0


def global_passes(depth: int, terminal_levels: int = 1) -> int:
    """Global-memory merge passes of a task whose LDS subtrees sit at `depth`
    and whose detection kernel merges `terminal_levels` levels itself:
    depth - terminal_levels tree levels, three per pass, two where two or four
    remain."""
    rem, passes = max(depth - terminal_levels, 0), 0
    while rem > 0:
        rem -= 2 if rem in (2, 4) else min(rem, 3)
        passes += 1
    return passes


def kernel_bytes(workload, pipe, ntiles: int, npeaks: int) -> dict:
    """Algorithmic HBM bytes per launch of each hot kernel (the minimum any
    implementation must move), keyed like the stage times:

    dedisperse  one launch for every DM segment: the uint8 filterbank once
                (nsamples * nchans) + the uint32 plane once (4 * ndm * n_out)
    prepare     per series: the float series once (4 n) + every downsampled copy
                (4 P_g per downsample group)
    subtree     per series and task with rows >= 2: read + write its rows x bins
                floats once (8 rows bins)
    levels      per series and task: 8 rows bins for each pass through global
                memory (global_passes(depth_t, terminal_levels_t) of them: a pass
                merges three tree levels, two where two or four remain), averaged
                over the launches
    detect      per series and task: the rows it merges from (depth 1, or depth 2
                where it merges two levels) once (4 rows bins; the input series
                row for 1-row tasks) + 20 B per peak
    """
    import numpy as np

    t = pipe.ffa_plan.tasks
    lay = pipe.ffa_layout
    n_out, ndm = pipe.out_nsamples, pipe.ndm
    rb = t["rows"].astype(np.float64) * t["bins"]
    depth = lay["task_depth"].astype(np.float64)
    series_per_tile = ndm / ntiles
    prep = 4.0 * n_out
    for f in np.unique(t["downsample_factor"]):
        if f != 1.0:
            prep += 4.0 * float(t["prepared_nsamples"][t["downsample_factor"] == f][0])
    subtree = 8.0 * float(rb[t["rows"] >= 2].sum())
    tlev = lay["terminal_levels"]
    passes = np.array([global_passes(int(d), int(k)) for d, k in zip(depth, tlev)], np.float64)
    levels = 8.0 * float((rb * passes).sum())
    detect = 4.0 * float(rb.sum())
    nlev = max(1, lay["level_launches"])
    dd_launches = 1 if workload.subband else len(workload.dm_segments)
    return {
        "dedisperse": (dd_launches, (workload.nsamples * workload.nchans + 4.0 * ndm * n_out)
                       / dd_launches),
        "ffa_prepare": (ntiles, prep * series_per_tile),
        "ffa_subtree": (ntiles, subtree * series_per_tile),
        "ffa_levels": (ntiles * nlev, levels * series_per_tile / nlev),
        "ffa_detect": (ntiles, detect * series_per_tile + 20.0 * npeaks / ntiles),
    }


KERNEL_NAMES = {
    "dedisperse": "dedisperse_subband_stream_kernel",
    "ffa_prepare": "ffa_prepare_fused_kernel",
    "ffa_subtree": "ffa_subtree_kernel",
    "ffa_levels": "ffa_level_kernel",
    "ffa_detect": "ffa_terminal_kernel<true>",
}


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by a copy kernel of
    known size), or None when no summary covers this kernel and workload."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as fh:
        for line in fh:
            if "=" in line and line.startswith("    "):
                k, v = line.strip().split("=", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def calibrate_copy(mib: int) -> None:
    """One device-to-device copy of `mib` MiB through the C ABI (the plain copy
    kernel whose FETCH_SIZE / WRITE_SIZE scale the pmc counters)."""
    import ctypes

    import gaffa_amd

    lib = ctypes.CDLL(gaffa_amd.HIP_LIBRARY)
    lib.gaffa_hip_malloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint64]
    lib.gaffa_hip_free.argtypes = [ctypes.c_void_p]
    lib.gaffa_hip_memcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                     ctypes.c_void_p]
    lib.gaffa_hip_memset.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64,
                                     ctypes.c_void_p]
    nbytes = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if lib.gaffa_hip_malloc(ctypes.byref(ptr), nbytes) != 0:
            raise RuntimeError("copy calibration: allocation failed")
    try:
        if lib.gaffa_hip_memset(a, 1, nbytes, None) != 0 or \
                lib.gaffa_hip_memcpy(b, a, nbytes, None) != 0 or \
                lib.gaffa_hip_device_synchronize() != 0:
            raise RuntimeError("copy calibration: copy failed")
    finally:
        lib.gaffa_hip_free(a)
        lib.gaffa_hip_free(b)
    print(f"[bench] copy calibration: {mib} MiB device-to-device", file=sys.stderr, flush=True)


def cpu_baseline(workload, filterbank, threads: int) -> dict:
    """gaffa's CPU path on one nominal group (32 DMs) of the same workload:
    subband dedispersion over the full time axis, then detrend + normalise +
    FFA search of the 32 series (find_dm_peaks_cpu), timed end to end."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle.oracle import LIBS, Oracle

    kind = "reference" if os.path.exists(LIBS["reference"]) else "port"
    cpu = Oracle(kind)
    seg = workload.dm_segments[0]
    ndm = min(128, seg.ndm)  # four nominal groups: about 10-15 s on 16 cores
    t0 = time.perf_counter()
    plane = cpu.dedisperse(filterbank, workload.frequencies, dm_low=seg.dm_low,
                           dm_step=seg.dm_step, ndm=ndm, tsamp=workload.tsamp, subband=True,
                           subband_channels=workload.subband_channels,
                           ndm_per_nominal=workload.ndm_per_nominal)
    t1 = time.perf_counter()
    opts = cpu.plan_options(workload.period_min, workload.period_max, workload.bins_min,
                            workload.bins_max, 1, workload.duty_cycle_max,
                            workload.width_trial_spacing)
    pp = cpu.preprocess_plan(window_samples=workload.window_samples(),
                             min_points=workload.min_points)
    dms = seg.dm_low + seg.dm_step * __import__("numpy").arange(ndm)
    peaks = cpu.find_dm_peaks(plane, dms, workload.tsamp, opts, pp, workload.snr_threshold)
    t2 = time.perf_counter()
    units = ndm * plane.shape[1]
    return {
        "value": units / (t2 - t0),
        "unit": UNIT,
        "cores": threads,
        "kind": kind,
        "sample": (f"the first {ndm} DM trials (DM {seg.dm_low:.2f}+{seg.dm_step:.4f}k) "
                   f"of the same filterbank over all {plane.shape[1]} output samples: subband "
                   f"dedispersion {t1 - t0:.2f} s + detrend/normalise/FFA {t2 - t1:.2f} s, "
                   f"{len(peaks)} peaks"),
    }


def scored_comparison(workload, pipe, search_ms: float) -> dict:
    """Plain pipeline: full fold (gaffa's pfold of the filterbank at the
    candidate's DM: aligned spectrum + 32 subints x 64 channels x 64 bins)
    of every candidate of the last search.  Scored pipeline: the learned
    scorer on every candidate (features and network on the GPU), full fold
    of those it keeps.  Detections: injected pulsars matched (frequency within
    a Fourier bin of 1/2, 1 or 2 times theirs; DM within the pulse width
    or two trials) by a folded candidate with S/N >= 8."""
    import numpy as np

    import gaffa_amd
    from gaffa_amd import scorer as sc

    cands = pipe.select_candidates()
    net, threshold = sc.load_trained()
    t0 = time.perf_counter()
    z = pipe.score_candidates(net, cands)
    score_ms = 1000.0 * (time.perf_counter() - t0)
    kept = [c for c, v in zip(cands, z) if v >= threshold]
    # the plain pipeline folds every candidate: timed on all of them, or on an
    # evenly spaced sample of 2000 and scaled when there are more
    sample = cands if len(cands) <= 2000 else cands[:: -(-len(cands) // 2000)]
    sample_ms, _ = pipe.fold_candidates_full(sample)
    plain_fold_ms = sample_ms * len(cands) / max(1, len(sample))
    scored_fold_ms, _ = pipe.fold_candidates_full(kept) if kept else (0.0, None)
    obs = pipe.out_nsamples * workload.tsamp
    dms = np.asarray(pipe.dms)
    f = workload.frequencies
    kdm = sc.DISPERSION_S * (f.min() ** -2 - f.max() ** -2)

    def detected(cs):
        found = set()
        for c in cs:
            if c.snr < 8.0:
                continue
            step = (dms[min(c.dm_index + 1, len(dms) - 1)] - dms[max(c.dm_index - 1, 0)]) / 2
            k = sc._match(c.period, c.dm, workload.pulsars, obs, kdm, step, sc.MAIN_RATIOS)
            if k >= 0:
                found.add(k)
        return found

    plain = detected(cands)
    scored = detected(kept)
    # the method without the network: fold only candidates with S/N >= 8
    strong = [c for c in cands if c.snr >= 8.0]
    by_snr = detected(strong)
    plain_ms = search_ms + plain_fold_ms
    scored_ms = search_ms + score_ms + scored_fold_ms
    return {
        "candidates": len(cands), "kept": len(kept), "threshold": threshold,
        "plain_fold_timed_on": len(sample),
        "fold_ms_per_candidate": plain_fold_ms / max(1, len(cands)),
        "plain_ms": plain_ms, "scored_ms": scored_ms, "score_ms": score_ms,
        "speedup": plain_ms / scored_ms,
        "injected": len(workload.pulsars), "plain_detections": len(plain),
        "scored_detections": len(plain & scored),
        "recall": len(plain & scored) / max(1, len(plain)),
        "snr8_kept": len(strong), "snr8_detections": len(plain & by_snr),
        "scored_and_snr8_kept": sum(1 for c in kept if c.snr >= 8.0),
    }


def bench_batched(args) -> int:
    """Batched FFA: float series of different lengths, each with its own
    octave plan, searched as tiles of one launch sequence each, replayed from
    a HIP graph, with the tile size chosen by the autotuner; the per-series
    launch loop (one sequence per series) is timed beside it."""
    import numpy as np

    from gaffa_amd.pipeline import batched_workload

    w = batched_workload(args.batched_series)
    t0 = time.perf_counter()
    b = w.search(series_per_tile=16)
    print(f"[bench] batched: {w.nseries} series, {int(w.lengths.sum())} samples, row "
          f"{b.row_length}, setup {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    best = b.autotune([8, 16, 32, 64])
    print(f"[bench] autotuned series per tile: {best}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        b.run(w.snr_threshold)
    total_ms, gpu_ms, npeaks, timing = 0.0, 0.0, 0, {}
    for i in range(args.steps):
        t1 = time.perf_counter()
        peaks, timing = b.run(w.snr_threshold)
        total_ms += 1000.0 * (time.perf_counter() - t1)
        gpu_ms += timing["gpu_ms"]
        npeaks = len(peaks)
        print(f"[bench] step {i}: {json.dumps(timing)}", file=sys.stderr, flush=True)
    loop_peaks, loop = b.run(w.snr_threshold, per_series=True)
    assert len(loop_peaks) == npeaks, "per-series loop found a different peak list"
    units = float(w.lengths.sum())
    ms = total_ms / args.steps
    plans = w.plans()
    rb = sum(float((p.tasks["rows"].astype(np.float64) * p.tasks["bins"]).sum()) for p in plans)
    nbytes = 4.0 * float(w.lengths.sum()) + 4.0 * rb  # series once + each task's rows once
    gbs = nbytes / (gpu_ms / args.steps * 1e-3) / 1e9
    print(json.dumps({
        "metric": METRIC, "value": units / (ms * 1e-3), "unit": UNIT, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded GPU N(0,1) series, 5% injected pulse trains)",
        "config": {"workload": (f"batched FFA: {w.nseries} series, lengths 2^20-2^22 at "
                                f"{w.tsamp} s, octave plans 0.1-10 s (bins 128-255), "
                                "each series its own plan"),
                   "series_per_tile": best, "tiles": timing.get("tiles"),
                   "launches": timing.get("launches"), "peaks": npeaks},
        "roofline": {"bound": "hbm", "kernel": "batched FFA launch sequence (all tiles)",
                     "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None},
        "per_series_loop": {"gpu_ms": loop["gpu_ms"], "launches": loop["launches"],
                            "speedup_of_batched": loop["gpu_ms"] / (gpu_ms / args.steps)},
    }), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["lband", "small", "large", "batched"],
                    default="lband")
    ap.add_argument("--batched-series", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--scored", action="store_true",
                    help="after the timed steps, compare the plain pipeline (full fold of every "
                         "candidate) with the scorer-pruned one (committed weights)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="after the timed steps, copy MIB MiB device-to-device (pmc scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the DM axis would "
              "shard across GPUs with no exchange); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np

    import gaffa_amd
    from gaffa_amd.pipeline import large_workload, lband_workload, small_workload

    if gaffa_amd.gpu_device_count() == 0:
        print("no HIP device visible", file=sys.stderr)
        return 1
    if args.workload == "batched":
        return bench_batched(args)
    w = {"lband": lband_workload, "small": small_workload,
         "large": large_workload}[args.workload]()
    print(f"[bench] workload {w.name}: {w.nchans} ch x {w.nsamples} samples, {w.ndm} DMs",
          file=sys.stderr, flush=True)
    pipe = w.pipeline()
    pipe.set_kernel_timing(True)
    if args.scored:
        pipe.enable_periodogram(True)
    pipe.generate(w.synth_spec())
    ndm, n_out = pipe.ndm, pipe.out_nsamples
    print(f"[bench] n_out {n_out}, FFA tasks {pipe.ffa_plan.task_count}, "
          f"series tile {pipe.ffa_tile}", file=sys.stderr, flush=True)

    for i in range(args.warmup):
        t = pipe.run()
        print(f"[bench] warmup {i}: {json.dumps(t)}", file=sys.stderr, flush=True)

    stage_sum: dict = {}
    t0 = time.perf_counter()
    for i in range(args.steps):
        t = pipe.run()  # ends with a stream synchronisation
        for k, v in t.items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
        print(f"[bench] step {i}: {json.dumps(t)}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    ms_per_step = 1000.0 * wall / args.steps
    units = ndm * n_out
    value = units / (wall / args.steps)
    stages = {k: v / args.steps for k, v in stage_sum.items()}
    npeaks = pipe.peak_count
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)

    # per-kernel rooflines: algorithmic bytes per launch / average launch time
    ntiles = int(stages.get("ffa_tiles", 0)) or -(-ndm // pipe.ffa_tile)
    kernels = {}
    for key, (launches, nbytes) in kernel_bytes(w, pipe, ntiles, npeaks).items():
        ms = stages.get(f"{key}_ms")
        if ms is None or launches == 0:
            continue
        launch_ms = ms / launches
        gbs = nbytes / (launch_ms * 1e-3) / 1e9
        kernels[key] = {"kernel": KERNEL_NAMES[key], "launches_per_step": launches,
                        "launch_ms": launch_ms, "bytes_per_launch": nbytes,
                        "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS,
                        "share_of_step": ms / stages["total_ms"]}
    dom_key = max(kernels, key=lambda k: kernels[k]["launch_ms"] * kernels[k]["launches_per_step"])
    dom = kernels[dom_key]

    result = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8->u32 sums (dedispersion), f32 (detrend, FFA, S/N)",
        "data": (f"synthetic (seeded GPU generator, {len(w.pulsars)} injected pulsars"
                 + (f", {len(w.rfi_channels)} RFI channels" if w.rfi_channels else "") + ")"),
        "config": {
            "workload": (f"{w.name}: {w.nchans} ch {w.fch1}/{w.foff} MHz, tsamp {w.tsamp}, "
                         f"{w.nsamples} samples; subband {w.subband_channels} ch/subband x "
                         f"{w.ndm_per_nominal} DMs/nominal over {ndm} DM trials "
                         f"{pipe.dms[0]:.1f}-{pipe.dms[-1]:.1f}; {w.median_seconds} s median; "
                         f"FFA {w.period_min}-{w.period_max} s, bins {w.bins_min}-{w.bins_max}, "
                         f"widths up to {int(w.duty_cycle_max * w.bins_min)}"),
            "ndm": ndm,
            "out_nsamples": n_out,
            "ffa_tasks": pipe.ffa_plan.task_count,
            "peaks": npeaks,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom["kernel"],
            "achieved": dom["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dom["frac"],
            "traffic": read_traffic(dom["kernel"].split("<")[0], w.name),
            "launch_ms": dom["launch_ms"],
            "bytes_per_launch": dom["bytes_per_launch"],
        },
        "stages_ms": {k: v for k, v in stages.items() if k.endswith("_ms")},
        "kernels": kernels,
    }
    if args.scored:
        result["scored_pipeline"] = scored_comparison(w, pipe, ms_per_step)
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        print(f"[bench] cpu baseline on {threads} threads", file=sys.stderr, flush=True)
        fb = pipe.download_filterbank(w.nsamples, w.nchans)
        result["cpu_baseline"] = cpu_baseline(w, fb, threads)
        del fb
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

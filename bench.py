"""Benchmark: the search hot path on one MI355X.

One step = subband dedispersion of a resident 1024-channel x 2^22-sample
uint8 filterbank over 1024 DM trials (0-500, smearing-tolerance steps), then
in-place uint32->float conversion + 2 s running-median detrend + normalise of
the DM x time plane, then the FFA search (0.1-10 s, boxcar widths 1-32) of
every DM trial with peak collection.  The filterbank is generated on the GPU
before timing (seeded, 8 injected pulsars).

Prints one JSON line: DM-trial*samples/s over the whole step (wall clock with
device synchronisation on both sides), the FFA stage's algorithmic-byte
roofline from HIP events recorded on the pipeline stream, and a CPU baseline
(gaffa's own CPU code, compiled from its sources, on a bounded sample).
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


def ffa_algorithmic_bytes(plan, n_out: int, nseries: int, npeaks: int) -> float:
    """Per launch sequence over `nseries` series: each prepare group reads the
    series and writes its prepared copy (factor-1 groups read in place), each
    task reads its rows*bins input once, each peak is written once (20 B)."""
    import numpy as np

    tasks = plan.tasks
    per_series = 0.0
    for f in np.unique(tasks["downsample_factor"]):
        if f != 1.0:
            p = int(tasks["prepared_nsamples"][tasks["downsample_factor"] == f][0])
            per_series += 4.0 * n_out + 4.0 * p
    per_series += 4.0 * float((tasks["rows"].astype(np.float64) * tasks["bins"]).sum())
    return per_series * nseries + 20.0 * npeaks


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
    ndm = 32
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
        "sample": (f"one nominal group ({ndm} DM trials, DM {seg.dm_low:.2f}+{seg.dm_step:.4f}k) "
                   f"of the same filterbank over all {plane.shape[1]} output samples: subband "
                   f"dedispersion {t1 - t0:.2f} s + detrend/normalise/FFA {t2 - t1:.2f} s, "
                   f"{len(peaks)} peaks"),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["lband", "small"], default="lband")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the DM axis would "
              "shard across GPUs with no exchange); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np

    import gaffa_amd
    from gaffa_amd.pipeline import lband_workload, small_workload

    if gaffa_amd.gpu_device_count() == 0:
        print("no HIP device visible", file=sys.stderr)
        return 1
    w = lband_workload() if args.workload == "lband" else small_workload()
    print(f"[bench] workload {w.name}: {w.nchans} ch x {w.nsamples} samples, {w.ndm} DMs",
          file=sys.stderr, flush=True)
    pipe = w.pipeline()
    pipe.generate(w.synth_spec())
    ndm, n_out = pipe.ndm, pipe.out_nsamples
    print(f"[bench] n_out {n_out}, FFA tasks {pipe.ffa_plan.task_count}, "
          f"series tile {pipe.ffa_tile}", file=sys.stderr, flush=True)

    for i in range(args.warmup):
        t = pipe.run()
        print(f"[bench] warmup {i}: {json.dumps(t)}", file=sys.stderr, flush=True)

    stage_sum = {"dedisperse_ms": 0.0, "preprocess_ms": 0.0, "ffa_ms": 0.0, "total_ms": 0.0}
    t0 = time.perf_counter()
    for i in range(args.steps):
        t = pipe.run()  # ends with a stream synchronisation
        for k in stage_sum:
            stage_sum[k] += t[k]
        print(f"[bench] step {i}: {json.dumps(t)}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    ms_per_step = 1000.0 * wall / args.steps
    units = ndm * n_out
    value = units / (wall / args.steps)
    stages = {k: v / args.steps for k, v in stage_sum.items()}
    npeaks = len(pipe.peaks)

    ffa_bytes = ffa_algorithmic_bytes(pipe.ffa_plan, n_out, ndm, npeaks)
    ffa_gbs = ffa_bytes / (stages["ffa_ms"] * 1e-3) / 1e9
    dd_bytes = w.nsamples * w.nchans + 4.0 * ndm * n_out
    dd_gbs = dd_bytes / (stages["dedisperse_ms"] * 1e-3) / 1e9
    pp_bytes = 8.0 * ndm * n_out
    pp_gbs = pp_bytes / (stages["preprocess_ms"] * 1e-3) / 1e9

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
        "data": "synthetic (seeded GPU generator, 8 injected pulsars)",
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
            "kernel": "FFA stage (prepare + subtree + level + terminal/detect launches)",
            "achieved": ffa_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": ffa_gbs / HBM_PEAK_GBS,
            "traffic": None,
        },
        "stages_ms": stages,
        "stage_roofline_frac": {
            "dedisperse": dd_gbs / HBM_PEAK_GBS,
            "preprocess": pp_gbs / HBM_PEAK_GBS,
            "ffa": ffa_gbs / HBM_PEAK_GBS,
        },
    }
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

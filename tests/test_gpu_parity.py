"""GPU parity: the HIP path (through the C ABI, via the C++ host layer or
ctypes) against the CPU oracle on the same seeded inputs.

Bars: dedispersed series bit-exact; detrended and normalised series
bit-exact (the FP64 statistics are summed in sample order); FFA rows and
peak lists bit-exact on identical inputs (same FP32 add tree, serial FP32
prefixes); end-to-end peak lists complete and identical in order and integer
fields, S/N within 1e-5 relative (the stated tolerance; equal in practice).
"""

import ctypes
import hashlib
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN
from oracle.oracle import Plan, channel_frequencies

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# -------------------------------------------------------------- synth ----
def test_synth_bytes_equal_oracle(gpu, port):
    from gaffa_amd.pipeline import small_workload

    say("synth: small workload slice")
    w = small_workload(nsamples=1 << 15)
    spec = w.synth_spec()
    got = gpu._synth_filterbank_gpu(spec)
    want = port.synth_filterbank(w.nsamples, w.nchans, w.fch1, w.foff, w.tsamp, w.seed,
                                 pulsars=w.pulsars)
    np.testing.assert_array_equal(got, want)
    got = gpu._synth_filterbank_gpu(gpu.SynthFilterbankSpec(
        nsamples=4096, nchans=64, fch1=1500.0, foff=-4.0, tsamp=256e-6, seed=1))
    assert hashlib.sha256(got[:1024].tobytes()).hexdigest() == \
        "c9f34b64d6ee966f3751a63fc3597e1b55c6fb7f8ef30f4f1bc45240ccecd2d5"


# ------------------------------------------------------- dedispersion ----
def test_basetest_dedispersion_bit_exact(gpu):
    from gaffa_amd import dedispersion
    from gaffa_amd.io import Filterbank

    say("dedispersion: basetest.fil 50 DMs")
    fb = Filterbank(os.path.join(GOLDEN, "basetest.fil"))
    g = np.load(os.path.join(GOLDEN, "basetest_dedispersion.npz"))
    multi = dedispersion.dedisperse_multi_dm(fb, dm_low=0.0, dm_step=1.0, ndm=50, backend="cuda")
    assert multi.backend == "cuda" and multi.data.dtype == np.uint32
    np.testing.assert_array_equal(multi.data, g["multi"])
    sub = dedispersion.dedisperse_subband(fb, dm_low=0.0, dm_step=1.0, ndm=50,
                                          subband_channels=16, ndm_per_nominal=8)
    np.testing.assert_array_equal(sub.data, g["subband"])
    # gaffa's degenerate identity: subband(1, 1) == brute force
    deg = dedispersion.dedisperse_subband(fb, dm_low=0.0, dm_step=1.0, ndm=4,
                                          subband_channels=1, ndm_per_nominal=1)
    # (4 DMs have a longer valid range than 50: compare the common prefix)
    np.testing.assert_array_equal(deg.data[:, : multi.nsamples], g["multi"][:4])
    single = dedispersion.dedisperse_single_dm(fb, dm=3.0, backend="gpu")
    np.testing.assert_array_equal(single.data[0, : multi.nsamples], g["multi"][3])


def test_small_brute_force_bit_exact(gpu, port):
    from gaffa_amd import dedispersion
    from gaffa_amd.io import Filterbank
    from gaffa_amd.pipeline import small_workload

    say("dedispersion: 64 ch x 2^17, 64 DMs brute force")
    w = small_workload()
    x = gpu._synth_filterbank_gpu(w.synth_spec())
    fb = Filterbank.from_array(x, tsamp=w.tsamp, frequency_mhz=w.frequencies)
    got = dedispersion.dedisperse_multi_dm(fb, dm_low=0.0, dm_step=1.0, ndm=64, backend="cuda")
    want = port.dedisperse(x, w.frequencies, dm_low=0.0, dm_step=1.0, ndm=64, tsamp=w.tsamp)
    assert got.shape == want.shape == (64, 130870)
    np.testing.assert_array_equal(got.data, want)


def test_subband_slice_bit_exact(gpu):
    from gaffa_amd import dedispersion
    from gaffa_amd.io import Filterbank

    say("dedispersion: 1024-ch slice, low- and high-DM subband segments")
    g = np.load(os.path.join(GOLDEN, "subband_slice.npz"))
    rng = np.random.default_rng(7)
    x = rng.integers(96, 160, size=(20000, 1024), dtype=np.uint8)
    fb = Filterbank.from_array(x, tsamp=64e-6, frequency_mhz=g["frequency"])
    for i, (lo, step, ndm) in enumerate(g["segments"]):
        r = dedispersion.dedisperse_subband(fb, dm_low=lo, dm_step=step, ndm=int(ndm),
                                            subband_channels=32, ndm_per_nominal=32)
        assert list(r.shape) == list(g[f"subband_{i}_shape"])
        digests = [hashlib.sha256(row.tobytes()).hexdigest() for row in r.data]
        assert digests == list(g[f"subband_{i}_digests"])
        np.testing.assert_array_equal(r.data[[0, int(ndm) - 1]], g[f"subband_{i}_rows"])


def test_subband_lband_slice_vs_oracle(gpu, port):
    from gaffa_amd import dedispersion
    from gaffa_amd.io import Filterbank
    from gaffa_amd.pipeline import lband_workload

    say("dedispersion: 1024-ch x 2^16 synthetic, 2 nominal groups per segment")
    w = lband_workload(nsamples=1 << 16, npulsars=2)
    x = gpu._synth_filterbank_gpu(w.synth_spec())
    fb = Filterbank.from_array(x, tsamp=w.tsamp, frequency_mhz=w.frequencies)
    for seg in (w.dm_segments[0], w.dm_segments[3]):
        r = dedispersion.dedisperse_subband(fb, dm_low=seg.dm_low, dm_step=seg.dm_step, ndm=64)
        want = port.dedisperse(x, w.frequencies, dm_low=seg.dm_low, dm_step=seg.dm_step, ndm=64,
                               tsamp=w.tsamp, subband=True, subband_channels=32,
                               ndm_per_nominal=32)
        np.testing.assert_array_equal(r.data, want)


def test_subband_large_slice_vs_oracle(gpu, port):
    """The 4096-channel configuration: 64 channels per subband (stage-1 sums
    of 64 x 255 still fit 16 bits; stage 2 flushes every 4 subbands), RFI
    channels in the input, and the first and last DM segments (DM 0 and
    ~990, delays up to 57 k samples)."""
    from gaffa_amd import dedispersion
    from gaffa_amd.io import Filterbank
    from gaffa_amd.pipeline import large_workload

    say("dedispersion: 4096-ch x 2^17 synthetic with RFI, nsub 64, DM 0 and ~990")
    w = large_workload(nsamples=1 << 17, npulsars=4)
    x = gpu._synth_filterbank_gpu(w.synth_spec())
    want_x = port.synth_filterbank(w.nsamples, w.nchans, w.fch1, w.foff, w.tsamp, w.seed,
                                   pulsars=w.pulsars, rfi_channels=w.rfi_channels,
                                   rfi_offset=w.rfi_offset, t_count=4096)
    np.testing.assert_array_equal(x[:4096], want_x)
    assert x[:, w.rfi_channels].mean() > x.mean() + 5  # the +40 channels are there
    fb = Filterbank.from_array(x, tsamp=w.tsamp, frequency_mhz=w.frequencies)
    # at DM ~990 a 32-trial nominal group spans ~1700 residual samples in the
    # lowest subband; 24 trials per group leave a short last group (64 = 24 +
    # 24 + 16); 48, 56 and 60 trials per group grow the residual windows until
    # the kernel's time chunk drops from 512 to 256, 64 and 16 samples
    top = w.dm_segments[-1]
    for seg, npn in ((w.dm_segments[0], 32), (top, 32), (top, 24), (top, 48), (top, 56),
                     (top, 60)):
        say(f"  DM {seg.dm_low:.1f} step {seg.dm_step:.4f}, {npn} trials per nominal group")
        r = dedispersion.dedisperse_subband(fb, dm_low=seg.dm_low, dm_step=seg.dm_step, ndm=64,
                                            subband_channels=64, ndm_per_nominal=npn)
        want = port.dedisperse(x, w.frequencies, dm_low=seg.dm_low, dm_step=seg.dm_step, ndm=64,
                               tsamp=w.tsamp, subband=True, subband_channels=64,
                               ndm_per_nominal=npn)
        np.testing.assert_array_equal(r.data, want)


def test_large_pipeline_segments_vs_oracle(gpu, port):
    """The pipeline at 4096 channels dedisperses every segment in one launch
    at the configured 32 trials per nominal group; the plane rows of the
    first, a middle and the last segment equal the oracle."""
    import dataclasses

    from gaffa_amd.pipeline import large_workload

    say("pipeline: 4096-ch x 2^17 dedispersion, segment nominal groups vs oracle")
    w = dataclasses.replace(large_workload(nsamples=1 << 17, npulsars=4), period_max=1.0)
    pipe = w.pipeline(ffa_workspace_bytes=2 << 30)
    assert pipe.segment_ndm_per_nominal == [32] * len(w.dm_segments)
    pipe.generate(w.synth_spec())
    pipe.dedisperse()
    plane = pipe.download_plane_u32()
    x = pipe.download_filterbank(w.nsamples, w.nchans)
    n_out = plane.shape[1]
    for k in (0, len(w.dm_segments) // 2, len(w.dm_segments) - 1):
        seg = w.dm_segments[k]
        say(f"  segment {k}: DM {seg.dm_low:.1f} step {seg.dm_step:.4f}")
        want = port.dedisperse(x, w.frequencies, dm_low=seg.dm_low, dm_step=seg.dm_step,
                               ndm=seg.ndm, tsamp=w.tsamp, subband=True, subband_channels=64,
                               ndm_per_nominal=32)
        np.testing.assert_array_equal(plane[k * seg.ndm:(k + 1) * seg.ndm], want[:, :n_out])


def test_c_abi_direct_call(gpu, port):
    """Drive gaffa_hip_dedisperse_multi_u8 through ctypes with raw device memory."""
    import gaffa_amd

    say("C ABI: malloc / memcpy / dedisperse_multi_u8 via ctypes")
    lib = ctypes.CDLL(gaffa_amd.HIP_LIBRARY)
    lib.gaffa_hip_last_error.restype = ctypes.c_char_p

    class Args(ctypes.Structure):
        _fields_ = [("samples", ctypes.c_void_p), ("nsamples", ctypes.c_uint64),
                    ("nchans", ctypes.c_uint64), ("chan_begin", ctypes.c_uint64),
                    ("chan_end", ctypes.c_uint64), ("delays", ctypes.c_void_p),
                    ("ndm", ctypes.c_uint64), ("out", ctypes.c_void_p),
                    ("out_nsamples", ctypes.c_uint64), ("out_stride", ctypes.c_uint64),
                    ("max_block_spread", ctypes.c_int32)]

    rng = np.random.default_rng(1)
    n, nch, ndm = 3000, 40, 7
    x = rng.integers(0, 256, size=(n, nch), dtype=np.uint8)
    freq = channel_frequencies(nch, 1400.0, -2.0)
    delays = port.multi_dm_delay_table(freq, 0.0, 5.0, ndm, freq.max(), 1e-3)
    n_out = n - int(delays.max())
    want = port.dedisperse(x, freq, dm_low=0.0, dm_step=5.0, ndm=ndm, tsamp=1e-3)

    def dmalloc(nbytes):
        p = ctypes.c_void_p()
        assert lib.gaffa_hip_malloc(ctypes.byref(p), ctypes.c_uint64(nbytes)) == 0
        return p

    d_x, d_d, d_o = dmalloc(x.nbytes), dmalloc(delays.nbytes), dmalloc(ndm * n_out * 4)
    try:
        assert lib.gaffa_hip_memcpy(d_x, x.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_uint64(x.nbytes), None) == 0
        dd = np.ascontiguousarray(delays, np.int32)
        assert lib.gaffa_hip_memcpy(d_d, dd.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_uint64(dd.nbytes), None) == 0
        a = Args(d_x, n, nch, 0, nch, d_d, ndm, d_o, n_out, n_out, int(delays.max()))
        rc = lib.gaffa_hip_dedisperse_multi_u8(ctypes.byref(a), None)
        assert rc == 0, lib.gaffa_hip_last_error()
        assert lib.gaffa_hip_dedisperse_check(None) == 0, lib.gaffa_hip_last_error()
        got = np.empty((ndm, n_out), np.uint32)
        assert lib.gaffa_hip_memcpy(got.ctypes.data_as(ctypes.c_void_p), d_o,
                                    ctypes.c_uint64(got.nbytes), None) == 0
        np.testing.assert_array_equal(got, want)
    finally:
        for p in (d_x, d_d, d_o):
            lib.gaffa_hip_free(p)


# ------------------------------------------------------ preprocessing ----
@pytest.mark.parametrize("name", ["small", "lband", "exact"])
def test_preprocess_vs_golden(gpu, port, name):
    say(f"preprocess: golden {name}")
    g = np.load(os.path.join(GOLDEN, "preprocess.npz"))
    window, min_points = (int(v) for v in g[f"{name}_window"])
    x = g[f"{name}_input"]
    u32 = x.astype(np.uint32)[None, :]
    det = gpu._preprocess_batch_gpu(u32, window_samples=window, min_points=min_points,
                                    normalise=False)
    np.testing.assert_array_equal(det[0], g[f"{name}_detrend"])
    full = gpu._preprocess_batch_gpu(u32, window_samples=window, min_points=min_points)
    # normalised bit for bit: the FP64 statistics are summed in sample order
    np.testing.assert_array_equal(full[0], g[f"{name}_preprocess"])
    # float32 input path
    full32 = gpu._preprocess_batch_gpu(x[None, :].copy(), window_samples=window,
                                       min_points=min_points)
    np.testing.assert_array_equal(full32, full)


def test_preprocess_batch_many_series(gpu, port):
    say("preprocess: 100 series x 2^16, window 3907/101")
    rng = np.random.default_rng(2)
    # two statistics workgroups of 64 series, the second one partial
    plane = rng.integers(30000, 31000, size=(100, 1 << 16)).astype(np.uint32)
    det = gpu._preprocess_batch_gpu(plane, window_samples=3907, min_points=101, normalise=False)
    full = gpu._preprocess_batch_gpu(plane, window_samples=3907, min_points=101)
    for i in (0, 11, 63, 64, 99):
        want = port.detrend(plane[i].astype(np.float32), 3907, 101)
        np.testing.assert_array_equal(det[i], want)
        np.testing.assert_array_equal(full[i], port.normalise(want))
    conv = gpu._preprocess_batch_gpu(plane, normalise=False)
    np.testing.assert_array_equal(conv, plane.astype(np.float32))


def test_preprocess_rejects_constant(gpu):
    say("preprocess: constant series rejected")
    plane = np.full((2, 5000), 7, np.uint32)
    with pytest.raises(ValueError, match="standard deviation"):
        gpu._preprocess_batch_gpu(plane)


# ---------------------------------------------------------------- FFA ----
def test_ffa_rows_and_search_vs_golden(gpu):
    from gaffa_amd import ffa

    say("ffa: golden series, rows of 3 tasks and the full search")
    g = np.load(os.path.join(GOLDEN, "ffa_search.npz"))
    plan = gpu.FfaPlan(g["tasks"], list(g["widths"]))
    series = g["series"]
    sel = [0, 37, len(g["tasks"]) - 1]
    got = np.concatenate([ffa.ffa_transform_rows(series[None, :], plan, k)[0].ravel()
                          for k in sel])
    np.testing.assert_array_equal(got, g["rows"])
    peaks = ffa.ffa_search(series, plan, snr_threshold=6.0, backend="cuda")
    want = g["peaks"]
    assert len(peaks) == len(want)
    for p, q in zip(peaks, want):
        assert (p.shift, p.phase, p.width, p.bins) == (q["shift"], q["phase"], q["width"], q["bins"])
        assert np.float32(p.snr) == q["snr"]
        assert p.period == q["period"] and p.frequency == q["frequency"]


@pytest.mark.parametrize("rows_bins", [(1, 5), (2, 7), (3, 4), (5, 4), (163, 181), (300, 181),
                                       (700, 256), (2679, 180), (2657, 180), (1203, 200),
                                       (1090, 256), (9000, 3), (5003, 2), (70000, 2)])
def test_ffa_transform_shapes_bit_exact(gpu, port, rows_bins):
    """Single-task plans over odd shapes: LDS-only trees (LDS depth 1),
    LDS depth 2 with the top two levels merged by the detection kernel
    (300 x 181), three global levels in one pass (700 x 256, 1203 x 200,
    1090 x 256, 5003 x 2), four global levels of which the detection kernel
    merges the top two (2679 x 180, 2657 x 180, 9000 x 3), seven levels
    (70000 x 2: depth 8, one three-level pass, a second, then two levels in
    the detection kernel), and level work items of 1, 2 and 3 rows
    (rows % 16 in {1, 2, 3})."""
    from oracle.oracle import TASK_DTYPE

    rows, bins = rows_bins
    say(f"ffa: transform rows={rows} bins={bins}")
    n = rows * bins + 3
    rng = np.random.default_rng(rows + bins)
    x = rng.normal(size=n).astype(np.float32)
    t = np.zeros(1, TASK_DTYPE)
    t[0] = (1.0, 1e-3, n, n, bins, rows, rows, bins * 1e-3, (bins + 1) * 1e-3)
    plan = gpu.FfaPlan(t, [1])
    got = gpu._ffa_transform_rows_gpu(x[None, :], plan, 0)[0]
    want = port.ffa_transform(x[: rows * bins], bins)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("rows_bins", [(167, 180), (2679, 256)])
def test_ffa_error_vs_fp64(gpu, port, rows_bins):
    """The FP32 transform (bit-exact to the CPU's) against the FP64 FFA on
    N(0,1) input: every output row is a sum of `rows` terms, and the error is
    stated relative to the largest output magnitude (tolerance 1e-5)."""
    from oracle.oracle import TASK_DTYPE

    rows, bins = rows_bins
    say(f"ffa: fp32 vs fp64 rows={rows} bins={bins}")
    n = rows * bins
    x = np.random.default_rng(rows).normal(size=n).astype(np.float32)
    t = np.zeros(1, TASK_DTYPE)
    t[0] = (1.0, 1e-3, n, n, bins, rows, rows, bins * 1e-3, (bins + 1) * 1e-3)
    got = gpu._ffa_transform_rows_gpu(x[None, :], gpu.FfaPlan(t, [1]), 0)[0].astype(np.float64)
    want = port.ffa_transform(x.astype(np.float64), bins)
    rel = float(np.abs(got - want).max() / np.abs(want).max())
    say(f"ffa: max |fp32 - fp64| / max |fp64| = {rel:.3e}")
    assert rel < 1e-5


def test_ffa_batch_search_bit_exact(gpu, port):
    """Several series, the small workload's plan (7 groups, 495 tasks)."""
    say("ffa: 6 series x 130870, small plan, threshold 5")
    rng = np.random.default_rng(8)
    n, tsamp = 130870, 256e-6
    batch = rng.normal(size=(6, n)).astype(np.float32)
    ph = (np.arange(n) * tsamp / 0.714) % 1.0
    batch[2] += (1.2 * np.exp(-0.5 * ((ph - 0.25) / 0.01) ** 2)).astype(np.float32)
    plan = gpu._make_riptide_ffa_plan(nsamples=n, tsamp=tsamp, period_min=0.2, period_max=2.0,
                                      duty_cycle_max=0.09, width_trial_spacing=2.0)
    got = gpu._ffa_search_batch_gpu(batch, plan, snr_threshold=5.0, series_tile=4)
    oplan = Plan(plan.tasks, np.array(plan.width_trials, np.uint64))
    for s in range(6):
        want = port.search_ffa(batch[s], oplan, 5.0)
        mine = [p for i, p in got if i == s]
        assert len(mine) == len(want), (s, len(mine), len(want))
        for p, q in zip(mine, want):
            assert (p.shift, p.phase, p.width, p.bins, np.float32(p.snr)) == \
                (q["shift"], q["phase"], q["width"], q["bins"], q["snr"])
    best = max((p for i, p in got if i == 2), key=lambda p: p.snr)
    assert abs(best.period - 0.714) < 0.003


@pytest.mark.parametrize("n", [(1 << 20) + 3, (1 << 22) - 1000])
def test_ffa_search_two_level_detection_bit_exact(gpu, port, n):
    """Search plans whose trees are deep enough that the detection kernel
    merges the top two levels itself: at 64 us, periods 0.1-0.45 s and bins
    180-256, 2^20 samples give LDS depth 2 for the second octave (no global
    pass at all) and 2^22 give depth 5 for the first (one three-level pass
    instead of 2 + 2), beside tasks that keep the one-level merge.  Every
    peak equals the oracle's."""
    say(f"ffa: two-level detection merge, n={n}")
    tsamp = 64e-6
    rng = np.random.default_rng(n % 1000)
    x = rng.normal(size=n).astype(np.float32)
    ph = (np.arange(n) * tsamp / 0.1703) % 1.0
    x += (0.6 * np.exp(-0.5 * ((ph - 0.4) / 0.01) ** 2)).astype(np.float32)
    plan = gpu._make_riptide_ffa_plan(nsamples=n, tsamp=tsamp, period_min=0.1, period_max=0.45,
                                      bins_min=180, bins_max=256, duty_cycle_max=0.18,
                                      width_trial_spacing=2.0)
    got = gpu._ffa_search_batch_gpu(x[None, :], plan, snr_threshold=6.0, series_tile=1)
    want = port.search_ffa(x, Plan(plan.tasks, np.array(plan.width_trials, np.uint64)), 6.0)
    assert len(got) == len(want) and len(want) > 0
    for (i, p), q in zip(got, want):
        assert i == 0
        assert (p.shift, p.phase, p.width, p.bins, np.float32(p.snr)) == \
            (q["shift"], q["phase"], q["width"], q["bins"], q["snr"])
    best = max((p for _, p in got), key=lambda p: p.snr)
    assert abs(best.period - 0.1703) < 0.001


def test_ffa_max_peaks_guard(gpu):
    say("ffa: max_peaks guard")
    from gaffa_amd import ffa

    x = np.random.default_rng(1).normal(size=1 << 14).astype(np.float32)
    plan = ffa.make_riptide_plan(nsamples=x.size, tsamp=1e-3, period_min=0.5, period_max=3.0)
    with pytest.raises(RuntimeError, match="max_peaks"):
        ffa.ffa_search(x, plan, snr_threshold=-1e6, max_peaks=3, backend="cuda")


# ------------------------------------------------------------ DM search ----
def test_dm_search_vs_golden(gpu):
    say("dm search: golden plane")
    g = np.load(os.path.join(GOLDEN, "dm_search.npz"))
    peaks = gpu._find_dm_peaks_gpu(g["plane"], list(g["dms"]), float(g["tsamp"][0]),
                                   period_min=0.4, period_max=2.0, duty_cycle_max=0.09,
                                   width_trial_spacing=2.0, window_samples=251, min_points=101,
                                   snr_threshold=6.0)
    want = g["peaks"]
    assert len(peaks) == len(want)
    for p, q in zip(peaks[:20], want[:20]):
        assert (p.dm_index, p.peak.shift, p.peak.width, p.peak.bins) == \
            (q["dm_index"], q["peak"]["shift"], q["peak"]["width"], q["peak"]["bins"])
        assert abs(p.peak.snr - q["peak"]["snr"]) <= 1e-5 * abs(q["peak"]["snr"])


def assert_dm_peaks_equal(got, want):
    """Complete DM peak lists (structured records): equal counts, identical
    DM index, period index, shift, phase, width index, width and bins in the
    same order, S/N within 1e-5 relative."""
    assert len(got) == len(want), (len(got), len(want))
    if len(want) == 0:
        return
    np.testing.assert_array_equal(got["dm_index"], want["dm_index"])
    for f in ("period_index", "shift", "phase", "width_index", "width", "bins"):
        np.testing.assert_array_equal(got["peak"][f], want["peak"][f], err_msg=f)
    np.testing.assert_allclose(got["peak"]["snr"], want["peak"]["snr"], rtol=1e-5, atol=0)
    np.testing.assert_array_equal(got["peak"]["period"], want["peak"]["period"])


CANDIDATE_FIELDS = ("dm", "dm_index", "period", "frequency", "width", "duty_cycle", "snr",
                    "peak_count", "dm_index_min", "dm_index_max", "frequency_min",
                    "frequency_max")
PEAK_FIELDS = ("period", "frequency", "width", "duty_cycle", "width_index", "period_index",
               "phase", "shift", "bins", "snr")


def assert_candidates_equal(got, want):
    """Candidate lists equal record by record, every field exact."""
    assert len(got) == len(want), (len(got), len(want))
    for k, (a, b) in enumerate(zip(got, want)):
        for f in CANDIDATE_FIELDS:
            assert getattr(a, f) == getattr(b, f), (k, f, getattr(a, f), getattr(b, f))
        pa, pb = a.best_peak, b.best_peak
        assert (pa.dm, pa.dm_index) == (pb.dm, pb.dm_index), k
        for f in PEAK_FIELDS:
            assert getattr(pa.peak, f) == getattr(pb.peak, f), (k, f)


def test_candidate_selection_device_equals_host(gpu):
    """gaffa's candidate selection on the device equals the host selection
    (the ranked path, itself equal to select_candidates_cpu) candidate for
    candidate, over frequency radii from 0 (every distinct trial frequency
    its own cluster) up, DM radii from 0 up, widths together or apart, and a
    max_candidates cut."""
    from gaffa_amd.pipeline import small_workload

    w = small_workload(nsamples=1 << 16)
    pipe = w.pipeline()
    pipe.generate(w.synth_spec())
    pipe.run()
    assert pipe.peak_count > 1000
    for radius in (0.0, 0.1, 2.0):
        for dm_radius in (0, 1, 50):
            for across in (True, False):
                for cut in (0, 7):
                    opts = dict(frequency_cluster_radius=radius, dm_cluster_radius=dm_radius,
                                cluster_across_widths=across, max_candidates=cut)
                    dev = pipe.select_candidates(**opts)
                    host = pipe.select_candidates(**opts, device=False)
                    assert len(dev) > 0
                    assert_candidates_equal(dev, host)
    # and the host ranked path equals gaffa's general order-based selection
    from gaffa_amd import _core as core

    general = core.select_candidates(pipe.peaks, pipe.out_nsamples * w.tsamp,
                                     frequency_cluster_radius=0.1, dm_cluster_radius=50)
    assert_candidates_equal(pipe.select_candidates(frequency_cluster_radius=0.1,
                                                   dm_cluster_radius=50), general)


@pytest.mark.parametrize("ffa_series_tile", [0, 7])
def test_small_pipeline_end_to_end(gpu, port, ffa_series_tile):
    """The 64-channel workload through the device-resident pipeline: the
    plane bit-exact and the complete peak list equal to the oracle's
    find_dm_peaks (with 7 series per FFA tile, ten tiles' sorted lists are
    merged)."""
    from gaffa_amd.pipeline import peak_records, small_workload

    say(f"pipeline: small workload end to end vs oracle, series tile {ffa_series_tile}")
    w = small_workload()
    pipe = w.pipeline(ffa_series_tile=ffa_series_tile)
    pipe.generate(w.synth_spec())
    pipe.dedisperse()
    plane = pipe.download_plane_u32()
    x = pipe.download_filterbank(w.nsamples, w.nchans)
    want_plane = port.dedisperse(x, w.frequencies, dm_low=0.0, dm_step=1.0, ndm=64, tsamp=w.tsamp)
    np.testing.assert_array_equal(plane, want_plane)
    pipe.preprocess()
    pipe.search()
    opts = port.plan_options(w.period_min, w.period_max, w.bins_min, w.bins_max, 1,
                             w.duty_cycle_max, w.width_trial_spacing)
    pp = port.preprocess_plan(window_samples=w.window_samples(), min_points=w.min_points)
    want = port.find_dm_peaks(plane, w.dms, w.tsamp, opts, pp, w.snr_threshold)
    got = peak_records(pipe)
    assert len(want) > 1000
    assert_dm_peaks_equal(got, want)
    # the injected pulsar (P 0.714 s, DM 30) is the top candidate
    top = pipe.peaks[0]
    assert abs(top.peak.period - 0.714) < 0.005 and abs(top.dm - 30.0) <= 3.0
    # the candidate stage of gaffa's DM-search benchmark on the same peaks,
    # against gaffa's own candidate code (the reference CPU build)
    from gaffa_amd.pipeline import CANDIDATE_OPTIONS, HARMONIC_OPTIONS, search_candidates
    from oracle.oracle import LIBS, Oracle

    if not os.path.exists(LIBS["reference"]):
        pytest.skip("reference CPU build (oracle/_ref) not present for the candidate stage")
    ref = Oracle("reference")
    cands = search_candidates(pipe, w)
    sel = ref.select_candidates(want, pipe.out_nsamples * w.tsamp, **CANDIDATE_OPTIONS)
    f = w.frequencies
    flagged, rel = ref.flag_harmonics(sel, w.nsamples * w.tsamp, f.min(), f.max(),
                                      **HARMONIC_OPTIONS)
    kept = flagged[rel["is_harmonic"] == 0]
    assert 0 < len(kept) < len(sel)
    assert [(c.dm_index, c.period, c.width, np.float32(c.snr), c.peak_count) for c in cands] == \
        [(int(k["dm_index"]), float(k["period"]), int(k["width"]), k["snr"], int(k["peak_count"]))
         for k in kept]


def _full_length_group_vs_oracle(port, w, segment: int, group: int):
    """Run the whole pipeline at the workload's full length on the GPU, then
    compare the complete peak list of one 32-trial nominal group with the
    oracle's subband dedispersion of that group's segment and find_dm_peaks
    of the group's rows, on the same filterbank bytes."""
    from gaffa_amd.pipeline import peak_records

    pipe = w.pipeline()
    pipe.generate(w.synth_spec())
    pipe.run()
    # the device candidate selection over every peak of the full search
    # equals the host's, at gaffa's benchmark options and with widths apart
    import time

    for opts in (dict(frequency_cluster_radius=0.1, dm_cluster_radius=50,
                      cluster_across_widths=True),
                 dict(frequency_cluster_radius=0.3, dm_cluster_radius=2,
                      cluster_across_widths=False)):
        t0 = time.perf_counter()
        dev = pipe.select_candidates(**opts)
        t1 = time.perf_counter()
        host = pipe.select_candidates(**opts, device=False)
        t2 = time.perf_counter()
        say(f"  candidates of {pipe.peak_count} peaks {opts}: {len(dev)} "
            f"(device {1e3 * (t1 - t0):.0f} ms, host {1e3 * (t2 - t1):.0f} ms)")
        assert_candidates_equal(dev, host)
    first = sum(s.ndm for s in w.dm_segments[:segment]) + group * w.ndm_per_nominal
    got = peak_records(pipe, first, first + w.ndm_per_nominal).copy()
    got["dm_index"] -= first
    x = pipe.download_filterbank(w.nsamples, w.nchans)
    n_out = pipe.out_nsamples  # set by the largest delay over all trials
    del pipe
    seg = w.dm_segments[segment]
    ndm = (group + 1) * w.ndm_per_nominal  # the segment's trials up to the group
    say(f"  oracle: subband dedispersion of {ndm} trials at DM {seg.dm_low:.2f}+"
        f"{seg.dm_step:.4f}k, FFA of trials {group * w.ndm_per_nominal}..{ndm - 1}")
    plane = port.dedisperse(x, w.frequencies, dm_low=seg.dm_low, dm_step=seg.dm_step, ndm=ndm,
                            tsamp=w.tsamp, subband=True, subband_channels=w.subband_channels,
                            ndm_per_nominal=w.ndm_per_nominal)
    del x
    # a sample's sum does not depend on the output length: the first n_out
    # samples of the oracle's rows are the pipeline's series
    assert plane.shape[1] >= n_out
    rows = np.ascontiguousarray(plane[group * w.ndm_per_nominal:, :n_out])
    del plane
    dms = seg.dm_low + seg.dm_step * np.arange(group * w.ndm_per_nominal, ndm)
    opts = port.plan_options(w.period_min, w.period_max, w.bins_min, w.bins_max, 1,
                             w.duty_cycle_max, w.width_trial_spacing)
    pp = port.preprocess_plan(window_samples=w.window_samples(), min_points=w.min_points)
    want = port.find_dm_peaks(rows, dms, w.tsamp, opts, pp, w.snr_threshold)
    say(f"  {len(got)} GPU peaks, {len(want)} oracle peaks")
    assert len(want) > 1000
    assert_dm_peaks_equal(got, want)


def test_lband_full_length_first_group_vs_oracle(gpu, port):
    """1024 channels x 2^22 samples, 1024 DM trials: the complete peak list of
    the first nominal group (DM trials 0..31) equals the oracle's."""
    from gaffa_amd.pipeline import lband_workload

    say("pipeline: lband at 2^22 samples, first nominal group vs oracle")
    _full_length_group_vs_oracle(port, lband_workload(), segment=0, group=0)


def test_large_full_length_last_group_vs_oracle(gpu, port):
    """4096 channels x 2^22 samples with RFI, 4096 DM trials up to DM ~1000:
    the complete peak list of the last nominal group (32 trials per group,
    as configured) equals the oracle's."""
    from gaffa_amd.pipeline import large_workload

    w = large_workload()
    say("pipeline: large at 2^22 samples, last nominal group vs oracle")
    last = w.dm_segments[-1].ndm // w.ndm_per_nominal - 1
    _full_length_group_vs_oracle(port, w, segment=len(w.dm_segments) - 1, group=last)


def _check_series_peaks(got, want, s):
    mine = [p for i, p in got if i == s]
    assert len(mine) == len(want), (s, len(mine), len(want))
    for p, q in zip(mine, want):
        assert (p.shift, p.phase, p.width, p.bins, np.float32(p.snr)) == \
            (q["shift"], q["phase"], q["width"], q["bins"], q["snr"])


def test_ffa_batched_heterogeneous_bit_exact(gpu, port):
    """Seven series of different lengths, each with its own octave plan, in
    uneven tiles of 3: one launch sequence per tile, replayed from a HIP graph,
    and the per-series launch loop, all equal to the oracle's search of the
    same bytes."""
    from gaffa_amd.pipeline import batched_workload

    say("ffa batch: 7 heterogeneous series, tiles of 3, graph + loop")
    w = batched_workload(7, seed=5, log2_min=15, log2_max=17, tsamp=256e-6, injected=0.5)
    plans = w.plans()
    b = gpu.FfaBatchedSearch(plans, series_per_tile=3, workspace_bytes=1 << 30)
    b.generate(w.seed, w.tsamp, list(w.period), list(w.phase0), list(w.sigma),
               list(w.amplitude))
    row = b.download_row()
    offs, lens = list(b.offsets), list(b.lengths)
    for s in range(7):
        x = row[offs[s]: offs[s] + lens[s]]
        assert abs(float(x.mean())) < 0.05 and abs(float(x.std()) - 1.0) < 0.2
        pad_end = offs[s + 1] if s + 1 < 7 else b.row_length
        assert not row[offs[s] + lens[s]: pad_end].any()
    first, timing = b.run(6.0)
    assert timing["tiles"] == 3
    replay, _ = b.run(6.0)
    loop, t_loop = b.run(6.0, per_series=True)
    assert t_loop["tiles"] == 7
    for s in range(7):
        x = row[offs[s]: offs[s] + lens[s]]
        want = port.search_ffa(x, Plan(plans[s].tasks, np.array(plans[s].width_trials,
                                                              np.uint64)), 6.0)
        for got in (first, replay, loop):
            _check_series_peaks(got, want, s)
    say(f"ffa batch: batched {timing['gpu_ms']:.2f} ms, loop {t_loop['gpu_ms']:.2f} ms")


def test_ffa_batched_full_length_slice_bit_exact(gpu, port):
    """Sixteen series of the batched workload's own shape: 2^20 to 2^22
    samples at 64 us, each with its own octave plan (840 tasks), half with
    injected pulse trains, in uneven tiles of 6; the batched graphs' and the
    per-series loop's peak lists equal the oracle's search of the same bytes,
    peak for peak."""
    from gaffa_amd.pipeline import batched_workload

    n = 16
    say(f"ffa batch: {n} series of 2^20..2^22 at 64 us, tiles of 6, vs oracle")
    w = batched_workload(n, seed=23, log2_min=20, log2_max=22, tsamp=64e-6, injected=0.5)
    plans = w.plans()
    b = gpu.FfaBatchedSearch(plans, series_per_tile=6, workspace_bytes=8 << 30)
    b.generate(w.seed, w.tsamp, list(w.period), list(w.phase0), list(w.sigma),
               list(w.amplitude))
    row = b.download_row()
    offs, lens = list(b.offsets), list(b.lengths)
    assert min(lens) >= 1 << 20 and max(lens) <= 1 << 22
    first, timing = b.run(6.0)
    assert timing["tiles"] == 3
    loop, _ = b.run(6.0, per_series=True)
    total = 0
    for s in range(n):
        x = row[offs[s]: offs[s] + lens[s]]
        want = port.search_ffa(x, Plan(plans[s].tasks, np.array(plans[s].width_trials,
                                                              np.uint64)), 6.0)
        total += len(want)
        for got in (first, loop):
            _check_series_peaks(got, want, s)
    assert total > 0
    say(f"ffa batch: {total} peaks equal, batched {timing['gpu_ms']:.1f} ms")


def test_ffa_batched_autotune_keeps_results(gpu):
    """The autotuner's knobs (series per tile, rows per level item, rows per
    terminal item) change the time only: every configuration, including odd
    row counts that split tasks unevenly, returns the same peaks; the chosen
    configuration is cached for a second batch of the same shape."""
    from gaffa_amd.pipeline import batched_workload

    say("ffa batch: autotune over tiles and rows per level / terminal item")
    w = batched_workload(5, seed=9, log2_min=15, log2_max=16, tsamp=256e-6, injected=1.0)
    b = w.search(series_per_tile=5, workspace_bytes=1 << 30)

    def key(peaks):
        return [(i, p.shift, p.phase, p.width, p.bins, p.snr) for i, p in peaks]

    before = key(b.run(6.0)[0])
    assert len(before) > 0
    for level_rows, terminal_rows in ((8, 16), (32, 32), (5, 7), (1, 1), (32, 3)):
        b.set_item_shape(level_rows=level_rows, terminal_rows=terminal_rows)
        assert key(b.run(6.0)[0]) == before, (level_rows, terminal_rows)
    b.set_item_shape()
    best = b.autotune([1, 2, 4])
    t = b.tuning
    assert best in (1, 2, 4) and b.series_per_tile == best and not t["cached"]
    assert key(b.run(6.0)[0]) == before
    again = w.search(series_per_tile=5, workspace_bytes=1 << 30)
    assert again.autotune([1, 2, 4]) == best and again.tuning["cached"]
    assert again.tuning["level_rows"] == t["level_rows"]
    assert key(again.run(6.0)[0]) == before


# ------------------------------------------------------------ folding ----
@pytest.mark.parametrize("dtype,nchans,outc", [(np.uint8, 16, 0), (np.uint16, 12, 4),
                                               (np.float32, 8, 2)])
def test_fold_spectrum_bit_exact(gpu, ref, dtype, nchans, outc):
    """Subint cubes equal gaffa's fold_spectrum_cpu bit for bit (same FP64
    order per bin); exposure equal; projections equal."""
    from gaffa_amd import pfold

    say(f"fold: spectrum {np.dtype(dtype).name} x {nchans} ch -> {outc or nchans}")
    rng = np.random.default_rng(nchans)
    ns, tsamp, period, nbin = 50000, 64e-6, 0.123456789, 64
    if dtype == np.float32:
        x = rng.normal(size=(ns, nchans)).astype(np.float32)
    else:
        x = rng.integers(0, 200, size=(ns, nchans)).astype(dtype)

    class Spec:
        data, tsamp = x, 64e-6

    r = pfold.fold_spectrum(Spec, period=period, nbin=nbin, tsubint=0.4,
                            nsubband=outc or None)
    sub = round(0.4 / tsamp)
    cube, expo = ref.fold_spectrum(x, period, tsamp, nbin, sub, output_channels=outc)
    np.testing.assert_array_equal(r.cube, cube)
    np.testing.assert_array_equal(r.exposure, expo)
    assert r.tsubint == sub * tsamp and r.nsubint == cube.shape[0]


def test_fold_profile_and_candidates(gpu, ref):
    """Profiles against gaffa's fold_time_series_cpu (its FP64 sum follows its
    OpenMP thread split, so within 1e-6), and pipeline candidate cubes equal a
    spectrum fold of the same row."""
    from gaffa_amd import pfold
    from gaffa_amd.pipeline import small_workload

    say("fold: profile + candidate cubes on the small pipeline")
    rng = np.random.default_rng(4)
    x = rng.integers(0, 5000, size=(3, 40000)).astype(np.uint32)

    class Ded:
        data, tsamp = x, 256e-6

    p = pfold.fold_profile(Ded, period=0.7137, nbin=100, dm_index=2)
    prof, expo = ref.fold_time_series(x[2], 0.7137, 256e-6, 100)
    np.testing.assert_allclose(p.profile, prof, rtol=1e-6)
    np.testing.assert_allclose(p.exposure, expo, rtol=1e-12)
    w = small_workload(nsamples=1 << 16)
    pipe = w.pipeline()
    pipe.generate(w.synth_spec())
    pipe.run()
    plane = pipe.download_plane_f32()
    cubes, expos = pipe.fold_candidates([30, 5], [0.714, 0.5], nbin=64, nsubint=32)
    assert cubes.shape == (2, 32, 64)
    n = pipe.out_nsamples
    sub = -(-n // 32)
    for j, (row, per) in enumerate(((30, 0.714), (5, 0.5))):
        c, e = ref.fold_spectrum(plane[row][:, None], per, w.tsamp, 64, sub)
        np.testing.assert_array_equal(cubes[j], c[:, 0, :])
        np.testing.assert_array_equal(expos[j], e)
    # the injected pulsar (P 0.714 s, DM 30) folds into a clear peak
    prof = cubes[0].mean(axis=0)
    assert prof.max() - np.median(prof) > 5 * prof.std() / 3


@pytest.mark.parametrize("name", ["lband", "small"])
def test_full_fold_of_candidates_vs_reference(gpu, ref, name):
    """gaffa's full fold of search candidates (aligned spectrum at the
    candidate's DM, 32 subints x 64 output channels x 64 bins): subband
    workloads fold from per-output-channel sums of the channel-major copy, the
    brute-force one from the aligned spectrum; both give gaffa's pfold cube
    (dedisperse_spectrum + fold_spectrum of the reference build) bit for bit."""
    from gaffa_amd.pipeline import lband_workload, small_workload

    say(f"fold: full folds of {name} candidates vs gaffa's pfold")
    w = lband_workload(nsamples=1 << 18, ndm=256) if name == "lband" else \
        small_workload(nsamples=1 << 16)
    pipe = w.pipeline()
    pipe.generate(w.synth_spec())
    pipe.run()
    cands = pipe.select_candidates(dm_cluster_radius=50)[:3]
    assert len(cands) == 3
    _, cubes = pipe.fold_candidates_full(cands, keep_cubes=True)
    x = pipe.download_filterbank(w.nsamples, w.nchans)
    for c, got in zip(cands, cubes):
        spec = ref.dedisperse_spectrum(x, w.frequencies, c.dm, w.tsamp)
        sub = -(-spec.shape[0] // 32)
        want, _ = ref.fold_spectrum(spec, c.period, w.tsamp, 64, sub, output_channels=64)
        np.testing.assert_array_equal(got, want)


# ------------------------------------------------------------ spectrum ----
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
def test_spectrum_bit_exact(gpu, port, dtype):
    say(f"spectrum: 1024 ch x 12000, DM 250, {np.dtype(dtype).name}")
    rng = np.random.default_rng(2)
    freq = channel_frequencies(1024, 1520.0, -0.293)
    x = (rng.integers(0, 250, size=(12000, 1024)) if dtype != np.float32
         else rng.normal(size=(12000, 1024))).astype(dtype)
    got = gpu._dedisperse_spectrum_array(x, frequency_mhz=list(freq), tsamp=64e-6, dm=250.0)
    want = port.dedisperse_spectrum(x, freq, 250.0, 64e-6)
    assert got.data.shape == want.shape and got.data.dtype == want.dtype
    np.testing.assert_array_equal(got.data, want)
    sub = gpu._dedisperse_spectrum_array(x, frequency_mhz=list(freq), tsamp=64e-6, dm=250.0,
                                         chan_begin=100, chan_end=612)
    np.testing.assert_array_equal(
        sub.data, port.dedisperse_spectrum(x, freq, 250.0, 64e-6, ref=freq.max(), cb=100,
                                           ce=612))
    # 513 channels fill no whole 32-bit words for 8/16-bit data: the staged path
    odd = gpu._dedisperse_spectrum_array(x, frequency_mhz=list(freq), tsamp=64e-6, dm=250.0,
                                         chan_begin=100, chan_end=613)
    np.testing.assert_array_equal(
        odd.data, port.dedisperse_spectrum(x, freq, 250.0, 64e-6, ref=freq.max(), cb=100,
                                           ce=613))


def test_spectrum_filterbank_and_fold(gpu, ref):
    from gaffa_amd import dedispersion, pfold
    from gaffa_amd.io import Filterbank

    say("spectrum: basetest.fil DM 10 + fold")
    fb = Filterbank(os.path.join(GOLDEN, "basetest.fil"))
    spec = dedispersion.dedisperse_spectrum(fb, dm=10.0, backend="cuda")
    want = ref.dedisperse_spectrum(fb.data[:, 0, :], fb.header.frequency_table, 10.0,
                                   fb.header.tsamp)
    np.testing.assert_array_equal(spec.data, want)
    assert spec.backend == "cuda" and spec.dm == 10.0
    r = pfold.fold_spectrum(spec, period=0.25, nbin=32, tsubint=0.5, nsubband=4)
    cube, expo = ref.fold_spectrum(want, 0.25, fb.header.tsamp, 32,
                                   round(0.5 / fb.header.tsamp), output_channels=4)
    np.testing.assert_array_equal(r.cube, cube)
    np.testing.assert_array_equal(r.exposure, expo)


@pytest.mark.parametrize("which,descending,chunk", [
    ("small", True, 1 << 14), ("small", False, 5000), ("lband", True, 1 << 15)])
def test_streamed_file_equals_resident(gpu, tmp_path, which, descending, chunk):
    """run_file streams a SIGPROC file to the device in pinned chunks and
    dedisperses each output range as soon as its lookahead has arrived; the
    searched plane and the peaks must equal the resident run() on the same
    samples, for files stored in either channel order."""
    from gaffa_amd.io import write_filterbank
    from gaffa_amd.pipeline import lband_workload, small_workload

    w = small_workload() if which == "small" else lband_workload(nsamples=1 << 19, npulsars=3)
    say(f"streaming: {which} {w.nchans}ch x {w.nsamples} descending={descending} chunk={chunk}")
    ref = w.pipeline(ffa_workspace_bytes=4 << 30)
    ref.generate(w.synth_spec())
    x = ref.download_filterbank(w.nsamples, w.nchans)
    ref.run()
    want_plane = ref.download_plane_f32()
    want = ref.peaks
    path = tmp_path / "obs.fil"
    if descending:  # the workloads' own (fch1, foff < 0) header
        write_filterbank(path, x[:, ::-1], tsamp=w.tsamp, fch1=w.fch1, foff=w.foff,
                         source_name="synthetic")
    else:
        write_filterbank(path, x, tsamp=w.tsamp, fch1=float(w.frequencies[0]), foff=-w.foff,
                         source_name="synthetic")
    del ref
    pipe = w.pipeline(ffa_workspace_bytes=4 << 30)
    st = pipe.run_file(str(path), chunk_samples=chunk)
    assert st["stream_bytes"] == w.nsamples * w.nchans
    assert st["stream_chunks"] == -(-w.nsamples // max(256, -(-chunk // 256) * 256))
    np.testing.assert_array_equal(pipe.download_filterbank(w.nsamples, w.nchans), x)
    np.testing.assert_array_equal(pipe.download_plane_f32(), want_plane)
    got = pipe.peaks
    assert len(got) == len(want) > 0
    for p, q in zip(got, want):
        assert (p.dm_index, p.peak.shift, p.peak.width, p.peak.bins, p.peak.snr) == \
            (q.dm_index, q.peak.shift, q.peak.width, q.peak.bins, q.peak.snr)

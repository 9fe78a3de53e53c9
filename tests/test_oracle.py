"""Pins the CPU oracle: reference known answers, the reference CPU build,
closed forms and the committed golden fixtures (CPU only)."""

import hashlib
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle.oracle import PEAK_DTYPE, Oracle, Plan, channel_frequencies

# ------------------------------------------------------------------ FFA ----
# riptide's hand-calculated transforms (test_ffa.cpp:44-105)
FFA_IN_8 = np.zeros((8, 8), np.float32)
FFA_IN_8[:, 7] = 1
FFA_OUT_8 = np.array([
    [0, 0, 0, 0, 0, 0, 0, 8], [0, 0, 0, 0, 0, 0, 4, 4], [0, 0, 0, 0, 0, 2, 4, 2],
    [0, 0, 0, 0, 2, 2, 2, 2], [0, 0, 0, 1, 2, 2, 2, 1], [0, 0, 1, 2, 1, 1, 2, 1],
    [0, 1, 1, 1, 2, 1, 1, 1], [1, 1, 1, 1, 1, 1, 1, 1]], np.float32)
ODD_IN = np.arange(1, 21, dtype=np.float32)
ODD_OUT = np.array([[45, 50, 55, 60], [46, 51, 56, 57], [50, 55, 56, 49], [52, 53, 54, 51],
                    [51, 52, 53, 54]], np.float32)


def test_ffa_transform_golden(port):
    np.testing.assert_array_equal(port.ffa_transform(FFA_IN_8.ravel(), 8), FFA_OUT_8)
    np.testing.assert_array_equal(port.ffa_transform(ODD_IN, 4), ODD_OUT)


def test_ffa_transform_rotation_and_padding(port):
    for shift in range(8):
        np.testing.assert_array_equal(
            port.ffa_transform(np.roll(FFA_IN_8, shift, axis=1).ravel(), 8),
            np.roll(FFA_OUT_8, shift, axis=1))
    for extra in range(1, 8):
        x = np.zeros((8, 8 + extra), np.float32)
        x[:, :8] = FFA_IN_8
        y = np.zeros_like(x)
        y[:, :8] = FFA_OUT_8
        np.testing.assert_array_equal(port.ffa_transform(x.ravel(), 8 + extra), y)


@pytest.mark.parametrize("rows,bins", [(1, 5), (2, 7), (3, 4), (13, 17), (64, 180), (257, 256)])
def test_ffa_transform_row0_is_plain_sum_and_f64_agrees(port, rows, bins):
    # Row 0 never shifts (h(0) = t(0) = 0 at every level): the column sums.
    rng = np.random.default_rng(rows * 1000 + bins)
    x = rng.integers(-20, 21, size=(rows, bins)).astype(np.float32)
    out = port.ffa_transform(x.ravel(), bins)
    np.testing.assert_array_equal(out[0], x.sum(axis=0))
    # integer-valued inputs make FP32 sums exact, so FP64 must agree exactly
    out64 = port.ffa_transform(x.astype(np.float64).ravel(), bins)
    np.testing.assert_array_equal(out64, out.astype(np.float64))
    # every row is a sum over all input rows: totals are preserved
    np.testing.assert_array_equal(out.sum(axis=1), np.full(rows, x.sum(), np.float32))


def test_ffa_transform_matches_reference_build(port, ref):
    rng = np.random.default_rng(3)
    for rows, bins in [(5, 4), (100, 181), (333, 256), (1000, 199)]:
        x = rng.normal(size=rows * bins).astype(np.float32)
        np.testing.assert_array_equal(port.ffa_transform(x, bins), ref.ffa_transform(x, bins))


def test_plan_known_answers(port):
    # test_ffa_plan.cpp:78-116
    p = port.make_plan(1000, 1.0, port.plan_options(240.0, 300.0, 240, 260, 2, 0.20, 1.5, 1000))
    assert list(p.widths) == [1, 2, 3, 4, 6, 9, 13, 19, 28, 42]
    p = port.make_plan(64, 0.25, port.plan_options(1.0, 4.0, 4, 5, 2, 0.5, 1.5, 1000))
    assert len(p.tasks) == 7 and list(p.widths) == [1, 2]
    first, last = p.tasks[0], p.tasks[-1]
    assert (first["bins"], first["rows"], first["rows_eval"]) == (4, 16, 12)
    assert first["period_begin"] == 1.0
    assert abs(first["period_end"] - 1.2244897959183674) < 1e-12
    assert (last["bins"], last["rows"], last["rows_eval"]) == (4, 4, 2)
    assert last["period_begin"] == 3.375
    assert abs(last["period_end"] - 3.681818181818182) < 1e-12


def test_plan_errors(port):
    bad = port.plan_options(1.0, 4.0, 4, 5, 2, 0.5, 1.5, 1)
    with pytest.raises(RuntimeError, match="max_tasks"):
        port.make_plan(64, 0.25, bad)
    with pytest.raises(ValueError, match="period_min must be >= tsamp"):
        port.make_plan(64, 0.25, port.plan_options(0.5, 4.0, 4, 5, 2, 0.5, 1.5, 1000))


def test_plans_match_golden(port):
    g = np.load(os.path.join(GOLDEN, "plans.npz"))
    for name in ("small", "lband", "large", "riptide_default"):
        n, ts, pmin, pmax, duty, spacing = g[f"{name}_args"]
        p = port.make_plan(int(n), ts, port.plan_options(pmin, pmax, 180, 256, 1, duty, spacing))
        np.testing.assert_array_equal(p.tasks, g[f"{name}_tasks"])
        np.testing.assert_array_equal(p.widths, g[f"{name}_widths"])
    small = g["small_tasks"]
    assert len(small) == 495 and len(set(small["downsample_factor"])) == 7
    assert len(g["lband_tasks"]) == 995


# --------------------------------------------------------- time series ----
def test_downsample_known_answers(port):
    # test_time_series.cpp:93-153
    assert port.downsampled_size(10, 2.0) == 5
    assert port.downsampled_size(10, 3.0) == 3
    assert port.downsampled_size(10, 2.5) == 4
    np.testing.assert_array_equal(port.downsample([1, 2, 3, 4], 2.0), [3, 7])
    np.testing.assert_allclose(port.downsample([1, 2, 3, 4, 5], 2.5), [4.5, 10.5], rtol=1e-7)
    np.testing.assert_allclose(port.downsample([1, 2, 3, 4, 5, 6], 2.25),
                               [1 + 2 + 0.25 * 3, 0.75 * 3 + 4 + 0.5 * 5], rtol=1e-7)
    assert abs(port.downsampled_variance(16, 2.0) - 2.0) < 1e-12
    assert abs(port.downsampled_variance(16, 4.0) - 10.0) < 1e-12
    assert abs(port.downsampled_variance(8, 2.5) - (2.5 - 1 / 3)) < 1e-12
    assert abs(port.downsampled_variance(5, 2.5) - 5 / 3) < 1e-12
    for n, f in [(0, 2.0), (4, 1.0), (4, 0.5), (4, 5.0), (4, math.inf)]:
        with pytest.raises(ValueError):
            port.downsampled_size(n, f)


def test_downsample_matches_reference(port, ref):
    rng = np.random.default_rng(4)
    x = rng.normal(size=100000).astype(np.float32)
    for f in (8.680555555555555, 12.394, 104.99, 622.98):
        np.testing.assert_array_equal(port.downsample(x, f), ref.downsample(x, f))


def test_normalise_constant(port):
    np.testing.assert_array_equal(port.normalise([2, 2, 2], reject_constant=False), [0, 0, 0])
    with pytest.raises(ValueError, match="standard deviation"):
        port.normalise([2, 2, 2])


# ------------------------------------------------------- preprocessing ----
def test_running_median_known_answers(port):
    # test_preprocessing.cpp:9-162
    assert port.running_median_window(5.0, 0.5) == 11
    assert port.running_median_window(1.25, 0.5) == 3
    assert port.running_median_window(0.1, 1.0) == 1
    np.testing.assert_array_equal(port.detrend([100, 0, 1, 2, 3], 3), [0, -1, 0, 0, 0])
    x = np.array([1, 1, 1, 9, 9, 9] * 2, np.float32)
    out = port.detrend(x, 7, 3)
    np.testing.assert_array_equal(out[[0, 1, 2, 3, 9, 10, 11]], [0, -1, -3, 4, 3, 1, 0])
    np.testing.assert_array_equal(port.detrend([10, 10, 20, 10, 10], 3), [0, 0, 10, 0, 0])
    with pytest.raises(ValueError, match="low-resolution"):
        port.detrend(np.arange(1, 11, dtype=np.float32), 11, 5)
    with pytest.raises(ValueError, match="finite"):
        port.detrend([1, np.nan, 3], 3)


def test_preprocess_matches_golden(port):
    g = np.load(os.path.join(GOLDEN, "preprocess.npz"))
    for name in ("small", "lband", "exact"):
        window, min_points = g[f"{name}_window"]
        x = g[f"{name}_input"]
        np.testing.assert_array_equal(port.detrend(x, window, min_points), g[f"{name}_detrend"])
        plan = port.preprocess_plan(window_samples=int(window), min_points=int(min_points))
        np.testing.assert_array_equal(port.preprocess(x, 1e-3, plan), g[f"{name}_preprocess"])


# ------------------------------------------------------- dedispersion ----
def test_dedispersion_known_answers(port):
    # test_dedispersion.cpp:94-314 (ref frequency 2000, tsamp 0.003111606)
    s = np.array([1, 2, 3, 4, 5, 6], np.uint8).reshape(2, 3)
    ts = 0.003111606
    out = port.dedisperse_single(s, [1000.0, 1200.0, 1500.0], dm=0.0, ref=2000.0, tsamp=ts)
    np.testing.assert_array_equal(out, [6, 15])
    out = port.dedisperse_single(s, [1000.0, 2000.0, 2000.0], dm=1.0, ref=2000.0, tsamp=ts)
    np.testing.assert_array_equal(out, [9])
    s8 = np.arange(1, 9, dtype=np.uint8).reshape(2, 4)
    out = port.dedisperse(s8, [1000.0, 1200.0, 1500.0, 2000.0], dm_low=0, dm_step=1, ndm=1,
                          ref=2000.0, tsamp=ts, subband=True, subband_channels=2,
                          ndm_per_nominal=32)
    np.testing.assert_array_equal(out, [[10, 26]])
    a = port.dedisperse(s, [1000.0, 2000.0, 2000.0], dm_low=0, dm_step=1, ndm=2, ref=2000.0,
                        tsamp=ts)
    b = port.dedisperse(s, [1000.0, 2000.0, 2000.0], dm_low=0, dm_step=1, ndm=2, ref=2000.0,
                        tsamp=ts, subband=True, subband_channels=1, ndm_per_nominal=1)
    np.testing.assert_array_equal(a, b)


def test_delay_formula_closed_form(port):
    # round(4148.808 * DM * (f^-2 - ref^-2) / tsamp), half away from zero
    for dm, f, ref_f, ts in [(30.0, 1250.0, 1500.0, 256e-6), (500.0, 1220.0, 1520.0, 64e-6)]:
        exact = 4148.808 * dm * (1 / f**2 - 1 / ref_f**2) / ts
        assert port.delay_bins(dm, f, ref_f, ts) == int(math.floor(exact + 0.5))


def test_basetest_matches_golden(port):
    from gaffa_amd.io import Filterbank

    fb = Filterbank(os.path.join(GOLDEN, "basetest.fil"))
    assert (fb.header.nchans, fb.header.nbits, fb.header.nsamples) == (128, 8, 8192)
    assert fb.header.frequency_table[0] < fb.header.frequency_table[-1]
    g = np.load(os.path.join(GOLDEN, "basetest_dedispersion.npz"))
    x = fb.data[:, 0, :]
    freq = np.array(fb.header.frequency_table)
    np.testing.assert_array_equal(freq, g["frequency"])
    multi = port.dedisperse(x, freq, dm_low=0.0, dm_step=1.0, ndm=50, tsamp=fb.header.tsamp)
    np.testing.assert_array_equal(multi, g["multi"])
    sub = port.dedisperse(x, freq, dm_low=0.0, dm_step=1.0, ndm=50, tsamp=fb.header.tsamp,
                          subband=True, subband_channels=16, ndm_per_nominal=8)
    np.testing.assert_array_equal(sub, g["subband"])
    # DM 0 is the plain channel sum
    np.testing.assert_array_equal(multi[0], x[: multi.shape[1]].astype(np.uint32).sum(axis=1))


def test_subband_slice_matches_golden(port):
    g = np.load(os.path.join(GOLDEN, "subband_slice.npz"))
    rng = np.random.default_rng(7)
    x = rng.integers(96, 160, size=(20000, 1024), dtype=np.uint8)
    assert hashlib.sha256(x.tobytes()).hexdigest() == str(g["samples_sha256"])
    for i, (lo, step, ndm) in enumerate(g["segments"]):
        out = port.dedisperse(x, g["frequency"], dm_low=lo, dm_step=step, ndm=int(ndm),
                              tsamp=64e-6, subband=True, subband_channels=32, ndm_per_nominal=32)
        assert list(out.shape) == list(g[f"subband_{i}_shape"])
        digests = [hashlib.sha256(r.tobytes()).hexdigest() for r in out]
        assert digests == list(g[f"subband_{i}_digests"])


# ------------------------------------------------- detection / search ----
def closed_form_snr(bins, width, best, total, stdnoise):
    b, w = np.float32(bins), np.float32(width)
    h = np.sqrt((b - w) / (b * w), dtype=np.float32)
    base = np.float32(w / (b - w) * h)
    return np.float32((np.float32(h + base) * np.float32(best) - base * np.float32(total))
                      / np.float32(stdnoise))


def _task(rows, bins, rows_eval=None):
    from oracle.oracle import TASK_DTYPE

    t = np.zeros(1, TASK_DTYPE)
    t[0] = (1.0, 0.25, rows * bins, rows * bins, bins, rows, rows_eval or rows, 0.25 * bins,
            0.25 * (bins + 1))
    return t


def test_detection_closed_form(port):
    # test_ffa_detection.cpp:134-213
    p = port.find_ffa_peaks(np.array([[0, 0, 5, 0]], np.float32), _task(1, 4), [1], 1.0, 0.0)
    assert len(p) == 1 and p[0]["phase"] == 2 and p[0]["period"] == 1.0
    assert abs(p[0]["snr"] - closed_form_snr(4, 1, 5, 5, 1)) <= 4 * np.spacing(np.float32(p[0]["snr"]))
    p = port.find_ffa_peaks(np.array([[4, 0, 0, 5]], np.float32), _task(1, 4), [2], 1.0, 0.0)
    assert p[0]["phase"] == 3
    p = port.find_ffa_peaks(np.array([[1, 1, 0, 1, 1, 0]], np.float32), _task(1, 6), [2], 1.0, 0.0)
    assert p[0]["phase"] == 0
    tr = np.zeros((4, 4), np.float32)
    tr[3, 1] = 10
    task = _task(4, 4)
    task["rows"] = 8
    p = port.find_ffa_peaks(tr, task, [1], 1.0, 0.0)
    p3 = p[p["shift"] == 3][0]
    assert p3["period"] == 0.25 * 4 * 4 / (4 - 3 / 7)


def test_search_matches_golden_and_reference(port, ref):
    g = np.load(os.path.join(GOLDEN, "ffa_search.npz"))
    plan = Plan(g["tasks"], g["widths"])
    peaks = port.search_ffa(g["series"], plan, 6.0)
    np.testing.assert_array_equal(peaks, g["peaks"])
    assert len(peaks) > 10 and abs(peaks[0]["period"] - 0.7311) < 0.002
    rows = port.ffa_rows(g["series"], g["tasks"][[0, 37, len(g["tasks"]) - 1]])
    np.testing.assert_array_equal(rows, g["rows"])
    np.testing.assert_array_equal(ref.search_ffa(g["series"], plan, 6.0), peaks)


def test_dm_search_matches_golden(port):
    g = np.load(os.path.join(GOLDEN, "dm_search.npz"))
    opts = port.plan_options(0.4, 2.0, 180, 256, 1, 0.09, 2.0)
    pp = port.preprocess_plan(window_samples=251, min_points=101)
    peaks = port.find_dm_peaks(g["plane"], g["dms"], float(g["tsamp"][0]), opts, pp, 6.0)
    np.testing.assert_array_equal(peaks, g["peaks"])
    assert peaks[0]["dm_index"] == 3


def test_peak_order(port):
    p = np.zeros(4, PEAK_DTYPE)
    p["snr"] = [5, 7, 7, np.inf]
    p["period"] = [1, 2, 1, 1]
    p["frequency"] = 1 / p["period"]
    s = port.sort_peaks(p)
    assert list(s["snr"][:3]) == [7, 7, 5] and s[0]["period"] == 1 and np.isinf(s[3]["snr"])


# --------------------------------------------------- synthetic inputs ----
def test_synth_closed_forms(port):
    x = port.synth_filterbank(4096, 64, 1500.0, -4.0, 256e-6, seed=1)
    assert x.dtype == np.uint8 and x.shape == (4096, 64)
    assert abs(x.mean() - 128.0) < 0.1
    assert abs(x.std() - 16.0) < 0.1
    # deterministic bytes (committed digest of the first 1024 x 64 block)
    assert hashlib.sha256(x[:1024].tobytes()).hexdigest() == SYNTH_SHA256


def test_synth_pulse_realigns_at_true_dm(port):
    freq = channel_frequencies(64, 1500.0, -4.0)
    pulsar = (0.2048, 30.0, 0.05, 3.0, 0.0)
    x = port.synth_filterbank(1 << 14, 64, 1500.0, -4.0, 256e-6, seed=2, pulsars=[pulsar])
    p = port.dedisperse(x, freq, dm_low=0.0, dm_step=30.0, ndm=2, tsamp=256e-6)
    n = p.shape[1]
    period = 800  # samples
    k = (n // period) * period
    prof0 = p[0][:k].reshape(-1, period).mean(axis=0)
    prof30 = p[1][:k].reshape(-1, period).mean(axis=0)
    # at the true DM the folded pulse is sharp and peaks at phase 0
    assert np.argmax(prof30) in (0, 1, period - 1)
    assert prof30.max() - np.median(prof30) > 2.0 * (prof0.max() - np.median(prof0))


SYNTH_SHA256 = "c9f34b64d6ee966f3751a63fc3597e1b55c6fb7f8ef30f4f1bc45240ccecd2d5"


def test_spectrum_matches_reference(port, ref):
    """Aligned spectrum: port == reference build for every dtype; each
    channel is the input column shifted by its own delay."""
    rng = np.random.default_rng(12)
    freq = channel_frequencies(48, 1500.0, -4.0)
    for dtype in (np.uint8, np.uint16, np.float32):
        x = (rng.integers(0, 250, size=(3000, 48)) if dtype != np.float32
             else rng.normal(size=(3000, 48))).astype(dtype)
        a = port.dedisperse_spectrum(x, freq, 37.5, 1e-3)
        b = ref.dedisperse_spectrum(x, freq, 37.5, 1e-3)
        np.testing.assert_array_equal(a, b)
        d = port.multi_dm_delay_table(freq, 37.5, 1.0, 1, freq.max(), 1e-3)[0]
        for c in (0, 17, 47):
            np.testing.assert_array_equal(a[:, c], x[d[c]: d[c] + a.shape[0], c])

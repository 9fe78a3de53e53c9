"""Test infrastructure only: ctypes access to the CPU oracle.

``Oracle("port")`` loads the independent restatement ``oracle/liboracle.so``;
``Oracle("reference")`` loads ``oracle/_ref/libgaffa_ref.so``, gaffa's own CPU
sources compiled by ``oracle/Makefile``.  Both expose the ABI in
``oracle/oracle_abi.h``.  Only tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only to check
or time the CPU path; the product package never does.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {
    "port": os.path.join(HERE, "liboracle.so"),
    "reference": os.path.join(HERE, "_ref", "libgaffa_ref.so"),
}

ERRORS = {
    1: ValueError,  # std::invalid_argument
    2: OverflowError,  # std::overflow_error
    3: IndexError,  # std::out_of_range
    4: RuntimeError,  # std::runtime_error
    5: RuntimeError,  # std::logic_error
    6: RuntimeError,
}


class Task(C.Structure):
    _fields_ = [
        ("downsample_factor", C.c_double),
        ("effective_tsamp", C.c_double),
        ("input_nsamples", C.c_uint64),
        ("prepared_nsamples", C.c_uint64),
        ("bins", C.c_uint64),
        ("rows", C.c_uint64),
        ("rows_eval", C.c_uint64),
        ("period_begin", C.c_double),
        ("period_end", C.c_double),
    ]

    def as_tuple(self):
        return tuple(getattr(self, f) for f, _ in self._fields_)


class Peak(C.Structure):
    _fields_ = [
        ("period", C.c_double),
        ("frequency", C.c_double),
        ("width", C.c_uint64),
        ("duty_cycle", C.c_double),
        ("width_index", C.c_uint64),
        ("period_index", C.c_uint64),
        ("phase", C.c_uint64),
        ("shift", C.c_uint64),
        ("bins", C.c_uint64),
        ("snr", C.c_float),
    ]


class DmPeak(C.Structure):
    _fields_ = [("dm", C.c_double), ("dm_index", C.c_uint64), ("peak", Peak)]


class PlanOptions(C.Structure):
    _fields_ = [
        ("period_min", C.c_double),
        ("period_max", C.c_double),
        ("bins_min", C.c_uint64),
        ("bins_max", C.c_uint64),
        ("min_periods", C.c_uint64),
        ("duty_cycle_max", C.c_double),
        ("width_trial_spacing", C.c_double),
        ("max_tasks", C.c_uint64),
    ]


class PreprocessPlan(C.Structure):
    _fields_ = [
        ("detrend", C.c_int32),
        ("window_samples", C.c_uint64),
        ("min_points", C.c_uint64),
        ("normalise", C.c_int32),
        ("reject_constant", C.c_int32),
    ]


class Pulsar(C.Structure):
    _fields_ = [
        ("period_s", C.c_double),
        ("dm", C.c_double),
        ("duty", C.c_double),
        ("amp_sigma", C.c_double),
        ("phase0", C.c_double),
    ]


PEAK_DTYPE = np.dtype(
    [
        ("period", "f8"),
        ("frequency", "f8"),
        ("width", "u8"),
        ("duty_cycle", "f8"),
        ("width_index", "u8"),
        ("period_index", "u8"),
        ("phase", "u8"),
        ("shift", "u8"),
        ("bins", "u8"),
        ("snr", "f4"),
    ],
    align=True,
)
assert PEAK_DTYPE.itemsize == C.sizeof(Peak)
DM_PEAK_DTYPE = np.dtype([("dm", "f8"), ("dm_index", "u8"), ("peak", PEAK_DTYPE)], align=True)
assert DM_PEAK_DTYPE.itemsize == C.sizeof(DmPeak)
CANDIDATE_DTYPE = np.dtype(
    [
        ("dm", "f8"), ("dm_index", "u8"), ("period", "f8"), ("frequency", "f8"),
        ("width", "u8"), ("duty_cycle", "f8"), ("snr", "f4"), ("peak_count", "u8"),
        ("dm_index_min", "u8"), ("dm_index_max", "u8"), ("frequency_min", "f8"),
        ("frequency_max", "f8"), ("best_peak", DM_PEAK_DTYPE),
    ],
    align=True,
)
HARMONIC_DTYPE = np.dtype(
    [
        ("is_harmonic", "i4"), ("parent_index", "u8"), ("numerator", "i4"),
        ("denominator", "i4"), ("frequency_error_bins", "f8"), ("phase_distance", "f8"),
        ("dm_distance", "f8"), ("expected_snr", "f8"), ("snr_distance", "f8"),
    ],
    align=True,
)


class HarmonicOptions(C.Structure):
    _fields_ = [("max_harmonic", C.c_uint64), ("denominator_max", C.c_uint64),
                ("frequency_tolerance_bins", C.c_double), ("phase_distance_max", C.c_double),
                ("dm_distance_max", C.c_double), ("use_snr_consistency", C.c_int32),
                ("snr_distance_max", C.c_double)]


TASK_DTYPE = np.dtype(
    [
        ("downsample_factor", "f8"),
        ("effective_tsamp", "f8"),
        ("input_nsamples", "u8"),
        ("prepared_nsamples", "u8"),
        ("bins", "u8"),
        ("rows", "u8"),
        ("rows_eval", "u8"),
        ("period_begin", "f8"),
        ("period_end", "f8"),
    ]
)
assert TASK_DTYPE.itemsize == C.sizeof(Task)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


@dataclass
class Plan:
    tasks: np.ndarray  # TASK_DTYPE structured array
    widths: np.ndarray  # uint64


class FbHeader(C.Structure):
    _fields_ = [("header_size", C.c_int64), ("nsamples", C.c_int64)] + \
        [(k, C.c_int32) for k in ("telescope_id", "machine_id", "data_type", "barycentric",
                                  "pulsarcentric", "ibeam", "nbeams", "npuls", "nbins", "nbits",
                                  "nifs", "nchans")] + \
        [(k, C.c_double) for k in ("az_start", "za_start", "src_raj", "src_dej", "tstart",
                                   "tsamp", "fch1", "foff", "refdm", "period")] + \
        [("uses_frequency_table", C.c_int32), ("rawdatafile", C.c_char * 256),
         ("source_name", C.c_char * 256)]


class Oracle:
    def __init__(self, kind: str = "port"):
        path = LIBS[kind]
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library {path} is not built (run make -C oracle)")
        self.kind = kind
        self.lib = C.CDLL(path)
        self.lib.orc_last_error.restype = C.c_char_p
        self.lib.orc_implementation.restype = C.c_char_p
        for name in dir(self.lib):
            pass

    # -- helpers --------------------------------------------------------
    def _call(self, fn, *args):
        rc = fn(*args)
        if rc != 0:
            msg = self.lib.orc_last_error().decode()
            raise ERRORS.get(rc, RuntimeError)(msg)

    @property
    def implementation(self) -> str:
        return self.lib.orc_implementation().decode()

    # -- SIGPROC reader (reference build only) ---------------------------
    def read_filterbank(self, path, preserve_file_order=False):
        """gaffa's read_filterbank: (header dict, frequency table, sample bytes)."""
        h = FbHeader()
        nbytes = C.c_uint64()
        fn = self.lib.orc_read_filterbank
        freq = np.empty(1 << 16, dtype=np.float64)
        self._call(fn, str(path).encode(), C.c_int32(1 if preserve_file_order else 0),
                   C.byref(h), _ptr(freq), C.c_uint64(freq.size), None, C.c_uint64(0),
                   C.byref(nbytes))
        data = np.empty(nbytes.value, dtype=np.uint8)
        self._call(fn, str(path).encode(), C.c_int32(1 if preserve_file_order else 0),
                   C.byref(h), _ptr(freq), C.c_uint64(freq.size), _ptr(data),
                   C.c_uint64(data.size), C.byref(nbytes))
        d = {k: getattr(h, k) for k, _ in FbHeader._fields_}
        d["rawdatafile"] = h.rawdatafile.decode()
        d["source_name"] = h.source_name.decode()
        d["uses_frequency_table"] = bool(h.uses_frequency_table)
        return d, freq[:h.nchans].copy(), data

    # -- dedispersion ---------------------------------------------------
    def delay_bins(self, dm, f, ref, tsamp) -> int:
        out = C.c_int32()
        self._call(self.lib.orc_delay_bins, C.c_double(dm), C.c_double(f), C.c_double(ref),
                   C.c_double(tsamp), C.byref(out))
        return out.value

    def multi_dm_delay_table(self, freq, dm_low, dm_step, ndm, ref, tsamp, chan_begin=0,
                             chan_end=None):
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        chan_end = freq.size if chan_end is None else chan_end
        count = C.c_uint64()
        args = (_ptr(freq), C.c_uint64(freq.size), C.c_double(dm_low), C.c_double(dm_step),
                C.c_uint64(ndm), C.c_double(ref), C.c_double(tsamp), C.c_uint64(chan_begin),
                C.c_uint64(chan_end))
        self._call(self.lib.orc_multi_dm_delay_table, *args, None, C.c_uint64(0), C.byref(count))
        out = np.empty(count.value, dtype=np.int32)
        self._call(self.lib.orc_multi_dm_delay_table, *args, _ptr(out), C.c_uint64(out.size),
                   C.byref(count))
        return out.reshape(ndm, chan_end - chan_begin)

    def _dtype_code(self, samples):
        return {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}[samples.dtype]

    def dedisperse(self, samples, freq, *, dm_low, dm_step, ndm, ref=None, tsamp,
                   chan_begin=0, chan_end=None, subband=False, subband_channels=32,
                   ndm_per_nominal=32):
        """samples: (nsamples, nchans) array; returns (ndm, n_out) uint32/float32."""
        samples = np.ascontiguousarray(samples)
        if samples.ndim == 3:
            samples = samples.reshape(samples.shape[0], samples.shape[2])
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        ref = float(freq.max()) if ref is None else ref
        chan_end = samples.shape[1] if chan_end is None else chan_end
        code = self._dtype_code(samples)
        nout = C.c_uint64()
        args = (C.c_int32(1 if subband else 0), C.c_int32(code), _ptr(samples),
                C.c_uint64(samples.shape[0]), C.c_uint64(samples.shape[1]), _ptr(freq),
                C.c_double(dm_low), C.c_double(dm_step), C.c_uint64(ndm), C.c_double(ref),
                C.c_double(tsamp), C.c_uint64(chan_begin), C.c_uint64(chan_end),
                C.c_uint64(subband_channels), C.c_uint64(ndm_per_nominal))
        # size query: the restatement computes the output length without the sums
        # only for the port; the reference always computes, so size from the formula
        n_out = self._multi_nout(samples.shape[0], freq, dm_low, dm_step, ndm, ref, tsamp,
                                 chan_begin, chan_end)
        out = np.empty((ndm, n_out), dtype=np.float32 if code == 2 else np.uint32)
        self._call(self.lib.orc_dedisperse, *args, _ptr(out), C.c_uint64(out.size), C.byref(nout))
        assert nout.value == n_out
        return out

    def _multi_nout(self, n, freq, dm_low, dm_step, ndm, ref, tsamp, cb, ce):
        dmax = dm_low + (ndm - 1) * dm_step
        d = self.delay_bins(dmax, float(np.min(freq[cb:ce])), ref, tsamp)
        if d < 0 or d >= n:
            # let the library raise its own error
            return 0
        return n - d

    def dedisperse_single(self, samples, freq, *, dm, ref=None, tsamp, chan_begin=0,
                          chan_end=None):
        samples = np.ascontiguousarray(samples)
        if samples.ndim == 3:
            samples = samples.reshape(samples.shape[0], samples.shape[2])
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        ref = float(freq.max()) if ref is None else ref
        chan_end = samples.shape[1] if chan_end is None else chan_end
        code = self._dtype_code(samples)
        out = np.empty(samples.shape[0], dtype=np.float32 if code == 2 else np.uint32)
        nout = C.c_uint64()
        self._call(self.lib.orc_dedisperse_single, C.c_int32(code), _ptr(samples),
                   C.c_uint64(samples.shape[0]), C.c_uint64(samples.shape[1]), _ptr(freq),
                   C.c_double(dm), C.c_double(ref), C.c_double(tsamp), C.c_uint64(chan_begin),
                   C.c_uint64(chan_end), _ptr(out), C.c_uint64(out.size), C.byref(nout))
        return out[: nout.value].copy()

    # -- time series ----------------------------------------------------
    def stats(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        m, v, s = C.c_double(), C.c_double(), C.c_double()
        self._call(self.lib.orc_time_series_stats, _ptr(x), C.c_uint64(x.size), C.byref(m),
                   C.byref(v), C.byref(s))
        return m.value, v.value, s.value

    def normalise(self, x, reject_constant=True):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty_like(x)
        self._call(self.lib.orc_normalise, _ptr(x), C.c_uint64(x.size),
                   C.c_int32(int(reject_constant)), _ptr(out))
        return out

    def downsampled_size(self, n, factor) -> int:
        out = C.c_uint64()
        self._call(self.lib.orc_downsampled_size, C.c_uint64(n), C.c_double(factor), C.byref(out))
        return out.value

    def downsampled_variance(self, n, factor) -> float:
        out = C.c_double()
        self._call(self.lib.orc_downsampled_variance, C.c_uint64(n), C.c_double(factor),
                   C.byref(out))
        return out.value

    def downsample(self, x, factor):
        x = np.ascontiguousarray(x, dtype=np.float32)
        m = self.downsampled_size(x.size, factor)
        out = np.empty(m, dtype=np.float32)
        count = C.c_uint64()
        self._call(self.lib.orc_downsample, _ptr(x), C.c_uint64(x.size), C.c_double(factor),
                   _ptr(out), C.c_uint64(m), C.byref(count))
        return out

    # -- preprocessing --------------------------------------------------
    def running_median_window(self, seconds, tsamp) -> int:
        out = C.c_uint64()
        self._call(self.lib.orc_running_median_window, C.c_double(seconds), C.c_double(tsamp),
                   C.byref(out))
        return out.value

    def detrend(self, x, window_samples, min_points=101):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty_like(x)
        self._call(self.lib.orc_detrend, _ptr(x), C.c_uint64(x.size), C.c_uint64(window_samples),
                   C.c_uint64(min_points), _ptr(out))
        return out

    @staticmethod
    def preprocess_plan(window_samples=None, min_points=101, normalise=True,
                        reject_constant=True) -> PreprocessPlan:
        return PreprocessPlan(int(window_samples is not None), window_samples or 0, min_points,
                              int(normalise), int(reject_constant))

    def preprocess(self, x, tsamp, plan: PreprocessPlan):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty_like(x)
        self._call(self.lib.orc_preprocess, _ptr(x), C.c_uint64(x.size), C.c_double(tsamp),
                   C.byref(plan), _ptr(out))
        return out

    # -- FFA ------------------------------------------------------------
    @staticmethod
    def plan_options(period_min, period_max, bins_min=180, bins_max=256, min_periods=1,
                     duty_cycle_max=0.20, width_trial_spacing=1.5, max_tasks=1_000_000):
        return PlanOptions(period_min, period_max, bins_min, bins_max, min_periods,
                           duty_cycle_max, width_trial_spacing, max_tasks)

    def make_plan(self, nsamples, tsamp, options: PlanOptions) -> Plan:
        tc, wc = C.c_uint64(), C.c_uint64()
        self._call(self.lib.orc_make_riptide_plan, C.c_uint64(nsamples), C.c_double(tsamp),
                   C.byref(options), None, C.c_uint64(0), C.byref(tc), None, C.c_uint64(0),
                   C.byref(wc))
        tasks = np.zeros(tc.value, dtype=TASK_DTYPE)
        widths = np.zeros(wc.value, dtype=np.uint64)
        self._call(self.lib.orc_make_riptide_plan, C.c_uint64(nsamples), C.c_double(tsamp),
                   C.byref(options), _ptr(tasks), C.c_uint64(tasks.size), C.byref(tc),
                   _ptr(widths), C.c_uint64(widths.size), C.byref(wc))
        return Plan(tasks, widths)

    def ffa_transform(self, x, bins):
        x = np.ascontiguousarray(x)
        rows = x.size // bins
        if x.dtype == np.float64:
            xin = np.ascontiguousarray(x[: rows * bins])
            out = np.empty(rows * bins, dtype=np.float64)
            self._call(self.lib.orc_ffa_transform_f64, _ptr(xin), C.c_uint64(rows),
                       C.c_uint64(bins), _ptr(out))
            return out.reshape(rows, bins)
        xin = np.ascontiguousarray(x[: rows * bins], dtype=np.float32)
        out = np.empty(rows * bins, dtype=np.float32)
        self._call(self.lib.orc_ffa_transform, _ptr(xin), C.c_uint64(rows), C.c_uint64(bins),
                   _ptr(out))
        return out.reshape(rows, bins)

    def ffa_rows(self, x, tasks):
        x = np.ascontiguousarray(x, dtype=np.float32)
        tasks = np.ascontiguousarray(tasks, dtype=TASK_DTYPE)
        count = C.c_uint64()
        self._call(self.lib.orc_ffa_rows, _ptr(x), C.c_uint64(x.size), _ptr(tasks),
                   C.c_uint64(tasks.size), None, C.c_uint64(0), C.byref(count))
        out = np.empty(count.value, dtype=np.float32)
        self._call(self.lib.orc_ffa_rows, _ptr(x), C.c_uint64(x.size), _ptr(tasks),
                   C.c_uint64(tasks.size), _ptr(out), C.c_uint64(out.size), C.byref(count))
        return out

    def task_stdnoise(self, task) -> float:
        t = np.ascontiguousarray(np.array(task, dtype=TASK_DTYPE).reshape(1))
        out = C.c_float()
        self._call(self.lib.orc_task_stdnoise, _ptr(t), C.byref(out))
        return out.value

    def find_ffa_peaks(self, transform, task, widths, stdnoise, snr_threshold=6.0,
                       max_peaks=0):
        tr = np.ascontiguousarray(transform, dtype=np.float32)
        rows, bins = tr.shape
        t = np.ascontiguousarray(np.array(task, dtype=TASK_DTYPE).reshape(1))
        w = np.ascontiguousarray(widths, dtype=np.uint64)
        args = (_ptr(tr), C.c_uint64(rows), C.c_uint64(bins), _ptr(t), _ptr(w),
                C.c_uint64(w.size), C.c_float(stdnoise), C.c_float(snr_threshold),
                C.c_uint64(max_peaks))
        count = C.c_uint64()
        self._call(self.lib.orc_find_ffa_peaks, *args, None, C.c_uint64(0), C.byref(count))
        out = np.zeros(count.value, dtype=PEAK_DTYPE)
        self._call(self.lib.orc_find_ffa_peaks, *args, _ptr(out), C.c_uint64(out.size),
                   C.byref(count))
        return out

    def search_ffa(self, x, plan: Plan, snr_threshold=6.0, max_peaks=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        tasks = np.ascontiguousarray(plan.tasks, dtype=TASK_DTYPE)
        w = np.ascontiguousarray(plan.widths, dtype=np.uint64)
        args = (_ptr(x), C.c_uint64(x.size), _ptr(tasks), C.c_uint64(tasks.size), _ptr(w),
                C.c_uint64(w.size), C.c_float(snr_threshold), C.c_uint64(max_peaks))
        # one pass with a generous buffer; retry with the exact size if needed
        out = np.zeros(4096, dtype=PEAK_DTYPE)
        count = C.c_uint64()
        rc = self.lib.orc_search_ffa(*args, _ptr(out), C.c_uint64(out.size), C.byref(count))
        if rc == 7:
            out = np.zeros(count.value, dtype=PEAK_DTYPE)
            rc = self.lib.orc_search_ffa(*args, _ptr(out), C.c_uint64(out.size), C.byref(count))
        if rc != 0:
            raise ERRORS.get(rc, RuntimeError)(self.lib.orc_last_error().decode())
        return out[: count.value].copy()

    def sort_peaks(self, peaks):
        p = np.ascontiguousarray(peaks, dtype=PEAK_DTYPE).copy()
        self._call(self.lib.orc_sort_peaks, _ptr(p), C.c_uint64(p.size))
        return p

    def find_dm_peaks(self, plane, dms, tsamp, options: PlanOptions,
                      preprocess: PreprocessPlan | None, snr_threshold=6.0, max_peaks=0):
        plane = np.ascontiguousarray(plane, dtype=np.uint32)
        ndm, ns = plane.shape
        dms = np.ascontiguousarray(dms, dtype=np.float64)
        pp = preprocess if preprocess is not None else PreprocessPlan(0, 0, 0, 0, 1)
        args = (_ptr(plane), C.c_uint64(ndm), C.c_uint64(ns), _ptr(dms), C.c_double(tsamp),
                C.byref(options), C.byref(pp), C.c_float(snr_threshold), C.c_uint64(max_peaks))
        out = np.zeros(1 << 16, dtype=DM_PEAK_DTYPE)
        count = C.c_uint64()
        rc = self.lib.orc_find_dm_peaks(*args, _ptr(out), C.c_uint64(out.size), C.byref(count))
        if rc == 7:
            out = np.zeros(count.value, dtype=DM_PEAK_DTYPE)
            rc = self.lib.orc_find_dm_peaks(*args, _ptr(out), C.c_uint64(out.size), C.byref(count))
        if rc != 0:
            raise ERRORS.get(rc, RuntimeError)(self.lib.orc_last_error().decode())
        return out[: count.value].copy()

    def dedisperse_spectrum(self, samples, freq, dm, tsamp, ref=None, cb=0, ce=None):
        samples = np.ascontiguousarray(samples)
        ns, nc = samples.shape
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        ref = float(freq.max()) if ref is None else ref
        ce = nc if ce is None else ce
        dtype = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1,
                 np.dtype(np.float32): 2}[samples.dtype]
        args = (C.c_int32(dtype), _ptr(samples), C.c_uint64(ns), C.c_uint64(nc), _ptr(freq),
                C.c_double(dm), C.c_double(ref), C.c_double(tsamp), C.c_uint64(cb),
                C.c_uint64(ce))
        no = C.c_uint64()
        self._call(self.lib.orc_dedisperse_spectrum, *args, None, C.c_uint64(0), C.byref(no))
        out = np.zeros((no.value, ce - cb), dtype=samples.dtype)
        self._call(self.lib.orc_dedisperse_spectrum, *args, _ptr(out), C.c_uint64(out.size),
                   C.byref(no))
        return out

    # -- folding (reference build only) ----------------------------------
    _FOLD_DTYPES = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.uint32): 2,
                    np.dtype(np.float32): 3}

    def fold_time_series(self, x, period, tsamp, nbin, start_sample=0):
        x = np.ascontiguousarray(x)
        prof = np.zeros(nbin, np.float32)
        expo = np.zeros(nbin, np.float64)
        self._call(self.lib.orc_fold_time_series, _ptr(x), C.c_int32(self._FOLD_DTYPES[x.dtype]),
                   C.c_uint64(x.size), C.c_double(period), C.c_double(tsamp), C.c_uint64(nbin),
                   C.c_uint64(start_sample), _ptr(prof), _ptr(expo))
        return prof, expo

    def fold_spectrum(self, x, period, tsamp, nbin, subint_samples, start_sample=0,
                      output_channels=0):
        x = np.ascontiguousarray(x)
        ns, nc = x.shape
        nsub = -(-ns // subint_samples)
        nout = output_channels or nc
        cube = np.zeros(nsub * nout * nbin, np.float32)
        expo = np.zeros(nsub * nbin, np.float64)
        count = C.c_uint64()
        self._call(self.lib.orc_fold_spectrum, _ptr(x), C.c_int32(self._FOLD_DTYPES[x.dtype]),
                   C.c_uint64(ns), C.c_uint64(nc), C.c_double(period), C.c_double(tsamp),
                   C.c_uint64(nbin), C.c_uint64(subint_samples), C.c_uint64(start_sample),
                   C.c_uint64(output_channels), _ptr(cube), _ptr(expo), C.c_uint64(cube.size),
                   C.byref(count))
        return cube.reshape(nsub, nout, nbin), expo.reshape(nsub, nbin)

    # -- candidates (reference build only) -------------------------------
    def select_candidates(self, peaks, observation_seconds, frequency_cluster_radius=0.1,
                          dm_cluster_radius=1, cluster_across_widths=True, max_candidates=0):
        peaks = np.ascontiguousarray(peaks, dtype=DM_PEAK_DTYPE)
        out = np.zeros(max(1, peaks.size), dtype=CANDIDATE_DTYPE)
        count = C.c_uint64()
        self._call(self.lib.orc_select_candidates, _ptr(peaks), C.c_uint64(peaks.size),
                   C.c_double(observation_seconds), C.c_double(frequency_cluster_radius),
                   C.c_uint64(dm_cluster_radius), C.c_int32(1 if cluster_across_widths else 0),
                   C.c_uint64(max_candidates), _ptr(out), C.c_uint64(out.size), C.byref(count))
        return out[: count.value].copy()

    def flag_harmonics(self, candidates, observation_seconds, frequency_low_mhz,
                       frequency_high_mhz, max_harmonic=16, denominator_max=5,
                       frequency_tolerance_bins=1.5, phase_distance_max=1.0,
                       dm_distance_max=3.0, use_snr_consistency=False, snr_distance_max=3.0):
        cands = np.ascontiguousarray(candidates, dtype=CANDIDATE_DTYPE)
        opts = HarmonicOptions(max_harmonic, denominator_max, frequency_tolerance_bins,
                               phase_distance_max, dm_distance_max,
                               1 if use_snr_consistency else 0, snr_distance_max)
        out = np.zeros(max(1, cands.size), dtype=CANDIDATE_DTYPE)
        rel = np.zeros(max(1, cands.size), dtype=HARMONIC_DTYPE)
        count = C.c_uint64()
        self._call(self.lib.orc_flag_harmonics, _ptr(cands), C.c_uint64(cands.size),
                   C.c_double(observation_seconds), C.c_double(frequency_low_mhz),
                   C.c_double(frequency_high_mhz), C.byref(opts), _ptr(out), _ptr(rel),
                   C.c_uint64(out.size), C.byref(count))
        return out[: count.value].copy(), rel[: count.value].copy()

    # -- synthetic data (port only) ---------------------------------------
    def synth_filterbank(self, nsamples, nchans, fch1, foff, tsamp, seed, pulsars=(),
                         rfi_channels=(), rfi_offset=0, t_begin=0, t_count=None):
        t_count = nsamples - t_begin if t_count is None else t_count
        ps = (Pulsar * max(1, len(pulsars)))(*[Pulsar(*p) for p in pulsars])
        rfi = np.ascontiguousarray(rfi_channels, dtype=np.uint64)
        out = np.empty((t_count, nchans), dtype=np.uint8)
        self._call(self.lib.orc_synth_filterbank, C.c_uint64(nsamples), C.c_uint64(nchans),
                   C.c_double(fch1), C.c_double(foff), C.c_double(tsamp), C.c_uint64(seed), ps,
                   C.c_uint64(len(pulsars)), _ptr(rfi) if rfi.size else None,
                   C.c_uint64(rfi.size), C.c_int32(rfi_offset), C.c_uint64(t_begin),
                   C.c_uint64(t_count), _ptr(out))
        return out


def channel_frequencies(nchans, fch1, foff):
    """Ascending channel frequencies, as the synthetic generator stores them."""
    c = np.arange(nchans, dtype=np.float64)
    if foff < 0:
        return fch1 + (nchans - 1 - c) * foff
    return fch1 + c * foff

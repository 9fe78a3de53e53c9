"""Search workloads: synthetic filterbank specs, DM plans and the GPU pipeline.

The workloads are the observation shapes the search hot path is tested and
measured on:

* ``small``   - 64 channels, 1500 MHz / -4 MHz, 256 us, 2^17 samples, one
  pulsar (P 0.714 s, DM 30, 3 % duty, 0.5 sigma per channel), brute-force
  dedispersion over DM 0..63, 1 s running median, FFA 0.2-2 s, widths 1..16.
* ``lband``   - 1024 channels, 1520 MHz / -0.293 MHz, 64 us, 2^22 samples,
  8 pulsars, subband dedispersion (32 subbands) over 1024 DM trials 0-500 with
  smearing-tolerance steps, 2 s running median, FFA 0.1-10 s, widths 1..32.

A DM plan is a list of uniform segments (whole nominal groups of 32 trials);
within a segment the step is constant so every segment is one of gaffa's
``MultiDmDedispersionPlan`` s.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._core import DM_PEAK_RECORD_BYTES, DmSegment, GpuSearchPipeline, SynthFilterbankSpec, \
    SynthPulsar

DISPERSION_S = 4148.808  # s MHz^2 / (pc cm^-3), dedispersion_delay.h:16

# gaffa's DmPeak record (dm_search.h:16-20 around ffa_peak.h:8-19), as laid
# out in memory by the host layer
DM_PEAK_DTYPE = np.dtype(
    [("dm", "f8"), ("dm_index", "u8"),
     ("peak", np.dtype([("period", "f8"), ("frequency", "f8"), ("width", "u8"),
                        ("duty_cycle", "f8"), ("width_index", "u8"), ("period_index", "u8"),
                        ("phase", "u8"), ("shift", "u8"), ("bins", "u8"), ("snr", "f4")],
                       align=True))],
    align=True)
assert DM_PEAK_DTYPE.itemsize == DM_PEAK_RECORD_BYTES

# The candidate stage of gaffa's GPU DM-search benchmark
# (benchmarks/cuda_dm_search_benchmark.cpp:105-136, 279-292): clustering over
# 0.1 Fourier bins and 50 DM trials across widths, then harmonic flagging up
# to the 32nd harmonic with denominators up to 100, harmonics removed.
CANDIDATE_OPTIONS = dict(frequency_cluster_radius=0.1, dm_cluster_radius=50,
                         cluster_across_widths=True)
HARMONIC_OPTIONS = dict(max_harmonic=32, denominator_max=100, frequency_tolerance_bins=1.5,
                        phase_distance_max=1.0, dm_distance_max=3.0)


def peak_records(pipe: GpuSearchPipeline, dm_begin: int = 0, dm_end: int = -1) -> np.ndarray:
    """The last search's peaks of DM trials [dm_begin, dm_end) as a structured
    array (DM_PEAK_DTYPE), in the list's order."""
    end = pipe.ndm if dm_end < 0 else dm_end
    return np.frombuffer(pipe.peak_records(dm_begin, end), dtype=DM_PEAK_DTYPE)


def search_candidates(pipe: GpuSearchPipeline, workload: "Workload") -> list:
    """Candidates of the last search as gaffa's benchmark makes them:
    clustered peaks with harmonics flagged and removed."""
    from ._core import flag_harmonics, remove_harmonics

    cands = pipe.select_candidates(**CANDIDATE_OPTIONS)
    if not cands:
        return []
    f = workload.frequencies
    flagged = flag_harmonics(cands, workload.nsamples * workload.tsamp, float(f.min()),
                             float(f.max()), **HARMONIC_OPTIONS)
    return remove_harmonics(flagged)


def profile_snr(profile: np.ndarray, widths=(1, 2, 4, 8, 16)) -> float:
    """Boxcar S/N of a folded pulse profile: the largest circular boxcar sum
    of the profile minus its median, over width w bins, divided by
    sqrt(w) times the noise sigma of one bin (1.4826 times the median
    absolute deviation, which a pulse of up to ~20 % duty barely moves)."""
    p = np.asarray(profile, np.float64).ravel()
    med = float(np.median(p))
    sigma = 1.4826 * float(np.median(np.abs(p - med)))
    if not np.isfinite(sigma) or sigma <= 0.0:
        return 0.0
    x = p - med
    c = np.concatenate([[0.0], np.cumsum(np.concatenate([x, x]))])
    n = len(x)
    best = -np.inf
    for w in widths:
        if w > n:
            break
        best = max(best, float(np.max(c[w:w + n] - c[:n])) / (sigma * math.sqrt(w)))
    return best


def channel_frequencies(nchans: int, fch1: float, foff: float) -> np.ndarray:
    """Ascending channel frequencies (gaffa's reader reverses foff < 0 files)."""
    c = np.arange(nchans, dtype=np.float64)
    if foff < 0:
        return fch1 + (nchans - 1 - c) * foff
    return fch1 + c * foff


def smearing_dm_segments(*, dm_max: float, ndm: int, tsamp: float, freq: np.ndarray,
                         segment_trials: int = 128) -> list[DmSegment]:
    """DM trials 0..dm_max whose step follows the smearing tolerance.

    The step at DM is proportional to the effective sample width
    sqrt(tsamp^2 + t_chan(DM)^2) over the band-crossing delay per unit DM,
    with t_chan the intra-channel smearing at the lowest frequency; it is held
    constant over each segment of ``segment_trials`` trials and the overall
    scale is solved so ``ndm`` trials end at ``dm_max``.
    """
    if ndm % segment_trials:
        raise ValueError("ndm must be a multiple of segment_trials")
    f_lo, f_hi = float(np.min(freq)), float(np.max(freq))
    chan_bw = float(np.min(np.diff(np.sort(freq)))) if len(freq) > 1 else 1.0
    k_band = DISPERSION_S * (f_lo ** -2 - f_hi ** -2)

    def step_at(dm: float) -> float:
        t_chan = 8.3e-6 * chan_bw * dm / (f_lo / 1000.0) ** 3
        return math.sqrt(tsamp ** 2 + t_chan ** 2) / k_band

    def build(scale: float):
        segs, dm = [], 0.0
        for _ in range(ndm // segment_trials):
            step = scale * step_at(dm)
            segs.append((dm, step))
            dm += segment_trials * step
        return segs, dm

    lo, hi = 1e-3, 1e3
    for _ in range(200):
        mid = math.sqrt(lo * hi)
        _, end = build(mid)
        if end > dm_max:
            hi = mid
        else:
            lo = mid
    segs, _ = build(lo)
    return [DmSegment(dm_low=d, dm_step=s, ndm=segment_trials) for d, s in segs]


def pulse_amplitude(snr: float, nchans: int, nsamples: int, duty: float) -> float:
    """Per-channel peak amplitude (noise sigma) of a Gaussian pulse of FWHM
    ``duty`` whose ideal matched-filter S/N over the whole band and
    observation is ``snr``: snr = A sqrt(nchans N sigma_phi sqrt(pi))."""
    sigma_phi = duty / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    return snr / math.sqrt(nchans * nsamples * sigma_phi * math.sqrt(math.pi))


@dataclass
class Workload:
    name: str
    nsamples: int
    nchans: int
    fch1: float
    foff: float
    tsamp: float
    seed: int
    pulsars: list = field(default_factory=list)  # (period_s, dm, duty, amp_sigma, phase0)
    rfi_channels: list = field(default_factory=list)
    rfi_offset: int = 0
    dm_segments: list = field(default_factory=list)
    subband: bool = True
    subband_channels: int = 32
    ndm_per_nominal: int = 32
    median_seconds: float = 1.0
    min_points: int = 101
    period_min: float = 0.2
    period_max: float = 2.0
    bins_min: int = 180
    bins_max: int = 256
    duty_cycle_max: float = 0.09
    width_trial_spacing: float = 2.0
    snr_threshold: float = 6.0

    @property
    def frequencies(self) -> np.ndarray:
        return channel_frequencies(self.nchans, self.fch1, self.foff)

    @property
    def ndm(self) -> int:
        return sum(s.ndm for s in self.dm_segments)

    @property
    def dms(self) -> np.ndarray:
        return np.concatenate([s.dm_low + np.arange(s.ndm) * s.dm_step for s in self.dm_segments])

    def window_samples(self) -> int:
        r = max(1, round(self.median_seconds / self.tsamp))
        return r + 1 if r % 2 == 0 else r

    def synth_spec(self) -> SynthFilterbankSpec:
        return SynthFilterbankSpec(
            nsamples=self.nsamples, nchans=self.nchans, fch1=self.fch1, foff=self.foff,
            tsamp=self.tsamp, seed=self.seed,
            pulsars=[SynthPulsar(*p) for p in self.pulsars],
            rfi_channels=list(self.rfi_channels), rfi_offset=self.rfi_offset)

    def pipeline(self, *, device_id: int = 0, ffa_workspace_bytes: int = 32 << 30,
                 ffa_series_tile: int = 0) -> GpuSearchPipeline:
        return GpuSearchPipeline(
            nsamples=self.nsamples, nchans=self.nchans, tsamp=self.tsamp,
            frequency_mhz=list(self.frequencies), dm_segments=list(self.dm_segments),
            subband=self.subband, subband_channels=self.subband_channels,
            ndm_per_nominal=self.ndm_per_nominal, window_samples=self.window_samples(),
            min_points=self.min_points, normalise=True, period_min=self.period_min,
            period_max=self.period_max, bins_min=self.bins_min, bins_max=self.bins_max,
            duty_cycle_max=self.duty_cycle_max, width_trial_spacing=self.width_trial_spacing,
            snr_threshold=self.snr_threshold, ffa_workspace_bytes=ffa_workspace_bytes,
            ffa_series_tile=ffa_series_tile, device_id=device_id)


def small_workload(nsamples: int = 1 << 17, seed: int = 20260101) -> Workload:
    """64 channels, one pulsar at DM 30, brute force over DM 0..63."""
    return Workload(
        name="small", nsamples=nsamples, nchans=64, fch1=1500.0, foff=-4.0, tsamp=256e-6,
        seed=seed, pulsars=[(0.714, 30.0, 0.03, 0.5, 0.25)],
        dm_segments=[DmSegment(dm_low=0.0, dm_step=1.0, ndm=64)], subband=False,
        median_seconds=1.0, period_min=0.2, period_max=2.0, duty_cycle_max=0.09,
        width_trial_spacing=2.0)


def lband_workload(nsamples: int = 1 << 22, seed: int = 20260202, ndm: int = 1024,
                   npulsars: int = 8) -> Workload:
    """1024 channels, 8 pulsars, subband dedispersion over 1024 trials 0-500."""
    rng = np.random.default_rng(seed)
    nchans, fch1, foff, tsamp = 1024, 1520.0, -0.293, 64e-6
    pulsars = []
    for _ in range(npulsars):
        period = float(np.exp(rng.uniform(np.log(0.1), np.log(5.0))))
        dm = float(rng.uniform(0.0, 500.0))
        duty = float(rng.uniform(0.01, 0.10))
        snr = float(rng.uniform(8.0, 30.0))
        pulsars.append((period, dm, duty, pulse_amplitude(snr, nchans, nsamples, duty),
                        float(rng.uniform(0.0, 1.0))))
    freq = channel_frequencies(nchans, fch1, foff)
    segs = smearing_dm_segments(dm_max=500.0, ndm=ndm, tsamp=tsamp, freq=freq)
    return Workload(
        name="lband", nsamples=nsamples, nchans=nchans, fch1=fch1, foff=foff, tsamp=tsamp,
        seed=seed, pulsars=pulsars, dm_segments=segs, subband=True, subband_channels=32,
        ndm_per_nominal=32, median_seconds=2.0, period_min=0.1, period_max=10.0,
        duty_cycle_max=0.18, width_trial_spacing=2.0)


def large_workload(nsamples: int = 1 << 22, seed: int = 20260505, ndm: int = 4096,
                   npulsars: int = 32) -> Workload:
    """4096 channels (1700 MHz, -0.195 MHz: an 800 MHz band), 32 pulsars,
    5 narrowband RFI channels at a constant +40, subband dedispersion with 64
    channels per subband over 4096 trials 0-1000."""
    rng = np.random.default_rng(seed)
    nchans, fch1, foff, tsamp = 4096, 1700.0, -0.195, 64e-6
    pulsars = []
    for _ in range(npulsars):
        period = float(np.exp(rng.uniform(np.log(0.05), np.log(10.0))))
        dm = float(rng.uniform(0.0, 1000.0))
        duty = float(rng.uniform(0.005, 0.10))
        snr = float(rng.uniform(6.0, 25.0))
        pulsars.append((period, dm, duty, pulse_amplitude(snr, nchans, nsamples, duty),
                        float(rng.uniform(0.0, 1.0))))
    rfi = sorted(int(c) for c in rng.choice(nchans, 5, replace=False))
    freq = channel_frequencies(nchans, fch1, foff)
    segs = smearing_dm_segments(dm_max=1000.0, ndm=ndm, tsamp=tsamp, freq=freq)
    return Workload(
        name="large", nsamples=nsamples, nchans=nchans, fch1=fch1, foff=foff, tsamp=tsamp,
        seed=seed, pulsars=pulsars, rfi_channels=rfi, rfi_offset=40, dm_segments=segs,
        subband=True, subband_channels=64, ndm_per_nominal=32, median_seconds=2.0,
        period_min=0.1, period_max=10.0, duty_cycle_max=0.18, width_trial_spacing=2.0)


@dataclass
class BatchedWorkload:
    """Many float series of different lengths, each searched with its own
    octave plan (one batch, heterogeneous row counts)."""
    nseries: int
    lengths: np.ndarray
    tsamp: float
    seed: int
    period: np.ndarray     # injected period (s) per series
    phase0: np.ndarray
    sigma: np.ndarray      # pulse sigma in phase units
    amplitude: np.ndarray  # per-sample peak amplitude, 0 = noise only
    period_min: float = 0.1
    period_max: float = 10.0
    bins: int = 128
    duty_cycle_max: float = 0.20
    width_trial_spacing: float = 1.5
    snr_threshold: float = 6.0

    def plans(self) -> list:
        from .ffa import make_octave_plan

        cache: dict = {}
        out = []
        for n in self.lengths:
            n = int(n)
            if n not in cache:
                cache[n] = make_octave_plan(nsamples=n, tsamp=self.tsamp,
                                            period_min=self.period_min,
                                            period_max=self.period_max, bins=self.bins,
                                            duty_cycle_max=self.duty_cycle_max,
                                            width_trial_spacing=self.width_trial_spacing)
            out.append(cache[n])
        return out

    def search(self, **kw):
        from ._core import FfaBatchedSearch

        b = FfaBatchedSearch(self.plans(), **kw)
        b.generate(self.seed, self.tsamp, list(self.period), list(self.phase0),
                   list(self.sigma), list(self.amplitude))
        return b


def batched_workload(nseries: int = 2048, seed: int = 20260303, *, tsamp: float = 64e-6,
                     log2_min: int = 20, log2_max: int = 22, injected: float = 0.05,
                     snr: tuple = (8.0, 30.0)) -> BatchedWorkload:
    """Float N(0,1) series with lengths uniform in [2^log2_min, 2^log2_max]
    samples; a fraction `injected` carries a Gaussian pulse train (period
    log-uniform 0.1-10 s, duty 1-10 %, integrated S/N uniform in `snr`)."""
    rng = np.random.default_rng(seed)
    lengths = rng.integers(1 << log2_min, (1 << log2_max) + 1, nseries)
    period = np.exp(rng.uniform(np.log(0.1), np.log(10.0), nseries))
    phase0 = rng.uniform(0.0, 1.0, nseries)
    duty = rng.uniform(0.01, 0.10, nseries)
    sigma = (duty / 2.3548).astype(np.float32)
    target = rng.uniform(snr[0], snr[1], nseries)
    # integrated S/N of a Gaussian pulse train on unit noise: A * sqrt(N sigma sqrt(pi))
    amp = target / np.sqrt(lengths * sigma * np.sqrt(np.pi))
    amp[rng.random(nseries) >= injected] = 0.0
    return BatchedWorkload(nseries=nseries, lengths=lengths, tsamp=tsamp, seed=seed,
                           period=period, phase0=phase0, sigma=sigma,
                           amplitude=amp.astype(np.float32))


__all__ = ["BatchedWorkload", "CANDIDATE_OPTIONS", "DM_PEAK_DTYPE", "DmSegment", "GpuSearchPipeline",
           "HARMONIC_OPTIONS", "SynthFilterbankSpec", "SynthPulsar", "Workload", "batched_workload",
           "channel_frequencies", "large_workload", "lband_workload", "peak_records", "profile_snr", "pulse_amplitude",
           "search_candidates", "small_workload", "smearing_dm_segments"]

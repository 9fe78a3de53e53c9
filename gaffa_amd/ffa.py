"""Fast-folding-algorithm search of preprocessed time series (gaffa's ``gaffa.ffa``).

``make_riptide_plan`` and ``ffa_search`` keep gaffa's signatures, defaults
and error messages (python/gaffa/ffa.py:12-133); the search runs on the GPU
(``backend="cuda"`` or ``"gpu"``).
``ffa_search_batch`` searches many equal-length series with one plan.
"""

from __future__ import annotations

import numpy as np

from ._core import (
    FfaBatchedSearch,
    FfaPeak,
    FfaPlan,
    _check_ffa_search,
    _ffa_search_batch_gpu,
    _ffa_search_gpu_host,
    _ffa_transform_rows_gpu,
    _make_riptide_ffa_plan,
)

_GPU_BACKENDS = ("cuda", "gpu")

def make_riptide_plan(
    *,
    nsamples: int,
    tsamp: float,
    period_min: float,
    period_max: float,
    bins_min: int = 180,
    bins_max: int = 256,
    min_periods: int = 1,
    duty_cycle_max: float = 0.20,
    width_trial_spacing: float = 1.5,
    max_tasks: int = 1_000_000,
) -> FfaPlan:
    """Build a reusable riptide-compatible FFA plan for series of ``nsamples``."""
    return _make_riptide_ffa_plan(
        nsamples=nsamples,
        tsamp=tsamp,
        period_min=period_min,
        period_max=period_max,
        bins_min=bins_min,
        bins_max=bins_max,
        min_periods=min_periods,
        duty_cycle_max=duty_cycle_max,
        width_trial_spacing=width_trial_spacing,
        max_tasks=max_tasks,
    )


def make_octave_plan(
    *,
    nsamples: int,
    tsamp: float,
    period_min: float = 0.1,
    period_max: float = 10.0,
    bins: int = 128,
    duty_cycle_max: float = 0.20,
    width_trial_spacing: float = 1.5,
) -> FfaPlan:
    """Period-octave plan: downsample factors f0 * 2^k with f0 = period_min /
    (tsamp * bins) and bins in [bins, 2 * bins) per octave.  It is gaffa's
    riptide plan generator with bins_max = 2 * bins - 1, whose cycle growth
    (bins_max + 1) / bins_min is then exactly 2 (ffa_plan.cpp:90-171)."""
    return make_riptide_plan(nsamples=nsamples, tsamp=tsamp, period_min=period_min,
                             period_max=period_max, bins_min=bins, bins_max=2 * bins - 1,
                             duty_cycle_max=duty_cycle_max,
                             width_trial_spacing=width_trial_spacing)


def _check_backend(backend: str, device_id: int) -> None:
    """gaffa's backend rules and messages (python/gaffa/ffa.py:116-133): "cpu"
    takes no device; "cuda" (here HIP on the GPU, "gpu" is a synonym) takes a
    non-negative ordinal.  gaffa_amd does not ship gaffa's CPU path (its CPU
    restatement is test-only), so backend="cpu" is refused, not computed."""
    if backend == "cpu":
        if device_id != 0:
            raise ValueError("device_id is only valid with backend='cuda'")
        raise ValueError("backend='cpu' is gaffa's CPU path, which gaffa_amd does not provide; "
                         "use backend='cuda' (or 'gpu')")
    if backend not in _GPU_BACKENDS:
        raise ValueError("backend must be 'cpu' or 'cuda' ('gpu' is accepted for 'cuda')")
    if device_id < 0:
        raise ValueError("device_id must be non-negative")


def ffa_search(
    time_series,
    plan: FfaPlan,
    *,
    snr_threshold: float = 6.0,
    max_peaks: int | None = None,
    backend: str = "cpu",
    device_id: int = 0,
) -> list[FfaPeak]:
    """Search one preprocessed float32 series; peaks in descending peak order.

    The signature and default are gaffa's: ``backend="cpu"`` is gaffa's
    default and raises ValueError here, so pass ``backend="cuda"`` (or
    ``"gpu"``) to search on the GPU."""
    _check_ffa_search(time_series, plan, max_peaks)
    _check_backend(backend, device_id)
    return _ffa_search_gpu_host(time_series, plan, device_id=device_id,
                                snr_threshold=snr_threshold, max_peaks=max_peaks)


def ffa_search_batch(
    batch,
    plan: FfaPlan,
    *,
    snr_threshold: float = 6.0,
    max_peaks: int | None = None,
    device_id: int = 0,
    series_tile: int = 0,
) -> list[tuple[int, FfaPeak]]:
    """Search every row of a C-contiguous (nseries, nsamples) float32 array.

    Returns ``(series_index, peak)`` pairs sorted by series, then peak order.
    """
    _check_backend("cuda", device_id)
    return _ffa_search_batch_gpu(batch, plan, device_id=device_id, snr_threshold=snr_threshold,
                                 max_peaks=max_peaks, series_tile=series_tile)


def ffa_transform_rows(batch, plan: FfaPlan, task: int, *, device_id: int = 0) -> np.ndarray:
    """The first ``rows_eval`` final FFA rows of ``task`` for every series."""
    return _ffa_transform_rows_gpu(batch, plan, task, device_id=device_id)


__all__ = ["FfaBatchedSearch", "FfaPeak", "FfaPlan", "ffa_search", "ffa_search_batch",
           "ffa_transform_rows", "make_octave_plan", "make_riptide_plan"]

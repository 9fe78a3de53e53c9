"""Pulsar folding of dedispersed time series and spectra (gaffa's ``gaffa.pfold``).

``fold_profile``, ``fold_spectrum`` and ``fold`` keep gaffa's signatures and
results (python/gaffa/pfold.py:17-149); the folds run on the GPU
(fractional phase overlap, FP64 sums in gaffa's order).
"""

from __future__ import annotations

from ._core import (
    DedispersedResult,
    FoldedProfile,
    FoldResult,
    _fold_dedispersed_profile,
    _fold_dedispersed_spectrum,
)


def fold_profile(dedispersed, *, period: float, nbin: int, dm_index: int = 0,
                 device_id: int = 0) -> FoldedProfile:
    """Fold row ``dm_index`` of a dedispersed time series into a profile."""
    if dm_index < 0:
        raise ValueError("dm_index must be non-negative")
    return _fold_dedispersed_profile(dedispersed.data, tsamp=dedispersed.tsamp, period=period,
                                     nbin=nbin, dm_index=dm_index, device_id=device_id)


def fold_spectrum(spectrum, *, period: float, nbin: int, tsubint: float = 10.0,
                  nsubband: int | None = None, device_id: int = 0) -> FoldResult:
    """Fold a (nsamples, nchans) dedispersed dynamic spectrum into subints."""
    if nsubband is not None and nsubband <= 0:
        raise ValueError("nsubband must be positive")
    return _fold_dedispersed_spectrum(spectrum.data, tsamp=spectrum.tsamp, period=period,
                                      nbin=nbin, tsubint=tsubint,
                                      output_channels=0 if nsubband is None else nsubband,
                                      device_id=device_id)


def fold(data, *, period: float, nbin: int, dm_index: int = 0, tsubint: float = 10.0,
         nsubband: int | None = None):
    """Fold a dedispersed time series (profile) or dynamic spectrum (cube)."""
    if isinstance(data, DedispersedResult):
        return fold_profile(data, period=period, nbin=nbin, dm_index=dm_index)
    if hasattr(data, "data") and hasattr(data, "tsamp") and getattr(data.data, "ndim", 0) == 2:
        return fold_spectrum(data, period=period, nbin=nbin, tsubint=tsubint, nsubband=nsubband)
    raise TypeError("fold expects DedispersedResult or DedispersedSpectrum")


__all__ = ["FoldResult", "FoldedProfile", "fold", "fold_profile", "fold_spectrum"]

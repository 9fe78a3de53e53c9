"""Python-facing dedispersion API (gaffa's ``gaffa.dedispersion``, on the GPU).

Functions take a :class:`gaffa_amd.io.Filterbank` and return host-resident
NumPy results; ``backend`` is ``"gpu"`` (``"cuda"`` is accepted as an alias
so gaffa scripts run unchanged).
"""

from __future__ import annotations

from ._core import (
    DedispersedResult,
    DedispersedSpectrum,
    dedisperse_multi_dm,
    dedisperse_single_dm,
    dedisperse_spectrum,
    dedisperse_subband,
    dedispersion_delay_bins,
)


__all__ = [
    "DedispersedResult",
    "DedispersedSpectrum",
    "dedisperse_multi_dm",
    "dedisperse_single_dm",
    "dedisperse_spectrum",
    "dedisperse_subband",
    "dedispersion_delay_bins",
]

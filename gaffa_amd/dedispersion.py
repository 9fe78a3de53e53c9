"""Python-facing dedispersion API (gaffa's ``gaffa.dedispersion``, on the GPU).

Functions take a :class:`gaffa_amd.io.Filterbank` and return host-resident
NumPy results.  ``backend`` takes gaffa's names with gaffa's defaults
(``"cpu"`` for the single-DM, multi-DM and spectrum forms, ``"cuda"`` for
subband): ``"cuda"`` (or its synonym ``"gpu"``) runs on the GPU, and
``"cpu"``, gaffa's own CPU path, raises ValueError because gaffa_amd does
not ship it.
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

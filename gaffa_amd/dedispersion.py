"""Python-facing dedispersion API (gaffa's ``gaffa.dedispersion``, on the GPU).

Functions take a :class:`gaffa_amd.io.Filterbank` and return host-resident
NumPy results; ``backend`` is ``"gpu"`` (``"cuda"`` is accepted as an alias
so gaffa scripts run unchanged).
"""

from __future__ import annotations

from ._core import (
    DedispersedResult,
    dedisperse_multi_dm,
    dedisperse_single_dm,
    dedisperse_subband,
    dedispersion_delay_bins,
)


def dedisperse_spectrum(filterbank, *, dm, backend="gpu", **_):
    """Aligned single-DM dynamic spectrum.

    Not part of the search hot path (it feeds spectrum folding); gaffa's
    ``dedisperse_spectrum`` has no GPU counterpart here yet.
    """
    raise NotImplementedError("dedisperse_spectrum is not provided by gaffa_amd")


__all__ = [
    "DedispersedResult",
    "dedisperse_multi_dm",
    "dedisperse_single_dm",
    "dedisperse_spectrum",
    "dedisperse_subband",
    "dedispersion_delay_bins",
]

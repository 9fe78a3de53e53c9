"""gaffa search hot path on AMD Instinct MI355X.

Dedispersion, running-median detrending + normalisation, the FFA transform
and the boxcar S/N periodogram run as hand-written HIP kernels for gfx950,
reached through the C ABI in ``include/gaffa_hip.h`` (``libgaffa_hip.so``)
and gaffa's C++ plan/result types (``_core``).  The native libraries are
built in-tree by ``python -m gaffa_amd.build``; importing the package
without them fails: there is no CPU fallback.
"""

from __future__ import annotations

import os

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from . import _core  # noqa: F401  (loads libgaffa_hip.so through its rpath)
except ImportError as exc:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "gaffa_amd native extension is not built (libgaffa_hip.so / _core); "
        "run `python -m gaffa_amd.build` first"
    ) from exc

from ._core import cuda_device_count, gpu_device_count, gpu_runtime_version  # noqa: E402
from . import dedispersion, ffa, io, pipeline, scorer  # noqa: E402
from .scorer import CandidateScorer  # noqa: E402

HIP_LIBRARY = os.path.join(_HERE, "libgaffa_hip.so")

__all__ = [
    "CandidateScorer",
    "HIP_LIBRARY",
    "cuda_device_count",
    "dedispersion",
    "ffa",
    "gpu_device_count",
    "gpu_runtime_version",
    "io",
    "pipeline",
    "scorer",
]

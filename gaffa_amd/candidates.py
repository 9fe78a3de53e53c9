"""Candidate selection and harmonic flagging over DM peak lists.

Host code (C++, in ``_core``) with gaffa's types, defaults and ordering
(include/gaffa/candidate.h, harmonic.h).  Peaks from a device-resident search
stay in C++: ``GpuSearchPipeline.select_candidates`` clusters them directly.
"""

from __future__ import annotations

from ._core import (
    Candidate,
    HarmonicCandidate,
    HarmonicRelation,
    flag_harmonics,
    remove_harmonics,
    select_candidates,
)

__all__ = ["Candidate", "HarmonicCandidate", "HarmonicRelation", "flag_harmonics",
           "remove_harmonics", "select_candidates"]

"""Learned candidate scorer (no gaffa counterpart).

``CandidateScorer`` holds a small 1D-conv + MLP network (~175k parameters,
bfloat16 weights and activations, FP32 accumulation on MFMA) that scores a
search candidate from its folded [32 subint][64 phase] cube and a 128-trial
periodogram patch, both normalised to zero mean and unit variance.  Forward,
backward and Adam run on the GPU through libgaffa_hip's scorer kernels.
"""

from __future__ import annotations

from ._core import CandidateScorer

__all__ = ["CandidateScorer"]

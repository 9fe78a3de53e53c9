"""Write tests/golden/scorer_heldout.npz: a held-out candidate set for the
scorer regression test (needs a GPU: the candidates come from the search).

Filterbanks: the scorer's training configuration (ScorerDataConfig: 256
channels, 2^18 samples at 256 us, 256 DM trials, half noise-only, up to eight
injected pulsars of S/N 4-30) at a seed never used in training or threshold
selection.  Every candidate of the S/N >= 8 fold list is kept: its folded
cube (normalised per candidate, stored as int8 in steps of 1/16), its
periodogram patch (float16), S/N, label, owning pulsar and harmonic
relation.

    python tests/golden/make_scorer_heldout.py [--filterbanks 120]
"""

import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

HELDOUT_SEED = 20260719  # training used 20260707..20260709


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--filterbanks", type=int, default=120)
    ap.add_argument("--out", default=os.path.join(HERE, "scorer_heldout.npz"))
    args = ap.parse_args()
    from gaffa_amd import scorer as sc

    s = sc.candidate_set(args.filterbanks, HELDOUT_SEED,
                         log=lambda m: print(m, file=sys.stderr, flush=True))
    keep = s.snr >= 8.0
    cubes = np.clip(np.rint(s.cubes[keep] * 16.0), -127, 127).astype(np.int8)
    np.savez_compressed(
        args.out, cubes_q16=cubes, patches=s.patches[keep].astype(np.float16),
        labels=s.labels[keep], snr=s.snr[keep], pulsar=s.pulsar[keep],
        related=s.related[keep], filterbanks=np.int64(args.filterbanks),
        seed=np.int64(HELDOUT_SEED), candidates_all=np.int64(len(s)))
    print(f"{int(keep.sum())} of {len(s)} candidates (S/N >= 8) from {args.filterbanks} "
          f"filterbanks -> {args.out}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

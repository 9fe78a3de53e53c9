"""Learned candidate scorer (no gaffa counterpart).

``CandidateScorer`` holds a small 1D-conv + MLP network (~175k parameters,
bfloat16 weights and activations, FP32 accumulation on MFMA) that scores a
search candidate from its folded [32 subint][64 phase] cube (normalised to
zero mean and unit variance) and a 128-trial periodogram patch (mean
removed, in S/N units).  Forward,
backward and Adam run on the GPU through libgaffa_hip's scorer kernels.

Training data is made by the search itself: synthetic uint8 N(128, 16)
filterbanks, half of them noise only, the rest with injected pulsars, go
through the GPU pipeline (dedispersion, detrend, FFA, candidate selection);
each candidate is labelled by whether it matches an injection (its
frequency within a Fourier bin of 1/2, 1 or 2 times the pulsar's, its DM
error smearing the pulse by no more than its width or two DM trials);
candidates at other small harmonic ratios (n / d, n, d <= 8) are related
to a pulsar and left out of training and of the noise counts.  The plain pipeline folds every
candidate above the S/N threshold; the scored pipeline folds only those the
network keeps.  ``evaluate`` measures both on a held-out seed: recall of the
plain pipeline's detections (pulsars with a matching candidate at S/N >= 8),
the false-positive rate on noise candidates and the fraction of candidates
still folded.
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass, replace

import numpy as np

from ._core import CandidateScorer
from .pipeline import DISPERSION_S, Workload, channel_frequencies, pulse_amplitude, \
    smearing_dm_segments

__all__ = ["CandidateScorer", "CandidateSet", "ScorerDataConfig", "candidate_set", "evaluate",
           "load_weights", "WEIGHTS_PATH"]

WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                            "scorer_weights.npy")


@dataclass(frozen=True)
class ScorerDataConfig:
    """Filterbanks the training candidates come from."""
    nchans: int = 256
    fch1: float = 1520.0
    foff: float = -300.0 / 256
    tsamp: float = 256e-6
    nsamples: int = 1 << 18
    dm_max: float = 500.0
    ndm: int = 256
    pulsars_per_filterbank: int = 8
    noise_only_fraction: float = 0.5
    period_range: tuple = (0.1, 10.0)    # log-uniform
    dm_range: tuple = (0.0, 500.0)
    duty_range: tuple = (0.005, 0.10)
    snr_range: tuple = (4.0, 30.0)
    snr_threshold: float = 6.0

    def workload(self, seed: int, pulsars: list) -> Workload:
        freq = channel_frequencies(self.nchans, self.fch1, self.foff)
        segs = smearing_dm_segments(dm_max=self.dm_max, ndm=self.ndm, tsamp=self.tsamp,
                                    freq=freq)
        return Workload(
            name="scorer", nsamples=self.nsamples, nchans=self.nchans, fch1=self.fch1,
            foff=self.foff, tsamp=self.tsamp, seed=seed, pulsars=pulsars, dm_segments=segs,
            subband=False, median_seconds=1.0, period_min=0.1, period_max=10.0,
            duty_cycle_max=0.18, width_trial_spacing=2.0, snr_threshold=self.snr_threshold)

    def band_delay_per_dm(self) -> float:
        f = channel_frequencies(self.nchans, self.fch1, self.foff)
        return DISPERSION_S * (f.min() ** -2 - f.max() ** -2)


@dataclass
class CandidateSet:
    cubes: np.ndarray        # (n, 32, 64) float32
    patches: np.ndarray      # (n, 128) float32
    labels: np.ndarray       # (n,) float32, 1 = matches an injected pulsar
    snr: np.ndarray          # (n,) candidate S/N
    pulsar: np.ndarray       # (n,) index into pulsar arrays, -1 = none
    related: np.ndarray      # (n,) bool: matches a pulsar at a ratio n/d <= 8
    pulsar_snr: np.ndarray   # (npulsars,) injected S/N
    pulsar_period: np.ndarray
    pulsar_dm: np.ndarray
    filterbanks: int = 0
    seconds: float = 0.0

    def __len__(self) -> int:
        return len(self.labels)


def _draw_pulsars(rng, cfg: ScorerDataConfig) -> list:
    out = []
    for _ in range(cfg.pulsars_per_filterbank):
        period = float(np.exp(rng.uniform(*np.log(cfg.period_range))))
        dm = float(rng.uniform(*cfg.dm_range))
        duty = float(rng.uniform(*cfg.duty_range))
        snr = float(rng.uniform(*cfg.snr_range))
        out.append((period, dm, duty, pulse_amplitude(snr, cfg.nchans, cfg.nsamples, duty),
                    float(rng.uniform(0.0, 1.0)), snr))
    return out


def _ratios(max_n: int = 8) -> np.ndarray:
    """Harmonic ratios a candidate may sit at: reduced n / d, n, d <= max_n."""
    return np.array(sorted({n / d for n in range(1, max_n + 1) for d in range(1, max_n + 1)
                            if math.gcd(n, d) == 1}))


# A candidate is a detection of a pulsar at its period, twice it or half it;
# other small ratios mark it as related (a weaker harmonic), neither signal
# nor noise for training.
MAIN_RATIOS = np.array([0.5, 1.0, 2.0])


def _match(period, dm, pulsars, obs_seconds, band_delay_per_dm, dm_step, ratios) -> int:
    """Index of the injected pulsar this candidate belongs to, or -1: its
    frequency within one Fourier bin (1 / observation) of n / d times the
    pulsar's (n, d <= 8), its DM error smearing the pulse across the band by
    no more than the pulse width or two DM trials, whichever is larger."""
    f = 1.0 / period
    for k, (p_period, p_dm, duty, *_rest) in enumerate(pulsars):
        smear = abs(dm - p_dm) * band_delay_per_dm
        if smear > max(duty * p_period, 2.0 * dm_step * band_delay_per_dm):
            continue
        if np.min(np.abs(f - ratios / p_period)) * obs_seconds <= 1.0:
            return k
    return -1


def candidate_set(n_filterbanks: int, seed: int, cfg: ScorerDataConfig = ScorerDataConfig(),
                  *, device_id: int = 0, log=None) -> CandidateSet:
    """Candidates (features, labels, S/N) from `n_filterbanks` synthetic
    filterbanks searched by the GPU pipeline."""
    import time

    rng = np.random.default_rng(seed)
    template = cfg.workload(seed, [])
    pipe = template.pipeline(device_id=device_id, ffa_workspace_bytes=8 << 30)
    pipe.enable_periodogram(True)
    obs = pipe.out_nsamples * cfg.tsamp
    kdm = cfg.band_delay_per_dm()
    dms = np.asarray(pipe.dms)
    ratios = _ratios()
    cubes, patches, labels, snrs, owner, related = [], [], [], [], [], []
    p_snr, p_period, p_dm = [], [], []
    t0 = time.perf_counter()
    for i in range(n_filterbanks):
        pulsars = [] if rng.random() < cfg.noise_only_fraction else _draw_pulsars(rng, cfg)
        w = replace(template, seed=int(rng.integers(1, 2 ** 62)),
                    pulsars=[p[:5] for p in pulsars])
        pipe.generate(w.synth_spec())
        pipe.run()
        cands = pipe.select_candidates()
        base = len(p_snr)
        for p in pulsars:
            p_period.append(p[0])
            p_dm.append(p[1])
            p_snr.append(p[5])
        if not cands:
            continue
        c, q = pipe.candidate_features(cands)
        cubes.append(c)
        patches.append(q)
        for cand in cands:
            step = dms[min(cand.dm_index + 1, len(dms) - 1)] - dms[max(cand.dm_index - 1, 0)]
            k = _match(cand.period, cand.dm, pulsars, obs, kdm, step / 2, MAIN_RATIOS)
            rel = k >= 0 or _match(cand.period, cand.dm, pulsars, obs, kdm, step / 2, ratios) >= 0
            labels.append(1.0 if k >= 0 else 0.0)
            owner.append(base + k if k >= 0 else -1)
            related.append(rel)
            snrs.append(cand.snr)
        if log is not None and (i + 1) % 50 == 0:
            log(f"[scorer] {i + 1}/{n_filterbanks} filterbanks, {len(labels)} candidates, "
                f"{time.perf_counter() - t0:.1f} s")
    return CandidateSet(
        cubes=np.concatenate(cubes) if cubes else np.zeros((0, 32, 64), np.float32),
        patches=np.concatenate(patches) if patches else np.zeros((0, 128), np.float32),
        labels=np.asarray(labels, np.float32), snr=np.asarray(snrs, np.float32),
        pulsar=np.asarray(owner, np.int64), related=np.asarray(related, bool),
        pulsar_snr=np.asarray(p_snr),
        pulsar_period=np.asarray(p_period), pulsar_dm=np.asarray(p_dm),
        filterbanks=n_filterbanks, seconds=time.perf_counter() - t0)


def balanced_indices(labels: np.ndarray, seed: int, related=None) -> np.ndarray:
    """An equal draw of positives and negatives (all of the smaller class),
    shuffled; negatives related to a pulsar (weaker harmonics) are left
    out."""
    rng = np.random.default_rng(seed)
    pos = np.nonzero(labels > 0)[0]
    neg = np.nonzero((labels == 0) & (~related if related is not None else True))[0]
    k = min(len(pos), len(neg))
    idx = np.concatenate([rng.choice(pos, k, replace=False), rng.choice(neg, k, replace=False)])
    rng.shuffle(idx)
    return idx


def evaluate(scores: np.ndarray, s: CandidateSet, threshold: float,
             detection_snr: float = 8.0) -> dict:
    """Plain pipeline vs scored pipeline on one candidate set.

    A pulsar is detected by the plain pipeline when a candidate matching it
    has S/N >= detection_snr; by the scored pipeline when such a candidate
    also scores >= threshold.  FPR: noise candidates kept / noise
    candidates.  folded: candidates kept / candidates (the plain pipeline
    folds them all)."""
    keep = scores >= threshold
    strong = s.snr >= detection_snr
    plain = set(s.pulsar[(s.pulsar >= 0) & strong].tolist())
    scored = set(s.pulsar[(s.pulsar >= 0) & strong & keep].tolist())
    noise = (s.labels == 0) & ~s.related
    return {
        "plain_detections": len(plain),
        "scored_detections": len(plain & scored),
        "recall": len(plain & scored) / max(1, len(plain)),
        "fpr": float(np.mean(keep[noise])) if noise.any() else 0.0,
        "folded_fraction": float(np.mean(keep)) if len(keep) else 0.0,
        "candidates": int(len(keep)),
    }


def recall_threshold(scores: np.ndarray, s: CandidateSet, recall: float = 0.99,
                     detection_snr: float = 8.0) -> float:
    """Largest threshold keeping `recall` of the plain pipeline's detections
    (pulsars with a matching candidate at S/N >= detection_snr): each such
    pulsar survives if its best-scoring strong candidate does (chosen on a
    validation set)."""
    best: dict = {}
    for i in np.nonzero((s.pulsar >= 0) & (s.snr >= detection_snr))[0]:
        k = int(s.pulsar[i])
        best[k] = max(best.get(k, -np.inf), float(scores[i]))
    if not best:
        return -np.inf
    vals = np.sort(np.array(list(best.values())))
    return float(vals[int(np.floor((1.0 - recall) * len(vals)))])


def budget_recall(scores: np.ndarray, s: CandidateSet, fraction: float) -> float:
    """Pulsars found (any matching candidate, any S/N) when only the best
    `fraction` of candidates by `scores` are folded, over the pulsars the
    whole list finds."""
    n = int(round(fraction * len(scores)))
    top = np.argsort(-scores, kind="stable")[:n]
    found = set(s.pulsar[top][s.pulsar[top] >= 0].tolist())
    total = set(s.pulsar[s.pulsar >= 0].tolist())
    return len(found) / max(1, len(total))


def load_weights(scorer: CandidateScorer, path: str = WEIGHTS_PATH) -> CandidateScorer:
    """Load parameters saved by ``main`` (float32 .npy, layout order)."""
    scorer.parameters = np.load(path, allow_pickle=False).astype(np.float32)
    return scorer


def load_trained(path: str = WEIGHTS_PATH, *, capacity: int = 1024, device_id: int = 0):
    """The committed scorer and its keep threshold (the logit chosen on the
    validation set, stored beside the weights)."""
    import json

    with open(os.path.splitext(path)[0] + ".json") as fh:
        meta = json.load(fh)
    net = load_weights(CandidateScorer(capacity=capacity, device_id=device_id), path)
    return net, float(meta["threshold"])


def snr_cut_baseline(s: CandidateSet, recall: float, detection_snr: float = 8.0) -> dict:
    """The method without the network: the highest S/N cut that keeps
    `recall` of the plain pipeline's detections, and what it folds."""
    best = None
    for cut in np.unique(s.snr):
        r = evaluate(s.snr, s, float(cut), detection_snr)
        if r["recall"] >= recall:
            best = (float(cut), r)
        else:
            break
    if best is None:
        return {"snr_cut": None}
    return {"snr_cut": best[0], **best[1]}


def train(train_set: CandidateSet, *, epochs: int = 6, batch: int = 256, lr: float = 1e-3,
          seed: int = 1, capacity: int = 1024, device_id: int = 0):
    """A scorer trained on a class-balanced draw of the set; returns
    (scorer, per-epoch loss)."""
    idx = balanced_indices(train_set.labels, seed, train_set.related)
    scorer = CandidateScorer(capacity=capacity, device_id=device_id)
    scorer.initialize(seed)
    losses = scorer.fit(train_set.cubes[idx], train_set.patches[idx], train_set.labels[idx],
                        epochs=epochs, batch=batch, learning_rate=lr, seed=seed)
    return scorer, losses


def main(argv=None) -> int:
    import argparse
    import json
    import sys
    import time

    ap = argparse.ArgumentParser(description="train the candidate scorer on injections and "
                                             "compare it with the plain pipeline")
    ap.add_argument("--train-filterbanks", type=int, default=400)
    ap.add_argument("--val-filterbanks", type=int, default=60)
    ap.add_argument("--eval-filterbanks", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=5)
    ap.add_argument("--seed", type=int, default=20260707)
    ap.add_argument("--recall", type=float, default=0.99)
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    tr = candidate_set(args.train_filterbanks, args.seed, log=log)
    log(f"[scorer] train: {len(tr)} candidates ({int(tr.labels.sum())} positive) in "
        f"{tr.seconds:.1f} s")
    va = candidate_set(args.val_filterbanks, args.seed + 1, log=log)
    ev = candidate_set(args.eval_filterbanks, args.seed + 2, log=log)
    t1 = time.perf_counter()
    scorer, losses = train(tr, epochs=args.epochs, seed=args.seed)
    t2 = time.perf_counter()
    log(f"[scorer] losses {losses}, {t2 - t1:.1f} s")
    thr = recall_threshold(scorer.forward(va.cubes, va.patches), va, args.recall)
    z = scorer.forward(ev.cubes, ev.patches)
    res = evaluate(z, ev, thr)
    base = snr_cut_baseline(ev, res["recall"])
    # the same folding budget spent by S/N rank and by score rank
    budget = base.get("folded_fraction") or 0.0
    ranking = {"budget": budget, "pulsars_by_snr": budget_recall(ev.snr, ev, budget),
               "pulsars_by_score": budget_recall(z, ev, budget)}
    # network and S/N cut together: fold what both keep
    both = np.where(ev.snr >= 8.0, z, -np.inf)
    combined = evaluate(both, ev, thr)
    out = {
        "train_candidates": len(tr), "train_positive": int(tr.labels.sum()),
        "train_filterbanks": tr.filterbanks, "data_seconds": t1 - t0, "train_seconds": t2 - t1,
        "epoch_loss": losses, "threshold": thr, "heldout": res, "snr_cut_at_equal_recall": base,
        "equal_budget": ranking, "scored_and_snr8": combined,
        "heldout_seed": args.seed + 2,
    }
    print(json.dumps(out), flush=True)
    if args.out:
        np.save(args.out, scorer.parameters)
        with open(os.path.splitext(args.out)[0] + ".json", "w") as fh:
            json.dump({"threshold": thr, "recall_target": args.recall, "seed": args.seed,
                       "train_filterbanks": args.train_filterbanks, "epochs": args.epochs,
                       "data": "ScorerDataConfig defaults", "heldout": res}, fh, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

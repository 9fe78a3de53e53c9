"""Candidate selection and harmonic flagging: the product's host code against
the reference's known answers (test_candidate.cpp, test_harmonic.cpp, restated)
and against the reference build on randomized peak lists (records equal)."""

import numpy as np
import pytest

from oracle.oracle import CANDIDATE_DTYPE, DM_PEAK_DTYPE


def peak(dm, dm_index, frequency, width, snr):
    p = np.zeros(1, DM_PEAK_DTYPE)
    p["dm"], p["dm_index"] = dm, dm_index
    q = p["peak"]
    q["period"], q["frequency"], q["width"] = 1.0 / frequency, frequency, width
    q["duty_cycle"], q["width_index"], q["bins"], q["snr"] = width / 100.0, width - 1, 100, snr
    p["peak"] = q
    return p


def select(core, peaks, obs=100.0, radius=0.1, dm_radius=1, across=True, cap=0):
    raw = core._select_candidates_records(np.concatenate(peaks) if isinstance(peaks, list)
                                          else peaks, obs, radius, dm_radius, across, cap)
    return np.frombuffer(raw, CANDIDATE_DTYPE)


def test_candidate_known_answers(core):
    c = select(core, [peak(10.0, 0, 1.0, 2, 7.0), peak(11.0, 1, 1.0005, 2, 10.0),
                      peak(12.0, 2, 1.0009, 2, 8.0)])
    assert len(c) == 1 and c[0]["peak_count"] == 3 and c[0]["dm_index"] == 1
    assert (c[0]["dm_index_min"], c[0]["dm_index_max"]) == (0, 2)
    assert c[0]["snr"] == np.float32(10.0)
    assert (c[0]["frequency_min"], c[0]["frequency_max"]) == (1.0, 1.0009)
    # frequency gap 0.002 Hz > 0.1 / 100 s
    c = select(core, [peak(10.0, 0, 1.0, 2, 7.0), peak(10.0, 0, 1.002, 2, 8.0)])
    assert len(c) == 2 and c[0]["snr"] == np.float32(8.0)
    # DM gap of 3 trials > radius 1
    c = select(core, [peak(10.0, 0, 1.0, 2, 7.0), peak(13.0, 3, 1.0005, 2, 9.0)])
    assert len(c) == 2 and c[0]["snr"] == np.float32(9.0)
    two_widths = [peak(10.0, 0, 1.0, 2, 7.0), peak(11.0, 1, 1.0005, 4, 9.0)]
    c = select(core, two_widths, across=False)
    assert len(c) == 2 and c[0]["snr"] == np.float32(9.0)
    c = select(core, two_widths)
    assert len(c) == 1 and c[0]["peak_count"] == 2 and c[0]["width"] == 4
    assert len(select(core, np.zeros(0, DM_PEAK_DTYPE))) == 0
    with pytest.raises(ValueError, match="observation_seconds"):
        select(core, two_widths, obs=0.0)
    with pytest.raises(ValueError, match="frequency_cluster_radius"):
        select(core, two_widths, radius=-1.0)
    bad = peak(10.0, 0, 1.0, 1, 5.0)
    bad["peak"]["frequency"] = np.inf
    with pytest.raises(ValueError, match="finite"):
        select(core, [bad])


def random_peaks(rng, n):
    """Clustered peak list: sources at a few frequencies and DMs, each spread
    over neighbouring trials and widths, plus noise peaks and exact S/N ties."""
    out = np.zeros(n, DM_PEAK_DTYPE)
    src_f = rng.uniform(0.2, 8.0, 40)
    src_dm = rng.integers(0, 200, 40)
    k = rng.integers(0, 40, n)
    noise = rng.random(n) < 0.3
    f = np.where(noise, rng.uniform(0.1, 10.0, n), src_f[k] * (1 + rng.normal(0, 2e-5, n)))
    dmi = np.where(noise, rng.integers(0, 256, n), np.clip(src_dm[k] + rng.integers(-3, 4, n),
                                                           0, 255))
    width = rng.choice([1, 2, 4, 8, 16], n)
    out["dm_index"] = dmi
    out["dm"] = dmi * 0.5
    pk = out["peak"]
    pk["frequency"], pk["period"] = f, 1.0 / f
    pk["width"], pk["duty_cycle"], pk["width_index"] = width, width / 256.0, np.log2(width)
    pk["bins"] = 256
    pk["phase"] = rng.integers(0, 256, n)
    pk["shift"] = rng.integers(0, 100, n)
    pk["snr"] = np.round(rng.uniform(6.0, 30.0, n) * 4) / 4  # many exact ties
    out["peak"] = pk
    return out


@pytest.mark.parametrize("across,dm_radius,radius", [(True, 1, 0.1), (False, 2, 0.5),
                                                     (True, 0, 0.0)])
def test_candidates_match_reference_build(core, ref, across, dm_radius, radius):
    rng = np.random.default_rng(dm_radius * 7 + int(across))
    peaks = random_peaks(rng, 20000)
    got = select(core, peaks, obs=268.0, radius=radius, dm_radius=dm_radius, across=across)
    want = ref.select_candidates(peaks, 268.0, radius, dm_radius, across)
    assert len(got) == len(want) > 100
    assert np.array_equal(got, want)  # every field (struct padding aside)
    capped = select(core, peaks, obs=268.0, radius=radius, dm_radius=dm_radius, across=across,
                    cap=17)
    assert np.array_equal(capped, want[:17])


def harmonics(core, cands, **kw):
    opts = dict(max_harmonic=16, denominator_max=5, frequency_tolerance_bins=1.5,
                phase_distance_max=1.0, dm_distance_max=3.0, use_snr_consistency=False,
                snr_distance_max=3.0)
    opts.update(kw)
    raw, rel = core._flag_harmonics_records(np.ascontiguousarray(cands, CANDIDATE_DTYPE), 100.0,
                                            1000.0, 1500.0, *opts.values())
    return np.frombuffer(raw, CANDIDATE_DTYPE), rel


def cand(dm, dm_index, frequency, snr):
    c = np.zeros(1, CANDIDATE_DTYPE)
    c["dm"], c["dm_index"], c["period"], c["frequency"] = dm, dm_index, 1.0 / frequency, frequency
    c["width"], c["duty_cycle"], c["snr"], c["peak_count"] = 2, 0.02, snr, 1
    c["dm_index_min"] = c["dm_index_max"] = dm_index
    c["frequency_min"] = c["frequency_max"] = frequency
    c["best_peak"] = peak(dm, dm_index, frequency, 2, snr)
    return c


def test_harmonic_known_answers(core):
    _, rel = harmonics(core, np.concatenate([cand(10.0, 0, 1.0, 30.0), cand(10.0, 0, 2.0, 20.0)]))
    assert not rel[0].is_harmonic and rel[1].is_harmonic and rel[1].parent_index == 0
    assert (rel[1].numerator, rel[1].denominator, rel[1].frequency_error_bins) == (2, 1, 0.0)
    _, rel = harmonics(core, np.concatenate([cand(10.0, 0, 1.0, 30.0), cand(10.0, 0, 1.5, 20.0)]),
                       denominator_max=4)
    assert rel[1].is_harmonic and (rel[1].numerator, rel[1].denominator) == (3, 2)
    for f in (10.0 / 9.0, 2.02, 2.01):  # high denominator, frequency, phase tolerance
        _, rel = harmonics(core, np.concatenate([cand(10.0, 0, 1.0, 30.0),
                                                 cand(10.0, 0, f, 20.0)]))
        assert not rel[1].is_harmonic
    _, rel = harmonics(core, np.concatenate([cand(10.0, 0, 1.0, 30.0),
                                             cand(1000010.0, 100, 2.0, 20.0)]))
    assert not rel[1].is_harmonic  # DM smearing distance
    _, rel = harmonics(core, np.concatenate([cand(10.0, 0, 1.0, 30.0), cand(10.0, 0, 1.5, 20.0),
                                             cand(10.0, 0, 2.25, 10.0)]),
                       max_harmonic=3, denominator_max=2)
    assert rel[1].is_harmonic and not rel[2].is_harmonic  # flagged ones never parent
    with pytest.raises(ValueError, match="max_harmonic"):
        harmonics(core, cand(10.0, 0, 1.0, 5.0), max_harmonic=1)


def test_harmonics_match_reference_build(core, ref):
    rng = np.random.default_rng(3)
    peaks = random_peaks(rng, 6000)
    cands = select(core, peaks, obs=100.0)
    # fold in exact harmonics of the strongest candidates
    extra = cands[:30].copy()
    extra["frequency"] = extra["frequency"] * rng.choice([2.0, 3.0, 1.5, 0.5, 4.0 / 3.0], 30)
    extra["period"] = 1.0 / extra["frequency"]
    extra["snr"] = extra["snr"] * 0.6
    allc = np.concatenate([cands, extra])
    got, rel = harmonics(core, allc)
    want, wrel = ref.flag_harmonics(allc, 100.0, 1000.0, 1500.0)
    assert np.array_equal(got, want)
    assert [r.is_harmonic for r in rel] == [bool(x) for x in wrel["is_harmonic"]]
    flagged = [i for i, r in enumerate(rel) if r.is_harmonic]
    assert len(flagged) >= 10
    for i in flagged:
        r, w = rel[i], wrel[i]
        assert (r.parent_index, r.numerator, r.denominator) == \
            (w["parent_index"], w["numerator"], w["denominator"])
        assert (r.frequency_error_bins, r.phase_distance, r.dm_distance, r.expected_snr,
                r.snr_distance) == (w["frequency_error_bins"], w["phase_distance"],
                                    w["dm_distance"], w["expected_snr"], w["snr_distance"])


@pytest.mark.parametrize("across,dm_radius,radius,grid", [(True, 50, 0.1, True), (False, 2, 0.5, True),
                                                          (True, 1, 0.1, False), (True, 0, 0.0, True)])
def test_ranked_selection_equals_general(core, across, dm_radius, radius, grid):
    """The pipeline's ranked fast path (counting sorts over dense period ranks)
    gives the same candidate records as the general selection, with many
    peaks sharing a trial period (as a search's do), DM and width."""
    rng = np.random.default_rng(31 + dm_radius)
    peaks = random_peaks(rng, 30000)
    if grid:  # frequencies on a trial grid: many peaks share each period
        trials = np.sort(rng.uniform(0.1, 10.0, 3000))
        f = trials[np.searchsorted(trials, peaks["peak"]["frequency"]).clip(0, 2999)]
        pk = peaks["peak"]
        pk["frequency"], pk["period"] = f, 1.0 / f
        peaks["peak"] = pk
    _, rank = np.unique(peaks["peak"]["period"], return_inverse=True)
    rank = rank.astype(np.uint32)
    want = select(core, peaks, obs=268.0, radius=radius, dm_radius=dm_radius, across=across)
    raw = core._select_candidates_ranked_records(peaks, rank, 268.0, radius, dm_radius, across, 0)
    got = np.frombuffer(raw, CANDIDATE_DTYPE)
    assert len(got) == len(want) > 10
    assert np.array_equal(got, want)
    raw = core._select_candidates_ranked_records(peaks, rank, 268.0, radius, dm_radius, across, 7)
    assert np.array_equal(np.frombuffer(raw, CANDIDATE_DTYPE), want[:7])

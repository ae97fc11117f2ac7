"""CPU tests of the long-baseline batch's host side: the star generator's
laws (sector counts, the one-year gap, injection kinds), the 10-minute
binning that feeds the coarse periodogram, the baseline-scaled feature bins
and the keep rule."""
import math

import numpy as np

import fbls_amd as F
from fbls_amd import longbase as L
from fbls_amd import proposal as PR
from fbls_amd import synth


def test_generator_layout_and_kinds():
    kinds, sectors = [], []
    for i in range(60):
        (t, f, e), kw = synth.long_baseline_star(i, seed0=7_000_000)
        k = kw["sectors"]
        assert 1 <= k <= 26
        sectors.append(k)
        kinds.append(kw["kind"])
        # two seasons of ceil(k/2) and floor(k/2) contiguous sectors
        first = (k + 1) // 2
        span = t[-1] - t[0]
        want = k * 27.4 + (365.0 if k > 1 else 0.0)
        assert want - 1.5 < span < want, (k, span)
        if k > 1:
            gaps = np.diff(t)
            assert gaps.max() > 365.0
            j = int(np.argmax(gaps))
            assert abs(t[j + 1] - t[0] - (first * 27.4 + 365.0)) < 1e-6
        assert np.all(e == kw["sigma"])
        if kw["kind"] == "none":
            assert kw.get("depth", 0.0) == 0.0
        else:
            assert 0.5 <= kw["period"] <= 100.0
            assert 1.0 / 24 - 1e-12 <= kw["duration"] <= 8.0 / 24 + 1e-12
        if kw["kind"] == "eb":
            assert 0.1 * kw["depth"] <= kw["secondary_depth"] <= 0.9 * kw["depth"]
    assert set(kinds) <= {"none", "transit", "eb"}
    assert 0.15 < kinds.count("transit") / len(kinds) < 0.45


def test_binned_state_preserves_window_sums():
    (t, f, e), _ = synth.long_baseline_star(3, seed0=7_000_000)
    cfg = L.search_config()
    s = F.family_state(t, f - 1.0, e, cfg.durations_days(), cfg.correlated_noise)
    b = L.binned_state(s)
    assert len(b.t) < len(s.t) / 4
    assert math.isclose(b.v.sum(), s.v.sum(), abs_tol=1e-9 * np.abs(s.v).sum())
    assert math.isclose(b.w.sum(), s.w.sum(), rel_tol=1e-12)
    assert np.all(np.diff(b.t) > 0)
    # every bin's time lies inside its 10-minute cell and sums its points
    width = 10.0 / 1440.0
    cell = np.floor(s.t / width)
    np.testing.assert_array_equal(np.unique(cell), np.floor(b.t / width))
    first = np.searchsorted(cell, np.unique(cell))
    np.testing.assert_allclose(b.w, np.add.reduceat(s.w, first), rtol=1e-12)
    # a window's sums over whole cells are unchanged
    lo, hi = 100.0, 101.5
    ms, mb = (s.t >= lo) & (s.t < hi), (b.t >= lo) & (b.t < hi)
    assert math.isclose(s.v[ms].sum(), b.v[mb].sum(), rel_tol=1e-9, abs_tol=1e-12)


def test_feature_bins_scale_with_baseline():
    pc = PR.ProposalConfig()
    ref = math.log(pc.period_max / pc.period_min) / pc.bins
    for t_last in (27.39, 420.0, 1077.0):
        c, e = L.feature_bins(t_last, pc)
        assert len(e) == len(c) + 1 and e[0] == 0.0 and e[-1] == np.inf
        step = np.diff(np.log(e[1:-1]))
        assert np.allclose(step, step[0], rtol=1e-9)
        assert step[0] <= ref * 27.4 / max(t_last, 27.4) * (1 + 1e-12)
        # every period of the search range falls in exactly one bin
        p = np.exp(np.linspace(math.log(0.5), math.log(min(100.0, t_last)), 1000))
        k = np.searchsorted(e, p, side="right") - 1
        assert np.all((k >= 0) & (k < len(c)))
        assert np.all((e[k] <= p) & (p < e[k + 1]))
        assert np.all((e[:-1][1:] < c[1:]) & (c[:-1] < e[1:][:-1]))
    # the single-sector star keeps the network's resolution
    c1, _ = L.feature_bins(27.4, pc)
    assert abs(np.log(c1[1] / c1[0]) - ref) / ref < 0.01


def test_keep_rule_dilates():
    p = np.zeros(20)
    p[5], p[0], p[19] = 0.5, 0.2, 0.3
    k = L.keep_from(p, 2, 0.25)
    assert list(np.flatnonzero(k)) == [3, 4, 5, 6, 7, 17, 18, 19]
    assert not L.keep_from(np.zeros(8), 2, 0.1).any()


def test_keep_mask_adds_raw_peaks_and_falls_back():
    p = np.zeros(30)
    p[5] = 0.5
    peaks = np.full(30, -np.inf, dtype=np.float32)
    peaks[20], peaks[25] = 90.0, 40.0
    # the network's bin alone, then with the periodogram's top bin
    assert list(np.flatnonzero(L.keep_mask(p, peaks, 1, (0.25, 0, (0.0, 0.0), 0)))) == [4, 5, 6]
    k = L.keep_mask(p, peaks, 1, (0.25, 0, (0.0, 0.0), 1))
    assert list(np.flatnonzero(k)) == [4, 5, 6, 19, 20, 21]
    k = L.keep_mask(p, peaks, 0, (0.25, 0, (0.0, 0.0), 2))
    assert list(np.flatnonzero(k)) == [5, 20, 25]
    # a three-element point is the same rule without raw bins
    assert (L.keep_mask(p, peaks, 1, (0.25, 0, (0.0, 0.0)))
            == L.keep_from(p, 1, 0.25)).all()
    # the largest probability inside the fallback band: every bin
    assert L.keep_mask(p, peaks, 1, (0.25, 0, (0.4, 0.6), 1)).all()
    assert not L.keep_mask(p, peaks, 1, (0.25, 0, (0.6, 0.9), 0))[10:].any()


def test_operating_point_reads_raw_topk():
    base = {"tau": np.asarray(0.05), "topk": np.asarray(8), "fallback": np.zeros(2)}
    assert L.operating_point(base) == (0.05, 8, (0.0, 0.0), 0)
    assert L.operating_point(dict(base, raw_topk=np.asarray(8)))[3] == 8


def test_merge_sweeps_sums_disjoint_sets():
    def part(stars, full, pruned, sec, kept, periods):
        return {"tau": 0.2, "topk": 0, "fallback": [0.0, 0.0], "raw_topk": 8,
                "stars": stars, "full_recovered": full, "pruned_recovered": pruned,
                "eb_full_recovered": 1, "eb_pruned_recovered": 1, "fallback_stars": 0,
                "templates_full": 10, "templates_pruned": 1, "seconds_full_scan": sec,
                "seconds_proposal": 1.0, "seconds_pruned_scan": 1.0,
                "kept_period_fraction": kept, "periods_full": periods}
    m = L.merge_sweeps([[part(10, 4, 4, 100.0, 0.01, 1000)],
                        [part(30, 6, 5, 300.0, 0.03, 3000)]])[0]
    assert (m["stars"], m["full_recovered"], m["pruned_recovered"]) == (40, 10, 9)
    assert m["recall"] == 0.9
    assert abs(m["kept_period_fraction"] - 0.025) < 1e-15
    assert m["speedup"] == (40 * 3600 / 4.0) / (40 * 3600 / 400.0)

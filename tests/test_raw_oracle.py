"""Pins the raw-mode half of the oracle (oracle/raw.c).

* fit_segment_noise_ex without binning and groups equals the pinned
  prepared-curve fit; binning follows its closed forms (inverse-variance
  means, 1/sqrt(sum w) errors) and never merges groups.
* biweight_flatten: constant curves flatten to zero, a slow trend is
  removed, an isolated outlier keeps its excess (the biweight rejects it
  from the location), and the location agrees with a direct numpy
  restatement on small windows.
* prepare_families: the reference's window-grouping test (curve.rs:609-616)
  and the clip / edge-trim rules.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import fbls_amd as F  # noqa: E402
import oracle as O  # noqa: E402


def test_fit_without_binning_matches_prepared_fit():
    rng = np.random.default_rng(1)
    t = np.sort(rng.uniform(0, 10, 3000)) + 2458000.0
    f = 1 + rng.normal(0, 1e-3, t.size)
    e = np.full(t.size, 1e-3)
    a = O.fit_segment_noise_ex(t, f, e, None, 0.0)
    b = O.fit_segment_noise(t, f, e)
    for k in ("time", "deficit", "variance"):
        np.testing.assert_array_equal(a[k], b[k])
    assert a["origin"] == b["origin"] and a["inflation"] == b["inflation"]


def test_binning_closed_form_and_groups():
    # two groups interleaved in time: bins never merge across groups
    t = np.array([0.0, 0.001, 0.002, 0.003, 0.0041, 0.0051, 0.3, 0.301, 0.31, 0.6, 0.601,
                  0.602, 0.9, 0.901]) + 100.0
    g = np.array([0, 0, 0, 0, 1, 1, 0, 0, 0, 1, 1, 1, 0, 0])
    f = np.array([1.0, 1.02, 0.98, 1.01, 2.0, 2.02, 1.0, 0.99, 1.01, 2.0, 2.02, 1.98, 1.0,
                  1.01])
    e = np.array([0.01, 0.02, 0.01, 0.01, 0.02, 0.02, 0.01, 0.01, 0.01, 0.02, 0.02, 0.02,
                  0.01, 0.01])
    r = O.fit_segment_noise_ex(t, f, e, g, bin_minutes=10.0)
    # 10-minute bins: group 0 has bins at 0, 0.3, 0.31, 0.9; group 1 at 0.004, 0.6
    assert len(r["time"]) == 6
    # first bin of group 0: inverse-variance mean of normalised flux
    m0 = np.sort(f[g == 0])[(np.sum(g == 0) - 1) // 2]
    sel = (g == 0) & (t < 100.01)
    w = 1 / (e[sel] / m0) ** 2
    want_f = np.sum((f[sel] / m0 - 1) * w) / w.sum()
    want_t = np.sum((t[sel] - t.min()) * w) / w.sum()
    k = int(np.argmin(np.abs(r["time"] - want_t)))
    assert r["time"][k] == pytest.approx(want_t, abs=1e-15)
    assert -r["deficit"][k] == pytest.approx(want_f, rel=1e-12, abs=1e-15)
    assert r["segments"] == 5  # g0 g1 g0 g0 g1 g0: every group change starts a segment


def test_biweight_properties():
    t = np.arange(3000) / 144.0
    assert np.all(O.biweight_flatten(t, np.zeros(t.size), 0.5) == 0.0)
    trend = 2e-3 * np.sin(t / 3.0)
    flat = O.biweight_flatten(t, trend, 0.5)
    assert np.max(np.abs(flat[200:-200])) < 2e-5
    rel = np.zeros(t.size)
    rel[1500] = 0.05
    flat = O.biweight_flatten(t, rel, 0.5)
    assert flat[1500] == pytest.approx(0.05, rel=1e-9)
    assert np.max(np.abs(np.delete(flat, 1500))) < 1e-12


def _biweight_direct(t, rel, window, c=5.0, iterations=10):
    f = 1 + rel
    out = np.empty(t.size)
    for i in range(t.size):
        lo = np.searchsorted(t, t[i] - window / 2, side="left")
        hi = np.searchsorted(t, t[i] + window / 2, side="right")
        v = f[lo:hi]
        loc = np.sort(v)[(len(v) - 1) // 2]
        for _ in range(iterations):
            dev = np.sort(np.abs(v - loc))[(len(v) - 1) // 2]
            scale = max(c * dev, 1e-12)
            u = (v - loc) / scale
            m = np.abs(u) < 1
            w = (1 - u[m] ** 2) ** 2
            if w.sum() > 0:
                loc = (w * v[m]).sum() / w.sum()
        out[i] = f[i] / loc - 1
    return out


def test_biweight_matches_direct_restatement():
    rng = np.random.default_rng(3)
    t = np.sort(rng.uniform(0, 3, 400))
    t[200:] += 1.0  # a gap splits the segments
    rel = rng.normal(0, 1e-3, t.size) + 1e-3 * t
    got = O.biweight_flatten(t, rel, 0.3)
    a = _biweight_direct(t[:200], rel[:200], 0.3)
    b = _biweight_direct(t[200:], rel[200:], 0.3)
    np.testing.assert_allclose(got, np.concatenate([a, b]), rtol=0, atol=1e-14)


def test_prepare_families_windows_and_rules():
    # curve.rs:609-616: durations 0.5, 1, 2, 12, 40 h -> windows 0.4 (three
    # durations), 1.5, 3.5
    d = np.array([0.5, 1.0, 2.0, 12.0, 40.0]) / 24.0
    rng = np.random.default_rng(5)
    t = np.arange(4000) / 720.0 + 2458000.0
    f = 1 + rng.normal(0, 5e-4, t.size)
    f[1000] += 0.05  # a positive outlier is clipped
    e = np.full(t.size, 5e-4)
    cfg = F.SearchConfig()
    fams = O.prepare_families(t, f, e, None, d, cfg)
    assert [fm["window"] for fm in fams] == [0.4, 1.5, 3.5]
    assert [list(fm["indices"]) for fm in fams] == [[0, 1, 2], [3], [4]]
    for fm in fams:
        # points within edge_trim windows of the edges are gone
        assert fm["time"].min() - t.min() >= 0.5 * fm["window"] - 1e-9
        assert t.max() - fm["time"].max() >= 0.5 * fm["window"] - 1e-9
        assert np.all(fm["flux"] <= 4.0 * 1.4826 * np.median(np.abs(fm["flux"])) + 1e-3)
    assert not np.any(np.isclose(fams[0]["time"], t[1000], atol=1e-9, rtol=0))


def _raw_star(seed=5, trend=0.004, period=3.7):
    from fbls_amd import synth

    t, f, e = F.synth_star(seed=seed, sectors=1, sigma=1e-3, period=period, depth=2e-3,
                           duration=2.5 / 24)
    slow = 1 + trend * np.sin(2 * np.pi * (t - t[0]) / 6.0)
    return t, 12000.0 * f * slow, 12000.0 * e


def _same(a, b):
    return a == b or (isinstance(a, float) and np.isnan(a) and np.isnan(b))


def test_search_transits_finds_injected_box_under_trend():
    t, raw, err = _raw_star()
    cfg = F.SearchConfig(passes=2, max_candidates=4, evidence_candidates=2)
    fams = O.prepare_families(t, raw, err, None, cfg.durations_days(), cfg)
    peaks, passes = O.search_transits(t, raw, err, None, cfg)
    assert len(fams) > 1
    assert peaks and abs(peaks[0]["period_days"] / 3.7 - 1) < 1e-3
    assert all(0 <= p["family"] < len(fams) for p in peaks)
    # the bank duration nearest the lead peak's is one its family searches
    d = cfg.durations_days()
    lead = peaks[0]
    assert int(np.argmin(np.abs(d * 24 - lead["duration_hours"]))) in fams[lead["family"]]["indices"]
    # pass 1 tests every family's period grid
    grids = 0
    for fam in fams:
        dm = d[fam["indices"]].min()
        base = O.fit_segment_noise(fam["time"], 1 + fam["flux"], fam["sigma"])["time"][-1]
        lo = max(cfg.period_min, 4 * dm * (1 + 1e-9))
        hi = min(cfg.period_max, base)
        grids += len(O.period_grid(lo, hi, base, dm, cfg.period_oversample))
    assert passes[0]["periods"] == grids


def test_search_transits_one_family_equals_prepared_search():
    # one detrending window for every duration: the raw search is the
    # prepared search of that family's curve
    t, raw, err = _raw_star(seed=9)
    cfg = F.SearchConfig(passes=2, max_candidates=3, evidence_candidates=2,
                         window_days=0.6, window_max_days=0.6)
    fams = O.prepare_families(t, raw, err, None, cfg.durations_days(), cfg)
    assert len(fams) == 1
    a, pa = O.search_transits(t, raw, err, None, cfg)
    b, pb = O.search_prepared(fams[0]["time"], fams[0]["flux"], fams[0]["sigma"], cfg)
    assert len(a) == len(b) and a
    for p, q in zip(a, b):
        assert all(_same(p[k], q[k]) for k in q), (p, q)
    assert [(s["rows"], s["periods"], s["accepted"]) for s in pa] == [
        (s["rows"], s["periods"], s["accepted"]) for s in pb]

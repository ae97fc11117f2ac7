"""Pins the multi-pass half of the oracle (oracle/pipeline.c): transit masks,
event counts, the rescoring sums, and the behaviour of the greedy selection
and pass loop on stars with known planets.

* transit_mask: the reference's own test (curve.rs:663-673).
* event_count, rescore_row, count_row: direct numpy restatements of the
  definitions (in-transit sets by explicit phase, per-cycle counts).
* search_prepared: two injected planets are both accepted in the first pass
  at their periods and epochs, no alias of an accepted peak is accepted
  (the 50 % overlap rule), pass two finds nothing and stops, the evidence
  prefers the planet model; min_events and max_candidates limit the output.
"""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import fbls_amd as F  # noqa: E402
import oracle as O  # noqa: E402


def test_transit_mask_reference_case():
    t = np.arange(1000) * 0.01
    m = O.transit_mask(t, 2.0, 1.0, 0.2, 0.0)
    direct = sum(abs(((x - 1.0 + 1.0) % 2.0) - 1.0) <= 0.1 for x in t)
    assert m.sum() == direct
    assert 100 <= direct <= 110


def test_transit_mask_margin_and_event_count():
    rng = np.random.default_rng(2)
    t = np.sort(rng.uniform(0.0, 30.0, 5000))
    P, ep, d = 3.3, 1.7, 0.25
    for margin in (0.0, 0.5):
        m = O.transit_mask(t, P, ep, d, margin)
        ph = (t - ep + P / 2) % P - P / 2
        np.testing.assert_array_equal(m, np.abs(ph) <= (1 + margin) * d / 2)
    inside = np.abs((t - ep + P / 2) % P - P / 2) <= d / 2
    cyc = np.round((t[inside] - ep) / P).astype(np.int64)
    _, counts = np.unique(cyc, return_counts=True)
    for k in (1, 3, 20, 40):
        assert O.event_count(t, P, ep, d, k) == int((counts >= k).sum())


def test_rescore_and_count_rows_direct():
    rng = np.random.default_rng(4)
    n = 4000
    t = np.sort(rng.uniform(0.0, 25.0, n))
    origin = 2458000.0
    v = rng.integers(-2 ** 20, 2 ** 20, n) * 2.0 ** -30
    w = rng.integers(1, 2 ** 20, n) * 2.0 ** -24
    ex = (rng.uniform(size=n) < 0.2).astype(np.uint8)
    P, ep, d = 2.9, origin + 0.8, 0.2
    x = (t + origin - ep) * (1.0 / P) + 0.5
    intr = np.abs((x - np.floor(x) - 0.5) * P) <= d / 2
    sel = intr & (ex == 0)
    got = O.rescore_row(t, v, w, ex, origin, P, ep, d)
    assert got[0] == v[sel].sum() and got[1] == w[sel].sum() and got[2] == w[sel].sum()
    ta = t + origin
    x = (ta - ep) * (1.0 / P) + 0.5
    intr = np.abs((x - np.floor(x) - 0.5) * P) <= d / 2
    c = O.count_row(ta, ex, P, ep, d)
    assert c[0] == intr.sum() and c[1] == (intr & (ex == 0)).sum()


def two_planet_star(seed=11):
    t, f, e = F.synth_star(seed=seed, sectors=1, sigma=8e-4, period=3.7, depth=2.5e-3,
                           duration=2.5 / 24)
    P2, ep2, d2, dep2 = 6.1, t[0] + 1.3, 3.0 / 24, 1.8e-3
    ph = ((t - ep2 + P2 / 2) % P2) - P2 / 2
    f = f - dep2 * (np.abs(ph) <= d2 / 2)
    return t, f - 1.0, e, (3.7, 6.1), (None, ep2)


def _alias(p1, p2):
    r = max(p1, p2) / min(p1, p2)
    return abs(r - round(r)) < 1e-3


def test_search_prepared_two_planets():
    t, f, e, periods, _ = two_planet_star()
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    peaks, passes = O.search_prepared(t, f, e, cfg)
    assert len(peaks) == 2
    got = sorted(p["period_days"] for p in peaks)
    assert got[0] == pytest.approx(3.7, rel=2e-3) and got[1] == pytest.approx(6.1, rel=2e-3)
    # the strongest first, both in pass 1; pass 2 accepts nothing and stops
    assert peaks[0]["snr"] >= peaks[1]["snr"]
    assert [p["pass"] for p in peaks] == [1, 1]
    assert [s["accepted"] for s in passes] == [2, 0]
    for a in peaks:
        for b in peaks:
            if a is not b:
                assert not _alias(a["period_days"], b["period_days"])
    for p in peaks:
        assert p["events"] >= cfg.min_events
        assert p["snr_without_strongest_event"] >= 0.5 * cfg.threshold
        assert p["model_preferred"] and p["planet_log_odds"] > 0
        assert math.isfinite(p["evidence_epoch_bjd"])


def test_search_prepared_limits():
    t, f, e, _, _ = two_planet_star()
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7, max_candidates=1,
                         model_comparison=False)
    peaks, passes = O.search_prepared(t, f, e, cfg)
    assert len(peaks) == 1 and math.isnan(peaks[0]["planet_log_odds"])
    # 40 events with >= 3 points each: only periods below ~0.7 d qualify, so
    # neither planet can be accepted
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7, min_events=40, passes=1)
    peaks, passes = O.search_prepared(t, f, e, cfg)
    assert passes[0]["accepted"] == len(peaks)
    for p in peaks:
        assert p["events"] >= 40 and p["period_days"] < 0.7

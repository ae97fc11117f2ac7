"""GPU parity of the single-event scan: fbls_single_event_search (the
segmented prefix and scan kernels) against the oracle's sequential
restatement on the same fixed-point fields, and the single-event stage of
fbls_search_prepared / fbls_search_transits against the oracle's search.

The prefix sums are exact integers on both sides and every template
expression keeps the oracle's operand order with contraction off; the
device appends survivors in hardware order and the host sorts them by
(phase bin, duration, shape), the oracle's loop order.  Rows, statistics
and selected events are bit-identical."""
import math

import numpy as np
import pytest

import fbls_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def _same(a, b):
    return a == b or (isinstance(a, float) and math.isnan(a) and math.isnan(b))


def _check_scan(gpu, t, f, e, cfg, label, capacity=131072):
    d = cfg.durations_days()
    s = F.family_state(t, f, e, d, cfg.correlated_noise)
    dev = F.DeviceReady.upload(gpu, s)
    v, w, _, _ = dev.fields()
    print(f"[events] {label}: n={len(s.t)} capacity={capacity}", flush=True)
    rows, stats = F.single_event_search(gpu, dev, d, s.rho, cfg, capacity)
    orows, ostats = O.single_event_search(s.t, v, w, s.baseline_norm, s.origin, d, s.rho, cfg,
                                          capacity)
    assert stats == ostats, (label, stats, ostats)
    assert len(rows) == len(orows)
    for r, o in zip(rows, orows):
        assert all(_same(r[k], o[k]) for k in o), (label, r, o)
    return rows, stats


def _dip(t, f, tc, dd, depth):
    return f - depth * (np.abs(t - tc) <= dd / 2)


def test_single_event_scan_dip(gpu):
    t, f, e = F.synth_star(seed=21, sectors=1, sigma=8e-4)
    f = _dip(t, f, t[0] + 11.37, 5.0 / 24, 4e-3)
    rows, stats = _check_scan(gpu, t, f - 1.0, e, F.SearchConfig(), "dip")
    assert rows and not stats["overflow"]


def test_single_event_scan_many_survivors_and_overflow(gpu):
    t, f, e = F.synth_star(seed=22, sectors=1, sigma=1e-3)
    cfg = F.SearchConfig(threshold=2.5)
    rows, stats = _check_scan(gpu, t, f - 1.0, e, cfg, "low threshold")
    assert stats["threshold_survivors"] > 300
    _, st2 = _check_scan(gpu, t, f - 1.0, e, cfg, "overflow", capacity=100)
    assert st2["overflow"]


def test_single_event_scan_four_sectors(gpu):
    from fbls_amd import synth

    t, f, e = synth.four_sector_red(seed=2)
    f = _dip(t, f, t[0] + 60.2, 6.0 / 24, 3e-3)
    rows, stats = _check_scan(gpu, t, f - 1.0, e, F.SearchConfig(), "four sectors")
    assert stats["phase_bins"] > 16384


def _events_equal(ev, oev, label):
    assert ev["event_stats"] == oev["event_stats"], label
    assert len(ev["events"]) == len(oev["events"]), (label, ev["events"], oev["events"])
    for a, b in zip(ev["events"], oev["events"]):
        assert all(_same(a[k], b[k]) for k in b), (label, a, b)


def test_search_prepared_single_events(gpu):
    t, f, e = F.synth_star(seed=21, sectors=1, sigma=8e-4, period=3.7, depth=2.5e-3,
                           duration=2.5 / 24)
    tc = t[0] + 11.37
    f = _dip(t, f, tc, 5.0 / 24, 4e-3)
    cfg = F.SearchConfig(period_min=1.0, period_max=8.0, passes=1, max_candidates=3,
                         evidence_candidates=1)
    print("[events] search_prepared with single events", flush=True)
    peaks, passes, rep = F.search_prepared(gpu, t, f - 1.0, e, cfg, report=True)
    opeaks, opasses, oev = O.search_prepared(t, f - 1.0, e, cfg, with_events=True)
    assert len(peaks) == len(opeaks)
    for p, o in zip(peaks, opeaks):
        assert all(_same(p[k], o[k]) for k in o)
    _events_equal(rep, oev, "prepared")
    assert rep["events"] and abs(rep["events"][0]["epoch_bjd"] - tc) < 0.1
    assert rep["nfam"] == 1 and len(rep["families"]) == 1
    assert rep["families"][0]["periods"] == passes[0]["periods"] == rep["periods_tested"]
    assert rep["observations"] == len(t)


def test_search_transits_single_events(gpu):
    t, f, e = F.synth_star(seed=23, sectors=1, sigma=1e-3, period=3.7, depth=2e-3,
                           duration=2.5 / 24)
    f = _dip(t, f, t[0] + 16.8, 6.0 / 24, 5e-3)
    slow = 1 + 0.004 * np.sin(2 * np.pi * (t - t[0]) / 6.0)
    raw, err = 12000.0 * f * slow, 12000.0 * e
    cfg = F.SearchConfig(passes=2, max_candidates=4, evidence_candidates=2)
    print("[events] search_transits with single events", flush=True)
    peaks, passes, rep = F.search_transits(gpu, t, raw, err, None, cfg, report=True)
    opeaks, opasses, oev = O.search_transits(t, raw, err, None, cfg, with_events=True)
    assert len(peaks) == len(opeaks)
    for p, o in zip(peaks, opeaks):
        assert all(_same(p[k], o[k]) for k in o)
    _events_equal(rep, oev, "transits")
    assert len(rep["event_stats"]) == rep["nfam"] > 1

"""GPU parity of the raw-mode path: the biweight detrend kernel
(fbls_biweight_flatten), the per-window families (fbls_prepare_families) and
the multi-family search (fbls_search_transits) against the oracle on the
same inputs.  The kernel follows the oracle's arithmetic order and uses
exact lower medians, so every value is bit-identical."""
import math

import numpy as np
import pytest

import fbls_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def _raw_star(seed, sectors=1, trend=0.004):
    t, f, e = F.synth_star(seed=seed, sectors=sectors, sigma=1e-3, period=3.7, depth=2e-3,
                           duration=2.5 / 24)
    slow = 1 + trend * np.sin(2 * np.pi * (t - t[0]) / 6.0)
    return t, 12000.0 * f * slow, 12000.0 * e


@pytest.mark.parametrize("case", ["sector", "wide", "gaps", "tiny", "empty"])
def test_biweight_matches_oracle(gpu, case):
    rng = np.random.default_rng(7)
    if case == "sector":  # 2-min cadence, 0.5-d windows (~360 values)
        t, raw, _ = _raw_star(3)
        rel, window = raw / np.median(raw) - 1, 0.5
    elif case == "wide":  # windows of ~4800 values (> 4096: read from global memory)
        t = np.linspace(0.0, 2.0, 6000)
        rel = rng.normal(0, 1e-3, t.size) + 2e-3 * np.sin(3 * t)
        window = 1.6
    elif case == "gaps":  # irregular sampling, gaps, outliers, repeated values
        t = np.sort(np.concatenate([rng.uniform(0, 3, 900), rng.uniform(4, 4.2, 300),
                                    rng.uniform(9, 12, 700)]))
        rel = rng.normal(0, 2e-3, t.size)
        rel[::37] += 0.05
        rel[100:140] = rel[100]
        window = 0.7
    elif case == "tiny":
        t = np.array([0.0, 0.01, 0.02, 5.0])
        rel = np.array([1e-3, -2e-3, 4e-3, 0.0])
        window = 0.5
    else:
        t, rel, window = np.zeros(0), np.zeros(0), 0.5
    print(f"[raw] biweight {case}: n={t.size} window={window}", flush=True)
    got = F.biweight_flatten(gpu, t, rel, window)
    want = O.biweight_flatten(t, rel, window)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("variant", ["err", "groups", "no_err"])
def test_prepare_families_matches_oracle(gpu, variant):
    t, raw, err = _raw_star(5)
    group = None
    if variant == "groups":
        group = (t > t[0] + 13.0).astype(np.int64)
        raw = raw * np.where(group == 1, 1.07, 1.0)
    if variant == "no_err":
        err = None
    cfg = F.SearchConfig()
    print(f"[raw] prepare_families {variant}: n={t.size}", flush=True)
    got = F.prepare_families(gpu, t, raw, err, group, cfg)
    want = O.prepare_families(t, raw, err, group, cfg.durations_days(), cfg)
    assert len(got) == len(want) > 1
    for g, w in zip(got, want):
        for k in ("time", "flux", "sigma", "indices"):
            np.testing.assert_array_equal(g[k], w[k])
        assert g["window"] == w["window"] and g["clipped"] == w["clipped"]


def _same(a, b):
    return a == b or (isinstance(a, float) and math.isnan(a) and math.isnan(b))


def test_search_transits_matches_oracle(gpu):
    t, raw, err = _raw_star(5)
    cfg = F.SearchConfig(passes=2, max_candidates=4, evidence_candidates=2)
    print(f"[raw] search_transits: n={t.size}", flush=True)
    peaks, passes = F.search_transits(gpu, t, raw, err, None, cfg)
    opeaks, opasses = O.search_transits(t, raw, err, None, cfg)
    assert peaks and abs(peaks[0]["period_days"] / 3.7 - 1) < 1e-3
    assert len(peaks) == len(opeaks)
    for p, o in zip(peaks, opeaks):
        bad = [k for k in o if not _same(p[k], o[k])]
        assert not bad, {k: (p[k], o[k]) for k in bad}
    assert [(s["pass"], s["rows"], s["periods"], s["accepted"]) for s in passes] == [
        (s["pass"], s["rows"], s["periods"], s["accepted"]) for s in opasses]

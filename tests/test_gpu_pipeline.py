"""GPU parity of the multi-pass search: fbls_search_prepared (GPU scans,
GPU rescoring, GPU evidence) against the oracle's search_prepared on the
same light curve.  Both quantise each pass's family state identically, so
every peak field and every pass statistic is bit-identical."""
import math

import numpy as np
import pytest

import fbls_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def _same(a, b):
    return a == b or (isinstance(a, float) and math.isnan(a) and math.isnan(b))


def _check(gpu, t, f, e, cfg, label):
    print(f"[pipeline] {label}: n={len(t)}", flush=True)
    peaks, passes = F.search_prepared(gpu, t, f, e, cfg)
    opeaks, opasses = O.search_prepared(t, f, e, cfg)
    assert len(peaks) == len(opeaks), (label, peaks, opeaks)
    for p, o in zip(peaks, opeaks):
        bad = [k for k in o if not _same(p[k], o[k])]
        assert not bad, (label, {k: (p[k], o[k]) for k in bad})
    assert len(passes) == len(opasses)
    for s, o in zip(passes, opasses):
        assert (s["pass"], s["rows"], s["periods"], s["accepted"]) == (
            o["pass"], o["rows"], o["periods"], o["accepted"])
    return peaks, passes


def test_pipeline_two_planets(gpu):
    t, f, e = F.synth_star(seed=11, sectors=1, sigma=8e-4, period=3.7, depth=2.5e-3,
                           duration=2.5 / 24)
    P2, ep2, d2, dep2 = 6.1, t[0] + 1.3, 3.0 / 24, 1.8e-3
    ph = ((t - ep2 + P2 / 2) % P2) - P2 / 2
    f = f - dep2 * (np.abs(ph) <= d2 / 2)
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    peaks, _ = _check(gpu, t, f - 1.0, e, cfg, "two planets")
    assert len(peaks) == 2


def test_pipeline_eclipsing_binary_and_noise(gpu):
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7, max_candidates=5)
    t, f, e = F.synth_star(seed=5, sectors=1, sigma=8e-4, period=2.3, depth=6e-3,
                           duration=3.0 / 24, secondary_depth=2e-3)
    _check(gpu, t, f - 1.0, e, cfg, "binary")
    t, f, e = F.synth_star(seed=6, sectors=1, sigma=1e-3)
    _check(gpu, t, f - 1.0, e, cfg, "noise only")


def test_pipeline_four_sector_red(gpu):
    from fbls_amd import synth

    t, f, e = synth.four_sector_red(seed=2)
    cfg = F.SearchConfig(period_min=0.5, period_max=20.0, max_candidates=4,
                         evidence_candidates=2)
    _check(gpu, t, f - 1.0, e, cfg, "four-sector")

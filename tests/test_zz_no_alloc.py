"""GPU: fbls_scan_bank_noalloc (the scan with no allocation, staging in a
fixed per-device array of the library) against fbls_scan_bank on the same
family: bit-identical records, scores, depths, winners and statistics, over
consecutive calls of decreasing size and through the range-split kernel."""
import numpy as np
import pytest

import fbls_amd as F

pytestmark = pytest.mark.gpu


def _family(gpu, cfg, seed, sectors, period):
    t, f, e = F.synth_star(seed=seed, sectors=sectors, sigma=1e-3, period=period,
                           depth=2e-3, duration=3.0 / 24)
    d = cfg.durations_days()
    s = F.family_state(t, f - 1.0, e, d, cfg.correlated_noise)
    return s, F.DeviceReady.upload(gpu, s), d


def _same(a, b, label):
    rec, score, depth, win, st = a
    rec2, score2, depth2, win2, st2 = b
    np.testing.assert_array_equal(rec2, rec, err_msg=label)
    np.testing.assert_array_equal(score2, score, err_msg=label)
    np.testing.assert_array_equal(depth2, depth, err_msg=label)
    if win is not None:
        np.testing.assert_array_equal(win2, win, err_msg=label)
    for key in ("phase_bins", "batches", "templates_scanned", "exact_tile_retries"):
        assert st2[key] == st[key], (label, key, st2[key], st[key])


def test_noalloc_equals_scan_bank_twice_smaller_second(gpu):
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    s, dev, d = _family(gpu, cfg, seed=1, sectors=1, period=3.7)
    bank = F.make_bank(cfg, d, s.rho)
    rng = np.random.default_rng(7)
    calls = [(np.sort(rng.uniform(0.5, 13.7, 3000)), 768, True),
             (np.sort(rng.uniform(2.0, 6.0, 700)), 384, False)]
    for k, (periods, nb, winner) in enumerate(calls):
        print(f"[noalloc] call {k}: np={len(periods)} nb={nb} n={len(s.t)}", flush=True)
        ref = F.scan_bank(gpu, dev, periods, bank, nb, winner=winner)
        got = F.scan_bank_noalloc(gpu, dev, periods, bank, nb, winner=winner)
        _same(ref, got, f"call {k}")
        assert np.isfinite(got[1]).any()
    dev.close()


def test_noalloc_range_split_kernel(gpu):
    """nb = 9000 with 13 durations needs more LDS than one period fits: the
    range-split kernel runs from the same scratch array."""
    cfg = F.SearchConfig(period_min=0.5, period_max=30.0)
    s, dev, d = _family(gpu, cfg, seed=3, sectors=2, period=12.5)
    bank = F.make_bank(cfg, d, s.rho)
    periods = np.linspace(11.0, 14.0, 48)
    print(f"[noalloc] split: np={len(periods)} nb=9000 n={len(s.t)}", flush=True)
    ref = F.scan_bank(gpu, dev, periods, bank, 9000)
    got = F.scan_bank_noalloc(gpu, dev, periods, bank, 9000)
    _same(ref, got, "split")
    dev.close()


def test_noalloc_rejects_more_periods_than_the_array_holds(gpu):
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    s, dev, d = _family(gpu, cfg, seed=1, sectors=1, period=3.7)
    bank = F.make_bank(cfg, d, s.rho)
    with pytest.raises(F.FblsError) as err:
        F.scan_bank_noalloc(gpu, dev, np.linspace(1.0, 2.0, 65537), bank, 64)
    assert err.value.code == 4 and "65536" in str(err.value)  # FBLS_E_ARG
    dev.close()

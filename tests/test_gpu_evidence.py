"""GPU parity of the evidence path: fbls_compare (HIP kernels) against the
oracle's compare_search on the same fixed-point fields.

Every interval sum is an exact integer on both sides, every expression keeps
the reference's operand order with contraction off, and exp / log are the
same routines on the device and in the oracle, so the contract is
bit-exact: rows (winner phase, duration, shape, SNR, depth, log Bayes
factors, diagnostics), the screen, the floors and every tile part.
"""
import math

import numpy as np
import pytest

import fbls_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def _same(a, b):
    return (a == b) or (isinstance(a, float) and math.isnan(a) and math.isnan(b))


def _state(gpu, cfg, **kw):
    t, f, e = F.synth_star(**kw)
    d = cfg.durations_days()
    s = F.family_state(t, f - 1.0, e, d, cfg.correlated_noise)
    dev = F.DeviceReady.upload(gpu, s)
    vq, wq, _, _ = dev.fields()
    return s, dev, vq, wq, d


def _check(gpu, s, dev, vq, wq, d, period, cfg, label):
    print(f"[evidence] {label}: P={period} n={len(s.t)}", flush=True)
    rows, dbg = F.compare_search(gpu, dev, period, d, s.rho, s.residual_energy,
                                 cfg, debug=True)
    orows, odbg = O.compare_search(s.t, vq, wq, s.baseline_norm, s.origin,
                                   s.residual_energy, d, s.rho, period, cfg,
                                   debug=True)
    assert dbg["nb"] == odbg["nb"] and dbg["tiles"] == odbg["tiles"]
    np.testing.assert_array_equal(dbg["screen"], odbg["screen"])
    assert len(rows) == len(orows), (label, rows, orows)
    for r, o in zip(rows, orows):
        bad = [k for k in r if not _same(r[k], o[k])]
        assert not bad, (label, bad, {k: (r[k], o[k]) for k in bad})
    if orows:
        np.testing.assert_array_equal(dbg["floors"][:3], odbg["floors"])
        assert dbg["floors"][3] == 1.0
        np.testing.assert_array_equal(dbg["parts"], odbg["parts"])
    return rows


@pytest.fixture(scope="module")
def single(gpu):
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    return (cfg,) + _state(gpu, cfg, seed=1, sectors=1, sigma=1e-3, period=3.7,
                           depth=2e-3, duration=2.5 / 24)


@pytest.mark.parametrize("period", [3.7, 3.700004, 7.4, 1.85, 1.3, 11.0, 0.4])
def test_compare_single_sector(gpu, single, period):
    cfg, s, dev, vq, wq, d = single
    rows = _check(gpu, s, dev, vq, wq, d, period, cfg, "single")
    if period in (3.7, 3.700004):
        assert rows and rows[0]["planet_log_odds"] > 0


def test_compare_exhaustive_and_custom_priors(gpu, single):
    cfg, s, dev, vq, wq, d = single
    import dataclasses

    c0 = dataclasses.replace(cfg, evidence_rtol=0.0)
    _check(gpu, s, dev, vq, wq, d, 3.7, c0, "rtol0")
    c1 = dataclasses.replace(cfg, stellar_radius_solar=0.7, target_flux_fraction=0.9,
                             planet_prior=0.02, ingress_fractions=[0.0, 0.3],
                             shape_layers=3, phase_oversample=3)
    _check(gpu, s, dev, vq, wq, d, 3.7, c1, "custom")


def test_compare_eclipsing_binary(gpu):
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    s, dev, vq, wq, d = _state(gpu, cfg, seed=5, sectors=1, sigma=8e-4, period=2.3,
                               depth=6e-3, duration=3.0 / 24, secondary_depth=2e-3)
    rows = _check(gpu, s, dev, vq, wq, d, 2.3, cfg, "binary")
    assert rows and rows[0]["log_bf_binary_noise"] > rows[0]["log_bf_planet_noise"]


def test_compare_four_sector_red(gpu):
    from fbls_amd import synth

    t, f, e = synth.four_sector_red(seed=2)
    cfg = F.SearchConfig(period_min=0.5, period_max=20.0)
    d = cfg.durations_days()
    s = F.family_state(t, f - 1.0, e, d, cfg.correlated_noise)
    dev = F.DeviceReady.upload(gpu, s)
    vq, wq, _, _ = dev.fields()
    for period in (11.3, 11.3 * 0.5, 17.9):
        _check(gpu, s, dev, vq, wq, d, period, cfg, "four-sector")


def test_compare_long_period_and_unsorted_upload(gpu):
    cfg = F.SearchConfig(period_min=0.5, period_max=100.0)
    t, f, e = F.synth_star(seed=9, sectors=2, sigma=5e-4, period=43.0, depth=3e-3,
                           duration=6.0 / 24, baseline_gap_days=200.0)
    d = cfg.durations_days()
    s = F.family_state(t, f - 1.0, e, d, cfg.correlated_noise)
    dev = F.DeviceReady.upload(gpu, s)
    vq, wq, _, _ = dev.fields()
    rows = _check(gpu, s, dev, vq, wq, d, 43.0, cfg, "long")
    # nb2 well beyond one LDS: 4 * ceil(4 * 43 / dmin / 4) bins per period
    perm = np.random.default_rng(0).permutation(len(s.t))
    dev2 = F.DeviceReady(gpu, s.t[perm], s.v[perm], s.w[perm], s.baseline_norm, s.origin)
    rows2 = F.compare_search(gpu, dev2, 43.0, d, s.rho, s.residual_energy, cfg)
    assert len(rows) == len(rows2)
    for a, b in zip(rows, rows2):
        assert all(_same(a[k], b[k]) for k in a)

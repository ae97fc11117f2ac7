"""GPU parity: the HIP scan against the FP64 oracle on the same inputs.

The device scans v and w in 32-bit fixed point (power-of-two scale), so the
oracle is fed the same fixed-point values (DeviceReady.fields()).  Every
interval sum is then exact on both sides and the contract is bit-exact:
winner indices (phase, width, shape, duration) and the records/scores.
A second comparison against the unquantised FP64 fields bounds the effect of
the fixed-point representation (scores within 1e-6 relative).
"""
import numpy as np
import pytest

import fbls_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def _star(seed=1, sectors=1, sigma=1e-3, period=3.7, depth=2e-3,
          duration=2.5 / 24, **kw):
    t, f, e = F.synth_star(seed=seed, sectors=sectors, sigma=sigma,
                           period=period, depth=depth, duration=duration, **kw)
    return t, f - 1.0, e


def _setup(gpu, cfg, seed=1, **kw):
    t, f, e = _star(seed=seed, **kw)
    d = cfg.durations_days()
    s = F.family_state(t, f, e, d, cfg.correlated_noise)
    dev = F.DeviceReady.upload(gpu, s)
    vq, wq, _, _ = dev.fields()
    return s, dev, vq, wq, d


def _compare(gpu, dev, s, vq, wq, d, periods, nb, cfg, label):
    print(f"[parity] {label}: np={len(periods)} nb={nb} n={len(s.t)}", flush=True)
    bank = F.make_bank(cfg, d, s.rho)
    rec, score, depth, win, st = F.scan_bank(gpu, dev, periods, bank, nb)
    orec, oscore, odepth, owin, ost = O.scan_bank(
        s.t, vq, wq, s.baseline_norm, periods, d, s.rho,
        cfg.ingress_fractions, cfg.shape_layers, cfg.shape_gate, nb)
    assert ost == 0
    bad = np.nonzero((win != owin).any(axis=1))[0]
    assert len(bad) == 0, (label, bad[:10], win[bad[:5]], owin[bad[:5]],
                           score[bad[:5]], oscore[bad[:5]])
    np.testing.assert_array_equal(score, oscore)
    np.testing.assert_array_equal(depth, odepth)
    np.testing.assert_array_equal(rec, orec)
    return st


@pytest.mark.parametrize("nb", [16, 96, 192, 384, 1000, 2631])
def test_scan_bank_bit_exact(gpu, nb):
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    s, dev, vq, wq, d = _setup(gpu, cfg)
    rng = np.random.default_rng(nb)
    periods = np.sort(rng.uniform(0.5, 13.7, 300))
    periods = np.concatenate([periods, [3.7, 3.7 * 2, 3.7 / 2]])
    st = _compare(gpu, dev, s, vq, wq, d, periods, nb, cfg, f"nb{nb}")
    print("retries", st["exact_tile_retries"])


def test_scan_bank_vs_unquantised(gpu):
    """Fixed point vs FP64 fields: scores agree to 1e-6 relative."""
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    s, dev, vq, wq, d = _setup(gpu, cfg)
    periods = np.linspace(0.5, 13.7, 200)
    nb = 384
    bank = F.make_bank(cfg, d, s.rho)
    rec, score, depth, win, st = F.scan_bank(gpu, dev, periods, bank, nb)
    orec, oscore, odepth, owin, ost = O.scan_bank(
        s.t, s.v, s.w, s.baseline_norm, periods, d, s.rho,
        cfg.ingress_fractions, cfg.shape_layers, cfg.shape_gate, nb)
    ok = np.isfinite(oscore)
    rel = np.abs(score[ok] - oscore[ok]) / np.abs(oscore[ok])
    print("max rel score error vs FP64 fields:", rel.max())
    assert rel.max() < 1e-6
    same = (win == owin).all(axis=1)
    assert same.mean() > 0.99


def test_signal_search_bench_star(gpu):
    """The bench workload end to end (grid, bands, prescan, rows, band
    statistics) against the oracle's signal_search on the same fields."""
    from fbls_amd import synth

    t, f, e = synth.four_sector_red(seed=2)
    cfg = F.SearchConfig(period_min=0.5, period_max=20.0)
    d = cfg.durations_days()
    s = F.family_state(t, f - 1.0, e, d, cfg.correlated_noise)
    dev = F.DeviceReady.upload(gpu, s)
    vq, wq, _, _ = dev.fields()
    print(f"[parity] signal_search 4-sector n={len(s.t)}", flush=True)
    rows, grid, bands = F.signal_search(gpu, dev, d, s.rho, cfg)
    orows, obands, og, ost = O.signal_search(
        s.t, vq, wq, s.baseline_norm, s.origin, d, s.rho,
        cfg.ingress_fractions, cfg, cfg.period_min, cfg.period_max,
        cfg.period_oversample)
    assert ost == 0
    assert len(grid) == og
    assert len(bands) == len(obands)
    for b, ob in zip(bands, obands):
        for k in ("periods", "prescan_survivors", "phase_bins", "batches",
                  "templates_scanned"):
            assert b[k] == ob[k], (k, b, ob)
    np.testing.assert_array_equal(rows, orows)
    best = rows[np.argmax(rows[:, 3])]
    assert abs(best[0] - 11.3) < 0.01 and best[3] > 7.1


@pytest.mark.parametrize("launch", [(0, 256, 2), (32768, 512, 1), (160 * 1024, 256, 1)])
def test_launch_shape_does_not_change_results(gpu, launch):
    """LDS budget, workgroup size and fold mode change the schedule only."""
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    s, dev, vq, wq, d = _setup(gpu, cfg, seed=5)
    periods = np.linspace(0.6, 13.0, 700)
    bank = F.make_bank(cfg, d, s.rho)
    print(f"[parity] launch {launch}", flush=True)
    ref = F.scan_bank(gpu, dev, periods, bank, 1536)
    gpu.set_launch(*launch)
    try:
        got = F.scan_bank(gpu, dev, periods, bank, 1536)
    finally:
        gpu.set_launch(0, 0, 0)
    for a, b in zip(ref[:4], got[:4]):
        np.testing.assert_array_equal(a, b)

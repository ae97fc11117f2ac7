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

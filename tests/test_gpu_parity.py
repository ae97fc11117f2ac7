"""GPU parity: the HIP scan against the FP64 oracle on the same inputs.

The device scans v and w as 32-bit fixed-point words (plain integers under a
power-of-two scale, or, for fields of wide dynamic range, a 27-bit mantissa
over a 5-bit shift), so the oracle is fed the same fixed-point values
(DeviceReady.fields()).  Every
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


def _wide_range_star(cfg):
    """The single-sector star with a 1e4-sigma flux spike and one point whose
    weight is 1e6 times the others' (flux error 1e-3 of theirs): both fields
    go to block mode."""
    t, f, e = _star(seed=1)
    f = f.copy()
    e = e.copy()
    f[5000] += 1e4 * e[5000]
    e[9000] *= 1e-3
    d = cfg.durations_days()
    s = F.family_state(t, f, e, d, cfg.correlated_noise)
    return s, d


def test_block_fixed_point_bit_exact_and_bounded(gpu):
    """Block-mode fields: the GPU equals the oracle on the same fixed-point
    values, and both stay within 1e-6 relative of the unquantised scores."""
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    s, d = _wide_range_star(cfg)
    assert F.quantize(s.v)[2] and F.quantize(s.w)[2]
    dev = F.DeviceReady.upload(gpu, s)
    vq, wq, _, _ = dev.fields()
    periods = np.concatenate([np.linspace(0.5, 13.7, 200), [3.7]])
    for nb in (96, 384):
        _compare(gpu, dev, s, vq, wq, d, periods, nb, cfg, f"block nb{nb}")
    bank = F.make_bank(cfg, d, s.rho)
    rec, score, depth, win, st = F.scan_bank(gpu, dev, periods, bank, 384)
    orec, oscore, odepth, owin, ost = O.scan_bank(
        s.t, s.v, s.w, s.baseline_norm, periods, d, s.rho,
        cfg.ingress_fractions, cfg.shape_layers, cfg.shape_gate, 384)
    ok = np.isfinite(oscore)
    rel = np.abs(score[ok] - oscore[ok]) / np.abs(oscore[ok])
    print("block mode: max rel score error vs FP64 fields:", rel.max(), flush=True)
    assert rel.max() < 1e-6


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


def test_signal_search_single_sector_star(gpu):
    """The single-sector workload (one 27.4-d sector at 2-min cadence, box
    P = 3.7 d, periods 0.5-13.7 d, the 13 default durations, prescan 2): the
    full signal search -- 13,065 grid periods in five octave bands (nb 192 to
    2631) -- against the oracle's signal_search: every row and every band
    statistic bit-exact."""
    from fbls_amd import synth

    t, f, e = synth.single_sector(seed=1)
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    d = cfg.durations_days()
    s = F.family_state(t, f - 1.0, e, d, cfg.correlated_noise)
    dev = F.DeviceReady.upload(gpu, s)
    vq, wq, _, _ = dev.fields()
    print(f"[parity] signal_search 1-sector n={len(s.t)}", flush=True)
    rows, grid, bands = F.signal_search(gpu, dev, d, s.rho, cfg)
    orows, obands, og, ost = O.signal_search(
        s.t, vq, wq, s.baseline_norm, s.origin, d, s.rho,
        cfg.ingress_fractions, cfg, cfg.period_min, cfg.period_max,
        cfg.period_oversample)
    assert ost == 0
    assert len(grid) == og == 13065
    assert [b["phase_bins"] for b in bands] == [192, 384, 768, 1536, 2631]
    assert len(bands) == len(obands)
    for b, ob in zip(bands, obands):
        for k in ("periods", "prescan_survivors", "phase_bins", "batches",
                  "templates_scanned"):
            assert b[k] == ob[k], (k, b, ob)
    np.testing.assert_array_equal(rows, orows)
    best = rows[np.argmax(rows[:, 3])]
    assert abs(best[0] - 3.7) < 0.01 and best[3] > 7.1


@pytest.mark.parametrize("launch", [(0, 256, 2), (32768, 512, 1), (160 * 1024, 256, 1),
                                    (0, 1024, 0), (128 * 1024, 1024, 0)])
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


@pytest.mark.parametrize("nb", [9000, 12000, 19200])
def test_range_split_kernel_bit_exact(gpu, nb):
    """Phase-bin counts whose histogram exceeds the LDS run on the range-split
    kernel; results stay bit-identical to the oracle."""
    cfg = F.SearchConfig(period_min=0.5, period_max=100.0)
    s, dev, vq, wq, d = _setup(gpu, cfg, seed=7, sectors=2, period=23.0,
                               depth=3e-3, duration=6.0 / 24)
    rng = np.random.default_rng(nb)
    lo = 4 * d[-1] * nb / (nb // 2)  # keep k = D nb / P well below nb / 2
    periods = np.sort(rng.uniform(max(lo, 20.0), 54.0, 120))
    periods = np.concatenate([periods, [23.0, 46.0]])
    _compare(gpu, dev, s, vq, wq, d, periods, nb, cfg, f"split nb{nb}")


def test_counts_from_weights_in_split_kernel(gpu):
    """Equal fixed-point weights: by default (fold mode 0) the fold skips the
    count atomics and derives each bin's count from its weight sum; folding
    the counts (mode 1) gives the same bits in the range-split kernel too,
    and both equal the oracle."""
    cfg = F.SearchConfig(period_min=0.5, period_max=100.0)
    s, dev, vq, wq, d = _setup(gpu, cfg, seed=9, sectors=2, period=31.0,
                               depth=3e-3, duration=6.0 / 24)
    assert len(np.unique(wq)) == 1
    periods = np.linspace(30.0, 54.0, 60)
    bank = F.make_bank(cfg, d, s.rho)
    _compare(gpu, dev, s, vq, wq, d, periods, 12000, cfg, "split, counts from weights")
    gpu.set_launch(0, 0, 1)
    try:
        got = F.scan_bank(gpu, dev, periods, bank, 12000)
    finally:
        gpu.set_launch(0, 0, 0)
    want = F.scan_bank(gpu, dev, periods, bank, 12000)
    for a, b in zip(want[:4], got[:4]):
        np.testing.assert_array_equal(a, b)


def test_signal_search_long_baseline(gpu):
    """A two-year baseline with a one-year gap at periods up to 100 d: the
    top bands use the range-split kernel (nb up to 19200)."""
    cfg = F.SearchConfig(period_min=80.0, period_max=100.0)
    t, f, e = F.synth_star(seed=11, sectors=2, sigma=1e-3, period=87.0,
                           depth=2e-3, duration=5.0 / 24,
                           baseline_gap_days=365.0)
    d = cfg.durations_days()
    s = F.family_state(t, f - 1.0, e, d, cfg.correlated_noise)
    dev = F.DeviceReady.upload(gpu, s)
    vq, wq, _, _ = dev.fields()
    print(f"[parity] long baseline n={len(s.t)} T={s.t[-1]:.0f}", flush=True)
    rows, grid, bands = F.signal_search(gpu, dev, d, s.rho, cfg)
    orows, obands, og, ost = O.signal_search(
        s.t, vq, wq, s.baseline_norm, s.origin, d, s.rho,
        cfg.ingress_fractions, cfg, cfg.period_min, cfg.period_max,
        cfg.period_oversample)
    assert ost == 0 and len(grid) == og
    assert max(b["phase_bins"] for b in bands) > 8000
    for b, ob in zip(bands, obands):
        for k in ("periods", "prescan_survivors", "phase_bins",
                  "templates_scanned"):
            assert b[k] == ob[k], (k, b, ob)
    np.testing.assert_array_equal(rows, orows)


def test_nonuniform_weights_and_unsorted_times(gpu):
    """Per-point errors (the general three-field fold) and a shuffled upload
    (the direct phase-bin path) give the oracle's results bit for bit."""
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    t, f, e = F.synth_star(seed=21, sectors=1, sigma=1e-3, period=2.9,
                           depth=3e-3, duration=2.0 / 24)
    rng = np.random.default_rng(21)
    e = e * rng.uniform(0.7, 1.4, len(e))
    d = cfg.durations_days()
    s = F.family_state(t, f - 1.0, e, d, cfg.correlated_noise)
    dev = F.DeviceReady.upload(gpu, s)
    vq, wq, _, _ = dev.fields()
    assert len(np.unique(wq)) > 1
    periods = np.linspace(0.6, 13.0, 500)
    for nb in (192, 1536):
        _compare(gpu, dev, s, vq, wq, d, periods, nb, cfg, f"nonuniform nb{nb}")
    perm = rng.permutation(len(s.t))
    shuffled = F.DeviceReady(gpu, s.t[perm], s.v[perm], s.w[perm],
                             s.baseline_norm, s.origin)
    bank = F.make_bank(cfg, d, s.rho)
    a = F.scan_bank(gpu, dev, periods, bank, 768)
    b = F.scan_bank(gpu, shuffled, periods, bank, 768)
    for x, y in zip(a[:4], b[:4]):
        np.testing.assert_array_equal(x, y)


def test_profile_stage_attribution(gpu, monkeypatch):
    """FBLS_PROFILE (scan.rs:152,222-229): unset, the fold / cumulative /
    scan seconds of every band stay 0, as in the reference's own report;
    set, the kernels' per-stage shader cycles share out each band's device
    time -- every stage positive, their sum the band's device seconds -- and
    the rows are those of the unprofiled search, also through the range-split
    kernel and scan_bank."""
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    s, dev, vq, wq, d = _setup(gpu, cfg, seed=5)
    print("[parity] profiled signal_search", flush=True)
    monkeypatch.delenv("FBLS_PROFILE", raising=False)
    rows0, grid0, bands0 = F.signal_search(gpu, dev, d, s.rho, cfg)
    for b in bands0:
        assert b["seconds_fold"] == b["seconds_cumulative"] == b["seconds_scan"] == 0.0
    monkeypatch.setenv("FBLS_PROFILE", "1")
    rows1, grid1, bands1 = F.signal_search(gpu, dev, d, s.rho, cfg)
    np.testing.assert_array_equal(rows0, rows1)
    np.testing.assert_array_equal(grid0, grid1)
    for b in bands1:
        stages = b["seconds_fold"] + b["seconds_cumulative"] + b["seconds_scan"]
        assert b["seconds_fold"] > 0 and b["seconds_cumulative"] > 0 and b["seconds_scan"] > 0
        assert abs(stages - b["seconds_device"]) <= 1e-9 * max(1.0, b["seconds_device"])
    # scan_bank through both kernels (12,000 bins need the range split)
    bank = F.make_bank(cfg, d, s.rho)
    for nb in (384, 12000):
        periods = np.linspace(3.5, 3.9, 24)
        monkeypatch.delenv("FBLS_PROFILE", raising=False)
        r0 = F.scan_bank(gpu, dev, periods, bank, nb)
        monkeypatch.setenv("FBLS_PROFILE", "1")
        r1 = F.scan_bank(gpu, dev, periods, bank, nb)
        for a, b in zip(r0[:4], r1[:4]):
            np.testing.assert_array_equal(a, b)
        st = r1[4]
        assert st["seconds_fold"] > 0 and st["seconds_cumulative"] > 0 and st["seconds_scan"] > 0

"""Pins the FP64 oracle (oracle/) before it is trusted as the GPU's checker.

Every oracle function is checked against something other than its own
formula: the reference's known-answer tests, the band statistics of the
reference's committed search report, an independent numpy brute force
(different summation structure: direct window sums and explicit staircase
weights, no prefix sums, no tiles-as-lanes), closed forms, and exact integer
arithmetic.
"""
import json
import math
import os

import numpy as np
import pytest

import fbls_amd as F
import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INF = float("inf")


# ------------------------------------------------------------------ grid

def test_grid_bounds_and_drift_kat():
    """grid.rs:34-39: endpoints exact, drift <= D/oversample (1 + 1e-9)."""
    g = O.period_grid(0.5, 20.0, 50.0, 0.02, 3.0)
    assert g[0] == 0.5 and g[-1] == 20.0
    drift = 50.0 * (g[1:] / g[:-1] - 1.0)
    assert np.all(drift <= 0.02 / 3.0 * (1.0 + 1e-9))
    # the count is the smallest that satisfies the drift bound
    n = len(g)
    step = math.log1p(0.02 / (3.0 * 50.0))
    assert (n - 2) * step < math.log(20.0 / 0.5) <= (n - 1) * step


def test_durations_rounding_matches_report():
    """config.rs:94 (0.5 * 1.25^i rounded half-even at 1e-6 h) reproduces
    the 19 durations of the reference's survey report."""
    d = json.load(open(os.path.join(GOLDEN, "search_bands_72668830.json")))
    hours = [round(0.5 * 1.25 ** i * 1e6) / 1e6 for i in range(19)]
    assert hours == d["config"]["durations_hours"]
    days = O.durations_days(hours[::-1] + hours[:3])
    assert np.array_equal(days, np.array(hours) / 24.0)


def _grid_of_count(pmin, pmax, dmin, count):
    """A baseline whose grid has exactly `count` periods (grid.rs:12-13)."""
    L = math.log(pmax / pmin)
    step = L / (count - 1.5)
    baseline = dmin / (3.0 * math.expm1(step))
    g = O.period_grid(pmin, pmax, baseline, dmin, 3.0)
    assert len(g) == count
    return g


def _fixture_families(pass_index=0):
    d = json.load(open(os.path.join(GOLDEN, "search_bands_72668830.json")))
    cfg_hours = d["config"]["durations_hours"]
    fams = []
    for f in d["passes"][pass_index]["families"]:
        # map the report's hours (d * 24 round trips) to the config's values
        hrs = sorted(min(cfg_hours, key=lambda c: abs(c - h))
                     for h in f["durations_hours"])
        days = O.durations_days(hrs)
        fams.append((f, days))
    return d, fams


@pytest.mark.parametrize("pass_index", [0, 1])
def test_band_plan_matches_reference_report(pass_index):
    """Grid, octave bands, phase bins, batches and templates_scanned of all
    11 families of the reference's search report (scan.rs:352,377-405,
    440-471), given the report's own prescan survivor counts."""
    d, fams = _fixture_families(pass_index)
    c = d["config"]
    ns = len(c["ingress_fractions"])
    for f, days in fams:
        pmin = max(c["period_min"], 4.0 * days[0] * (1.0 + 1e-9))
        pmax = c["period_max"]
        grid = _grid_of_count(pmin, pmax, days[0], f["periods"])
        bands = O.plan_bands(grid, days, c["phase_oversample"],
                             c["prescan_factor"])
        assert len(bands) == len(f["bands"]), f["family"]
        for b, ref in zip(bands, f["bands"]):
            assert b["count"] == ref["periods"]
            surv = ref["prescan_survivors"]
            if b["nb_coarse"]:
                templates = (b["count"] * b["nd"] * ns * b["nb_coarse"]
                             + surv * b["nd"] * ns * b["nb"])
                batches = (O.bank_batches(b["count"], b["nb_coarse"],
                                          c["period_batch"], c["max_fold_mib"])
                           + O.bank_batches(surv, b["nb"], c["period_batch"],
                                            c["max_fold_mib"]))
                phase_bins = b["nb"] if surv else b["nb_coarse"]
            else:
                templates = b["count"] * b["nd"] * ns * b["nb"]
                batches = O.bank_batches(b["count"], b["nb"],
                                         c["period_batch"], c["max_fold_mib"])
                phase_bins = b["nb"]
            assert phase_bins == ref["phase_bins"], (f["family"], b, ref)
            assert templates == ref["templates_scanned"], (f["family"], b, ref)
            assert batches == ref["batches"], (f["family"], b, ref)
    assert sum(f["periods"] for f, _ in fams) == d["periods_tested"]


# ----------------------------------------------------------------- stats

def test_median_lower_and_mad():
    """stats.rs:4-18: lower middle element; 1.4826 * MAD."""
    assert O.median_lower([3.0, 1.0, 2.0, 4.0]) == 2.0
    assert O.median_lower([5.0, np.nan, 1.0, 3.0]) == 3.0
    assert math.isnan(O.median_lower([np.nan]))
    x = [1.0, 2.0, 3.0, 4.0, 100.0]
    assert O.mad_scale(x) == 1.4826 * 1.0


def test_noise_model_kat():
    """curve.rs:636-661 (noise_model_normalizes_and_inflates)."""
    n = 500
    t = 1000.0 + np.arange(n) * 0.01
    f = 2000.0 + 3.0 * np.sin(np.arange(n) * 1.7)
    e = np.ones(n)
    m = O.fit_segment_noise(t, f, e, 0.75)
    assert m["origin"] == 1000.0
    assert m["time"][0] == 0.0
    median = np.sort(f)[(n - 1) // 2]
    assert abs(m["deficit"][0] - (1.0 - f[0] / median)) < 1e-12
    assert m["inflation"] > 2.0
    assert m["segments"] == 1


def test_prepare_ready_closed_forms():
    rng = np.random.default_rng(0)
    n = 1000
    deficit = rng.normal(0, 1e-3, n)
    var = rng.uniform(0.5e-6, 2e-6, n)
    r = O.prepare_ready(deficit, var)
    p = 1.0 / var
    assert r["baseline_norm"] == math.fsum(p) or \
        abs(r["baseline_norm"] - math.fsum(p)) < 1e-12 * math.fsum(p)
    # v is p-weighted mean-subtracted: sum v = 0 up to rounding
    assert abs(math.fsum(r["v"])) < 1e-9 * np.abs(r["v"]).sum()
    mean = math.fsum(p * deficit) / math.fsum(p)
    assert np.allclose(r["v"], p * deficit - p * mean, rtol=1e-12, atol=1e-9)
    with pytest.raises(ValueError):
        O.prepare_ready(deficit[:5], var[:5])


def test_correlated_noise_white_is_one():
    """White noise at the stated sigma: measured scatter ~1, floored at 1."""
    rng = np.random.default_rng(1)
    t = np.arange(20000) / 720.0
    s = np.full(len(t), 1e-3)
    r = rng.normal(0, 1e-3, len(t))
    d = np.array([0.5, 1.0, 2.0, 4.0]) / 24.0
    rho = O.correlated_noise_factors(t, r, s, d)
    assert np.all(rho >= 1.0) and np.all(rho < 1.15)


def test_correlated_noise_scaled_residuals():
    """Residuals 3x the stated sigma in every cell: rho ~ 3 (robust scatter
    of standardised cell means), and rho scales linearly with that factor."""
    rng = np.random.default_rng(2)
    t = np.arange(30000) / 720.0
    s = np.full(len(t), 1e-3)
    base = rng.normal(0, 1e-3, len(t))
    d = np.array([1.0]) / 24.0
    r1 = O.correlated_noise_factors(t, 3 * base, s, d)[0]
    r2 = O.correlated_noise_factors(t, 6 * base, s, d)[0]
    assert 2.7 < r1 < 3.3
    assert abs(r2 / r1 - 2.0) < 1e-9


def test_correlated_noise_unsupported_inherits():
    """A duration with < 20 cells inherits sqrt(D / D_anchor) scaling from
    the nearest supported one (curve.rs:504-538)."""
    rng = np.random.default_rng(3)
    t = np.arange(3000) / 720.0  # ~4.2 d
    s = np.full(len(t), 1e-3)
    r = rng.normal(0, 1e-3, len(t))
    d = np.array([1.0 / 24.0, 4.0])  # 4 d: one cell only
    rho = O.correlated_noise_factors(t, r, s, d)
    assert abs(rho[1] - max(1.0, rho[0] * math.sqrt(4.0 / (1.0 / 24.0)))) < 1e-12


# ------------------------------------------------------------ quantizer

def test_quantize_exact_and_bounded():
    rng = np.random.default_rng(4)
    x = rng.normal(0, 3e5, 10000)
    q, e, block = O.quantize(x)
    assert not block  # Gaussian: max |x| ~ 5 mean |x|
    assert np.abs(q).max() <= 2 ** 30
    assert np.abs(q).max() > 2 ** 29  # the scale uses the range
    assert np.all(np.abs(q * 2.0 ** -e - x) <= 2.0 ** (-e - 1))
    # power-of-two scale: the dequantised values are exact binary fractions
    assert np.array_equal(np.ldexp(q.astype(np.float64), -e) * 2.0 ** e, q)


def _fixed_sum_bound(x, idx):
    """2^-26 (sum over the window of |x| + |window| mean |x|)."""
    return 2.0 ** -26 * (np.abs(x[idx]).sum() + len(idx) * np.abs(x).mean())


@pytest.mark.parametrize("case", ["clean", "spike_1e4", "spike_1e6", "weights_1e6"])
def test_fixed_point_error_bound(case):
    """Every value's error is at most 2^-26 max(|x|, mean |x|) and window sums
    are exact: a plain Gaussian field, one 1e4- and one 1e6-sigma spike, and
    weights spread over six decades (1/sigma^2 with one point 1e6 times the
    rest) each stay within the bound, which is checked against the exact
    rational value of the window sums (Python integers)."""
    from fractions import Fraction

    rng = np.random.default_rng(12)
    n = 20000
    if case.startswith("weights"):
        x = rng.uniform(0.8, 1.2, n) * 1e6
        x[777] *= 1e6
    else:
        x = rng.normal(0.0, 1.0, n) * 3e3
        if case != "clean":
            x[4321] = float(case.split("_")[1]) * 3e3
    q, e, block = O.quantize(x)
    assert block == (case != "clean")
    ints = O.fixed_values(q, block)
    assert np.all(ints == np.round(ints)) and np.abs(ints).sum() < 2.0 ** 53
    got = np.ldexp(ints, -e)
    per = 2.0 ** -26 * np.maximum(np.abs(x), np.abs(x).mean())
    assert np.all(np.abs(got - x) <= per)
    # windows: a box over the spike, a long run and scattered points
    for idx in (np.arange(4000, 4600), np.arange(0, n, 3), rng.choice(n, 500, replace=False)):
        exact = sum(Fraction(float(v)) for v in x[idx])
        fixed = sum(Fraction(float(v)) for v in got[idx])
        s = float(np.sum(got[idx]))  # FP64 sum of the fixed values: exact
        assert Fraction(s) == fixed
        assert abs(float(fixed - exact)) <= _fixed_sum_bound(x, idx)


@pytest.mark.parametrize("spike", [0.0, 1e4, 1e5])
def test_fixed_point_score_error(spike):
    """Box scores on the fixed-point fields against the unquantised FP64
    fields, on a single-sector star with a flux spike of `spike` sigma and one
    weight 1e6 times the rest: within 1e-6 relative (the bound of the value
    errors, 2^-26 max(|x|, mean |x|), through the box sums)."""
    t, f, e = F.synth_star(seed=1, sectors=1, sigma=1e-3, period=3.7, depth=2e-3,
                           duration=2.5 / 24)
    f, e = f - 1.0, e.copy()
    if spike:
        f[5000] += spike * e[5000]
        e[9000] *= 1e-3
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    d = cfg.durations_days()
    s = F.family_state(t, f, e, d, cfg.correlated_noise)
    vq, wq = O.quantized(s.v), O.quantized(s.w)
    periods = np.concatenate([np.linspace(0.5, 13.7, 150), [3.7]])
    args = (s.baseline_norm, periods, d, s.rho, cfg.ingress_fractions, cfg.shape_layers,
            cfg.shape_gate, 384)
    _, sq, _, wn_q, _ = O.scan_bank(s.t, vq, wq, *args)
    _, sx, _, wn_x, _ = O.scan_bank(s.t, s.v, s.w, *args)
    ok = np.isfinite(sx)
    rel = np.abs(sq[ok] - sx[ok]) / np.abs(sx[ok])
    assert rel.max() < 1e-6, rel.max()
    assert (wn_q == wn_x).all(axis=1).mean() > 0.99


def test_fixed_point_block_words():
    """Block words: a 27-bit mantissa over a 5-bit shift, the smallest shift
    that keeps the mantissa below 2^26 (checked by brute force)."""
    x = np.array([1.0, -3.0, 2.5e7, -7.77e9, 123456789.0, 0.0] + [1.0] * 200)
    q, e, block = O.quantize(x)
    assert block
    for word, xv in zip(q[:6], x[:6]):
        s = int(word) & 31
        m = (int(word) - s) // 32
        r = round(xv * 2.0 ** e)
        assert abs(m) < 2 ** 26
        assert all(abs(round(r / 2 ** k)) >= 2 ** 26 for k in range(s))
        assert m == round(r / 2 ** s)


# ---------------------------------------------------------- phase cell

def test_phase_cell_integer_brute_force():
    """floor(fl(t * scale)) mod nb, checked with Python's IEEE multiply and
    exact integer modulo (kernels.rs:314-327)."""
    rng = np.random.default_rng(5)
    for _ in range(20000):
        t = float(rng.uniform(0, 800))
        nb = int(rng.integers(8, 20000))
        P = float(rng.uniform(0.3, 100))
        scale = nb / P
        want = math.floor(t * scale) % nb
        assert O.phase_cell(t, scale, nb) == want
    # exact bin boundaries: t * scale an integer
    for nb in (8, 96, 4097):
        scale = float(nb)
        for k in range(0, 3 * nb, 7):
            assert O.phase_cell(k / nb, scale, nb) == (k % nb if (k / nb) * scale >= k else (k - 1) % nb)


# ------------------------------------------------------------------ rng

def test_rng_stream_restated():
    """survey/numeric.rs:345-364 restated with Python integers."""
    seed = 12345
    s = (seed * 0x9E3779B97F4A7C15) % 2 ** 64 | 1
    r = O.Rng(seed)
    mask = 2 ** 64 - 1
    for _ in range(1000):
        x = s
        x ^= (x << 13) & mask
        x ^= x >> 7
        x ^= (x << 17) & mask
        s = x
        assert r.uniform() == (x >> 11) / float(1 << 53)


# ------------------------------------------------------ scan brute force

def _staircase_counts(j, k, factors_s, layers, nb):
    """Number of nested layers covering each bin of template (j, k)."""
    c = np.zeros(2 * nb, dtype=np.int64)
    for layer in range(layers):
        width = max(1, int(math.floor(k * factors_s[layer] + 0.5)))
        start = j + (k - width) // 2
        c[start:start + width] += 1
    return c[:nb] + c[nb:]


def brute_scan(t, vq, ev, wq, ew, C, P, durations, rho, ingress, layers, gate,
               nb):
    """Independent restatement of one period's scan (kernels.rs:608-845 with
    f64 ranking): integer per-bin sums, direct window sums, explicit
    staircase weights, global argmax by (merit desc, key asc, di asc)."""
    n = len(t)
    scale = nb / P
    b = np.floor(t * scale).astype(np.int64) % nb
    Hv = np.zeros(nb, dtype=np.int64)
    Hw = np.zeros(nb, dtype=np.int64)
    np.add.at(Hv, b, vq.astype(np.int64))
    np.add.at(Hw, b, wq.astype(np.int64))
    Hn = np.bincount(b, minlength=nb)
    nb1 = nb + 1
    L = float(layers)
    factors = [[1.0 - 2.0 * ing * (l / (layers - 1) if layers > 1 else 0.0)
                for l in range(layers)] for ing in ingress]
    cands = []
    for di, D in enumerate(durations):
        raw_k = max(1, int(math.floor(D * scale + 0.5)))
        k = min(raw_k, nb // 2)
        r = rho[di]
        box = []
        gate_tiles = set()
        for j in range(nb):
            idx = (j + np.arange(k)) % nb
            V = int(Hv[idx].sum()) * 2.0 ** -ev
            W = int(Hw[idx].sum()) * 2.0 ** -ew
            N = int(Hn[idx].sum())
            usable = raw_k < nb // 2 and 3 <= N <= n - 3
            box.append((V, W, usable))
            if usable and V > 0:
                q = W - W * W / C
                m = (V / r) / math.sqrt(max(q, 1e-300))
                m = m if (q > 0 and m < INF) else -INF
                if m > -INF:
                    cands.append((m, k * nb1 + j, di))
                if m >= gate:
                    gate_tiles.add(j // 128)
        for si in range(1, len(ingress)):
            for j in range(nb):
                if j // 128 not in gate_tiles:
                    continue
                V0, W0, usable = box[j]
                if ingress[si] == 0.0:
                    a, hp, norm = V0, W0, W0
                else:
                    c = _staircase_counts(j, k, factors[si], layers, nb)
                    a = int((c * Hv).sum()) * 2.0 ** -ev / L
                    hp = int((c * Hw).sum()) * 2.0 ** -ew / L
                    norm = int((c * c * Hw).sum()) * 2.0 ** -ew / (L * L)
                if usable and a > 0:
                    q = norm - hp * hp / C
                    m = (a / r) / math.sqrt(max(q, 1e-300))
                    if q > 0 and m < INF:
                        cands.append((m, (si * nb1 + k) * nb1 + j, di))
    if not cands:
        return (0, 1, 0, 0), -INF
    m, key, di = max(cands, key=lambda x: (x[0], -x[1], -x[2]))
    return (key % nb1, (key // nb1) % nb1, key // (nb1 * nb1), di), m


def _small_star(seed, n=1500, depth=4e-3, period=1.37, gap=True, sigma=1e-3):
    rng = np.random.default_rng(seed)
    t = np.arange(n) / 96.0  # 15-min cadence, ~15.6 d
    if gap:
        t = t[(t < 6.0) | (t > 7.0)]
    y = rng.normal(0, sigma, len(t))
    ph = ((t - 0.4) % period)
    y += np.where(ph < 0.12, depth, 0.0)  # deficit
    w = np.full(len(t), 1.0 / sigma ** 2) * rng.uniform(0.8, 1.2, len(t))
    C = float(w.sum())
    mean = float((w * y).sum()) / C
    v = w * y - w * mean
    vq, ev, bv = O.quantize(v)
    wq, ew, bw = O.quantize(w)
    assert not bv and not bw  # plain words: the value is q * 2^-e
    return t, vq, ev, wq, ew, C


@pytest.mark.parametrize("nb", [8, 16, 50, 128, 200, 256])
def test_scan_bank_vs_brute_force(nb):
    t, vq, ev, wq, ew, C = _small_star(nb)
    vv = vq * 2.0 ** -ev
    ww = wq * 2.0 ** -ew
    durations = np.array([0.06, 0.09, 0.12, 0.18])
    rho = np.array([1.0, 1.03, 1.0, 1.2])
    ingress = [0.0, 0.2, 0.45]
    rng = np.random.default_rng(nb)
    periods = np.concatenate([[1.37, 2.74], rng.uniform(0.8, 3.0, 10)])
    gate = 0.6 * 7.1
    rec, score, depth, win, st = O.scan_bank(t, vv, ww, C, periods, durations,
                                             rho, ingress, 4, gate, nb)
    assert st == 0
    for i, P in enumerate(periods):
        w_idx, m = brute_scan(t, vq, ev, wq, ew, C, P, durations, rho,
                              ingress, 4, gate, nb)
        assert tuple(win[i]) == w_idx, (nb, P, tuple(win[i]), w_idx)
        assert score[i] == m, (nb, P, score[i], m)


def test_scan_bank_exact_ties_take_smallest_key():
    """Noise-free data with empty phase ranges: many templates have identical
    sums; the winner must be the smallest key (kernels.rs:154-198)."""
    t = np.concatenate([np.arange(0, 3.0, 0.01), np.arange(5.0, 8.0, 0.01)])
    y = np.zeros(len(t))
    y[(t % 1.0) < 0.05] = 1e-3
    w = np.full(len(t), 1e6)
    C = float(w.sum())
    v = w * y - w * (float((w * y).sum()) / C)
    vq, ev, bv = O.quantize(v)
    wq, ew, bw = O.quantize(w)
    assert not bv and not bw  # plain words: the value is q * 2^-e
    vv, ww = vq * 2.0 ** -ev, wq * 2.0 ** -ew
    durations = np.array([0.04, 0.05, 0.06])
    rho = np.ones(3)
    periods = np.array([1.0, 2.0, 0.5])
    for nb in (40, 100, 160):
        rec, score, depth, win, st = O.scan_bank(t, vv, ww, C, periods,
                                                 durations, rho,
                                                 [0.0, 0.2, 0.45], 4, 4.26, nb)
        for i, P in enumerate(periods):
            w_idx, m = brute_scan(t, vq, ev, wq, ew, C, P, durations, rho,
                                  [0.0, 0.2, 0.45], 4, 4.26, nb)
            assert tuple(win[i]) == w_idx and score[i] == m


def test_box_closed_form_depth_and_snr():
    """Noiseless box on a bin-aligned ephemeris: depth = V / q equals the
    injected depth and SNR = delta sqrt(W (1 - W / C)) / rho."""
    nb = 200
    P = 2.0
    t = np.arange(0, 20.0, P / nb / 3.0)  # 3 points per bin
    delta = 5e-4
    k = 6
    j0 = 37
    b = np.floor(t * (nb / P)).astype(int) % nb
    y = np.where((b >= j0) & (b < j0 + k), delta, 0.0)
    w = np.full(len(t), 4e6)
    C = float(w.sum())
    v = w * y - w * (float((w * y).sum()) / C)
    D = k * P / nb  # duration whose width is exactly k bins
    rho = 1.1
    rec, score, depth, win, st = O.scan_bank(
        t, v, w, C, np.array([P]), np.array([D]), np.array([rho]), [0.0], 4,
        4.26, nb)
    W = float(w[(b >= j0) & (b < j0 + k)].sum())
    assert tuple(win[0][:3]) == (j0, k, 0)
    assert abs(depth[0] - delta) < 1e-12 * delta
    snr = delta * math.sqrt(W * (1.0 - W / C)) / rho
    assert abs(score[0] - snr) < 1e-12 * snr


def test_staircase_record_matches_explicit_template():
    """The winner's record (a, norm, d) equals the explicit dot products of
    the staircase template h = (covering layers) / L with the per-bin sums."""
    t, vq, ev, wq, ew, C = _small_star(7, depth=8e-3)
    vv, ww = vq * 2.0 ** -ev, wq * 2.0 ** -ew
    durations = np.array([0.09, 0.12, 0.16])
    rho = np.array([1.0, 1.1, 1.0])
    nb = 120
    periods = np.linspace(1.3, 1.45, 40)
    rec, score, depth, win, st = O.scan_bank(t, vv, ww, C, periods, durations,
                                             rho, [0.0, 0.2, 0.45], 4, 4.26, nb)
    shapes = 0
    for i in range(len(periods)):
        phase, width, si, di = (int(x) for x in win[i])
        if si == 0 or not np.isfinite(score[i]):
            continue
        shapes += 1
        scale = nb / periods[i]
        b = np.floor(t * scale).astype(np.int64) % nb
        Hv = np.bincount(b, weights=vv, minlength=nb)
        Hw = np.bincount(b, weights=ww, minlength=nb)
        ing = [0.0, 0.2, 0.45][si]
        f = [1.0 - 2.0 * ing * l / 3.0 for l in range(4)]
        c = _staircase_counts(phase, width, f, 4, nb)
        h = c / 4.0
        a = float((h * Hv).sum())
        hp = float((h * Hw).sum())
        norm = float((h * h * Hw).sum())
        assert abs(rec[i, 4] - a / rho[di]) < 1e-9 * abs(a / rho[di])
        assert abs(rec[i, 5] - norm) < 1e-9 * norm
        assert abs(rec[i, 6] - hp * hp / C) < 1e-9 * hp * hp / C
    assert shapes > 0


def test_scan_periods_is_prescan_then_full():
    """scan.rs:381-405: the coarse pass at nb / 2, survivors at >= 0.7 thr,
    the full pass at nb, rows at >= thr."""
    t, vq, ev, wq, ew, C = _small_star(11, depth=3e-3)
    vv, ww = vq * 2.0 ** -ev, wq * 2.0 ** -ew

    class Cfg:
        shape_layers = 4
        phase_oversample = 4
        period_batch = 4096
        prescan_factor = 2
        threshold = 7.1
        shape_gate = 0.6 * 7.1
        max_fold_mib = 1024.0
        prescan_fraction = 0.7

    durations = np.array([0.06, 0.09, 0.12])
    rho = np.ones(3)
    ing = [0.0, 0.2, 0.45]
    periods = np.linspace(1.0, 1.9, 300)
    rows, stats, st = O.scan_periods(t, vv, ww, C, 100.0, periods, durations,
                                     rho, ing, Cfg)
    nb = max(8, math.ceil(4 * periods[-1] / durations.min()))
    _, sc, _, _, _ = O.scan_bank(t, vv, ww, C, periods, durations, rho, ing,
                                 4, Cfg.shape_gate, nb // 2)
    surv = periods[np.isfinite(sc) & (sc >= 0.7 * 7.1)]
    assert stats.prescan_survivors == len(surv) and len(surv) > 0
    rec, sc2, dp, win, _ = O.scan_bank(t, vv, ww, C, surv, durations, rho, ing,
                                       4, Cfg.shape_gate, nb)
    keep = np.isfinite(sc2) & (sc2 >= 7.1)
    assert len(rows) == keep.sum() and keep.sum() > 0
    r = rec[keep]
    np.testing.assert_array_equal(rows[:, 0], r[:, 0])
    np.testing.assert_array_equal(rows[:, 1], (r[:, 2] + r[:, 3] / 2.0) / r[:, 1] + 100.0)
    np.testing.assert_array_equal(rows[:, 3], sc2[keep])
    assert stats.templates_scanned == (300 * 3 * 3 * (nb // 2)
                                       + len(surv) * 3 * 3 * nb)

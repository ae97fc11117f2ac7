"""Pins the evidence half of the oracle (oracle/evidence.c), which restates
src/search/evidence.rs and src/search/math.rs of the reference.

* erfc: the reference's own known-answer test (math.rs:183-210, 16 values,
  relative 1e-14), with the oracle's exp and with glibc's.
* exp / log (the oracle's own, shared with the HIP path): within 1 ulp of
  glibc on random and special inputs.
* log_add: identities.  positive_amplitude_evidence: the closed form of the
  half-normal integral, -0.5 ln(1 + tau^2 q) + z^2/2 + ln 2 + log Phi(z)
  with scipy's log_ndtr (including the asymptotic branch z < -26), and
  direct quadrature of the integral.  subspace_upper: the bound property
  against a brute-force maximum over nonnegative depth pairs.
* parameters: the constants of evidence.rs:665-689.
* compare_search: an independent numpy brute force (explicit staircase
  weight vectors dotted with the folded fields, scipy special functions,
  logaddexp over every template) reproduces the exhaustive oracle's winners
  and log Bayes factors; pruning with the default rtol moves them by less
  than the reference's margin; glibc exp/log leave every key unchanged.
"""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

scipy_special = pytest.importorskip("scipy.special")
scipy_integrate = pytest.importorskip("scipy.integrate")


@pytest.fixture(autouse=True)
def _own_math():
    O.use_libm(False)
    yield
    O.use_libm(False)


ERFC_KAT = [  # math.rs:186-202
    (0.0, 1.0), (0.5, 0.4795001221869535), (1.0, 0.15729920705028513),
    (2.0, 0.004677734981047266), (3.0, 2.209049699858544e-05),
    (5.0, 1.5374597944280347e-12), (10.0, 2.088487583762545e-45),
    (-1.0, 1.8427007929497148), (-3.0, 1.9999779095030014),
    (0.3, 0.6713732405408726), (0.46, 0.5153446099983203),
    (0.47, 0.5062549491139179), (3.99, 1.6739211364520824e-08),
    (4.01, 1.4196939421371057e-08), (15.0, 7.212994172451209e-100),
    (26.0, 5.663192408856145e-296)]


@pytest.mark.parametrize("libm", [False, True])
def test_erfc_reference_values(libm):
    O.use_libm(libm)
    for x, want in ERFC_KAT:
        got = O.erfc(x)
        assert abs((got - want) / want) < 1e-14, (x, got, want)


def _ulps(a, b):
    if a == b or (math.isnan(a) and math.isnan(b)):
        return 0
    ia = int(np.float64(a).view(np.int64))
    ib = int(np.float64(b).view(np.int64))
    return abs(ia - ib)


def test_exp_log_within_one_ulp_of_glibc():
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(-745, 709.7, 4000), rng.uniform(-1, 1, 4000),
                         rng.uniform(-40, 40, 4000), [0.0, -0.0, 1.0, -1.0,
                                                      709.78, -745.13, 1e-300]])
    worst = max(_ulps(O.exp(x), math.exp(x)) for x in xs)
    assert worst <= 1
    ys = np.concatenate([np.exp(rng.uniform(-700, 700, 4000)), rng.uniform(0.5, 2, 4000),
                         rng.uniform(0.999, 1.001, 4000), [5e-324, 2.2250738585072014e-308,
                                                           1.0, 2.0, 0.5, 1e308]])
    worst = max(_ulps(O.log(y), math.log(y)) for y in ys)
    assert worst <= 1
    assert O.exp(710.0) == math.inf and O.exp(-746.0) == 0.0
    assert O.log(0.0) == -math.inf and math.isnan(O.log(-1.0))
    assert O.log(math.inf) == math.inf and O.exp(-math.inf) == 0.0


def test_log_add_identities():
    ninf = -math.inf
    assert O.log_add(ninf, ninf) == ninf
    for x in (-700.0, -3.5, 0.0, 2.25, 650.0):
        assert O.log_add(x, ninf) == x and O.log_add(ninf, x) == x
        assert abs(O.log_add(x, x) - (x + math.log(2.0))) <= 4e-16 * max(1.0, abs(x))
    rng = np.random.default_rng(3)
    for _ in range(300):
        a, b = rng.uniform(-30, 30, 2)
        assert abs(O.log_add(a, b) - np.logaddexp(a, b)) <= 2e-15 * max(1.0, abs(a), abs(b))
        assert O.log_add(a, b) == O.log_add(b, a)


def _pae_closed(a, q, tau):
    q = max(q, 0.0)
    prec = 1.0 + tau * tau * q
    z = tau * a / math.sqrt(prec)
    return -0.5 * math.log(prec) + 0.5 * z * z + math.log(2.0) + float(
        scipy_special.log_ndtr(z))


def test_positive_amplitude_evidence_closed_form():
    rng = np.random.default_rng(11)
    cases = [(0.0, 0.0, 0.01), (5.0, 1e4, 0.01), (-5.0, 1e4, 0.1), (1e3, 1e6, 0.01),
             (-3e4, 1e6, 0.1), (-1e6, 1e6, 0.01), (2e4, -5.0, 0.01)]
    cases += [(rng.normal(0, 3e3), 10 ** rng.uniform(2, 8), rng.choice([0.01, 0.1]))
              for _ in range(400)]
    for a, q, tau in cases:
        got = O.positive_amplitude_evidence(a, q, tau)
        want = _pae_closed(a, q, tau)
        # the closed form cancels z^2/2 against log Phi(z): allow its rounding
        z2 = (tau * a) ** 2 / (1.0 + tau * tau * max(q, 0.0))
        tol = 1e-12 * max(1.0, abs(want)) + 4.0 * 2.2e-16 * z2
        assert abs(got - want) <= tol, (a, q, tau, got, want)
    # the asymptotic branch is taken for z < -26
    a, q, tau = -3e4, 1e6, 0.1
    assert tau * a / math.sqrt(1 + tau * tau * q) < -26


def test_positive_amplitude_evidence_quadrature():
    for a, q, tau in [(30.0, 1e4, 0.01), (-20.0, 5e3, 0.1), (300.0, 2e5, 0.01),
                      (0.0, 1e3, 0.05)]:
        def f(d):
            return (2.0 / (tau * math.sqrt(2 * math.pi))) * math.exp(
                -d * d / (2 * tau * tau) + a * d - q * d * d / 2)
        val, _ = scipy_integrate.quad(f, 0.0, math.inf, epsabs=0, epsrel=1e-13, limit=400)
        want = math.log(val)
        got = O.positive_amplitude_evidence(a, q, tau)
        assert abs(got - want) < 1e-9 * max(1.0, abs(want)), (a, q, tau, got, want)


def test_subspace_upper_bounds_the_two_template_gain():
    rng = np.random.default_rng(5)
    g = np.linspace(0.0, 1.0, 161)
    d1, d2 = np.meshgrid(g, g, indexing="ij")
    for _ in range(200):
        q, r = 10 ** rng.uniform(-1, 2, 2)
        cross = rng.uniform(-0.95, 0.95) * math.sqrt(q * r)
        a, b = rng.normal(0, 3, 2)
        # maximiser scale: both depths up to 1.2 x the unconstrained optimum
        det = q * r - cross * cross
        s1 = abs((a * r - b * cross) / det) + abs(a / q) + 1e-3
        s2 = abs((b * q - a * cross) / det) + abs(b / r) + 1e-3
        x, y = d1 * 1.2 * s1, d2 * 1.2 * s2
        gain = a * x + b * y - 0.5 * (q * x * x + 2 * cross * x * y + r * y * y)
        best = float(gain.max())
        ub = O.subspace_upper(a, b, q, r, cross, 1.0)
        assert ub >= best - 1e-12, (a, b, q, r, cross, ub, best)
        if a <= 0 and b <= 0:
            assert ub == pytest.approx(1e-6)
        # interior optimum: the bound is the exact maximum (up to its margin)
        if det > 1e-6 * q * r and a * r - b * cross > 0 and b * q - a * cross > 0:
            exact = 0.5 * (a * a * r + b * b * q - 2 * a * b * cross) / det
            assert (ub - 1e-6) / 1.000001 == pytest.approx(exact, rel=1e-12)
    # single-template degenerate cases (math.rs:164-173)
    assert O.subspace_upper(2.0, 0.0, 4.0, 0.0, 0.0, 1.0) == pytest.approx(0.5 * 4 / 4 * 1.000001 + 1e-6)
    assert O.subspace_upper(0.0, 3.0, 0.0, 9.0, 0.0, 1.0) == pytest.approx(0.5 * 9 / 9 * 1.000001 + 1e-6)


class Cfg:
    """SearchConfig fields of the evidence path (config.rs defaults)."""

    def __init__(self, **kw):
        self.ingress_fractions = [0.0, 0.2, 0.45]
        self.shape_layers = 4
        self.phase_oversample = 4
        self.threshold = 7.1
        self.evidence_rtol = 1e-12
        self.planet_prior = 0.01
        self.binary_prior = 0.01
        self.planet_depth_scale = 0.01
        self.binary_depth_scale = 0.1
        self.diluted_binary_depth_scale = 0.01
        self.stellar_radius_solar = None
        self.planet_radius_scale_rj = 1.0
        self.target_flux_fraction = 1.0
        self.__dict__.update(kw)


def test_parameters():
    p = O.evidence_parameters(Cfg())
    want = [0.01, 0.1, 0.01, math.log(0.01), math.log(0.01), math.log(0.98)]
    want += [0.0, 0.75, 0.10, 0.01, 0.15, 0.10, 0.1, 0.08, 0.20, 0.5, 0.015, 0.25,
             1.0, 0.005, 0.35]
    want += [0.25, 0.5, 0.75, 1.0, 4.0 / 3.0, 2.0, 4.0, 0.001, 0.05]
    assert len(p) == 30
    np.testing.assert_allclose(p, want, rtol=1e-15, atol=0)
    p = O.evidence_parameters(Cfg(stellar_radius_solar=0.5, target_flux_fraction=0.8))
    assert p[0] == pytest.approx((0.10045 / 0.5) ** 2 * 0.8, rel=1e-15)
    assert p[1] == pytest.approx(0.1 * 0.8, rel=1e-15)
    assert p[2] == 0.01


# ------------------------------------------------------------ brute force

def _star(seed, n_days=6.0, period=1.0, depth=3e-3, dur=0.06, secondary=0.0,
          odd_even=1.0):
    """Small prepared star: white noise, box transits, optional secondary
    eclipse at phase 0.5 and alternating primary depths."""
    rng = np.random.default_rng(seed)
    t = np.arange(0.0, n_days, 1.0 / 288.0)  # 5-min cadence
    sigma = 1e-3
    y = rng.normal(0.0, sigma, t.size)
    epoch = 0.37
    ph = ((t - epoch) / period) % 1.0
    cyc = np.floor((t - epoch) / period)
    intr = (ph < dur / period)
    dep = np.where(cyc % 2 == 0, depth, depth * odd_even)
    y = y - intr * dep
    if secondary:
        ph2 = ((t - epoch - 0.5 * period) / period) % 1.0
        y = y - (ph2 < dur / period) * secondary
    w = np.full(t.size, 1.0 / sigma ** 2)
    deficit = -y
    c = w.sum()
    mean = (w * deficit).sum() / c
    v = w * deficit - w * mean
    energy = float((deficit * v).sum())
    return t, v, w, c, energy


def _norm_logpdf_scale(a, q, tau):
    q = np.maximum(q, 0.0)
    prec = 1.0 + tau * tau * q
    z = tau * a / np.sqrt(prec)
    return -0.5 * np.log(prec) + 0.5 * z * z + np.log(2.0) + scipy_special.log_ndtr(z)


def _weights(nb2, start, width, fractions, layers):
    """Explicit staircase h (value and square) over the 2nb bins: layer l
    covers [start + (width - inner_l) // 2, + inner_l) circularly."""
    h = np.zeros(nb2)
    for f in fractions:
        inner = max(int(math.floor(width * f + 0.5)), 1)
        s0 = start + (width - inner) // 2
        idx = (s0 + np.arange(inner)) % nb2
        h[idx] += 1.0 / layers
    return h


def _brute(t, v, w, c, energy, durations, rho, period, cfg):
    """Exhaustive evaluation of every template (no pruning)."""
    order = np.argsort(durations, kind="stable")
    keep = [i for i in order if durations[i] < period / 4]
    dur = np.array([durations[i] for i in keep])
    rh = np.array([rho[i] for i in keep])
    nd, ns, L = len(dur), len(cfg.ingress_fractions), cfg.shape_layers
    n = t.size
    nb = max(8, 4 * int(math.ceil(cfg.phase_oversample * period / dur[0] / 4.0)))
    nb2, scale = 2 * nb, nb / period
    cell = np.floor(t * scale).astype(np.int64) % nb2
    fv = np.bincount(cell, v, nb2)
    fw = np.bincount(cell, w, nb2)
    fn = np.bincount(cell, None, nb2).astype(float)
    dlm = -0.5 * ((n - 1) * np.log(rh * rh) + energy / (rh * rh))
    dlm -= dlm.max()
    ing = np.array(cfg.ingress_fractions)
    lp_s, lb_s = -4 * ing, 2 * ing
    pri_p = lp_s - np.logaddexp.reduce(lp_s)
    pri_b = lb_s - np.logaddexp.reduce(lb_s)
    P = O.evidence_parameters(cfg)
    tau_p, tau_b, tau_d = P[0], P[1], P[2]
    fracs = [[1 - 2 * x * l / (max(L, 2) - 1) for l in range(L)] for x in ing]
    # orbit bank
    phases, logw = [0.25, 0.5, 0.75], np.log([0.1, 0.8, 0.1])
    bank = np.full((nd, 3), -np.inf)
    vc = 0
    for di in range(nd):
        k = max(math.floor(dur[di] * scale + 0.5), 1.0)
        vc += k < nb // 4
        for o in range(3):
            shift = math.floor(phases[o] * nb + 0.5)
            if k < nb // 4 and k < nb // 2 and shift >= k and shift + k <= nb:
                bank[di, o] = logw[o]
    wnorm = math.log(vc) - np.logaddexp.reduce(bank.ravel())
    orb = bank + wnorm
    lp_all, lb_all = [], []
    best_j = (-np.inf, None, None)
    best_s = (-np.inf, None, None)
    for di in range(nd):
        raw_k = max(int(math.floor(dur[di] * scale + 0.5)), 1)
        if not raw_k < nb2 // 8:
            continue
        k = min(raw_k, nb2 // 8 - 1)
        rho2 = rh[di] ** 2
        for si in range(ns):
            for j in range(nb):
                he = _weights(nb2, j, k, fracs[si], L)
                ho = _weights(nb2, j + nb, k, fracs[si], L)
                cnt = fn[(j + np.arange(k)) % nb2].sum() + fn[(j + nb + np.arange(k)) % nb2].sum()
                ae, ao = he @ fv, ho @ fv
                pe, po = he @ fw, ho @ fw
                ne, no = (he * he) @ fw, (ho * ho) @ fw
                qe, qo, qeo = ne - pe * pe / c, no - po * po / c, -pe * po / c
                ap, qp = ae + ao, qe + qo + 2 * qeo
                snr = ap / math.sqrt(max(qp * rho2, 1e-300))
                depth = ap / max(qp, 1e-300)
                valid = (3 <= cnt <= n - 3) and ap > 0 and qp > 0 and snr >= cfg.threshold
                ep_terms, eb_terms, comps = [], [], []
                for o in range(3):
                    if not np.isfinite(orb[di, o]):
                        continue
                    off = math.floor(phases[o] * nb + 0.5)
                    hs = (_weights(nb2, j + off, k, fracs[si], L)
                          + _weights(nb2, j + off + nb, k, fracs[si], L))
                    ass, ps, nsum = hs @ fv, hs @ fw, (hs * hs) @ fw
                    qs, qps = nsum - ps * ps / c, -(pe + po) * ps / c
                    pa, pb, pq, pr, px = ap / rho2, ass / rho2, qp / rho2, qs / rho2, qps / rho2
                    for ri in range(5):
                        ratio, wp, wb = P[6 + 3 * ri], P[7 + 3 * ri], P[8 + 3 * ri]
                        aa = pa + ratio * pb
                        qq = pq + 2 * ratio * px + ratio * ratio * pr
                        ep_terms.append(_norm_logpdf_scale(aa, qq, tau_p) + math.log(wp) + orb[di, o])
                        eb = (np.logaddexp(_norm_logpdf_scale(aa, qq, tau_b),
                                           _norm_logpdf_scale(aa, qq, tau_d))
                              + math.log(0.25 * wb) + orb[di, o])
                        eb_terms.append(eb)
                        comps.append((eb, ratio, phases[o], 1.0, 1.0))
                ba, bb, bq, br, bx = ae / rho2, ao / rho2, qe / rho2, qo / rho2, qeo / rho2
                for ri in range(7):
                    ratio = P[21 + ri]
                    aa = ba + ratio * bb
                    qq = bq + 2 * ratio * bx + ratio * ratio * br
                    eb = (np.logaddexp(_norm_logpdf_scale(aa, qq, tau_b),
                                       _norm_logpdf_scale(aa, qq, tau_d))
                          + math.log(0.25 / 7.0))
                    eb_terms.append(eb)
                    comps.append((eb, 0.0, 0.5, ratio, 2.0))
                lp = np.logaddexp.reduce(ep_terms) + pri_p[si] + dlm[di]
                lb = np.logaddexp.reduce(eb_terms) + pri_b[si] + dlm[di]
                lp_all.append(lp)
                lb_all.append(lb)
                key = (si * nd + di) * nb + j
                comp = max(comps, key=lambda x: x[0])  # first of equal maxima
                vals = dict(snr=snr, depth=depth, lp=lp, lb=lb, comp=comp, key=key,
                            j=j, di=di, si=si)
                if valid and (lp > best_j[0] or (lp == best_j[0] and key < best_j[1])):
                    best_j = (lp, key, vals)
                if valid and (snr > best_s[0] or (snr == best_s[0] and key < best_s[1])):
                    best_s = (snr, key, vals)
    in_bank = [max(math.floor(d * scale + 0.5), 1.0) < nb // 4 for d in dur]
    log_mass = math.log(nb) + np.logaddexp.reduce([dlm[i] for i in range(nd) if in_bank[i]])
    lp_tot = np.logaddexp.reduce(lp_all) - log_mass
    lb_tot = np.logaddexp.reduce(lb_all) - log_mass
    return dict(nb=nb, scale=scale, dur=dur, lp=lp_tot, lb=lb_tot,
                joint=best_j[2], signal=best_s[2], P=P)


@pytest.mark.parametrize("kind", ["planet", "binary"])
def test_compare_search_matches_brute_force(kind):
    if kind == "planet":
        t, v, w, c, energy = _star(21, depth=3e-3)
    else:
        t, v, w, c, energy = _star(22, depth=4e-3, secondary=1.5e-3, odd_even=0.6)
    durations = np.array([0.5, 0.75, 1.0, 1.5, 2.0]) / 24.0
    rho = np.array([1.0, 1.01, 1.03, 1.02, 1.05])
    period = 1.0
    cfg = Cfg(evidence_rtol=0.0)
    rows = O.compare_search(t, v, w, c, 0.0, energy, durations, rho, period, cfg)
    ref = _brute(t, v, w, c, energy, durations, rho, period, cfg)
    assert rows, "the injected signal must pass the screen"
    by_flag = {r["selection_flags"]: r for r in rows}
    jrow = by_flag.get(1) or by_flag.get(3)
    srow = by_flag.get(2) or by_flag.get(3)
    for row, win in ((jrow, ref["joint"]), (srow, ref["signal"])):
        assert (row["phase"], row["duration_index"], row["shape"]) == (
            win["j"], win["di"], win["si"])
        assert row["snr"] == pytest.approx(win["snr"], rel=1e-9)
        assert row["depth"] == pytest.approx(win["depth"], rel=1e-9)
        width = max(math.floor(ref["dur"][win["di"]] * ref["scale"] + 0.5), 1.0)
        assert row["epoch"] == pytest.approx((win["j"] + width / 2) / ref["scale"], rel=1e-15)
        assert row["log_bf_planet_noise"] == pytest.approx(ref["lp"], rel=1e-9)
        assert row["log_bf_binary_noise"] == pytest.approx(ref["lb"], rel=1e-9)
        P = ref["P"]
        joint = P[3] + ref["lp"] - np.logaddexp(P[5], P[4] + ref["lb"])
        assert row["planet_log_odds"] == pytest.approx(joint, rel=1e-9, abs=1e-9)
        _, ratio, phase, odd, mult = win["comp"]
        assert row["binary_secondary_ratio"] == ratio
        assert row["binary_secondary_phase"] == phase
        assert row["binary_odd_even_ratio"] == odd
        assert row["binary_period_days"] == mult * period
    if kind == "binary":
        assert jrow["log_bf_binary_noise"] > jrow["log_bf_planet_noise"]
    else:
        assert jrow["planet_log_odds"] > 0


def _single_sector_state():
    import fbls_amd as F
    from fbls_amd import synth

    t, f, e = synth.single_sector(seed=1)
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    d = cfg.durations_days()
    s = F.family_state(t, f - 1.0, e, d, cfg.correlated_noise)
    return s, d, cfg


def test_pruning_within_margin_and_libm_invariance():
    s, d, cfg = _single_sector_state()
    args = (s.t, s.v, s.w, s.baseline_norm, s.origin, s.residual_energy, d, s.rho)
    pruned, dbg = O.compare_search(*args, 3.7, cfg, debug=True)
    assert dbg["parts"][:, :, 13].sum() > 0, "pruning must omit something here"
    cfg0 = Cfg(evidence_rtol=0.0)
    full = O.compare_search(*args, 3.7, cfg0)
    assert len(pruned) == len(full) > 0
    for a, b in zip(pruned, full):
        for k in ("phase", "duration_index", "shape", "selection_flags"):
            assert a[k] == b[k]
        for k in ("log_bf_planet_noise", "log_bf_binary_noise"):
            # omitted mass < rtol of the total (evidence.rs:865-869)
            assert abs(a[k] - b[k]) <= 1e-11
    O.use_libm(True)
    glibc = O.compare_search(*args, 3.7, cfg)
    O.use_libm(False)
    assert len(glibc) == len(pruned)
    for a, b in zip(pruned, glibc):
        for k in ("phase", "duration_index", "shape", "selection_flags"):
            assert a[k] == b[k]
        for k, x in a.items():
            if isinstance(x, float) and math.isfinite(x):
                assert b[k] == pytest.approx(x, rel=1e-12, abs=1e-12), k


def test_screen_below_threshold_gives_no_rows():
    s, d, cfg = _single_sector_state()
    args = (s.t, s.v, s.w, s.baseline_norm, s.origin, s.residual_energy, d, s.rho)
    assert O.compare_search(*args, 1.3, cfg) == []
    # no duration below a quarter period
    assert O.compare_search(*args, 0.05, cfg) == []

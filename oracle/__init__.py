"""ctypes view of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker of the HIP path and never
as the thing measured or shipped.  The oracle restates the reference
(isakruas/fbls) search path in C; see oracle/fbls_oracle.h for the functions
and the reference lines each one follows, and tests/test_oracle.py for how
each is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.c_double
_sz = ctypes.c_size_t
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_P = ctypes.POINTER


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class ScanStats(ctypes.Structure):
    _fields_ = [("periods", _u64), ("prescan_survivors", _u64),
                ("phase_bins", _u64), ("batches", _u64),
                ("templates_scanned", _u64)]


class Row(ctypes.Structure):
    _fields_ = [("period", _d), ("epoch", _d), ("duration", _d),
                ("score", _d), ("depth", _d)]


class BankCfg(ctypes.Structure):
    _fields_ = [("layers", _u32), ("phase_oversample", _u32),
                ("period_batch", _u32), ("prescan_factor", _u32),
                ("threshold", _d), ("shape_gate", _d),
                ("max_fold_mib", _d), ("prescan_fraction", _d)]


class SearchCfg(ctypes.Structure):
    _fields_ = [("period_min", _d), ("period_max", _d),
                ("period_oversample", _d)]


class EvidenceCfg(ctypes.Structure):
    _fields_ = [("ingress", _P(_d)), ("ns", _sz), ("layers", _u32),
                ("phase_oversample", _u32), ("threshold", _d),
                ("evidence_rtol", _d), ("planet_prior", _d),
                ("binary_prior", _d), ("planet_depth_scale", _d),
                ("binary_depth_scale", _d),
                ("diluted_binary_depth_scale", _d),
                ("stellar_radius_solar", _d), ("planet_radius_scale_rj", _d),
                ("target_flux_fraction", _d)]


class EvidenceRow(ctypes.Structure):
    _fields_ = [(k, _d) for k in (
        "period", "epoch", "duration", "snr", "depth", "planet_log_odds",
        "log_bf_planet_noise", "log_bf_binary_noise",
        "binary_secondary_ratio", "binary_secondary_phase",
        "binary_odd_even_ratio", "binary_period_days", "ingress_fraction")] + [
        ("selection_flags", _u32), ("phase", _u32), ("duration_index", _u32),
        ("shape", _u32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OrcPeak(ctypes.Structure):
    _fields_ = ([(k, _d) for k in (
        "period_days", "epoch_bjd", "duration_hours", "snr", "snr_folded",
        "snr_without_strongest_event", "depth", "depth_ppm")]
        + [(k, _u64) for k in ("events", "family", "pass_", "row")]
        + [(k, _d) for k in (
            "planet_log_odds", "log_bf_planet_noise", "log_bf_binary_noise",
            "binary_secondary_ratio", "binary_secondary_phase",
            "binary_odd_even_ratio", "binary_period_days", "ingress_fraction",
            "evidence_epoch_bjd", "evidence_duration_hours")]
        + [("model_preferred", _u32), ("binary_diagnostics_complete", _u32)])

    def as_dict(self):
        d = {k.rstrip("_"): getattr(self, k) for k, _ in self._fields_}
        d["model_preferred"] = bool(d["model_preferred"])
        d["binary_diagnostics_complete"] = bool(d["binary_diagnostics_complete"])
        return d


class OrcPassStats(ctypes.Structure):
    _fields_ = [("pass_", _u64), ("rows", _u64), ("periods", _u64), ("accepted", _u64)]

    def as_dict(self):
        return {k.rstrip("_"): getattr(self, k) for k, _ in self._fields_}


class OrcPipelineCfg(ctypes.Structure):
    _fields_ = [("durations", _P(_d)), ("nd", _sz), ("ingress", _P(_d)), ("ns", _sz),
                ("correlated", ctypes.c_int), ("bank", BankCfg), ("search", SearchCfg),
                ("max_candidates", _sz), ("passes", _sz), ("min_events", _sz),
                ("evidence_candidates", _sz), ("model_comparison", ctypes.c_int),
                ("planet_log_odds_min", _d), ("evidence", EvidenceCfg),
                ("single_events", ctypes.c_int)]


class EventStats(ctypes.Structure):
    _fields_ = [("templates_scanned", _u64), ("threshold_survivors", _u64),
                ("overflow", _u32), ("phase_bins", _u32)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["overflow"] = bool(d["overflow"])
        return d


class SingleEvent(ctypes.Structure):
    _fields_ = [(k, _d) for k in ("epoch_bjd", "duration_hours", "snr", "depth",
                                  "depth_ppm")] + [("points", _u64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class EventsOut(ctypes.Structure):
    _fields_ = [("events", _P(SingleEvent)), ("events_cap", _sz), ("nevents", _sz),
                ("stats", _P(EventStats)), ("stats_cap", _sz), ("nstats", _sz)]


class PrepCfg(ctypes.Structure):
    _fields_ = [(k, _d) for k in ("bin_minutes", "window_min", "window_factor",
                                  "window_max", "clip_sigma", "edge_trim")]


def set_threads(n):
    """Worker threads of the oracle's parallel loops (0: the OpenMP default)."""
    lib().orc_set_threads(int(n))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_durations_days.restype = _sz
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_period_grid_count.restype = _sz
        L.orc_period_grid_count.argtypes = [_d, _d, _d, _d, _d]
        L.orc_period_grid.restype = None
        L.orc_period_grid.argtypes = [_d, _d, _d, _d, _d, _P(_d), _sz]
        L.orc_median_lower.restype = _d
        L.orc_mad_scale.restype = _d
        L.orc_phase_cell.restype = _u32
        L.orc_phase_cell.argtypes = [_d, _d, _u32]
        L.orc_bank_batches.restype = _u64
        L.orc_bank_batches.argtypes = [_sz, _u32, _u32, _d]
        L.orc_rng_new.restype = _u64
        L.orc_rng_new.argtypes = [_u64]
        L.orc_rng_uniform.restype = _d
        L.orc_rng_uniform.argtypes = [_P(_u64)]
        L.orc_plan_bands.restype = _sz
        L.orc_rng_normal.restype = _d
        L.orc_rng_normal.argtypes = [_P(_u64), _d, _d]
        L.orc_quantize.restype = ctypes.c_int
        L.orc_fixed_value.restype = _d
        L.orc_fixed_value.argtypes = [ctypes.c_int32, ctypes.c_int]
        for f in ("orc_exp", "orc_log", "orc_erfc"):
            getattr(L, f).restype = _d
            getattr(L, f).argtypes = [_d]
        L.orc_log_add.restype = _d
        L.orc_log_add.argtypes = [_d, _d]
        L.orc_positive_amplitude_evidence.restype = _d
        L.orc_positive_amplitude_evidence.argtypes = [_d, _d, _d]
        L.orc_subspace_upper.restype = _d
        L.orc_subspace_upper.argtypes = [_d] * 6
        L.orc_evidence_use_libm.argtypes = [ctypes.c_int]
        L.orc_evidence_use_libm.restype = None
        L.orc_evidence_parameters.argtypes = [_P(EvidenceCfg), _P(_d)]
        L.orc_evidence_parameters.restype = None
        L.orc_search_prepared.argtypes = [
            _P(_d), _P(_d), _P(_d), _sz, _P(OrcPipelineCfg), _P(OrcPeak), _sz,
            _P(_sz), _P(OrcPassStats), _P(_sz), _P(EventsOut)]
        L.orc_single_event_search.argtypes = [
            _P(_d), _P(_d), _P(_d), _sz, _d, _d, _P(_d), _P(_d), _sz, _P(_d), _sz, _sz, _d,
            _u32, _sz, _P(Row), _P(_sz), _P(EventStats)]
        L.orc_single_event_search.restype = None
        L.orc_select_events.argtypes = [_P(Row), _sz, _P(_d), _sz, _d, _sz,
                                        _P(_P(ctypes.c_uint8)), _sz, _P(SingleEvent)]
        L.orc_select_events.restype = _sz
        L.orc_select_peaks.restype = _sz
        L.orc_select_peaks.argtypes = [
            _sz, _P(_d), _P(_d), _P(_d), _P(_sz), _P(_d), _P(_d), _P(_d), _P(_d), _P(_sz),
            _P(Row), _P(_sz), _sz, _P(_d), _sz, _d, _sz, _sz, _P(ctypes.c_uint8), _sz,
            _P(ctypes.c_uint8), _P(OrcPeak), _sz]
        L.orc_coarse_features.argtypes = [_P(_d), _P(_d), _P(_d), _sz, _d, _P(_d), _sz,
                                          _P(_d), _sz, _u32, _u32, _P(_d)]
        L.orc_coarse_features.restype = None
        L.orc_fit_segment_noise_ex.argtypes = [
            _P(_d), _P(_d), _P(_d), _P(ctypes.c_int64), _sz, _d, _d, _P(_d), _P(_d), _P(_d),
            _P(_sz), _P(_d), _P(_d), _P(_sz)]
        L.orc_biweight_flatten.argtypes = [_P(_d), _P(_d), _sz, _d, _d, _d, _sz, _P(_d)]
        L.orc_biweight_flatten.restype = None
        L.orc_prepare_families.argtypes = [
            _P(_d), _P(_d), _P(_d), _P(ctypes.c_int64), _sz, _P(_d), _sz, _P(PrepCfg),
            _P(_sz), _P(_sz), _P(_d), _P(_d), _P(_d), _P(_d), _P(_sz), _P(ctypes.c_uint8)]
        L.orc_search_transits.argtypes = [
            _P(_d), _P(_d), _P(_d), _P(ctypes.c_int64), _sz, _P(PrepCfg), _P(OrcPipelineCfg),
            _P(OrcPeak), _sz, _P(_sz), _P(OrcPassStats), _P(_sz), _P(EventsOut)]
        L.orc_transit_mask.argtypes = [_P(_d), _sz, _d, _d, _d, _d, _P(ctypes.c_uint8)]
        L.orc_transit_mask.restype = None
        L.orc_event_count.argtypes = [_P(_d), _sz, _d, _d, _d, _sz]
        L.orc_event_count.restype = _sz
        L.orc_rescore_row.argtypes = [_P(_d), _P(_d), _P(_d), _P(ctypes.c_uint8), _sz,
                                      _d, _d, _d, _d, _P(_d)]
        L.orc_rescore_row.restype = None
        L.orc_count_row.argtypes = [_P(_d), _P(ctypes.c_uint8), _sz, _d, _d, _d, _P(_u32)]
        L.orc_count_row.restype = None
        L.orc_compare_search_ex.argtypes = [
            _P(_d), _P(_d), _P(_d), _sz, _d, _d, _d, _P(_d), _P(_d), _sz, _d,
            _P(EvidenceCfg), _P(EvidenceRow), _P(_sz), _P(_d), _sz, _P(_d),
            _sz, _P(_d), _P(_sz), _P(_sz)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_P(_d))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def durations_days(hours):
    h = _f64(hours)
    out = np.empty_like(h)
    m = lib().orc_durations_days(_dp(h), _sz(len(h)), _dp(out))
    return out[:m]


def period_grid(minimum, maximum, baseline, dmin, oversample):
    L = lib()
    n = L.orc_period_grid_count(minimum, maximum, baseline, dmin, oversample)
    out = np.empty(n)
    L.orc_period_grid(minimum, maximum, baseline, dmin, oversample, _dp(out), n)
    return out


def median_lower(x):
    x = _f64(x)
    return lib().orc_median_lower(_dp(x), _sz(len(x)))


def mad_scale(x):
    x = _f64(x)
    return lib().orc_mad_scale(_dp(x), _sz(len(x)))


def fit_segment_noise(time, flux, err, gap=0.75):
    time, flux = _f64(time), _f64(flux)
    n = len(time)
    t, d, v = np.empty(n), np.empty(n), np.empty(n)
    m, org, infl, seg = _sz(), _d(), _d(), _sz()
    e = _f64(err) if err is not None else None
    st = lib().orc_fit_segment_noise(
        _dp(time), _dp(flux), _dp(e) if e is not None else None, _sz(n),
        _d(gap), _dp(t), _dp(d), _dp(v), ctypes.byref(m), ctypes.byref(org),
        ctypes.byref(infl), ctypes.byref(seg))
    if st:
        raise ValueError(f"fit_segment_noise status {st}")
    k = m.value
    return dict(time=t[:k], deficit=d[:k], variance=v[:k], origin=org.value,
                inflation=infl.value, segments=seg.value)


def prepare_ready(deficit, variance):
    deficit, variance = _f64(deficit), _f64(variance)
    n = len(deficit)
    v, w = np.empty(n), np.empty(n)
    c, e = _d(), _d()
    st = lib().orc_prepare_ready(_dp(deficit), _dp(variance), _sz(n), _dp(v),
                                 _dp(w), ctypes.byref(c), ctypes.byref(e))
    if st:
        raise ValueError(f"prepare_ready status {st}")
    return dict(v=v, w=w, baseline_norm=c.value, residual_energy=e.value)


def correlated_noise_factors(time, residual, sigma, durations):
    time, residual, sigma = _f64(time), _f64(residual), _f64(sigma)
    durations = _f64(durations)
    rho = np.empty(len(durations))
    lib().orc_correlated_noise_factors(_dp(time), _dp(residual), _dp(sigma),
                                       _sz(len(time)), _dp(durations),
                                       _sz(len(durations)), _dp(rho))
    return rho


def family_state(time, flux, sigma, durations, correlated=True):
    """curve.rs:541-569 for a prepared curve (flux relative to zero)."""
    model = fit_segment_noise(time, 1.0 + _f64(flux), sigma, 0.75)
    ready = prepare_ready(model["deficit"], model["variance"])
    if correlated:
        rho = correlated_noise_factors(model["time"], -model["deficit"],
                                       np.sqrt(model["variance"]), durations)
    else:
        rho = np.ones(len(durations))
    return dict(t=model["time"], v=ready["v"], w=ready["w"],
                baseline_norm=ready["baseline_norm"], origin=model["origin"],
                rho=rho, inflation=model["inflation"],
                segments=model["segments"])


def quantize(x):
    """(words, e, block) of the fixed-point representation (orc_quantize)."""
    x = _f64(x)
    q = np.empty(len(x), dtype=np.int32)
    block = ctypes.c_int(0)
    e = lib().orc_quantize(_dp(x), _sz(len(x)), q.ctypes.data_as(_P(ctypes.c_int32)),
                           ctypes.byref(block))
    return q, e, bool(block.value)


def fixed_values(q, block):
    """The integers (units of 2^-e) the words stand for."""
    q = np.asarray(q, dtype=np.int64)
    if not block:
        return q.astype(np.float64)
    s = q & 31
    return np.ldexp(((q - s) // 32).astype(np.float64), s.astype(np.int32))


def quantized(x):
    """The values the device sees for x: fixed_values(q) * 2^-e."""
    q, e, block = quantize(x)
    return np.ldexp(fixed_values(q, block), -e)


def phase_cell(t, scale, nb):
    return lib().orc_phase_cell(t, scale, nb)


def scan_bank(t, v, w, baseline_norm, periods, durations, rho, ingress,
              layers, gate, nb):
    """kernels.rs fold/prefix/scan/merge at nb bins, f64 ranking everywhere.
    Returns (records[np,9], score[np], depth[np], winner[np,4], status)."""
    t, v, w = _f64(t), _f64(v), _f64(w)
    periods = _f64(periods)
    durations, rho, ingress = _f64(durations), _f64(rho), _f64(ingress)
    np_ = len(periods)
    rec = np.empty((np_, 9))
    score = np.empty(np_)
    depth = np.empty(np_)
    win = np.empty((np_, 4), dtype=np.uint32)
    st = lib().orc_scan_bank(
        _dp(t), _dp(v), _dp(w), _sz(len(t)), _d(baseline_norm), _dp(periods),
        _sz(np_), _dp(durations), _dp(rho), _sz(len(durations)), _dp(ingress),
        _sz(len(ingress)), _u32(layers), _d(gate), _u32(nb), _dp(rec),
        _dp(score), _dp(depth), win.ctypes.data_as(_P(_u32)))
    return rec, score, depth, win, st


def _bank_cfg(cfg):
    b = BankCfg()
    b.layers = cfg.shape_layers
    b.phase_oversample = cfg.phase_oversample
    b.period_batch = cfg.period_batch
    b.prescan_factor = cfg.prescan_factor
    b.threshold = cfg.threshold
    b.shape_gate = cfg.shape_gate
    b.max_fold_mib = cfg.max_fold_mib
    b.prescan_fraction = cfg.prescan_fraction
    return b


def scan_periods(t, v, w, baseline_norm, origin, periods, durations, rho,
                 ingress, cfg):
    t, v, w, periods = _f64(t), _f64(v), _f64(w), _f64(periods)
    durations, rho, ingress = _f64(durations), _f64(rho), _f64(ingress)
    rows = (Row * max(1, len(periods)))()
    nrows = _sz()
    stats = ScanStats()
    b = _bank_cfg(cfg)
    st = lib().orc_scan_periods(
        _dp(t), _dp(v), _dp(w), _sz(len(t)), _d(baseline_norm), _d(origin),
        _dp(periods), _sz(len(periods)), _dp(durations), _dp(rho),
        _sz(len(durations)), _dp(ingress), _sz(len(ingress)),
        ctypes.byref(b), rows, ctypes.byref(nrows), ctypes.byref(stats))
    out = np.array([(r.period, r.epoch, r.duration, r.score, r.depth)
                    for r in rows[:nrows.value]]).reshape(-1, 5)
    return out, stats, st


def signal_search(t, v, w, baseline_norm, origin, durations, rho, ingress,
                  cfg, period_min, period_max, period_oversample=3.0):
    t, v, w = _f64(t), _f64(v), _f64(w)
    durations, rho, ingress = _f64(durations), _f64(rho), _f64(ingress)
    b = _bank_cfg(cfg)
    s = SearchCfg(period_min, period_max, period_oversample)
    maximum = min(period_max, t[-1])
    minimum = max(period_min, 4.0 * durations.min() * (1.0 + 1e-9))
    cap = 1
    if maximum > minimum:
        cap = lib().orc_period_grid_count(minimum, maximum, t[-1],
                                          durations.min(), period_oversample)
    rows = (Row * cap)()
    stats = (ScanStats * 64)()
    nrows, nbands, gcount = _sz(), _sz(), _sz()
    st = lib().orc_signal_search(
        _dp(t), _dp(v), _dp(w), _sz(len(t)), _d(baseline_norm), _d(origin),
        _dp(durations), _dp(rho), _sz(len(durations)), _dp(ingress),
        _sz(len(ingress)), ctypes.byref(b), ctypes.byref(s), rows, _sz(cap),
        ctypes.byref(nrows), stats, ctypes.byref(nbands), ctypes.byref(gcount))
    out = np.array([(r.period, r.epoch, r.duration, r.score, r.depth)
                    for r in rows[:nrows.value]]).reshape(-1, 5)
    band_stats = [dict(periods=x.periods, prescan_survivors=x.prescan_survivors,
                       phase_bins=x.phase_bins, batches=x.batches,
                       templates_scanned=x.templates_scanned)
                  for x in stats[:nbands.value]]
    return out, band_stats, gcount.value, st


class Band(ctypes.Structure):
    _fields_ = [("lo", _d), ("first", _sz), ("count", _sz), ("nd", _sz),
                ("nb", _u32), ("nb_coarse", _u32)]


def plan_bands(grid, sorted_durations, phase_oversample=4, prescan_factor=2):
    grid, sd = _f64(grid), _f64(sorted_durations)
    out = (Band * 64)()
    m = lib().orc_plan_bands(_dp(grid), _sz(len(grid)), _dp(sd), _sz(len(sd)),
                             _u32(phase_oversample), _u32(prescan_factor), out,
                             _sz(64))
    return [dict(lo=b.lo, first=b.first, count=b.count, nd=b.nd, nb=b.nb,
                 nb_coarse=b.nb_coarse) for b in out[:min(m, 64)]]


def bank_batches(np_, nb, period_batch, max_fold_mib):
    return lib().orc_bank_batches(np_, nb, period_batch, max_fold_mib)


class Rng:
    """survey/numeric.rs:345-364 restated in C."""

    def __init__(self, seed):
        self.state = _u64(lib().orc_rng_new(seed))

    def uniform(self):
        return lib().orc_rng_uniform(ctypes.byref(self.state))

    def normal(self, mean, sigma):
        return lib().orc_rng_normal(ctypes.byref(self.state), mean, sigma)


# ------------------------------------------------------------ evidence

def exp(x):
    return lib().orc_exp(float(x))


def log(x):
    return lib().orc_log(float(x))


def erfc(x):
    return lib().orc_erfc(float(x))


def log_add(x, y):
    return lib().orc_log_add(float(x), float(y))


def positive_amplitude_evidence(a, q, tau):
    return lib().orc_positive_amplitude_evidence(float(a), float(q), float(tau))


def subspace_upper(a, b, q, r, cross, rho2=1.0):
    return lib().orc_subspace_upper(float(a), float(b), float(q), float(r),
                                    float(cross), float(rho2))


def use_libm(on):
    """Switch the evidence restatement between its own exp/log and glibc's."""
    lib().orc_evidence_use_libm(1 if on else 0)


def evidence_cfg(cfg):
    """EvidenceCfg from a SearchConfig-like object; the ingress array is
    kept alive on the returned structure."""
    ing = np.ascontiguousarray(cfg.ingress_fractions, dtype=np.float64)
    c = EvidenceCfg(
        ingress=_dp(ing), ns=len(ing), layers=cfg.shape_layers,
        phase_oversample=cfg.phase_oversample, threshold=cfg.threshold,
        evidence_rtol=cfg.evidence_rtol, planet_prior=cfg.planet_prior,
        binary_prior=cfg.binary_prior,
        planet_depth_scale=cfg.planet_depth_scale,
        binary_depth_scale=cfg.binary_depth_scale,
        diluted_binary_depth_scale=cfg.diluted_binary_depth_scale,
        stellar_radius_solar=(cfg.stellar_radius_solar
                              if cfg.stellar_radius_solar else 0.0),
        planet_radius_scale_rj=cfg.planet_radius_scale_rj,
        target_flux_fraction=cfg.target_flux_fraction)
    c._keep = ing
    return c


def evidence_parameters(cfg):
    c = evidence_cfg(cfg)
    out = np.empty(30)
    lib().orc_evidence_parameters(ctypes.byref(c), _dp(out))
    return out


def compare_search(t, v, w, baseline_norm, origin, residual_energy,
                   durations, rho, period, cfg, debug=False):
    """compare_search (evidence.rs:693-1003) for one family at one period.
    Returns a list of row dicts; with debug, also
    {screen, parts, floors, tiles, nb}."""
    t, v, w = _f64(t), _f64(v), _f64(w)
    d, r = _f64(durations), _f64(rho)
    c = evidence_cfg(cfg)
    rows = (EvidenceRow * 2)()
    nrows = _sz()
    cap = 4 * (int(cfg.phase_oversample * period / max(d.min(), 1e-9)) + 8)
    screen = np.zeros(cap * 4)
    parts = np.zeros(cap * 30)
    floors = np.zeros(3)
    tiles, nb = _sz(), _sz()
    lib().orc_compare_search_ex(
        _dp(t), _dp(v), _dp(w), len(t), baseline_norm, origin,
        residual_energy, _dp(d), _dp(r), len(d), period, ctypes.byref(c),
        rows, ctypes.byref(nrows), _dp(screen), screen.size, _dp(parts),
        parts.size, _dp(floors), ctypes.byref(tiles), ctypes.byref(nb))
    out = [rows[i].as_dict() for i in range(nrows.value)]
    if not debug:
        return out
    k = tiles.value
    return out, {"screen": screen[:4 * k].reshape(k, 4),
                 "parts": parts[:30 * k].reshape(k, 2, 15),
                 "floors": floors, "tiles": k, "nb": nb.value}


# ------------------------------------------------------------ pipeline

def transit_mask(time, period, epoch, duration, margin=0.0):
    time = _f64(time)
    out = np.zeros(len(time), dtype=np.uint8)
    lib().orc_transit_mask(_dp(time), len(time), period, epoch, duration, margin,
                           out.ctypes.data_as(_P(ctypes.c_uint8)))
    return out.astype(bool)


def event_count(time, period, epoch, duration, minimum_points=3):
    time = _f64(time)
    return lib().orc_event_count(_dp(time), len(time), period, epoch, duration,
                                 minimum_points)


def rescore_row(time, v, w, excluded, origin, period, epoch, duration):
    time, v, w = _f64(time), _f64(v), _f64(w)
    ex = np.ascontiguousarray(excluded, dtype=np.uint8)
    out = np.zeros(3)
    lib().orc_rescore_row(_dp(time), _dp(v), _dp(w), ex.ctypes.data_as(_P(ctypes.c_uint8)),
                          len(time), origin, period, epoch, duration, _dp(out))
    return out


def count_row(time_all, excluded, period, epoch, duration):
    t = _f64(time_all)
    ex = np.ascontiguousarray(excluded, dtype=np.uint8)
    out = np.zeros(2, dtype=np.uint32)
    lib().orc_count_row(_dp(t), ex.ctypes.data_as(_P(ctypes.c_uint8)), len(t), period,
                        epoch, duration, out.ctypes.data_as(_P(_u32)))
    return out


def _pipeline_cfg(cfg, keep):
    d = _f64(cfg.durations_days())
    ing = _f64(cfg.ingress_fractions)
    ev = evidence_cfg(cfg)
    c = OrcPipelineCfg(
        durations=_dp(d), nd=len(d), ingress=_dp(ing), ns=len(ing),
        correlated=int(bool(cfg.correlated_noise)), bank=_bank_cfg(cfg),
        search=SearchCfg(cfg.period_min, cfg.period_max, cfg.period_oversample),
        max_candidates=cfg.max_candidates, passes=cfg.passes,
        min_events=cfg.min_events, evidence_candidates=cfg.evidence_candidates,
        model_comparison=int(bool(cfg.model_comparison)),
        planet_log_odds_min=cfg.planet_log_odds_min, evidence=ev,
        single_events=int(bool(getattr(cfg, "single_events", False))))
    keep.extend([d, ing])
    return c


def _events_out(cfg, nfam_cap, keep):
    cap = max(cfg.max_candidates, 1)
    ev = (SingleEvent * cap)()
    st = (EventStats * max(nfam_cap, 1))()
    keep.extend([ev, st])
    return EventsOut(ev, cap, 0, st, len(st), 0)


def _events_result(out):
    return dict(events=[out.events[i].as_dict() for i in range(min(out.nevents, out.events_cap))],
                event_stats=[out.stats[i].as_dict() for i in range(min(out.nstats, out.stats_cap))])


def search_prepared(time, flux, err, cfg, with_events=False):
    """search_prepared + search_families (pipeline.rs:41-221, no single
    events): (peaks in acceptance order, pass statistics)."""
    time, flux, err = _f64(time), _f64(flux), _f64(err)
    keep = []
    c = _pipeline_cfg(cfg, keep)
    cap = max(cfg.max_candidates, 1)
    peaks = (OrcPeak * cap)()
    passes = (OrcPassStats * max(cfg.passes, 1))()
    npk, npa = _sz(), _sz()
    eo = _events_out(cfg, 1, keep)
    st = lib().orc_search_prepared(_dp(time), _dp(flux), _dp(err), len(time),
                                   ctypes.byref(c), peaks, cap, ctypes.byref(npk),
                                   passes, ctypes.byref(npa), ctypes.byref(eo))
    if st:
        raise ValueError(f"search_prepared status {st}")
    out = ([peaks[i].as_dict() for i in range(min(npk.value, cap))],
           [passes[i].as_dict() for i in range(npa.value)])
    return out + (_events_result(eo),) if with_events else out


def coarse_features(t, v, w, baseline_norm, periods, durations, nb_min=16, nb_max=1024):
    """Coarse periodogram of one star: [period, duration]."""
    t, v, w = _f64(t), _f64(v), _f64(w)
    periods, durations = _f64(periods), _f64(durations)
    out = np.empty((len(periods), len(durations)))
    lib().orc_coarse_features(_dp(t), _dp(v), _dp(w), len(t), baseline_norm, _dp(periods),
                              len(periods), _dp(durations), len(durations), nb_min, nb_max,
                              _dp(out))
    return out


# ------------------------------------------------------------ raw mode

def _i64p(a):
    return None if a is None else a.ctypes.data_as(_P(ctypes.c_int64))


def fit_segment_noise_ex(time, flux, err=None, group=None, bin_minutes=0.0, gap=0.75):
    time, flux = _f64(time), _f64(flux)
    n = len(time)
    e = _f64(err) if err is not None else None
    g = np.ascontiguousarray(group, dtype=np.int64) if group is not None else None
    t, d, v = np.empty(n), np.empty(n), np.empty(n)
    m, org, infl, seg = _sz(), _d(), _d(), _sz()
    st = lib().orc_fit_segment_noise_ex(
        _dp(time), _dp(flux), _dp(e) if e is not None else None, _i64p(g), n, bin_minutes,
        gap, _dp(t), _dp(d), _dp(v), ctypes.byref(m), ctypes.byref(org), ctypes.byref(infl),
        ctypes.byref(seg))
    if st:
        raise ValueError(f"fit_segment_noise status {st}")
    k = m.value
    return dict(time=t[:k], deficit=d[:k], variance=v[:k], origin=org.value,
                inflation=infl.value, segments=seg.value)


def biweight_flatten(time, rel, window, gap=0.5, c=5.0, iterations=10):
    time, rel = _f64(time), _f64(rel)
    out = np.empty(len(time))
    lib().orc_biweight_flatten(_dp(time), _dp(rel), len(time), window, gap, c, iterations,
                               _dp(out))
    return out


def prep_cfg(cfg):
    return PrepCfg(cfg.bin_minutes, cfg.window_days, cfg.window_factor, cfg.window_max_days,
                   cfg.clip_sigma, cfg.edge_trim_windows)


def prepare_families(time, flux, err, group, durations, cfg):
    """prepare_families (curve.rs:387-425): list of family dicts."""
    time, flux = _f64(time), _f64(flux)
    n = len(time)
    e = _f64(err) if err is not None else None
    g = np.ascontiguousarray(group, dtype=np.int64) if group is not None else None
    d = _f64(durations)
    nd = len(d)
    pc = prep_cfg(cfg)
    nf = _sz()
    fn = np.zeros(nd, dtype=np.uintp)
    ft, ff, fs = np.zeros(nd * n), np.zeros(nd * n), np.zeros(nd * n)
    fw = np.zeros(nd)
    fc = np.zeros(nd, dtype=np.uintp)
    fm = np.zeros(nd * nd, dtype=np.uint8)
    st = lib().orc_prepare_families(
        _dp(time), _dp(flux), _dp(e) if e is not None else None, _i64p(g), n, _dp(d), nd,
        ctypes.byref(pc), ctypes.byref(nf), fn.ctypes.data_as(_P(_sz)), _dp(ft), _dp(ff),
        _dp(fs), _dp(fw), fc.ctypes.data_as(_P(_sz)), fm.ctypes.data_as(_P(ctypes.c_uint8)))
    if st:
        raise ValueError(f"prepare_families status {st}")
    out = []
    for k in range(nf.value):
        m = int(fn[k])
        out.append(dict(time=ft[k * n:k * n + m], flux=ff[k * n:k * n + m],
                        sigma=fs[k * n:k * n + m], window=float(fw[k]), clipped=int(fc[k]),
                        indices=np.nonzero(fm[k * nd:(k + 1) * nd])[0]))
    return out


def search_transits(time, flux, err, group, cfg, with_events=False):
    """search_transits (pipeline.rs:27-37, 70-221, no single events) on raw
    flux: (peaks in acceptance order, pass statistics)."""
    time, flux = _f64(time), _f64(flux)
    e = _f64(err) if err is not None else None
    g = np.ascontiguousarray(group, dtype=np.int64) if group is not None else None
    keep = []
    c = _pipeline_cfg(cfg, keep)
    pc = prep_cfg(cfg)
    cap = max(cfg.max_candidates, 1)
    peaks = (OrcPeak * cap)()
    passes = (OrcPassStats * max(cfg.passes, 1))()
    npk, npa = _sz(), _sz()
    eo = _events_out(cfg, len(cfg.durations_days()), keep)
    st = lib().orc_search_transits(_dp(time), _dp(flux), _dp(e) if e is not None else None,
                                   _i64p(g), len(time), ctypes.byref(pc), ctypes.byref(c),
                                   peaks, cap, ctypes.byref(npk), passes, ctypes.byref(npa),
                                   ctypes.byref(eo))
    if st:
        raise ValueError(f"search_transits status {st}")
    out = ([peaks[i].as_dict() for i in range(min(npk.value, cap))],
           [passes[i].as_dict() for i in range(npa.value)])
    return out + (_events_result(eo),) if with_events else out


def single_event_search(t, v, w, C, origin, durations, rho, cfg, capacity=131072):
    """events.rs:18-153 on one family's fields: (rows as dicts, stats)."""
    t, v, w = _f64(t), _f64(v), _f64(w)
    d, r = _f64(durations), _f64(rho)
    ing = _f64(cfg.ingress_fractions)
    rows = (Row * max(capacity, 1))()
    nr = _sz()
    st = EventStats()
    lib().orc_single_event_search(_dp(t), _dp(v), _dp(w), len(t), C, origin, _dp(d), _dp(r),
                                  len(d), _dp(ing), len(ing), cfg.shape_layers, cfg.threshold,
                                  cfg.phase_oversample, capacity, rows, ctypes.byref(nr),
                                  ctypes.byref(st))
    keys = ("period", "epoch", "duration", "score", "depth")
    return [{k: getattr(rows[i], k) for k in keys} for i in range(nr.value)], st.as_dict()


def select_peaks(fams, rows, family, time_all, threshold, limit, min_events,
                 masks=(), excluded=None):
    """peaks.rs:308-487 on a given table.  fams: dicts with t (relative), v,
    w (fixed-point values), C, origin, dur, rho; rows: dicts with period,
    epoch, duration, score, depth; family: the family of each row.
    Returns (peaks, excluded after the selection)."""
    ft = np.concatenate([_f64(f["t"]) for f in fams])
    fv = np.concatenate([_f64(f["v"]) for f in fams])
    fw = np.concatenate([_f64(f["w"]) for f in fams])
    foff = np.cumsum([0] + [len(f["t"]) for f in fams]).astype(np.uintp)
    fdur = np.concatenate([_f64(f["dur"]) for f in fams])
    frho = np.concatenate([_f64(f["rho"]) for f in fams])
    fdoff = np.cumsum([0] + [len(f["dur"]) for f in fams]).astype(np.uintp)
    fC = _f64([f["C"] for f in fams])
    forg = _f64([f["origin"] for f in fams])
    tab = (Row * max(len(rows), 1))()
    for i, r in enumerate(rows):
        tab[i] = Row(r["period"], r["epoch"], r["duration"], r["score"], r["depth"])
    fam_idx = np.ascontiguousarray(family, dtype=np.uintp)
    time_all = _f64(time_all)
    n_all = len(time_all)
    ms = np.ascontiguousarray(np.array(masks, dtype=np.uint8).reshape(-1, n_all))
    ex = (np.zeros(n_all, np.uint8) if excluded is None
          else np.ascontiguousarray(excluded, dtype=np.uint8).copy())
    out = (OrcPeak * max(limit, 1))()
    sz = _P(ctypes.c_size_t)
    u8 = _P(ctypes.c_uint8)
    n = lib().orc_select_peaks(
        len(fams), _dp(ft), _dp(fv), _dp(fw), foff.ctypes.data_as(sz), _dp(fC), _dp(forg),
        _dp(fdur), _dp(frho), fdoff.ctypes.data_as(sz), tab, fam_idx.ctypes.data_as(sz),
        len(rows), _dp(time_all), n_all, threshold, limit, min_events, ms.ctypes.data_as(u8),
        len(ms), ex.ctypes.data_as(u8), out, limit)
    return [out[i].as_dict() for i in range(n)], ex.astype(bool)


def select_events(rows, time, threshold, limit, masks=()):
    """peaks.rs:490-550: rows are dicts with epoch, duration, score, depth."""
    tab = (Row * max(len(rows), 1))()
    for i, r in enumerate(rows):
        tab[i] = Row(float("inf"), r["epoch"], r["duration"], r["score"], r["depth"])
    time = _f64(time)
    ms = [np.ascontiguousarray(m, dtype=np.uint8) for m in masks]
    mp = (_P(ctypes.c_uint8) * max(len(ms), 1))(*[m.ctypes.data_as(_P(ctypes.c_uint8)) for m in ms])
    out = (SingleEvent * max(limit, 1))()
    n = lib().orc_select_events(tab, len(rows), _dp(time), len(time), threshold, limit, mp,
                                len(ms), out)
    return [out[i].as_dict() for i in range(n)]

"""MI355X-native FBLS search hot path.

Python face of ``libfbls_hip.so`` (C ABI in ``include/fbls.h``).  The names
mirror the reference's search interface (isakruas/fbls):

* ``SearchConfig``          src/config.rs:48-131
* ``durations_days``        src/config.rs:177-182
* ``period_grid``           src/search/grid.rs:5-27
* ``family_state``          src/prep/curve.rs:541-569 (prepared curves)
* ``Gpu``                   src/search/scan.rs:39-99
* ``DeviceReady.upload``    src/search/scan.rs:101-125
* ``scan_bank``             src/search/scan.rs:144-357
* ``scan_periods``          src/search/scan.rs:363-408
* ``signal_search``         src/search/scan.rs:413-473
* ``search_prepared``       src/pipeline.rs:41-221 (passes, selection, evidence)

Every scan runs on the GPU through the HIP kernels; there is no CPU
fallback, and a missing library or device raises.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import time as _time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfbls_hip.so")

_d = ctypes.c_double
_sz = ctypes.c_size_t
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int32
_P = ctypes.POINTER
_vp = ctypes.c_void_p

FBLS_OK = 0
ERRORS = {1: "FBLS_E_FOLD_BUDGET", 2: "FBLS_E_NONPOSITIVE_PRECISION",
          3: "FBLS_E_HIP", 4: "FBLS_E_ARG", 5: "FBLS_E_OOM", 6: "FBLS_E_DATA"}

# Every symbol include/fbls.h declares.
EXPORTS = (
    "fbls_ctx_create", "fbls_ctx_destroy", "fbls_mem_info",
    "fbls_ctx_set_launch", "fbls_family_upload", "fbls_family_free",
    "fbls_family_fields", "fbls_scan_bank", "fbls_scan_periods",
    "fbls_signal_search", "fbls_period_grid_count", "fbls_period_grid",
    "fbls_durations_days", "fbls_family_state", "fbls_quantize",
    "fbls_synth_star", "fbls_last_error", "fbls_plan_bands",
    "fbls_bank_batches", "fbls_ctx_launch_totals", "fbls_compare",
    "fbls_evidence_parameters", "fbls_family_upload_packed",
    "fbls_signal_search_batch", "fbls_search_prepared", "fbls_cnn_create",
    "fbls_cnn_destroy", "fbls_cnn_param_count", "fbls_cnn_get_params",
    "fbls_cnn_set_params", "fbls_cnn_forward", "fbls_cnn_loss_grad",
    "fbls_cnn_train_step", "fbls_coarse_features", "fbls_ctx_set_graphs",
    "fbls_autotune", "fbls_biweight_flatten", "fbls_prepare_families",
    "fbls_search_transits", "fbls_single_event_search", "fbls_read_inputs",
    "fbls_format_f64", "fbls_json_reformat", "fbls_cli", "fbls_coarse_inputs",
    "fbls_cnn_set_fused", "fbls_signal_search_bins", "fbls_propose",
    "fbls_signal_search_bins_ragged", "fbls_scan_bank_noalloc", "fbls_family_upload_list",
    "fbls_family_state_many", "fbls_family_prepare_upload_many",
)


class FblsError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code


class Bank(ctypes.Structure):
    _fields_ = [("durations", _P(_d)), ("rho", _P(_d)), ("nd", _sz),
                ("ingress", _P(_d)), ("ns", _sz), ("layers", _u32),
                ("threshold", _d), ("shape_gate", _d),
                ("phase_oversample", _u32), ("period_batch", _u32),
                ("max_fold_mib", _d), ("prescan_factor", _u32),
                ("prescan_fraction", _d)]


class ScanStats(ctypes.Structure):
    _fields_ = [("periods", _u64), ("prescan_survivors", _u64),
                ("phase_bins", _u64), ("batches", _u64),
                ("templates_scanned", _u64), ("exact_tile_retries", _u64),
                ("seconds", _d), ("seconds_fold", _d),
                ("seconds_cumulative", _d), ("seconds_scan", _d),
                ("seconds_device", _d)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LaunchTotals(ctypes.Structure):
    _fields_ = [("launches", _u64), ("periods", _u64), ("fold_pairs", _u64),
                ("seconds", _d)]


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


class PipelineCfgC(ctypes.Structure):
    pass


class EvidenceDebug(ctypes.Structure):
    _fields_ = [("screen", _P(_d)), ("screen_cap", _sz), ("parts", _P(_d)),
                ("parts_cap", _sz), ("floors", _d * 4), ("tiles", _sz),
                ("nb", _sz)]


PEAK_FIELDS_D1 = ("period_days", "epoch_bjd", "duration_hours", "snr",
                  "snr_folded", "snr_without_strongest_event", "depth",
                  "depth_ppm")
PEAK_FIELDS_U = ("events", "family", "pass_", "row")
PEAK_FIELDS_D2 = ("planet_log_odds", "log_bf_planet_noise",
                  "log_bf_binary_noise", "binary_secondary_ratio",
                  "binary_secondary_phase", "binary_odd_even_ratio",
                  "binary_period_days", "ingress_fraction",
                  "evidence_epoch_bjd", "evidence_duration_hours")


class Peak(ctypes.Structure):
    _fields_ = ([(k, _d) for k in PEAK_FIELDS_D1] + [(k, _u64) for k in PEAK_FIELDS_U]
                + [(k, _d) for k in PEAK_FIELDS_D2]
                + [("model_preferred", _u32), ("binary_diagnostics_complete", _u32)])

    def as_dict(self):
        d = {k.rstrip("_"): getattr(self, k) for k, _ in self._fields_}
        d["model_preferred"] = bool(d["model_preferred"])
        d["binary_diagnostics_complete"] = bool(d["binary_diagnostics_complete"])
        return d


class PassStats(ctypes.Structure):
    _fields_ = [("pass_", _u64), ("rows", _u64), ("periods", _u64),
                ("accepted", _u64), ("seconds_scan", _d), ("seconds_select", _d)]

    def as_dict(self):
        return {k.rstrip("_"): getattr(self, k) for k, _ in self._fields_}


class CnnCfg(ctypes.Structure):
    _fields_ = [("in_channels", _u32), ("width", _u32), ("layers", _u32),
                ("kernel", _u32), ("pos_weight", ctypes.c_float)]


class TuneEntry(ctypes.Structure):
    _fields_ = [("lds_bytes", _u32), ("block", _u32), ("seconds", _d)]


class PrepCfg(ctypes.Structure):
    _fields_ = [(k, _d) for k in ("bin_minutes", "window_min", "window_factor",
                                  "window_max", "clip_sigma", "edge_trim")]


class Row(ctypes.Structure):
    _fields_ = [("period", _d), ("epoch", _d), ("duration", _d),
                ("score", _d), ("depth", _d)]


class SearchCfgC(ctypes.Structure):
    _fields_ = [("period_min", _d), ("period_max", _d),
                ("period_oversample", _d)]


PipelineCfgC._fields_ = [
    ("durations", _P(_d)), ("nd", _sz), ("correlated", ctypes.c_int),
    ("bank", Bank), ("search", SearchCfgC), ("max_candidates", _sz),
    ("passes", _sz), ("min_events", _sz), ("evidence_candidates", _sz),
    ("model_comparison", ctypes.c_int), ("planet_log_odds_min", _d),
    ("evidence", EvidenceCfg), ("single_events", ctypes.c_int)]


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


class FamilyPass(ctypes.Structure):
    _fields_ = ([(k, _u64) for k in ("pass_", "family", "observations", "clipped",
                                     "segments", "periods", "nd")]
                + [("window_days", _d), ("error_inflation", _d),
                   ("durations_hours", _d * 64), ("rho", _d * 64),
                   ("band_first", _u64), ("nbands", _u64)])


class SearchReport(ctypes.Structure):
    _fields_ = [("events", _P(SingleEvent)), ("events_cap", _sz), ("nevents", _sz),
                ("event_stats", _P(EventStats)), ("event_stats_cap", _sz),
                ("nevent_stats", _sz),
                ("period_table", _P(Row)), ("period_table_cap", _sz),
                ("period_table_rows", _sz), ("period_table_family", _P(_u64)),
                ("families", _P(FamilyPass)), ("families_cap", _sz), ("nfamilies", _sz),
                ("bands", _P(ScanStats)), ("bands_cap", _sz), ("nbands", _sz),
                ("observations", _u64), ("periods_tested", _u64), ("nfam", _u64),
                ("baseline_days", _d), ("seconds_preparation", _d), ("seconds_search", _d),
                ("seconds_scan", _d), ("seconds_select", _d), ("seconds_evidence", _d),
                ("seconds_single_events", _d)]


class BandC(ctypes.Structure):
    _fields_ = [("lo", _d), ("first", _sz), ("count", _sz), ("nd", _sz),
                ("nb", _u32), ("nb_coarse", _u32)]


class SynthCfg(ctypes.Structure):
    _fields_ = [("seed", _u64), ("sectors", _u32), ("sector_days", _d),
                ("gap_days", _d), ("cadence_days", _d), ("t0", _d),
                ("sigma", _d), ("red_fraction", _d), ("ar_phi", _d),
                ("period", _d), ("epoch", _d), ("depth", _d),
                ("duration", _d), ("secondary_depth", _d),
                ("baseline_gap_days", _d)]


_lib = None


def lib():
    """Load libfbls_hip.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIB_PATH)
        L.fbls_last_error.restype = ctypes.c_char_p
        L.fbls_period_grid_count.restype = _sz
        L.fbls_period_grid_count.argtypes = [_d, _d, _d, _d, _d]
        L.fbls_period_grid.argtypes = [_d, _d, _d, _d, _d, _P(_d), _sz]
        L.fbls_durations_days.restype = _sz
        L.fbls_durations_days.argtypes = [_P(_d), _sz, _P(_d)]
        L.fbls_quantize.argtypes = [_P(_d), _sz, _P(_i32), _P(ctypes.c_int)]
        L.fbls_ctx_create.argtypes = [ctypes.c_int, _P(_vp)]
        L.fbls_ctx_destroy.argtypes = [_vp]
        L.fbls_ctx_destroy.restype = None
        L.fbls_family_free.argtypes = [_vp]
        L.fbls_family_free.restype = None
        L.fbls_mem_info.argtypes = [_vp, _P(_sz), _P(_sz)]
        L.fbls_ctx_set_launch.argtypes = [_vp, _u32, _u32, _u32]
        L.fbls_ctx_launch_totals.argtypes = [_vp, _P(LaunchTotals), ctypes.c_int]
        L.fbls_ctx_set_graphs.argtypes = [_vp, ctypes.c_int, _P(_u64)]
        L.fbls_biweight_flatten.argtypes = [_vp, _P(_d), _P(_d), _sz, _d, _d, _d, _u32, _P(_d)]
        L.fbls_prepare_families.argtypes = [
            _vp, _P(_d), _P(_d), _P(_d), _P(ctypes.c_int64), _sz, _P(_d), _sz, _P(PrepCfg),
            _P(_sz), _P(_sz), _P(_d), _P(_d), _P(_d), _P(_d), _P(_sz), _P(ctypes.c_uint8)]
        L.fbls_autotune.argtypes = [_vp, _vp, _P(_d), _P(_d), _sz, _P(Bank), _P(SearchCfgC),
                                    _u32, _P(TuneEntry), _sz, _P(_sz), _P(TuneEntry)]
        L.fbls_compare.argtypes = [_vp, _vp, _d, _P(_d), _P(_d), _sz, _d,
                                   _P(EvidenceCfg), _P(EvidenceRow), _P(_sz),
                                   _P(EvidenceDebug)]
        L.fbls_evidence_parameters.argtypes = [_P(EvidenceCfg), _P(_d)]
        L.fbls_scan_bank_noalloc.argtypes = [_vp, _vp, _P(_d), _sz, _P(Bank), _u32, _P(_d),
                                             _P(_d), _P(_d), _P(_u32), _P(ScanStats)]
        _fp = _P(ctypes.c_float)
        L.fbls_cnn_create.argtypes = [_vp, _P(CnnCfg), _u64, _P(_vp)]
        L.fbls_cnn_destroy.argtypes = [_vp]
        L.fbls_cnn_destroy.restype = None
        L.fbls_cnn_param_count.argtypes = [_vp]
        L.fbls_cnn_param_count.restype = _sz
        L.fbls_cnn_get_params.argtypes = [_vp, _fp]
        L.fbls_cnn_set_params.argtypes = [_vp, _fp]
        L.fbls_cnn_forward.argtypes = [_vp, _fp, _u32, _u32, _fp]
        L.fbls_cnn_set_fused.argtypes = [_vp, ctypes.c_int]
        L.fbls_cnn_loss_grad.argtypes = [_vp, _fp, _fp, _u32, _u32, _fp, _fp]
        L.fbls_cnn_train_step.argtypes = [_vp, _fp, _fp, _u32, _u32, ctypes.c_float,
                                          ctypes.c_float, _fp]
        L.fbls_search_prepared.argtypes = [
            _vp, _P(_d), _P(_d), _P(_d), _sz, _P(PipelineCfgC), _P(Peak), _sz,
            _P(_sz), _P(PassStats), _sz, _P(_sz), _P(SearchReport)]
        L.fbls_search_transits.argtypes = [
            _vp, _P(_d), _P(_d), _P(_d), _P(ctypes.c_int64), _sz, _P(PrepCfg),
            _P(PipelineCfgC), _P(Peak), _sz, _P(_sz), _P(PassStats), _sz, _P(_sz),
            _P(SearchReport)]
        L.fbls_read_inputs.argtypes = [_P(ctypes.c_char_p), _sz, _P(_d), _P(_d), _P(_d),
                                       _P(ctypes.c_int64), _sz, _P(_sz), _P(ctypes.c_int)]
        L.fbls_format_f64.argtypes = [_d, ctypes.c_char_p, _sz]
        L.fbls_format_f64.restype = _sz
        L.fbls_json_reformat.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _sz, _P(_sz)]
        L.fbls_cli.argtypes = [ctypes.c_int, _P(ctypes.c_char_p)]
        L.fbls_single_event_search.argtypes = [
            _vp, _vp, _P(_d), _P(_d), _sz, _P(_d), _sz, _u32, _d, _u32, _sz, _P(Row),
            _P(_sz), _P(EventStats)]
        L.fbls_family_upload.argtypes = [_vp, _P(_d), _P(_d), _P(_d), _sz, _d,
                                         _d, _P(_vp)]
        L.fbls_family_fields.argtypes = [_vp, _P(_d), _P(_d),
                                         _P(ctypes.c_int), _P(ctypes.c_int)]
        L.fbls_scan_bank.argtypes = [_vp, _vp, _P(_d), _sz, _P(Bank), _u32,
                                     _P(_d), _P(_d), _P(_d), _P(_u32),
                                     _P(ScanStats)]
        L.fbls_scan_periods.argtypes = [_vp, _vp, _P(_d), _sz, _P(Bank),
                                        _P(Row), _P(_sz), _P(ScanStats)]
        L.fbls_signal_search.argtypes = [
            _vp, _vp, _P(_d), _P(_d), _sz, _P(Bank), _P(SearchCfgC), _P(Row),
            _sz, _P(_sz), _P(ScanStats), _sz, _P(_sz), _P(_d), _sz, _P(_sz)]
        L.fbls_family_upload_packed.argtypes = [
            _vp, _sz, _P(_sz), _P(_d), _P(_d), _P(_d), _P(_d), _P(_d), _P(_vp)]
        L.fbls_family_upload_list.argtypes = [
            _vp, _sz, _P(_P(_d)), _P(_P(_d)), _P(_P(_d)), _P(_sz), _P(_d), _P(_d), _P(_vp)]
        L.fbls_signal_search_batch.argtypes = [
            _vp, _P(_vp), _sz, _P(_d), _P(_d), _sz, _P(Bank), _P(SearchCfgC),
            _P(Row), _sz, _P(_sz), _P(ScanStats), _P(ctypes.c_uint8)]
        L.fbls_coarse_features.argtypes = [_vp, _P(_vp), _sz, _P(_d), _sz, _P(_d), _sz,
                                           _u32, _u32, _P(_d)]
        L.fbls_propose.argtypes = [_vp, _vp, _P(_vp), _sz, _P(_d), _sz, _P(_d), _sz, _u32,
                                   _u32, _P(ctypes.c_float), _P(ctypes.c_float)]
        L.fbls_signal_search_bins.argtypes = [
            _vp, _P(_vp), _sz, _P(_d), _P(_d), _sz, _P(Bank), _P(SearchCfgC),
            _P(Row), _sz, _P(_sz), _P(ScanStats), _P(ctypes.c_uint8), _sz, _P(_d)]
        L.fbls_signal_search_bins_ragged.argtypes = [
            _vp, _P(_vp), _sz, _P(_d), _P(_d), _sz, _P(Bank), _P(SearchCfgC),
            _P(Row), _sz, _P(_sz), _P(ScanStats), _P(ctypes.c_uint8), _P(_sz), _P(_d)]
        L.fbls_coarse_inputs.argtypes = [_vp, _P(_vp), _sz, _P(_d), _sz, _P(_d), _sz,
                                           _u32, _u32, _P(ctypes.c_float)]
        L.fbls_family_state.argtypes = [
            _P(_d), _P(_d), _P(_d), _sz, _P(_d), _sz, ctypes.c_int, _P(_d),
            _P(_d), _P(_d), _P(_sz), _P(_d), _P(_d), _P(_d), _P(_d), _P(_sz),
            _P(_d)]
        L.fbls_family_state_many.argtypes = [
            _sz, _P(_P(_d)), _P(_P(_d)), _P(_P(_d)), _P(_sz), _d, _P(_d), _sz, ctypes.c_int,
            _P(_P(_d)), _P(_P(_d)), _P(_P(_d)), _P(_sz), _P(_d), _P(_d), _P(_d), _P(_d),
            _P(_sz), _P(_d), _P(ctypes.c_int)]
        L.fbls_family_prepare_upload_many.argtypes = [
            _vp, _sz, _P(_P(_d)), _P(_P(_d)), _P(_P(_d)), _P(_sz), _d, _P(_d), _sz,
            ctypes.c_int, _P(_sz), _P(_d), _P(_d), _P(_d), _P(_d), _P(_sz), _P(_d), _P(_d),
            _P(ctypes.c_int), _P(_vp)]
        L.fbls_synth_star.argtypes = [_P(SynthCfg), _P(_d), _P(_d), _P(_d),
                                      _sz, _P(_sz)]
        L.fbls_plan_bands.argtypes = [_P(_d), _sz, _P(_d), _sz, _u32, _u32,
                                      _P(BandC), _sz, _P(_sz)]
        L.fbls_bank_batches.restype = _u64
        L.fbls_bank_batches.argtypes = [_sz, _u32, _u32, _d]
        _lib = L
    return _lib


def _check(status):
    if status != FBLS_OK:
        raise FblsError(status, lib().fbls_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _dp(a):
    return a.ctypes.data_as(_P(_d)) if a is not None else None


# ------------------------------------------------------------------ config

def _default_hours():
    # config.rs:94: 0.5 * 1.25^i rounded half-to-even at 1e-6 h
    return [round(0.5 * 1.25 ** i * 1e6) / 1e6 for i in range(13)]


@dataclasses.dataclass
class SearchConfig:
    """Search constants of the hot path (config.rs:48-131)."""
    period_min: float = 0.3
    period_max: float = 10.0
    durations_hours: list = dataclasses.field(default_factory=_default_hours)
    threshold: float = 7.1
    correlated_noise: bool = True
    ingress_fractions: list = dataclasses.field(
        default_factory=lambda: [0.0, 0.2, 0.45])
    shape_layers: int = 4
    phase_oversample: int = 4
    shape_gate_fraction: float = 0.6
    prescan_factor: int = 2
    prescan_fraction: float = 0.7
    period_oversample: float = 3.0
    period_batch: int = 4096
    max_fold_mib: float = 1024.0
    # raw-mode preparation (config.rs:11-16,54-59)
    bin_minutes: float = 10.0
    window_days: float = 0.4
    window_factor: float = 3.0
    window_max_days: float = 3.5
    edge_trim_windows: float = 0.5
    clip_sigma: float = 4.0
    # passes and selection (config.rs:25-34,68-77)
    max_candidates: int = 20
    passes: int = 3
    min_events: int = 2
    model_comparison: bool = True
    evidence_candidates: int = 32
    planet_log_odds_min: float = 0.0
    single_events: bool = True
    # model comparison (config.rs:33-42,76-85)
    evidence_rtol: float = 1e-12
    planet_prior: float = 0.01
    binary_prior: float = 0.01
    planet_depth_scale: float = 0.01
    binary_depth_scale: float = 0.1
    diluted_binary_depth_scale: float = 0.01
    stellar_radius_solar: float = None
    planet_radius_scale_rj: float = 1.0
    target_flux_fraction: float = 1.0

    def validate(self):
        """config.rs:91-131 (the fields of the search path)."""
        if not (0.0 < self.period_min < self.period_max):
            raise ValueError("require 0 < period_min < period_max")
        if not self.durations_hours or any(
                not (np.isfinite(d) and d > 0) for d in self.durations_hours):
            raise ValueError("durations_hours must contain positive finite values")
        if min(self.durations_hours) / 24.0 >= self.period_min / 4.0:
            raise ValueError("minimum duration must be below one quarter of period_min")
        if not self.ingress_fractions or any(
                not (0.0 <= x < 0.5) for x in self.ingress_fractions):
            raise ValueError("ingress fractions must be in [0, 0.5)")
        if (self.shape_layers < 1 or self.phase_oversample < 1
                or self.period_batch < 1 or self.prescan_factor < 1):
            raise ValueError("invalid integer search parameter")
        if not (0.0 < self.prescan_fraction <= 1.0):
            raise ValueError("prescan_fraction must be in (0, 1]")

    def durations_days(self):
        return durations_days(self.durations_hours)

    @property
    def shape_gate(self):
        # scan.rs:433-437
        if self.ingress_fractions[0] == 0.0:
            return self.shape_gate_fraction * self.threshold
        return 0.0


# ------------------------------------------------------------ host helpers

def durations_days(hours):
    h = _f64(hours)
    out = np.empty_like(h)
    m = lib().fbls_durations_days(_dp(h), len(h), _dp(out))
    return out[:m]


def period_grid(minimum, maximum, baseline, minimum_duration, oversample):
    L = lib()
    n = L.fbls_period_grid_count(minimum, maximum, baseline, minimum_duration,
                                 oversample)
    out = np.empty(n)
    _check(L.fbls_period_grid(minimum, maximum, baseline, minimum_duration,
                              oversample, _dp(out), n))
    return out


def plan_bands(grid, sorted_durations, phase_oversample=4, prescan_factor=2):
    """Octave bands of one family's grid (scan.rs:440-471, 377-381)."""
    grid, sd = _f64(grid), _f64(sorted_durations)
    n = _sz()
    L = lib()
    _check(L.fbls_plan_bands(_dp(grid), len(grid), _dp(sd), len(sd),
                             phase_oversample, prescan_factor, None, 0,
                             ctypes.byref(n)))
    out = (BandC * max(1, n.value))()
    _check(L.fbls_plan_bands(_dp(grid), len(grid), _dp(sd), len(sd),
                             phase_oversample, prescan_factor, out, n.value,
                             ctypes.byref(n)))
    return [dict(lo=b.lo, first=b.first, count=b.count, nd=b.nd, nb=b.nb,
                 nb_coarse=b.nb_coarse) for b in out[:n.value]]


def bank_batches(np_, nb, period_batch=4096, max_fold_mib=1024.0):
    """scan.rs:177-183: batches of one scan_bank call (0: over budget)."""
    return lib().fbls_bank_batches(np_, nb, period_batch, max_fold_mib)


def quantize(x):
    """(words, e, block) of the device's fixed-point fields (fbls_quantize)."""
    x = _f64(x)
    q = np.empty(len(x), dtype=np.int32)
    block = ctypes.c_int(0)
    e = lib().fbls_quantize(_dp(x), len(x), q.ctypes.data_as(_P(_i32)), ctypes.byref(block))
    return q, e, bool(block.value)


@dataclasses.dataclass
class FamilyState:
    t: np.ndarray
    v: np.ndarray
    w: np.ndarray
    baseline_norm: float
    origin: float
    durations: np.ndarray
    rho: np.ndarray
    inflation: float
    segments: int
    residual_energy: float = 0.0


def family_state(time, flux, sigma, durations, correlated=True):
    """curve.rs:541-569 for a prepared curve (flux relative to zero)."""
    time, flux, sigma = _f64(time), _f64(flux), _f64(sigma)
    durations = _f64(durations)
    n = len(time)
    t, v, w = np.empty(n), np.empty(n), np.empty(n)
    rho = np.empty(len(durations))
    m, c, org, infl, seg, en = _sz(), _d(), _d(), _d(), _sz(), _d()
    _check(lib().fbls_family_state(
        _dp(time), _dp(flux), _dp(sigma), n, _dp(durations), len(durations),
        int(bool(correlated)), _dp(t), _dp(v), _dp(w), ctypes.byref(m),
        ctypes.byref(c), ctypes.byref(org), _dp(rho), ctypes.byref(infl),
        ctypes.byref(seg), ctypes.byref(en)))
    k = m.value
    if k < n:  # points were dropped: keep only the prepared ones
        t, v, w = t[:k].copy(), v[:k].copy(), w[:k].copy()
    return FamilyState(t, v, w, c.value, org.value, durations, rho, infl.value, seg.value,
                       en.value)


def family_states(curves, durations, correlated=True, flux_offset=0.0):
    """family_state of many curves ((time, flux, sigma) each) in one call,
    prepared on the library's own threads (fbls_family_state_many); with
    flux_offset=-1.0 the fluxes are normalised to 1 and the result equals
    family_state(time, flux - 1.0, sigma, ...).  The states' arrays are
    views into three shared buffers."""
    k = len(curves)
    if k == 0:
        return []
    durations = _f64(durations)
    nd = len(durations)
    ins = [(_f64(a), _f64(b), _f64(c)) for a, b, c in curves]  # kept alive for the call
    n = np.array([len(a[0]) for a in ins], dtype=np.uintp)
    at = np.concatenate(([0], np.cumsum(n)))
    T, V, W = np.empty(int(at[-1])), np.empty(int(at[-1])), np.empty(int(at[-1]))
    PD = _P(_d)

    def ptrs(arrs):
        return (PD * k)(*[_dp(x) for x in arrs])

    def parts(buf):
        return (PD * k)(*[ctypes.cast(buf.ctypes.data + 8 * int(at[i]), PD) for i in range(k)])

    m, c, org, infl, en = np.empty(k, np.uintp), np.empty(k), np.empty(k), np.empty(k), np.empty(k)
    seg = np.empty(k, np.uintp)
    rho = np.empty((k, nd))
    _check(lib().fbls_family_state_many(
        k, ptrs([a[0] for a in ins]), ptrs([a[1] for a in ins]), ptrs([a[2] for a in ins]),
        n.ctypes.data_as(_P(_sz)), float(flux_offset), _dp(durations), nd,
        int(bool(correlated)), parts(T), parts(V), parts(W), m.ctypes.data_as(_P(_sz)),
        _dp(c), _dp(org), _dp(rho), _dp(infl), seg.ctypes.data_as(_P(_sz)), _dp(en), None))
    out = []
    for i in range(k):
        a, e = int(at[i]), int(at[i]) + int(m[i])
        out.append(FamilyState(T[a:e], V[a:e], W[a:e], float(c[i]), float(org[i]), durations,
                               rho[i], float(infl[i]), int(seg[i]), float(en[i])))
    return out


def synth_star(**kw):
    """Seeded synthetic light curve: returns (time, flux, flux_err)."""
    cfg = SynthCfg(seed=1, sectors=1, sector_days=27.4, gap_days=1.0,
                   cadence_days=1.0 / 720.0, t0=2458000.0, sigma=1e-3,
                   red_fraction=0.0, ar_phi=0.0, period=0.0, epoch=np.nan,
                   depth=0.0, duration=0.0, secondary_depth=0.0,
                   baseline_gap_days=0.0)
    for k, v in kw.items():
        setattr(cfg, k, v)
    n = _sz()
    L = lib()
    _check(L.fbls_synth_star(ctypes.byref(cfg), None, None, None, 0,
                             ctypes.byref(n)))
    t, f, e = np.empty(n.value), np.empty(n.value), np.empty(n.value)
    _check(L.fbls_synth_star(ctypes.byref(cfg), _dp(t), _dp(f), _dp(e),
                             n.value, ctypes.byref(n)))
    return t, f, e


# ----------------------------------------------------------------- device

class Gpu:
    """Device context: HIP stream, events and scratch (scan.rs:39-99)."""

    def __init__(self, device=0):
        self._h = _vp()
        _check(lib().fbls_ctx_create(device, ctypes.byref(self._h)))
        self.device = device

    def memory(self):
        fr, tot = _sz(), _sz()
        _check(lib().fbls_mem_info(self._h, ctypes.byref(fr), ctypes.byref(tot)))
        return fr.value, tot.value

    def set_launch(self, lds_bytes=0, block=0, fold_mode=0):
        """Launch-shape override (fbls_ctx_set_launch): LDS bytes per
        workgroup, workgroup size, fold mode (0 default, 1 counts always
        folded, 2 one atomic per observation).  Results never change."""
        _check(lib().fbls_ctx_set_launch(self._h, lds_bytes, block, fold_mode))

    def set_graphs(self, enable=None):
        """Evidence calls through a replayed hipGraph (True) or eager
        launches (False); returns the number of graph submissions so far."""
        n = _u64()
        _check(lib().fbls_ctx_set_graphs(self._h, -1 if enable is None else int(bool(enable)),
                                         ctypes.byref(n)))
        return n.value

    def launch_totals(self, reset=False):
        """Kernel launches on this context since the last reset: count,
        periods, fold pairs and HIP-event seconds."""
        out = LaunchTotals()
        _check(lib().fbls_ctx_launch_totals(self._h, ctypes.byref(out),
                                            1 if reset else 0))
        return {k: getattr(out, k) for k, _ in out._fields_}

    def close(self):
        if self._h:
            lib().fbls_ctx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceReady:
    """A family's fields resident in HBM (scan.rs:101-125)."""

    def __init__(self, gpu, t, v, w, baseline_norm, origin):
        t, v, w = _f64(t), _f64(v), _f64(w)
        self.gpu = gpu
        self.n = len(t)
        self.baseline_norm = baseline_norm
        self.origin = origin
        self.t_last = float(t[-1])
        self._h = _vp()
        _check(lib().fbls_family_upload(gpu._h, _dp(t), _dp(v), _dp(w),
                                        len(t), baseline_norm, origin,
                                        ctypes.byref(self._h)))

    @classmethod
    def upload(cls, gpu, state: FamilyState):
        return cls(gpu, state.t, state.v, state.w, state.baseline_norm,
                   state.origin)

    @classmethod
    def upload_packed(cls, gpu, states):
        """Many families in one device allocation (ragged layout), read from
        each state's own arrays (fbls_family_upload_list)."""
        k = len(states)
        arrs = [(_f64(s.t), _f64(s.v), _f64(s.w)) for s in states]  # kept alive for the call
        PD = _P(_d)
        tp = (PD * k)(*[_dp(a[0]) for a in arrs])
        vp = (PD * k)(*[_dp(a[1]) for a in arrs])
        wp = (PD * k)(*[_dp(a[2]) for a in arrs])
        n = np.array([len(a[0]) for a in arrs], dtype=np.uintp)
        c = _f64([s.baseline_norm for s in states])
        o = _f64([s.origin for s in states])
        hs = (_vp * k)()
        _check(lib().fbls_family_upload_list(
            gpu._h, k, tp, vp, wp, n.ctypes.data_as(_P(_sz)), _dp(c), _dp(o), hs))
        out = []
        for i, st in enumerate(states):
            d = cls.__new__(cls)
            d.gpu = gpu
            d.n = len(st.t)
            d.baseline_norm = st.baseline_norm
            d.origin = st.origin
            d.t_last = float(st.t[-1])
            d._h = _vp(hs[i])
            out.append(d)
        return out

    @classmethod
    def prepare_upload(cls, gpu, curves, durations, correlated=True, flux_offset=0.0):
        """family_states of the curves ((time, flux, sigma) each) and their
        upload in one call (fbls_family_prepare_upload_many): the families
        on the device, and a dict of the preparation's per-star outputs
        (rho as a stars x durations array; n, baseline_norm, origin,
        inflation, segments, residual_energy)."""
        k = len(curves)
        durations = _f64(durations)
        nd = len(durations)
        if k == 0:
            return [], {"rho": np.empty((0, nd))}
        ins = [(_f64(a), _f64(b), _f64(c)) for a, b, c in curves]  # kept alive for the call
        PD = _P(_d)

        def ptrs(arrs):
            return (PD * k)(*[_dp(x) for x in arrs])

        n = np.array([len(a[0]) for a in ins], dtype=np.uintp)
        m, seg = np.empty(k, np.uintp), np.empty(k, np.uintp)
        c, org, infl, en, tl = np.empty(k), np.empty(k), np.empty(k), np.empty(k), np.empty(k)
        rho = np.empty((k, nd))
        hs = (_vp * k)()
        _check(lib().fbls_family_prepare_upload_many(
            gpu._h, k, ptrs([a[0] for a in ins]), ptrs([a[1] for a in ins]),
            ptrs([a[2] for a in ins]), n.ctypes.data_as(_P(_sz)), float(flux_offset),
            _dp(durations), nd, int(bool(correlated)), m.ctypes.data_as(_P(_sz)), _dp(c),
            _dp(org), _dp(rho), _dp(infl), seg.ctypes.data_as(_P(_sz)), _dp(en), _dp(tl), None,
            hs))
        out = []
        for i in range(k):
            d = cls.__new__(cls)
            d.gpu = gpu
            d.n = int(m[i])
            d.baseline_norm = float(c[i])
            d.origin = float(org[i])
            d._h = _vp(hs[i])
            d.t_last = float(tl[i])
            out.append(d)
        info = {"rho": rho, "n": m.astype(np.int64), "baseline_norm": c, "origin": org,
                "inflation": infl, "segments": seg.astype(np.int64), "residual_energy": en}
        return out, info

    def fields(self):
        """The fixed-point values the device scans (exact doubles)."""
        v, w = np.empty(self.n), np.empty(self.n)
        ev, ew = ctypes.c_int(), ctypes.c_int()
        _check(lib().fbls_family_fields(self._h, _dp(v), _dp(w),
                                        ctypes.byref(ev), ctypes.byref(ew)))
        return v, w, ev.value, ew.value

    def close(self):
        if self._h:
            lib().fbls_family_free(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- evidence

def evidence_cfg(cfg: SearchConfig):
    """fbls_evidence_cfg of a SearchConfig (the ingress array is kept alive
    on the returned structure)."""
    ing = _f64(cfg.ingress_fractions)
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


def evidence_parameters(cfg: SearchConfig):
    """parameters (evidence.rs:665-689): the 30 model constants."""
    c = evidence_cfg(cfg)
    out = np.empty(30)
    _check(lib().fbls_evidence_parameters(ctypes.byref(c), _dp(out)))
    return out


def compare_search(gpu, dev, period, durations, rho, residual_energy,
                   cfg: SearchConfig, debug=False):
    """compare_search (evidence.rs:693-1003) on the GPU: the joint-evidence
    and SNR winners at one trial period as row dicts (merged, flags 3, when
    identical).  With debug, also {screen, parts, floors, tiles, nb}."""
    d, r = _f64(durations), _f64(rho)
    c = evidence_cfg(cfg)
    rows = (EvidenceRow * 2)()
    nrows = _sz()
    dbg = None
    if debug:
        cap = 4 * (int(cfg.phase_oversample * period / max(d.min(), 1e-9)) + 8)
        screen, parts = np.zeros(cap * 4), np.zeros(cap * 30)
        dbg = EvidenceDebug(screen=_dp(screen), screen_cap=screen.size,
                            parts=_dp(parts), parts_cap=parts.size)
    _check(lib().fbls_compare(gpu._h, dev._h, float(period), _dp(d), _dp(r),
                              len(d), float(residual_energy), ctypes.byref(c),
                              rows, ctypes.byref(nrows),
                              ctypes.byref(dbg) if dbg is not None else None))
    out = [rows[i].as_dict() for i in range(nrows.value)]
    if not debug:
        return out
    k = dbg.tiles
    return out, {"screen": screen[:4 * k].reshape(k, 4),
                 "parts": parts[:30 * k].reshape(k, 2, 15),
                 "floors": np.array(dbg.floors[:]), "tiles": k, "nb": dbg.nb}


def make_bank(cfg: SearchConfig, durations, rho, max_fold_mib=None):
    durations, rho = _f64(durations), _f64(rho)
    ing = _f64(cfg.ingress_fractions)
    b = Bank(durations=_dp(durations), rho=_dp(rho), nd=len(durations),
             ingress=_dp(ing), ns=len(ing), layers=cfg.shape_layers,
             threshold=cfg.threshold, shape_gate=cfg.shape_gate,
             phase_oversample=cfg.phase_oversample,
             period_batch=cfg.period_batch,
             max_fold_mib=cfg.max_fold_mib if max_fold_mib is None else max_fold_mib,
             prescan_factor=cfg.prescan_factor,
             prescan_fraction=cfg.prescan_fraction)
    b._keep = (durations, rho, ing)  # keep the arrays alive with the struct
    return b


def scan_bank(gpu, dev, periods, bank, nb, winner=True):
    """Best template per period at nb bins.  Returns records (np x 9), score,
    depth, winner (np x 4 {phase, width, shape, duration index}) and stats."""
    periods = _f64(periods)
    n = len(periods)
    rec = np.empty((n, 9))
    score, depth = np.empty(n), np.empty(n)
    win = np.empty((n, 4), dtype=np.uint32) if winner else None
    st = ScanStats()
    _check(lib().fbls_scan_bank(
        gpu._h, dev._h, _dp(periods), n, ctypes.byref(bank), nb, _dp(rec),
        _dp(score), _dp(depth),
        win.ctypes.data_as(_P(_u32)) if win is not None else None,
        ctypes.byref(st)))
    return rec, score, depth, win, st.as_dict()


def scan_bank_noalloc(gpu, dev, periods, bank, nb, winner=True):
    """scan_bank through the entry point that allocates nothing (its staging
    is a fixed per-device array of the library: at most 65536 periods per
    call).  Same returns as scan_bank."""
    periods = _f64(periods)
    n = len(periods)
    rec = np.empty((n, 9))
    score, depth = np.empty(n), np.empty(n)
    win = np.empty((n, 4), dtype=np.uint32) if winner else None
    st = ScanStats()
    _check(lib().fbls_scan_bank_noalloc(
        gpu._h, dev._h, _dp(periods), n, ctypes.byref(bank), nb, _dp(rec),
        _dp(score), _dp(depth),
        win.ctypes.data_as(_P(_u32)) if win is not None else None,
        ctypes.byref(st)))
    return rec, score, depth, win, st.as_dict()


def _row_buffer(n):
    """An uninitialised (n, 5) float64 array laid out as n fbls_row structs,
    and its pointer."""
    buf = np.empty((max(n, 1), 5))
    return buf, buf.ctypes.data_as(_P(Row))


def scan_periods(gpu, dev, periods, bank):
    periods = _f64(periods)
    buf, rows = _row_buffer(len(periods))
    nrows = _sz()
    st = ScanStats()
    _check(lib().fbls_scan_periods(gpu._h, dev._h, _dp(periods), len(periods),
                                   ctypes.byref(bank), rows,
                                   ctypes.byref(nrows), ctypes.byref(st)))
    return buf[:nrows.value].copy(), st.as_dict()


def signal_search(gpu, dev, durations, rho, cfg: SearchConfig):
    """One family: returns (rows [k x 5], grid, band stats)."""
    durations, rho = _f64(durations), _f64(rho)
    maximum = min(cfg.period_max, dev.t_last)
    minimum = max(cfg.period_min, 4.0 * durations.min() * (1.0 + 1e-9))
    cap = 1
    if maximum > minimum:
        cap = lib().fbls_period_grid_count(minimum, maximum, dev.t_last,
                                           float(durations.min()),
                                           cfg.period_oversample)
    bank = make_bank(cfg, durations, rho)
    scfg = SearchCfgC(cfg.period_min, cfg.period_max, cfg.period_oversample)
    buf, rows = _row_buffer(cap)
    stats = (ScanStats * 64)()
    grid = np.empty(cap)
    nrows, nbands, gcount = _sz(), _sz(), _sz()
    _check(lib().fbls_signal_search(
        gpu._h, dev._h, _dp(durations), _dp(rho), len(durations),
        ctypes.byref(bank), ctypes.byref(scfg), rows, cap,
        ctypes.byref(nrows), stats, 64, ctypes.byref(nbands), _dp(grid), cap,
        ctypes.byref(gcount)))
    return (buf[:nrows.value].copy(), grid[:gcount.value],
            [s.as_dict() for s in stats[:nbands.value]])


def signal_search_batch(gpu, devs, durations, rhos, cfg: SearchConfig,
                        rows_cap=None, keep=None):
    """Many stars: returns ([rows k_s x 5 per star], [stats per star]).
    keep: optional list of per-star boolean arrays over each star's grid;
    only those periods are scanned (each exactly as in the full search)."""
    durations = _f64(durations)
    rho = np.ascontiguousarray(np.asarray(rhos, dtype=np.float64).reshape(
        len(devs), len(durations)))
    bank = make_bank(cfg, durations, rho[0])
    scfg = SearchCfgC(cfg.period_min, cfg.period_max, cfg.period_oversample)
    hs = (_vp * len(devs))(*[d._h.value for d in devs])
    offs = np.zeros(len(devs) + 1, dtype=np.uintp)
    stats = (ScanStats * len(devs))()
    if rows_cap is None:  # rows per star <= its grid count
        dmin = float(durations.min())
        rows_cap = 0
        for d in devs:
            hi = min(cfg.period_max, d.t_last)
            lo = max(cfg.period_min, 4.0 * dmin * (1.0 + 1e-9))
            if hi > lo:
                rows_cap += lib().fbls_period_grid_count(lo, hi, d.t_last, dmin,
                                                         cfg.period_oversample)
    buf, rows = _row_buffer(rows_cap)
    kp = None
    if keep is not None:
        kflat = np.ascontiguousarray(np.concatenate(
            [np.asarray(k, dtype=np.uint8) for k in keep]), dtype=np.uint8)
        kp = kflat.ctypes.data_as(_P(ctypes.c_uint8))
    _check(lib().fbls_signal_search_batch(
        gpu._h, hs, len(devs), _dp(durations), _dp(rho), len(durations),
        ctypes.byref(bank), ctypes.byref(scfg), rows, rows_cap,
        offs.ctypes.data_as(_P(_sz)), stats, kp))
    return ([buf[int(offs[i]):int(offs[i + 1])] for i in range(len(devs))],
            [s.as_dict() for s in stats])


def signal_search_bins(gpu, devs, durations, rhos, cfg: SearchConfig, bin_keep, edges,
                       rows_cap=None):
    """Many stars, each restricted to the grid periods of its kept bins
    (fbls_signal_search_bins; bin k is edges[k] <= P < edges[k + 1]).
    Returns ([rows k_s x 5 per star], [stats per star])."""
    durations = _f64(durations)
    rho = np.ascontiguousarray(np.asarray(rhos, dtype=np.float64).reshape(
        len(devs), len(durations)))
    bk = np.ascontiguousarray(bin_keep, dtype=np.uint8).reshape(len(devs), -1)
    edges = _f64(edges)
    assert len(edges) == bk.shape[1] + 1
    bank = make_bank(cfg, durations, rho[0])
    scfg = SearchCfgC(cfg.period_min, cfg.period_max, cfg.period_oversample)
    hs = (_vp * len(devs))(*[d._h.value for d in devs])
    offs = np.zeros(len(devs) + 1, dtype=np.uintp)
    stats = (ScanStats * len(devs))()
    if rows_cap is None:
        dmin = float(durations.min())
        rows_cap = 0
        for d in devs:
            hi = min(cfg.period_max, d.t_last)
            lo = max(cfg.period_min, 4.0 * dmin * (1.0 + 1e-9))
            if hi > lo:
                rows_cap += lib().fbls_period_grid_count(lo, hi, d.t_last, dmin,
                                                         cfg.period_oversample)
    buf, rows = _row_buffer(rows_cap)
    _check(lib().fbls_signal_search_bins(
        gpu._h, hs, len(devs), _dp(durations), _dp(rho), len(durations),
        ctypes.byref(bank), ctypes.byref(scfg), rows, rows_cap,
        offs.ctypes.data_as(_P(_sz)), stats, bk.ctypes.data_as(_P(ctypes.c_uint8)),
        bk.shape[1], _dp(edges)))
    return ([buf[int(offs[i]):int(offs[i + 1])] for i in range(len(devs))],
            [s.as_dict() for s in stats])


def signal_search_bins_ragged(gpu, devs, durations, rhos, cfg: SearchConfig, bin_keeps,
                              edges, rows_cap=None):
    """fbls_signal_search_bins with bins of its own for every star:
    bin_keeps[s] (booleans) and edges[s] (len(bin_keeps[s]) + 1 ascending
    values) restrict star s.  One call, the same merged launches.  Returns
    ([rows k_s x 5 per star], [stats per star])."""
    durations = _f64(durations)
    rho = np.ascontiguousarray(np.asarray(rhos, dtype=np.float64).reshape(
        len(devs), len(durations)))
    offs = np.zeros(len(devs) + 1, dtype=np.uintp)
    offs[1:] = np.cumsum([len(k) for k in bin_keeps])
    bk = np.ascontiguousarray(np.concatenate([np.asarray(k, dtype=np.uint8) for k in bin_keeps]))
    ed = _f64(np.concatenate([np.asarray(e, dtype=np.float64) for e in edges]))
    for k, e in zip(bin_keeps, edges):
        assert len(e) == len(k) + 1
    bank = make_bank(cfg, durations, rho[0])
    scfg = SearchCfgC(cfg.period_min, cfg.period_max, cfg.period_oversample)
    hs = (_vp * len(devs))(*[d._h.value for d in devs])
    roffs = np.zeros(len(devs) + 1, dtype=np.uintp)
    stats = (ScanStats * len(devs))()
    if rows_cap is None:
        dmin = float(durations.min())
        rows_cap = 0
        for d in devs:
            hi = min(cfg.period_max, d.t_last)
            lo = max(cfg.period_min, 4.0 * dmin * (1.0 + 1e-9))
            if hi > lo:
                rows_cap += lib().fbls_period_grid_count(lo, hi, d.t_last, dmin,
                                                         cfg.period_oversample)
    buf, rows = _row_buffer(rows_cap)
    _check(lib().fbls_signal_search_bins_ragged(
        gpu._h, hs, len(devs), _dp(durations), _dp(rho), len(durations),
        ctypes.byref(bank), ctypes.byref(scfg), rows, rows_cap,
        roffs.ctypes.data_as(_P(_sz)), stats, bk.ctypes.data_as(_P(ctypes.c_uint8)),
        offs.ctypes.data_as(_P(_sz)), _dp(ed)))
    return ([buf[int(roffs[i]):int(roffs[i + 1])] for i in range(len(devs))],
            [s.as_dict() for s in stats])


def pipeline_cfg(cfg: SearchConfig):
    """fbls_pipeline_cfg of a SearchConfig (arrays kept alive on it)."""
    d = cfg.durations_days()
    bank = make_bank(cfg, d, np.ones(len(d)))
    ev = evidence_cfg(cfg)
    c = PipelineCfgC(
        durations=_dp(d), nd=len(d), correlated=int(bool(cfg.correlated_noise)),
        bank=bank, search=SearchCfgC(cfg.period_min, cfg.period_max,
                                     cfg.period_oversample),
        max_candidates=cfg.max_candidates, passes=cfg.passes,
        min_events=cfg.min_events, evidence_candidates=cfg.evidence_candidates,
        model_comparison=int(bool(cfg.model_comparison)),
        planet_log_odds_min=cfg.planet_log_odds_min, evidence=ev,
        single_events=int(bool(cfg.single_events)))
    c._keep = (d, bank, ev)
    return c


def _new_report(cfg: SearchConfig, table_cap=0):
    """A fbls_search_report with room for every family (at most one per
    duration), every band, the selected events and table_cap rows of the
    pass-1 period table."""
    nd = len(cfg.durations_days())
    fam_cap = nd * max(cfg.passes, 1)
    r = SearchReport()
    keep = dict(events=(SingleEvent * max(cfg.max_candidates, 1))(),
                stats=(EventStats * nd)(), fams=(FamilyPass * fam_cap)(),
                bands=(ScanStats * (fam_cap * 64))(),
                table=(Row * max(table_cap, 1))(),
                table_family=np.zeros(max(table_cap, 1), dtype=np.uint64))
    r.events, r.events_cap = keep["events"], cfg.max_candidates
    r.event_stats, r.event_stats_cap = keep["stats"], nd
    r.families, r.families_cap = keep["fams"], fam_cap
    r.bands, r.bands_cap = keep["bands"], fam_cap * 64
    r.period_table, r.period_table_cap = keep["table"], table_cap
    r.period_table_family = keep["table_family"].ctypes.data_as(_P(_u64))
    r._keep = keep
    return r


def _report_dict(r: SearchReport):
    k = r._keep
    fams = []
    for i in range(min(r.nfamilies, r.families_cap)):
        f = r.families[i]
        nd = min(int(f.nd), 64)
        fams.append(dict(
            pass_=int(f.pass_), family=int(f.family), window_days=f.window_days,
            observations=int(f.observations), clipped=int(f.clipped),
            error_inflation=f.error_inflation, segments=int(f.segments),
            durations_hours=list(f.durations_hours[:nd]), periods=int(f.periods),
            rho=list(f.rho[:nd]),
            bands=[r.bands[b].as_dict() for b in range(int(f.band_first),
                                                        int(f.band_first + f.nbands))
                   if b < r.bands_cap]))
    nt = min(r.period_table_rows, r.period_table_cap)
    table = [dict(period=k["table"][i].period, epoch=k["table"][i].epoch,
                  duration=k["table"][i].duration, score=k["table"][i].score,
                  depth=k["table"][i].depth) for i in range(nt)]
    return dict(
        events=[r.events[i].as_dict() for i in range(min(r.nevents, r.events_cap))],
        event_stats=[r.event_stats[i].as_dict()
                     for i in range(min(r.nevent_stats, r.event_stats_cap))],
        families=fams, period_table=table,
        period_table_family=k["table_family"][:nt].astype(int).tolist(),
        period_table_rows=int(r.period_table_rows), observations=int(r.observations),
        periods_tested=int(r.periods_tested), nfam=int(r.nfam),
        baseline_days=r.baseline_days, seconds_preparation=r.seconds_preparation,
        seconds_search=r.seconds_search, seconds_scan=r.seconds_scan,
        seconds_select=r.seconds_select, seconds_evidence=r.seconds_evidence,
        seconds_single_events=r.seconds_single_events)


def single_event_search(gpu, dev, durations, rho, cfg: SearchConfig, capacity=131072):
    """fbls_single_event_search of one uploaded family: (rows as dicts, stats)."""
    durations, rho = _f64(durations), _f64(rho)
    ing = _f64(cfg.ingress_fractions)
    rows = (Row * max(capacity, 1))()
    nr = _sz()
    st = EventStats()
    _check(lib().fbls_single_event_search(
        gpu._h, dev._h, _dp(durations), _dp(rho), len(durations), _dp(ing), len(ing),
        cfg.shape_layers, cfg.threshold, cfg.phase_oversample, capacity, rows,
        ctypes.byref(nr), ctypes.byref(st)))
    keys = ("period", "epoch", "duration", "score", "depth")
    return [{k: getattr(rows[i], k) for k in keys} for i in range(nr.value)], st.as_dict()


def autotune(gpu, dev, durations, rho, cfg: SearchConfig, repeats=3):
    """fbls_autotune: the fastest launch shape for this star's search, kept
    on the context.  Returns (best, table) as dicts."""
    durations, rho = _f64(durations), _f64(rho)
    bank = make_bank(cfg, durations, rho)
    scfg = SearchCfgC(cfg.period_min, cfg.period_max, cfg.period_oversample)
    table = (TuneEntry * 32)()
    n = _sz()
    best = TuneEntry()
    _check(lib().fbls_autotune(gpu._h, dev._h, _dp(durations), _dp(rho), len(durations),
                               ctypes.byref(bank), ctypes.byref(scfg), repeats, table, 32,
                               ctypes.byref(n), ctypes.byref(best)))
    as_d = lambda e: {"lds_bytes": e.lds_bytes, "block": e.block, "seconds": e.seconds}
    return as_d(best), [as_d(table[i]) for i in range(min(n.value, 32))]


def search_prepared(gpu, time, flux, flux_err, cfg: SearchConfig, report=False,
                    table_cap=0):
    """search_prepared + search_families (pipeline.rs:41-238) on the GPU:
    returns (peaks in acceptance order as dicts, pass statistics), plus the
    report dict (single events, family and band statistics, the first
    table_cap rows of the pass-1 period table, timings) when report."""
    time, flux, flux_err = _f64(time), _f64(flux), _f64(flux_err)
    c = pipeline_cfg(cfg)
    cap = max(cfg.max_candidates, 1)
    peaks = (Peak * cap)()
    passes = (PassStats * max(cfg.passes, 1))()
    npk, npa = _sz(), _sz()
    rep = _new_report(cfg, table_cap)
    _check(lib().fbls_search_prepared(
        gpu._h, _dp(time), _dp(flux), _dp(flux_err), len(time), ctypes.byref(c),
        peaks, cap, ctypes.byref(npk), passes, len(passes), ctypes.byref(npa),
        ctypes.byref(rep)))
    out = ([peaks[i].as_dict() for i in range(min(npk.value, cap))],
           [passes[i].as_dict() for i in range(min(npa.value, len(passes)))])
    return out + (_report_dict(rep),) if report else out


def by_signal(peaks):
    """signal_candidates order (pipeline.rs:239-240)."""
    return sorted(peaks, key=lambda p: -p["snr"])


def by_model(peaks):
    """candidates order (pipeline.rs:241-259)."""
    def key(p):
        ev = p["log_bf_planet_noise"] if p["model_preferred"] else float("nan")
        v = ev if ev == ev else p["planet_log_odds"]
        return (0 if p["model_preferred"] else 1, -(v if v == v else -np.inf))
    return sorted(peaks, key=key)


# ----------------------------------------------------- band-proposal network

def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fptr(a):
    return a.ctypes.data_as(_P(ctypes.c_float))


class Cnn:
    """The band-proposal network on the GPU (fbls_cnn_*)."""

    def __init__(self, gpu, in_channels=8, width=64, layers=6, kernel=7,
                 pos_weight=20.0, seed=1):
        self.gpu = gpu
        self.cfg = CnnCfg(in_channels, width, layers, kernel, pos_weight)
        self._h = _vp()
        _check(lib().fbls_cnn_create(gpu._h, ctypes.byref(self.cfg), seed,
                                     ctypes.byref(self._h)))
        self.count = lib().fbls_cnn_param_count(self._h)

    def params(self):
        out = np.empty(self.count, dtype=np.float32)
        _check(lib().fbls_cnn_get_params(self._h, _fptr(out)))
        return out

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.count
        _check(lib().fbls_cnn_set_params(self._h, _fptr(p)))

    def set_fused(self, enable):
        """Fused conv + GELU inference kernels (default) or the training
        kernels; the logits are bit-identical."""
        _check(lib().fbls_cnn_set_fused(self._h, 1 if enable else 0))

    def forward(self, x):
        x = _f32(x)
        B, L, _ = x.shape
        out = np.empty((B, L), dtype=np.float32)
        _check(lib().fbls_cnn_forward(self._h, _fptr(x), B, L, _fptr(out)))
        return out

    def loss_grad(self, x, y):
        x, y = _f32(x), _f32(y)
        B, L, _ = x.shape
        loss = ctypes.c_float()
        g = np.empty(self.count, dtype=np.float32)
        _check(lib().fbls_cnn_loss_grad(self._h, _fptr(x), _fptr(y), B, L,
                                        ctypes.byref(loss), _fptr(g)))
        return loss.value, g

    def train_step(self, x, y, lr=1e-3, weight_decay=1e-4):
        x, y = _f32(x), _f32(y)
        B, L, _ = x.shape
        loss = ctypes.c_float()
        _check(lib().fbls_cnn_train_step(self._h, _fptr(x), _fptr(y), B, L, lr,
                                         weight_decay, ctypes.byref(loss)))
        return loss.value

    def close(self):
        if self._h:
            lib().fbls_cnn_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def coarse_inputs(gpu, devs, periods, durations, nb_min=16, nb_max=1024):
    """The network input (fbls_coarse_inputs): clip(SNR, -5, 40) / 8 as
    float32, [star, period, duration]."""
    periods, durations = _f64(periods), _f64(durations)
    hs = (_vp * len(devs))(*[d._h.value for d in devs])
    out = np.empty((len(devs), len(periods), len(durations)), dtype=np.float32)
    _check(lib().fbls_coarse_inputs(gpu._h, hs, len(devs), _dp(periods), len(periods),
                                    _dp(durations), len(durations), nb_min, nb_max,
                                    out.ctypes.data_as(_P(ctypes.c_float))))
    return out


def propose(gpu, net, devs, periods, durations, nb_min=16, nb_max=1024, peaks=False):
    """fbls_propose: the network's logits [star, period bin] from the coarse
    periodogram computed and consumed on the device; with peaks, also the
    periodogram's largest SNR over the durations per bin (logits, peaks)."""
    periods, durations = _f64(periods), _f64(durations)
    hs = (_vp * len(devs))(*[d._h.value for d in devs])
    out = np.empty((len(devs), len(periods)), dtype=np.float32)
    pk = np.empty((len(devs), len(periods)), dtype=np.float32) if peaks else None
    _check(lib().fbls_propose(gpu._h, net._h, hs, len(devs), _dp(periods), len(periods),
                              _dp(durations), len(durations), nb_min, nb_max,
                              out.ctypes.data_as(_P(ctypes.c_float)),
                              pk.ctypes.data_as(_P(ctypes.c_float)) if peaks else None))
    return (out, pk) if peaks else out


def coarse_features(gpu, devs, periods, durations, nb_min=16, nb_max=1024):
    """Coarse periodogram (fbls_coarse_features): [star, period, duration]."""
    periods, durations = _f64(periods), _f64(durations)
    hs = (_vp * len(devs))(*[d._h.value for d in devs])
    out = np.empty((len(devs), len(periods), len(durations)))
    _check(lib().fbls_coarse_features(gpu._h, hs, len(devs), _dp(periods), len(periods),
                                      _dp(durations), len(durations), nb_min, nb_max,
                                      _dp(out)))
    return out


# ------------------------------------------------------------ raw mode

def biweight_flatten(gpu, time, rel, window, gap=0.5, c=5.0, iterations=10):
    """fbls_biweight_flatten: (1 + rel) / biweight location - 1 on the GPU."""
    time, rel = _f64(time), _f64(rel)
    out = np.empty(len(time))
    _check(lib().fbls_biweight_flatten(gpu._h, _dp(time), _dp(rel), len(time), window, gap,
                                       c, iterations, _dp(out)))
    return out


def prep_cfg(cfg: SearchConfig):
    return PrepCfg(cfg.bin_minutes, cfg.window_days, cfg.window_factor, cfg.window_max_days,
                   cfg.clip_sigma, cfg.edge_trim_windows)


def prepare_families(gpu, time, flux, err, group, cfg: SearchConfig):
    """fbls_prepare_families: list of family dicts (time, flux, sigma,
    window, clipped, indices of the durations it searches)."""
    time, flux = _f64(time), _f64(flux)
    n = len(time)
    e = _f64(err) if err is not None else None
    g = np.ascontiguousarray(group, dtype=np.int64) if group is not None else None
    d = cfg.durations_days()
    nd = len(d)
    pc = prep_cfg(cfg)
    nf = _sz()
    fn = np.zeros(nd, dtype=np.uintp)
    ft, ff, fs = np.zeros(nd * n), np.zeros(nd * n), np.zeros(nd * n)
    fw = np.zeros(nd)
    fc = np.zeros(nd, dtype=np.uintp)
    fm = np.zeros(nd * nd, dtype=np.uint8)
    _check(lib().fbls_prepare_families(
        gpu._h, _dp(time), _dp(flux), _dp(e) if e is not None else None,
        g.ctypes.data_as(_P(ctypes.c_int64)) if g is not None else None, n, _dp(d), nd,
        ctypes.byref(pc), ctypes.byref(nf), fn.ctypes.data_as(_P(_sz)), _dp(ft), _dp(ff),
        _dp(fs), _dp(fw), fc.ctypes.data_as(_P(_sz)), fm.ctypes.data_as(_P(ctypes.c_uint8))))
    out = []
    for k in range(nf.value):
        m = int(fn[k])
        out.append(dict(time=ft[k * n:k * n + m], flux=ff[k * n:k * n + m],
                        sigma=fs[k * n:k * n + m], window=float(fw[k]), clipped=int(fc[k]),
                        indices=np.nonzero(fm[k * nd:(k + 1) * nd])[0]))
    return out


def search_transits(gpu, time, flux, err, group, cfg: SearchConfig, report=False,
                    table_cap=0):
    """search_transits (pipeline.rs:27-37, 70-238) on raw flux: the families
    of prepare_families searched together.  Returns (peaks in acceptance
    order as dicts, pass statistics), plus the report dict when report."""
    time, flux = _f64(time), _f64(flux)
    e = _f64(err) if err is not None else None
    g = np.ascontiguousarray(group, dtype=np.int64) if group is not None else None
    c = pipeline_cfg(cfg)
    pc = prep_cfg(cfg)
    cap = max(cfg.max_candidates, 1)
    peaks = (Peak * cap)()
    passes = (PassStats * max(cfg.passes, 1))()
    npk, npa = _sz(), _sz()
    rep = _new_report(cfg, table_cap)
    _check(lib().fbls_search_transits(
        gpu._h, _dp(time), _dp(flux), _dp(e) if e is not None else None,
        g.ctypes.data_as(_P(ctypes.c_int64)) if g is not None else None, len(time),
        ctypes.byref(pc), ctypes.byref(c), peaks, cap, ctypes.byref(npk), passes, len(passes),
        ctypes.byref(npa), ctypes.byref(rep)))
    out = ([peaks[i].as_dict() for i in range(min(npk.value, cap))],
           [passes[i].as_dict() for i in range(min(npa.value, len(passes)))])
    return out + (_report_dict(rep),) if report else out


# ------------------------------------------------------------ formats

def read_inputs(paths):
    """fbls_read_inputs (main.rs:50-132): dict of time, flux, err (None when
    the files carry no flux_err) and group."""
    enc = [os.fsencode(p) for p in paths]
    arr = (ctypes.c_char_p * max(len(enc), 1))(*enc)
    n, he = _sz(), ctypes.c_int()
    _check(lib().fbls_read_inputs(arr, len(enc), None, None, None, None, 0, ctypes.byref(n),
                                  ctypes.byref(he)))
    m = n.value
    t, f, e = np.empty(m), np.empty(m), np.empty(m)
    g = np.empty(m, dtype=np.int64)
    _check(lib().fbls_read_inputs(arr, len(enc), _dp(t), _dp(f), _dp(e),
                                  g.ctypes.data_as(_P(ctypes.c_int64)), m, ctypes.byref(n),
                                  ctypes.byref(he)))
    return dict(time=t, flux=f, err=e if he.value else None, group=g)


def format_f64(x):
    """The report's text of a double (fbls_format_f64)."""
    buf = ctypes.create_string_buffer(64)
    lib().fbls_format_f64(float(x), buf, 64)
    return buf.value.decode()


def json_reformat(text):
    """fbls_json_reformat: JSON text re-printed in the report layout."""
    raw = text.encode() if isinstance(text, str) else text
    n = _sz()
    _check(lib().fbls_json_reformat(raw, None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    _check(lib().fbls_json_reformat(raw, buf, n.value + 1, ctypes.byref(n)))
    return buf.value.decode()


def cli(args):
    """fbls_cli in this process: args[0] is "search" or "batch"; returns the
    exit status."""
    enc = [os.fsencode(a) for a in args]
    arr = (ctypes.c_char_p * max(len(enc), 1))(*enc)
    return lib().fbls_cli(len(enc), arr)


CLI_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fbls")

// Search configuration of the command-line host (see config.h).
#include "config.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fbls {

SearchConfig::SearchConfig() {
  // config.rs:94: 0.5 h x 1.25^i, rounded half to even at 1e-6 h
  for (int i = 0; i < 13; i++)
    durations_hours.push_back(std::nearbyint(0.5 * std::pow(1.25, i) * 1e6) / 1e6);
}

std::string SearchConfig::validate() const {
  if (!(0.0 < period_min && period_min < period_max))
    return "require 0 < period_min < period_max";
  if (durations_hours.empty() ||
      std::any_of(durations_hours.begin(), durations_hours.end(),
                  [](double d) { return !(std::isfinite(d) && d > 0.0); }))
    return "durations_hours must contain positive finite values";
  const double dmin = *std::min_element(durations_hours.begin(), durations_hours.end());
  if (dmin / 24.0 >= period_min / 4.0)
    return "minimum duration must be below one quarter of period_min";
  if (ingress_fractions.empty() ||
      std::any_of(ingress_fractions.begin(), ingress_fractions.end(),
                  [](double x) { return !(x >= 0.0 && x < 0.5); }))
    return "ingress fractions must be in [0, 0.5)";
  if (shape_layers < 1 || phase_oversample < 1 || period_batch < 1 || passes < 1 ||
      prescan_factor < 1)
    return "invalid integer search parameter";
  if (!(0.0 < prescan_fraction && prescan_fraction <= 1.0))
    return "prescan_fraction must be in (0, 1]";
  return "";
}

std::vector<double> SearchConfig::durations_days() const {
  std::vector<double> d;
  for (double h : durations_hours) d.push_back(h / 24.0);
  std::sort(d.begin(), d.end());
  d.erase(std::unique(d.begin(), d.end()), d.end());
  return d;
}

json::Value SearchConfig::to_json() const {
  using json::Value;
  Value v = Value::object();
  auto u = [](size_t x) { return Value::unsigned_(x); };
  auto f = [](double x) { return Value::number(x); };
  v["period_min"] = f(period_min);
  v["period_max"] = f(period_max);
  v["durations_hours"] = json::numbers(durations_hours.data(), durations_hours.size());
  v["threshold"] = f(threshold);
  v["bin_minutes"] = f(bin_minutes);
  v["window_days"] = f(window_days);
  v["window_factor"] = f(window_factor);
  v["window_max_days"] = f(window_max_days);
  v["edge_trim_windows"] = f(edge_trim_windows);
  v["clip_sigma"] = f(clip_sigma);
  v["correlated_noise"] = Value::boolean(correlated_noise);
  v["ingress_fractions"] = json::numbers(ingress_fractions.data(), ingress_fractions.size());
  v["shape_layers"] = u(shape_layers);
  v["phase_oversample"] = u(phase_oversample);
  v["shape_gate_fraction"] = f(shape_gate_fraction);
  v["prescan_factor"] = u(prescan_factor);
  v["prescan_fraction"] = f(prescan_fraction);
  v["period_oversample"] = f(period_oversample);
  v["max_candidates"] = u(max_candidates);
  v["period_batch"] = u(period_batch);
  v["max_fold_mib"] = f(max_fold_mib);
  v["passes"] = u(passes);
  v["min_events"] = u(min_events);
  v["single_events"] = Value::boolean(single_events);
  v["model_comparison"] = Value::boolean(model_comparison);
  v["evidence_candidates"] = u(evidence_candidates);
  v["evidence_rtol"] = f(evidence_rtol);
  v["planet_log_odds_min"] = f(planet_log_odds_min);
  v["planet_prior"] = f(planet_prior);
  v["binary_prior"] = f(binary_prior);
  v["planet_depth_scale"] = f(planet_depth_scale);
  v["binary_depth_scale"] = f(binary_depth_scale);
  v["diluted_binary_depth_scale"] = f(diluted_binary_depth_scale);
  v["stellar_radius_solar"] =
      stellar_radius_solar > 0.0 ? f(stellar_radius_solar) : Value::null();
  v["planet_radius_scale_rj"] = f(planet_radius_scale_rj);
  v["target_flux_fraction"] = f(target_flux_fraction);
  return v;
}

CabiConfig::CabiConfig(const SearchConfig& cfg)
    : durations(cfg.durations_days()), ingress(cfg.ingress_fractions) {
  std::memset(&pipeline, 0, sizeof(pipeline));
  std::memset(&prep, 0, sizeof(prep));
  fbls_pipeline_cfg& p = pipeline;
  p.durations = durations.data();
  p.nd = durations.size();
  p.correlated = cfg.correlated_noise;
  p.bank.ingress = ingress.data();
  p.bank.ns = ingress.size();
  p.bank.layers = (uint32_t)cfg.shape_layers;
  p.bank.threshold = cfg.threshold;
  // scan.rs:433-437
  p.bank.shape_gate = ingress[0] == 0.0 ? cfg.shape_gate_fraction * cfg.threshold : 0.0;
  p.bank.phase_oversample = (uint32_t)cfg.phase_oversample;
  p.bank.period_batch = (uint32_t)cfg.period_batch;
  p.bank.max_fold_mib = cfg.max_fold_mib;
  p.bank.prescan_factor = (uint32_t)cfg.prescan_factor;
  p.bank.prescan_fraction = cfg.prescan_fraction;
  p.search.period_min = cfg.period_min;
  p.search.period_max = cfg.period_max;
  p.search.period_oversample = cfg.period_oversample;
  p.max_candidates = cfg.max_candidates;
  p.passes = cfg.passes;
  p.min_events = cfg.min_events;
  p.evidence_candidates = cfg.evidence_candidates;
  p.model_comparison = cfg.model_comparison;
  p.planet_log_odds_min = cfg.planet_log_odds_min;
  p.single_events = cfg.single_events;
  fbls_evidence_cfg& e = p.evidence;
  e.ingress = ingress.data();
  e.ns = ingress.size();
  e.layers = (uint32_t)cfg.shape_layers;
  e.phase_oversample = (uint32_t)cfg.phase_oversample;
  e.threshold = cfg.threshold;
  e.evidence_rtol = cfg.evidence_rtol;
  e.planet_prior = cfg.planet_prior;
  e.binary_prior = cfg.binary_prior;
  e.planet_depth_scale = cfg.planet_depth_scale;
  e.binary_depth_scale = cfg.binary_depth_scale;
  e.diluted_binary_depth_scale = cfg.diluted_binary_depth_scale;
  e.stellar_radius_solar = cfg.stellar_radius_solar;
  e.planet_radius_scale_rj = cfg.planet_radius_scale_rj;
  e.target_flux_fraction = cfg.target_flux_fraction;
  prep.bin_minutes = cfg.bin_minutes;
  prep.window_min = cfg.window_days;
  prep.window_factor = cfg.window_factor;
  prep.window_max = cfg.window_max_days;
  prep.clip_sigma = cfg.clip_sigma;
  prep.edge_trim = cfg.edge_trim_windows;
}

}  // namespace fbls

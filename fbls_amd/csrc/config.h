// The search configuration of the command-line host (isakruas/fbls
// src/config.rs:5-139): every field with its default, the validation, the
// sorted unique durations, its report form and the C-ABI configurations
// built from it.
#pragma once
#include <string>
#include <vector>

#include "../../include/fbls.h"
#include "json.h"

namespace fbls {

struct SearchConfig {
  double period_min = 0.3, period_max = 10.0;
  std::vector<double> durations_hours;  // config.rs:94 defaults, see below
  double threshold = 7.1;
  double bin_minutes = 10.0, window_days = 0.4, window_factor = 3.0, window_max_days = 3.5;
  double edge_trim_windows = 0.5, clip_sigma = 4.0;
  bool correlated_noise = true;
  std::vector<double> ingress_fractions{0.0, 0.2, 0.45};
  size_t shape_layers = 4, phase_oversample = 4;
  double shape_gate_fraction = 0.6;
  size_t prescan_factor = 2;
  double prescan_fraction = 0.7, period_oversample = 3.0;
  size_t max_candidates = 20, period_batch = 4096;
  double max_fold_mib = 1024.0;
  size_t passes = 3, min_events = 2;
  bool single_events = true, model_comparison = true;
  size_t evidence_candidates = 32;
  double evidence_rtol = 1e-12, planet_log_odds_min = 0.0;
  double planet_prior = 0.01, binary_prior = 0.01;
  double planet_depth_scale = 0.01, binary_depth_scale = 0.1, diluted_binary_depth_scale = 0.01;
  double stellar_radius_solar = 0.0;  // <= 0: none (null in the report)
  double planet_radius_scale_rj = 1.0, target_flux_fraction = 1.0;

  SearchConfig();
  // config.rs:91-131; empty string when valid
  std::string validate() const;
  // config.rs:134-139
  std::vector<double> durations_days() const;
  json::Value to_json() const;
};

// The C-ABI configurations of one search; the arrays they point at live
// here.
struct CabiConfig {
  std::vector<double> durations, ingress;
  fbls_pipeline_cfg pipeline;
  fbls_prep_cfg prep;
  explicit CabiConfig(const SearchConfig& cfg);
  CabiConfig(const CabiConfig&) = delete;
  CabiConfig& operator=(const CabiConfig&) = delete;
};

}  // namespace fbls

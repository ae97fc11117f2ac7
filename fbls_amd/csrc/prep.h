// Host-side (CPU) pieces of the hot path's inputs: period grid, durations,
// the diagonal noise model of a prepared curve, fixed-point quantisation and
// the synthetic light-curve generator.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/fbls.h"

namespace fbls {

// Fixed-point fields.  Every word stands for an integer in units of 2^-e,
// and the integers of one array sum in absolute value to below 2^53, so any
// sum of them is exact in FP64 whatever the order.
//  plain: word = rint(x * 2^e), e = 30 - exponent(max |x|), |word| <= 2^30;
//    absolute error <= 2^-30 max |x| per value.
//  block: word = mantissa * 32 + shift, value = mantissa * 2^shift, a 27-bit
//    signed mantissa over a 5-bit shift, e = 52 - exponent(sum |x|);
//    relative error <= 2^-26 per value (absolute below 2^-52 sum |x| for
//    the values under 2^26 units).
// plain when max |x| <= 16 mean |x|, block otherwise, so that in both
// modes |error| <= 2^-26 max(|x|, mean |x|) for every value.
struct Fixed {
  int e = 0;
  bool block = false;
};
Fixed quantize(const double* x, size_t n, int32_t* q);
// The integer a word stands for (units of 2^-e).
inline double fixed_value(int32_t word, bool block) {
  return block ? (double)(word >> 5) * (double)(1ull << (word & 31)) : (double)word;
}

// grid.rs:5-27
size_t period_grid_count(double minimum, double maximum, double baseline,
                         double minimum_duration, double oversample);
void period_grid(double minimum, double maximum, double* out, size_t count);
// Elements [i0, i1) of the same grid (each element is computed alone).
void period_grid_range(double minimum, double maximum, double* out, size_t count, size_t i0,
                       size_t i1);
// Element i of period_grid(minimum, maximum, ., count), evaluated alone.
struct GridFormula {
  double minimum = 0.0, maximum = 0.0, a = 0.0, b = 0.0;
  size_t count = 0;
  GridFormula() = default;
  GridFormula(double mn, double mx, size_t n);
  double operator()(size_t i) const;
};

// config.rs:177-182
size_t durations_days(const double* hours, size_t n, double* out);

// scan.rs:440-471 (octave bands) and scan.rs:377-381 (nb per band)
std::vector<fbls_band> plan_bands(const double* grid, size_t count,
                                  const double* sorted_durations, size_t nd,
                                  uint32_t phase_oversample,
                                  uint32_t prescan_factor);
// plan_bands over a grid given by its formula: band edges by binary search.
std::vector<fbls_band> plan_bands(const GridFormula& grid, const double* sd, size_t nd,
                                  uint32_t phase_oversample, uint32_t prescan_factor);

// curve.rs:541-569 for a prepared curve (flux relative to 0, errors given).
int family_state(const double* time, const double* flux, const double* sigma,
                 size_t n, const double* durations, size_t nd, int correlated,
                 double* t, double* v, double* w, size_t* out_n,
                 double* baseline_norm, double* origin, double* rho,
                 double* inflation, size_t* segments, double* residual_energy,
                 std::string* msg, double flux_offset = 0.0);

// One detrended family of the raw path (curve.rs:23-41 PreparedCurve).
struct Family {
  std::vector<double> time, flux, sigma;  // absolute time, relative flux, sigma
  double window = 0.0;
  size_t clipped = 0;
  std::vector<size_t> indices;            // bank durations of this family
};
// The biweight detrend (curve.rs:204-248): out[i] = (1 + rel[i]) /
// location_i - 1 over sorted times t, windows per gap-separated segment.
using Flatten = std::function<int(const std::vector<double>& t, const std::vector<double>& rel,
                                  double window, double gap, double c, int iterations,
                                  std::vector<double>* out)>;
// curve.rs:355-377
std::vector<std::pair<double, std::vector<size_t>>> detrending_windows(
    const double* durations, size_t nd, double minimum, double factor, double maximum);
// curve.rs:387-425
int prepare_families(const double* time, const double* flux, const double* err,
                     const int64_t* group, size_t n, const double* durations, size_t nd,
                     const fbls_prep_cfg* cfg, const Flatten& flatten,
                     std::vector<Family>* out, std::string* msg);

int synth_star(const fbls_synth_cfg* cfg, double* time, double* flux,
               double* flux_err, size_t cap, size_t* n, std::string* msg);

}  // namespace fbls

/*
 * FP64 CPU oracle of the FBLS per-family period scan.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (fbls_amd/, include/fbls.h) never links or calls it.
 *
 * Every function restates the reference implementation (isakruas/fbls) line
 * by line; citations are path:line in that repository.  Summation order is
 * sequential and deterministic.  Compile with -ffp-contract=off: every
 * multiply-add below must round twice, as the reference's Rust code does.
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   orc_period_grid          grid.rs:34-39 KAT + search.json band fixture
 *   orc_durations_days       config.rs:94 rounding + search.json config hours
 *   orc_median_lower/mad     closed forms on small vectors
 *   orc_fit_segment_noise    curve.rs:636-661 KAT (noise_model_normalizes_and_inflates)
 *   orc_prepare_ready        closed forms (sum v = 0 up to rounding, C = sum p)
 *   orc_correlated_noise     closed forms on constructed residuals
 *   orc_phase_cell           exact integer brute force (long double / i128)
 *   orc_scan_bank            brute force over all (j, k) templates in a
 *                            separate pure-Python restatement, the closed
 *                            form box depth/SNR on noiseless injections,
 *                            staircase sums vs an explicit dot product
 *   orc_scan_periods         prescan survivor logic vs two orc_scan_bank calls
 *   orc_signal_search        search.json fixture (nb, periods, batches,
 *                            templates_scanned per band)
 *   orc_rng_*                survey/numeric.rs:345-364 restated; stream KAT
 */
#ifndef FBLS_ORACLE_H
#define FBLS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* config.rs:94 / config.rs:177-182 */
size_t orc_durations_days(const double* hours, size_t n, double* out_days);

/* grid.rs:5-27 */
size_t orc_period_grid_count(double minimum, double maximum, double baseline,
                             double minimum_duration, double oversample);
void orc_set_threads(int n);
void orc_period_grid(double minimum, double maximum, double baseline,
                     double minimum_duration, double oversample, double* out,
                     size_t count);

/* stats.rs:4-18 */
double orc_median_lower(const double* values, size_t n);
double orc_mad_scale(const double* values, size_t n);

/* curve.rs:89-198 without binning (bin_minutes = 0) and without groups.
 * Outputs are sorted, origin-relative; returns 0 or a nonzero error code. */
int orc_fit_segment_noise(const double* time, const double* flux,
                          const double* err /* may be NULL */, size_t n,
                          double gap_days, double* out_time,
                          double* out_deficit, double* out_variance,
                          size_t* out_n, double* origin, double* inflation,
                          size_t* segments);

/* curve.rs:427-450 */
int orc_prepare_ready(const double* deficit, const double* variance, size_t n,
                      double* v, double* w, double* baseline_norm,
                      double* residual_energy);

/* curve.rs:456-539 */
void orc_correlated_noise_factors(const double* time, const double* residual,
                                  const double* sigma, size_t n,
                                  const double* durations, size_t nd,
                                  double* rho);

/* The build's fixed-point input representation; returns e, the words
 * standing for integers in units of 2^-e.  When max |x| <= 16 mean |x|
 * (*block = 0): q = rint(x * 2^e), e = 30 - ceil-exponent(max |x|), so
 * |q| <= 2^30.  Otherwise (*block = 1): e = 52 - ceil-exponent(sum |x|),
 * r = rint(x * 2^e), and the word m * 32 + s holds r to 26 significant bits
 * as m * 2^s (|m| < 2^26, the smallest such s). */
int orc_quantize(const double* x, size_t n, int32_t* q, int* block);
/* The integer (units of 2^-e) a word of orc_quantize stands for. */
double orc_fixed_value(int32_t word, int block);

/* kernels.rs:314-327: floor(t * scale) mod nb */
uint32_t orc_phase_cell(double t, double scale, uint32_t nb);

/* kernels.rs:464-845,1113-1155 with the f64 ("exact") ranking everywhere.
 * records: np x 9 {P, scale, phase, width, a, norm, d, ingress, rho};
 * winner: np x 4 u32 {phase, width, shape, duration index} (may be NULL).
 * Returns 0, or 2 when a valid template had nonpositive precision
 * (scan.rs:335-337). */
int orc_scan_bank(const double* t, const double* v, const double* w, size_t n,
                  double baseline_norm, const double* periods, size_t np,
                  const double* durations, const double* rho, size_t nd,
                  const double* ingress, size_t ns, uint32_t layers,
                  double gate, uint32_t nb, double* records, double* score,
                  double* depth, uint32_t* winner);

typedef struct {
  uint64_t periods, prescan_survivors, phase_bins, batches, templates_scanned;
} orc_scan_stats;

typedef struct {
  double period, epoch, duration, score, depth;
} orc_row;

typedef struct {
  uint32_t layers, phase_oversample, period_batch, prescan_factor;
  double threshold, shape_gate, max_fold_mib, prescan_fraction;
} orc_bank_cfg;

/* scan.rs:363-408.  rows must hold np entries; returns status. */
int orc_scan_periods(const double* t, const double* v, const double* w,
                     size_t n, double baseline_norm, double origin,
                     const double* periods, size_t np, const double* durations,
                     const double* rho, size_t nd, const double* ingress,
                     size_t ns, const orc_bank_cfg* cfg, orc_row* rows,
                     size_t* nrows, orc_scan_stats* stats);

typedef struct {
  double period_min, period_max, period_oversample;
} orc_search_cfg;

/* scan.rs:413-473 for one family.  durations/rho unsorted as given.
 * rows must hold grid-count entries; stats at most 64 bands. */
int orc_signal_search(const double* t, const double* v, const double* w,
                      size_t n, double baseline_norm, double origin,
                      const double* durations, const double* rho, size_t nd,
                      const double* ingress, size_t ns,
                      const orc_bank_cfg* bank, const orc_search_cfg* scfg,
                      orc_row* rows, size_t rows_cap, size_t* nrows,
                      orc_scan_stats* stats, size_t* nbands,
                      size_t* grid_count);

/* Band plan of signal_search (scan.rs:440-471) and scan_periods' nb
 * (scan.rs:377-381).  Per band: lo, first index, count, nd, nb, nb_coarse
 * (0 without prescan).  Returns the number of bands (<= cap written). */
typedef struct {
  double lo;
  size_t first, count, nd;
  uint32_t nb, nb_coarse;
} orc_band;
size_t orc_plan_bands(const double* grid, size_t count,
                      const double* sorted_durations, size_t nd,
                      uint32_t phase_oversample, uint32_t prescan_factor,
                      orc_band* out, size_t cap);

/* batch count of scan_bank's memory budget, scan.rs:177-183 */
uint64_t orc_bank_batches(size_t np, uint32_t nb, uint32_t period_batch,
                          double max_fold_mib);

/* ------------------------------------------------------------ evidence
 * src/search/evidence.rs and src/search/math.rs (oracle/evidence.c). */

/* SearchConfig fields read by compare_search (config.rs:33-42,53-85).
 * stellar_radius_solar <= 0 means None. */
typedef struct {
  const double* ingress;
  size_t ns;
  uint32_t layers, phase_oversample;
  double threshold, evidence_rtol;
  double planet_prior, binary_prior, planet_depth_scale, binary_depth_scale,
      diluted_binary_depth_scale, stellar_radius_solar, planet_radius_scale_rj,
      target_flux_fraction;
} orc_evidence_cfg;

/* EvidenceRow (evidence.rs:637-653) plus the decoded winner indices. */
typedef struct {
  double period, epoch, duration, snr, depth, planet_log_odds,
      log_bf_planet_noise, log_bf_binary_noise, binary_secondary_ratio,
      binary_secondary_phase, binary_odd_even_ratio, binary_period_days,
      ingress_fraction;
  uint32_t selection_flags, phase, duration_index, shape;
} orc_evidence_row;

double orc_exp(double x);
double orc_log(double x);
void orc_evidence_use_libm(int on);
double orc_erfc(double x);                                   /* math.rs:8-100 */
double orc_log_add(double x, double y);                      /* math.rs:103-110 */
double orc_positive_amplitude_evidence(double a, double q, double tau); /* math.rs:115-140 */
double orc_subspace_upper(double a, double b, double q, double r,
                          double cross, double rho2);        /* math.rs:145-177 */
void orc_evidence_parameters(const orc_evidence_cfg* cfg, double* p30); /* evidence.rs:665-689 */
/* evidence.rs:693-1003.  rows holds 2 entries; *nrows in 0..2. */
int orc_compare_search(const double* t, const double* v, const double* w,
                       size_t n, double baseline_norm, double origin,
                       double residual_energy, const double* durations,
                       const double* rho, size_t nd, double period,
                       const orc_evidence_cfg* cfg, orc_evidence_row* rows,
                       size_t* nrows);
/* The same with the intermediate device arrays of the reference copied out
 * (each may be NULL; sizes in doubles): screen = tiles x 4 after
 * comparison_screen, parts = tiles x 2 x 15 after the full joint_tiles
 * (zeros when the screen is below threshold), floors = {floor_p, floor_b,
 * pilot key}; *tiles_out = ceil(nb / 64), *nb_out = nb. */
int orc_compare_search_ex(const double* t, const double* v, const double* w,
                          size_t n, double baseline_norm, double origin,
                          double residual_energy, const double* durations,
                          const double* rho, size_t nd, double period,
                          const orc_evidence_cfg* cfg, orc_evidence_row* rows,
                          size_t* nrows, double* screen, size_t screen_cap,
                          double* parts, size_t parts_cap, double* floors,
                          size_t* tiles_out, size_t* nb_out);

/* ------------------------------------------------------------ pipeline
 * Multi-pass prepared search with peak selection (oracle/pipeline.c). */

/* curve.rs:572-585 */
void orc_transit_mask(const double* time, size_t n, double period,
                      double epoch, double duration, double margin,
                      uint8_t* out);
/* peaks.rs:66-83 */
size_t orc_event_count(const double* time, size_t n, double period,
                       double epoch, double duration, size_t minimum_points);
/* kernels.rs:934-989 for one row: {a, hp, norm} */
void orc_rescore_row(const double* time, const double* v, const double* w,
                     const uint8_t* excluded, size_t n, double origin,
                     double period, double epoch, double duration, double* out3);
/* kernels.rs:992-1042 for one row: {in transit, in transit and not excluded} */
void orc_count_row(const double* time_all, const uint8_t* excluded, size_t n,
                   double period, double epoch, double duration,
                   uint32_t* out2);

/* Peak (peaks.rs:19-45); None is NaN, flags are 0/1. */
typedef struct {
  double period_days, epoch_bjd, duration_hours, snr, snr_folded,
      snr_without_strongest_event, depth, depth_ppm;
  uint64_t events, family, pass, row;
  double planet_log_odds, log_bf_planet_noise, log_bf_binary_noise,
      binary_secondary_ratio, binary_secondary_phase, binary_odd_even_ratio,
      binary_period_days, ingress_fraction, evidence_epoch_bjd,
      evidence_duration_hours;
  uint32_t model_preferred, binary_diagnostics_complete;
} orc_peak;

typedef struct {
  uint64_t pass, rows, periods, accepted;
} orc_pass_stats;

typedef struct {
  const double* durations; /* days, sorted unique (config.rs durations_days) */
  size_t nd;
  const double* ingress;
  size_t ns;
  int correlated;
  orc_bank_cfg bank;       /* bank.threshold is the selection threshold */
  orc_search_cfg search;
  size_t max_candidates, passes, min_events, evidence_candidates;
  int model_comparison;
  double planet_log_odds_min;
  orc_evidence_cfg evidence;
  int single_events;
} orc_pipeline_cfg;

/* events.rs:11-16 (EventStats) plus the bin count; peaks.rs:47-55 */
typedef struct {
  uint64_t templates_scanned, threshold_survivors;
  uint32_t overflow, phase_bins;
} orc_event_stats;
typedef struct {
  double epoch_bjd, duration_hours, snr, depth, depth_ppm;
  uint64_t points;
} orc_single_event;
/* single events of a search: selected events (capacity events_cap) and the
 * scan statistics of each family (capacity stats_cap); counts are totals */
typedef struct {
  orc_single_event* events;
  size_t events_cap, nevents;
  orc_event_stats* stats;
  size_t stats_cap, nstats;
} orc_events_out;

/* events.rs:18-153: every template at the single trial period whose score
 * reaches the threshold, in (phase bin, ascending duration, shape) order.
 * rows: capacity x {period, epoch, duration, score, depth}. */
void orc_single_event_search(const double* t, const double* v, const double* w, size_t n,
                             double C, double origin, const double* durations,
                             const double* rho, size_t nd, const double* ingress, size_t ns,
                             size_t layers, double threshold, uint32_t phase_oversample,
                             size_t capacity, orc_row* rows, size_t* nrows,
                             orc_event_stats* stats);
/* peaks.rs:308-487 on a caller's table (the selection step alone): see
 * pipeline.c for the family layout */
size_t orc_select_peaks(size_t nfam, const double* ft, const double* fv, const double* fw,
                        const size_t* foff, const double* fC, const double* forigin,
                        const double* fdur, const double* frho, const size_t* fdoff,
                        const orc_row* table, const size_t* family, size_t nrows,
                        const double* time_all, size_t n_all, double threshold, size_t limit,
                        size_t min_events, const uint8_t* masks_in, size_t n_masks_in,
                        uint8_t* excluded, orc_peak* out, size_t cap);
/* peaks.rs:490-550: masks n_masks x n_time; out capacity limit */
size_t orc_select_events(const orc_row* table, size_t nrows, const double* time,
                         size_t n_time, double threshold, size_t limit,
                         uint8_t* const* masks, size_t n_masks, orc_single_event* out);

/* pipeline.rs:41-221 (search_prepared without single events): peaks in
 * acceptance order (capacity cap), pass statistics (capacity passes). */
int orc_search_prepared(const double* time, const double* flux,
                        const double* err, size_t n, const orc_pipeline_cfg* cfg,
                        orc_peak* peaks, size_t cap, size_t* npeaks,
                        orc_pass_stats* pstats, size_t* npasses, orc_events_out* events);

/* Coarse periodogram of the band-proposal network (oracle/proposal.c):
 * out[i][g] for np periods and ng durations; see fbls_coarse_features. */
void orc_coarse_features(const double* t, const double* v, const double* w, size_t n,
                         double C, const double* periods, size_t np,
                         const double* durations, size_t ng, uint32_t nb_min,
                         uint32_t nb_max, double* out);

/* ------------------------------------------------------------ raw mode
 * oracle/raw.c: curve.rs:89-198 with groups and binning, and the biweight
 * detrend of curve.rs:204-248. */
int orc_fit_segment_noise_ex(const double* time, const double* flux, const double* err,
                             const int64_t* group, size_t n, double bin_minutes,
                             double gap, double* out_t, double* out_def, double* out_var,
                             size_t* out_n, double* origin, double* inflation,
                             size_t* segments);
void orc_biweight_flatten(const double* time, const double* rel, size_t n, double window,
                          double gap, double c, size_t iterations, double* out);
typedef struct {
  double bin_minutes, window_min, window_factor, window_max, clip_sigma, edge_trim;
} orc_prep_cfg;
int orc_prepare_families(const double* time, const double* flux, const double* err,
                         const int64_t* group, size_t n, const double* durations, size_t nd,
                         const orc_prep_cfg* cfg, size_t* nfam, size_t* fam_n,
                         double* fam_time, double* fam_flux, double* fam_sigma,
                         double* fam_window, size_t* fam_clipped, uint8_t* fam_members);

/* pipeline.rs:27-37 then 70-221 (search_transits without single events):
 * the families of orc_prepare_families searched together. */
int orc_search_transits(const double* time, const double* flux, const double* err,
                        const int64_t* group, size_t n, const orc_prep_cfg* prep,
                        const orc_pipeline_cfg* cfg, orc_peak* peaks, size_t cap,
                        size_t* npeaks, orc_pass_stats* pstats, size_t* npasses,
                        orc_events_out* events);

/* survey/numeric.rs:345-364 */
uint64_t orc_rng_new(uint64_t seed);
double orc_rng_uniform(uint64_t* state);
double orc_rng_normal(uint64_t* state, double mean, double sigma);

#ifdef __cplusplus
}
#endif
#endif

/*
 * fbls.h - C ABI of the MI355X-native FBLS search hot path.
 *
 * The reference (isakruas/fbls, Rust + cuda-oxide) crosses into device code
 * through the generated `gpu::LoadedModule` methods and the host drivers in
 * src/search/scan.rs.  This library replaces that seam at scan_bank
 * granularity: the caller keeps grid, band, prescan and selection policy if
 * it wants to, or calls the full family driver fbls_signal_search.
 *
 * Conventions
 *  - Every call is synchronous on return (scan.rs:330-333 copies results to
 *    the host before returning).
 *  - The caller owns all host arrays and sizes the outputs; the library owns
 *    device memory behind opaque handles.  No device pointer crosses the ABI.
 *  - One fbls_ctx per host thread (scan.rs:47-64 keeps one `Gpu` per thread);
 *    distinct contexts may be used concurrently, one context must not be.
 *  - Status codes: FBLS_OK, or one of FBLS_E_*; the message is available from
 *    fbls_last_error() on the calling thread.  No C++ exception leaves an
 *    entry point (host allocation failure returns FBLS_E_OOM).  Infinite scores are data
 *    (score -inf when no template is usable, kernels.rs:823-827).
 *  - Functions marked [host] run on the CPU only and need no GPU.
 */
#ifndef FBLS_H
#define FBLS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FBLS_OK = 0,
  FBLS_E_FOLD_BUDGET = 1,            /* scan.rs:180-182 */
  FBLS_E_NONPOSITIVE_PRECISION = 2,  /* scan.rs:335-337 */
  FBLS_E_HIP = 3,                    /* device/runtime failure */
  FBLS_E_ARG = 4,                    /* invalid argument / configuration */
  FBLS_E_OOM = 5,                    /* device or host allocation failed */
  FBLS_E_DATA = 6                    /* too few observations, bad flux, ... */
};

#define FBLS_RECORD_STRIDE 9  /* {P, scale, phase, width, a, norm, d, ingress, rho} */
#define FBLS_WINNER_STRIDE 4  /* {phase, width, shape, duration index} */

typedef struct fbls_ctx fbls_ctx;
typedef struct fbls_family fbls_family;

/* Template bank of one band; mirrors `Bank` (scan.rs:127-139). */
typedef struct {
  const double* durations; /* nd, ascending (scan.rs:418-424 sorts them) */
  const double* rho;       /* nd correlated-noise factors */
  size_t nd;
  const double* ingress;   /* ns ingress fractions, [0] = 0 is the box */
  size_t ns;
  uint32_t layers;         /* staircase layers (config.rs:105) */
  double threshold;        /* SNR threshold (config.rs:96) */
  double shape_gate;       /* tile gate for non-box shapes (scan.rs:433-437) */
  uint32_t phase_oversample;
  uint32_t period_batch;   /* only enters the `batches` statistic */
  double max_fold_mib;     /* fold budget; 0 selects min(1024, free/2) */
  uint32_t prescan_factor;
  double prescan_fraction;
} fbls_bank;

/* `ScanStats` (scan.rs:25-37) */
typedef struct {
  uint64_t periods, prescan_survivors, phase_bins, batches, templates_scanned,
      exact_tile_retries;
  double seconds, seconds_fold, seconds_cumulative, seconds_scan;
  /* device time of the scan kernels, from HIP events */
  double seconds_device;
} fbls_scan_stats;

/* `Row` (scan.rs:16-23) */
typedef struct {
  double period, epoch, duration, score, depth;
} fbls_row;

/* Search options of `signal_search` (scan.rs:413-437, config.rs:48-77). */
typedef struct {
  double period_min, period_max, period_oversample;
} fbls_search_cfg;

/* -- context ------------------------------------------------------------ */

/* `Gpu::new` (scan.rs:66-77): device context, stream and kernels. */
int fbls_ctx_create(int device, fbls_ctx** out);
void fbls_ctx_destroy(fbls_ctx* ctx);
/* `Gpu::memory` (scan.rs:80-90) */
int fbls_mem_info(fbls_ctx* ctx, size_t* free_bytes, size_t* total_bytes);
/* Launch shape override for tuning: LDS bytes per workgroup (0 = auto),
 * workgroup size (256, 512 or 1024, 0 = auto) and fold mode (0 = auto:
 * run-length pre-aggregation in registers, and a family whose fixed-point
 * weights are all equal folds no counts, each bin's count being its weight
 * sum divided by that weight; 1 = run-length, counts always folded; 2 = one
 * LDS atomic per observation and period).  Results never depend on them. */
int fbls_ctx_set_launch(fbls_ctx* ctx, uint32_t lds_bytes, uint32_t block,
                        uint32_t fold_mode);

/* Running totals of the scan kernel launches made on this context, timed
 * with HIP events on the context's stream (the profile counters of
 * `seconds_device`, scan.rs:354-356, summed over launches).  fold_pairs =
 * sum over launched periods of the family's observation count.  reset != 0
 * zeroes the totals after reading. */
typedef struct {
  uint64_t launches, periods, fold_pairs;
  double seconds;
} fbls_launch_totals;
int fbls_ctx_launch_totals(fbls_ctx* ctx, fbls_launch_totals* out, int reset);

/* Launch-shape autotuner: times fbls_signal_search of `fam` under each
 * candidate (LDS budget per workgroup 16-160 KiB, which sets the periods per
 * workgroup, x workgroup size 256/512/1024; best of `repeats` after a warm-up),
 * fails with FBLS_E_HIP if any shape changes the rows, and keeps the
 * fastest on the context (as fbls_ctx_set_launch).  table (capacity
 * table_cap, may be NULL) receives every measured shape. */
typedef struct {
  uint32_t lds_bytes, block;
  double seconds;  /* kernel seconds of one search */
} fbls_tune_entry;
int fbls_autotune(fbls_ctx* ctx, const fbls_family* fam, const double* durations,
                  const double* rho, size_t nd, const fbls_bank* bank,
                  const fbls_search_cfg* cfg, uint32_t repeats, fbls_tune_entry* table,
                  size_t table_cap, size_t* ntable, fbls_tune_entry* best);

/* hipGraph replay of the evidence call (fbls_compare: the table upload, six
 * kernels and the copies back are captured per call and updated into one
 * cached executable graph).  enable: 1 on (default), 0 off (eager launches),
 * negative leaves it unchanged; *graph_launches (may be NULL) counts graph
 * submissions on this context.  Results never depend on it.  The HIP
 * runtime fails allocations made by other threads while a stream of the
 * process is being captured (and then the capture too), so a process that
 * uses several contexts from concurrent threads turns graphs off on each. */
int fbls_ctx_set_graphs(fbls_ctx* ctx, int enable, uint64_t* graph_launches);

/* -- family upload ------------------------------------------------------ */

/* `DeviceReady::upload` (scan.rs:110-125).  t: origin-relative times >= 0;
 * v, w: precision-weighted fields of `prepare_ready` (curve.rs:427-450);
 * baseline_norm: C = sum of w.  v and w are stored as 32-bit fixed point
 * with a power-of-two scale per array (|q| <= 2^30), which makes every bin
 * and interval sum exact. */
int fbls_family_upload(fbls_ctx* ctx, const double* t, const double* v,
                       const double* w, size_t n, double baseline_norm,
                       double origin, fbls_family** out);
void fbls_family_free(fbls_family* fam);
/* Packed upload of many families (stars) into one device allocation (a
 * ragged CSR layout): family k holds observations offsets[k] ..
 * offsets[k+1] - 1 of t, v, w, with baseline_norm[k] and origin[k].  Writes
 * nfam handles to out; each is freed with fbls_family_free (the allocation
 * goes with the last).  The per-star `DeviceReady::upload` of `fbls batch`
 * (batch.rs:76, scan.rs:110-125) for a whole batch at once. */
int fbls_family_upload_packed(fbls_ctx* ctx, size_t nfam, const size_t* offsets,
                              const double* t, const double* v, const double* w,
                              const double* baseline_norm, const double* origin,
                              fbls_family** out);
/* The same upload from one array per star: family k holds n[k] observations
 * at t[k], v[k], w[k] (no packed copy needed); same handles and ownership as
 * fbls_family_upload_packed.  The per-star `DeviceReady::upload` of `fbls
 * batch` (batch.rs:76, scan.rs:110-125) for a whole batch at once. */
int fbls_family_upload_list(fbls_ctx* ctx, size_t nfam, const double* const* t,
                            const double* const* v, const double* const* w, const size_t* n,
                            const double* baseline_norm, const double* origin,
                            fbls_family** out);
/* The fixed-point values the device scans, as doubles (exact), and the
 * power-of-two exponents: v_dev = round(v * 2^ev) * 2^-ev. */
int fbls_family_fields(const fbls_family* fam, double* v_out, double* w_out,
                       int* ev, int* ew);

/* -- the scan ----------------------------------------------------------- */

/* `scan_bank` (scan.rs:144-357): best template per period at nb phase bins.
 * records: np x 9, score/depth: np, winner: np x 4 or NULL. */
int fbls_scan_bank(fbls_ctx* ctx, const fbls_family* fam,
                   const double* periods, size_t np, const fbls_bank* bank,
                   uint32_t nb, double* records, double* score, double* depth,
                   uint32_t* winner, fbls_scan_stats* stats);

/* fbls_scan_bank without any allocation: no hipMalloc, hipFree, pinned or
 * library allocation inside the call.  No counterpart in the reference; for
 * callers that cannot allocate (tight serving loops, harnesses that forbid
 * allocation).  Same arguments and results as fbls_scan_bank (bit-identical
 * records, scores, depths, winners and statistics); the family is uploaded
 * beforehand.  The launch table, periods, scales, counters and staged
 * outputs live in a fixed 8 MiB device array of the library on each device
 * (held until the process ends), which limits a call to np <= 65536
 * periods (larger np: FBLS_E_ARG).  Every call on one device shares that
 * array and keeps the last call's contents; calls from several threads of
 * one process on one device are serialized by the library, since calls that
 * overlapped on the array would race.  Synchronous on return. */
int fbls_scan_bank_noalloc(fbls_ctx* ctx, const fbls_family* fam,
                           const double* periods, size_t np, const fbls_bank* bank,
                           uint32_t nb, double* records, double* score, double* depth,
                           uint32_t* winner, fbls_scan_stats* stats);

/* `scan_periods` (scan.rs:363-408): coarse prescan then full scan of the
 * survivors; rows (capacity np) receive periods whose score >= threshold. */
int fbls_scan_periods(fbls_ctx* ctx, const fbls_family* fam,
                      const double* periods, size_t np, const fbls_bank* bank,
                      fbls_row* rows, size_t* nrows, fbls_scan_stats* stats);

/* `signal_search` (scan.rs:413-473) for one family: period grid, octave
 * bands, durations below a quarter of each band.  durations/rho may be
 * unsorted.  rows capacity rows_cap (>= grid count); stats capacity
 * stats_cap bands.  grid may be NULL; else capacity grid_cap. */
int fbls_signal_search(fbls_ctx* ctx, const fbls_family* fam,
                       const double* durations, const double* rho, size_t nd,
                       const fbls_bank* bank, const fbls_search_cfg* cfg,
                       fbls_row* rows, size_t rows_cap, size_t* nrows,
                       fbls_scan_stats* stats, size_t stats_cap,
                       size_t* nbands, double* grid, size_t grid_cap,
                       size_t* grid_count);

/* `signal_search` of many stars (one family each) as `fbls batch` runs
 * them (batch.rs:61-106), with the work of all stars merged: the coarse
 * passes of every band of up to 2^21 grid periods' worth of stars share one
 * launch, their fine passes a second.  durations: the nd bank durations
 * (any order, shared); rho: nfam x nd factors, in the order of durations.
 * rows: the rows of star k are rows[row_offsets[k] .. row_offsets[k+1]);
 * row_offsets holds nfam + 1 entries and is filled even when rows_cap is
 * too small (then FBLS_E_ARG; rows_cap >= the sum of the stars' grid counts
 * always suffices).  star_stats (nfam, may be NULL): the band
 * statistics summed per star (phase_bins = the largest).  keep (may be
 * NULL): one flag per grid period of every star, concatenated in star order
 * (each star's grid as fbls_period_grid); only flagged periods are scanned,
 * each exactly as in the full search (same bands, nb and prescan), so the
 * rows are the full search's rows at those periods. */
int fbls_signal_search_batch(fbls_ctx* ctx, const fbls_family* const* fams,
                             size_t nfam, const double* durations,
                             const double* rho, size_t nd,
                             const fbls_bank* bank, const fbls_search_cfg* cfg,
                             fbls_row* rows, size_t rows_cap,
                             size_t* row_offsets, fbls_scan_stats* star_stats,
                             const uint8_t* keep);

/* fbls_signal_search_batch restricted by period bins instead of a
 * per-period mask: bin k covers the periods edges[k] <= P < edges[k + 1]
 * (edges ascending, nbins + 1 values) and star s scans only the grid
 * periods of its bins with bin_keep[s * nbins + k] set, each exactly as in
 * the full search.  Only those periods are evaluated (the bands come from a
 * binary search on the grid), so the host cost follows the kept periods.
 * No counterpart in the reference: the learned proposal's pruned search. */
int fbls_signal_search_bins(fbls_ctx* ctx, const fbls_family* const* fams, size_t nfam,
                            const double* durations, const double* rho, size_t nd,
                            const fbls_bank* bank, const fbls_search_cfg* cfg,
                            fbls_row* rows, size_t rows_cap, size_t* row_offsets,
                            fbls_scan_stats* star_stats, const uint8_t* bin_keep,
                            size_t nbins, const double* edges);

/* fbls_signal_search_bins with bins of its own for every star (stars of
 * different baselines have coarse periodograms of different resolution):
 * star s has the bins [bin_offsets[s], bin_offsets[s + 1]) of bin_keep
 * (bin_offsets: nfam + 1 values from 0) and their bin_offsets[s + 1] -
 * bin_offsets[s] + 1 ascending edges at edges + bin_offsets[s] + s.  One
 * call scans every star in the same merged launches.  No counterpart in the
 * reference: the learned proposal's pruned search of long-baseline stars. */
int fbls_signal_search_bins_ragged(fbls_ctx* ctx, const fbls_family* const* fams, size_t nfam,
                                   const double* durations, const double* rho, size_t nd,
                                   const fbls_bank* bank, const fbls_search_cfg* cfg,
                                   fbls_row* rows, size_t rows_cap, size_t* row_offsets,
                                   fbls_scan_stats* star_stats, const uint8_t* bin_keep,
                                   const size_t* bin_offsets, const double* edges);

/* -- evidence (model comparison) ---------------------------------------- */

/* SearchConfig fields read by `compare_search` (config.rs:33-42,53-85).
 * stellar_radius_solar <= 0 means None. */
typedef struct {
  const double* ingress;      /* ns ingress fractions (shapes) */
  size_t ns;
  uint32_t layers;            /* staircase layers */
  uint32_t phase_oversample;
  double threshold;           /* SNR threshold */
  double evidence_rtol;       /* pruning tolerance, 0 = exhaustive */
  double planet_prior, binary_prior;
  double planet_depth_scale, binary_depth_scale, diluted_binary_depth_scale;
  double stellar_radius_solar, planet_radius_scale_rj, target_flux_fraction;
} fbls_evidence_cfg;

/* `EvidenceRow` (evidence.rs:637-653) plus the decoded winner: phase bin,
 * index into the period's sorted durations below P/4, shape index. */
typedef struct {
  double period, epoch, duration, snr, depth, planet_log_odds,
      log_bf_planet_noise, log_bf_binary_noise, binary_secondary_ratio,
      binary_secondary_phase, binary_odd_even_ratio, binary_period_days,
      ingress_fraction;
  uint32_t selection_flags; /* 1 joint winner, 2 SNR winner, 3 both */
  uint32_t phase, duration_index, shape;
} fbls_evidence_row;

/* Intermediate device arrays of one comparison, for inspection (every
 * pointer may be NULL; capacities in doubles).  screen: tiles x 4 after the
 * amplitude screen; parts: tiles x 2 x 15 after the joint pass (untouched
 * when the screen is below threshold); floors: {floor_p, floor_b, pilot
 * key, passed}.  tiles = ceil(nb / 64). */
typedef struct {
  double* screen;
  size_t screen_cap;
  double* parts;
  size_t parts_cap;
  double floors[4];
  size_t tiles, nb;
} fbls_evidence_debug;

/* `compare_search` (evidence.rs:693-1003): joint planet / binary / noise
 * evidence of one family at one trial period.  durations/rho: the family's
 * bank (any order; those below period/4 are used, sorted).
 * residual_energy: E of `prepare_ready` (curve.rs:440).  rows receives up
 * to 2 rows (*nrows): the joint-evidence winner and the SNR winner, merged
 * into one with flags 3 when identical; none when no template reaches the
 * threshold.  debug may be NULL. */
int fbls_compare(fbls_ctx* ctx, const fbls_family* fam, double period,
                 const double* durations, const double* rho, size_t nd,
                 double residual_energy, const fbls_evidence_cfg* cfg,
                 fbls_evidence_row rows[2], size_t* nrows,
                 fbls_evidence_debug* debug);

/* [host] `parameters` (evidence.rs:665-689): the 30 model constants. */
int fbls_evidence_parameters(const fbls_evidence_cfg* cfg, double out[30]);

/* -- multi-pass search ---------------------------------------------------- */

/* `Peak` (src/select/peaks.rs:19-45).  Options that are None are NaN. */
typedef struct {
  double period_days, epoch_bjd, duration_hours, snr, snr_folded,
      snr_without_strongest_event, depth, depth_ppm;
  uint64_t events, family, pass, row;
  double planet_log_odds, log_bf_planet_noise, log_bf_binary_noise,
      binary_secondary_ratio, binary_secondary_phase, binary_odd_even_ratio,
      binary_period_days, ingress_fraction, evidence_epoch_bjd,
      evidence_duration_hours;
  uint32_t model_preferred, binary_diagnostics_complete;
} fbls_peak;

/* One pass of the search (the `passes[]` statistics of pipeline.rs:143). */
typedef struct {
  uint64_t pass, rows, periods, accepted;
  double seconds_scan, seconds_select;
} fbls_pass_stats;

/* SearchConfig fields of the prepared search (config.rs:5-88). */
typedef struct {
  const double* durations;   /* days, sorted unique (`durations_days`) */
  size_t nd;
  int correlated;            /* correlated_noise */
  fbls_bank bank;            /* shapes, layers, threshold (also the selection
                                threshold), gate, prescan, budget; its
                                durations/rho are ignored */
  fbls_search_cfg search;
  size_t max_candidates, passes, min_events, evidence_candidates;
  int model_comparison;
  double planet_log_odds_min;
  fbls_evidence_cfg evidence;
  int single_events;         /* run the single-event scan (pipeline.rs:225-238) */
} fbls_pipeline_cfg;

/* `EventStats` (events.rs:11-16) of one family, plus its bin count. */
typedef struct {
  uint64_t templates_scanned, threshold_survivors;
  uint32_t overflow, phase_bins;
} fbls_event_stats;

/* `SingleEvent` (src/select/peaks.rs:47-55). */
typedef struct {
  double epoch_bjd, duration_hours, snr, depth, depth_ppm;
  uint64_t points;
} fbls_single_event;

/* One family in one pass (`statistics.passes[].families[]`, pipeline.rs:
 * 112-118).  window_days is NaN for prepared input; durations_hours and
 * rho list the first min(nd, 64) durations of the family. */
typedef struct {
  uint64_t pass, family, observations, clipped, segments, periods, nd;
  double window_days, error_inflation;
  double durations_hours[64], rho[64];
  uint64_t band_first, nbands;  /* the family's bands in fbls_search_report.bands */
} fbls_family_pass;

/* Everything the schema-3 report holds beyond peaks and pass totals
 * (pipeline.rs:262-288).  The caller provides the arrays (NULL and 0 skip
 * one); the call sets every count to the number available, writes at most
 * the capacity, and fills the scalars. */
typedef struct {
  fbls_single_event* events;      size_t events_cap;      size_t nevents;
  fbls_event_stats* event_stats;  size_t event_stats_cap; size_t nevent_stats;
  fbls_row* period_table;         size_t period_table_cap; size_t period_table_rows;
  uint64_t* period_table_family;  /* period_table_cap entries, or NULL */
  fbls_family_pass* families;     size_t families_cap;    size_t nfamilies;
  fbls_scan_stats* bands;         size_t bands_cap;       size_t nbands;
  uint64_t observations, periods_tested, nfam;
  double baseline_days, seconds_preparation, seconds_search, seconds_scan,
      seconds_select, seconds_evidence, seconds_single_events;
} fbls_search_report;

/* `search_prepared` + `search_families` (pipeline.rs:41-238): keeps finite
 * points with positive errors sorted by time, builds the family state, then
 * per pass runs the signal search, the greedy peak selection with GPU
 * rescoring (peaks.rs:308-487) and the masking of the lead peak; then the
 * model comparison of the evidence_candidates strongest peaks and, when
 * cfg->single_events, the single-event scan of every family with the
 * greedy event selection (peaks.rs:490-550).  Peaks in acceptance order
 * (capacity peaks_cap, *npeaks is the total); pass statistics capacity
 * passes_cap; report (may be NULL) receives the single events and the
 * statistics of the schema-3 report. */
int fbls_search_prepared(fbls_ctx* ctx, const double* time, const double* flux,
                         const double* err, size_t n, const fbls_pipeline_cfg* cfg,
                         fbls_peak* peaks, size_t peaks_cap, size_t* npeaks,
                         fbls_pass_stats* passes, size_t passes_cap, size_t* npasses,
                         fbls_search_report* report);

/* `single_event_search` (events.rs:18-153, kernels.rs:849-930): every
 * template at the trial period baseline + 2 max(D) + 1e-6 (nb = max(8,
 * ceil(phase_oversample x trial / min(D))) bins) whose score reaches the
 * threshold, as rows with period +inf, ordered by (phase bin, duration in
 * ascending order, shape).  The family's times must ascend.  More than
 * `capacity` survivors set stats->overflow and return no rows. */
int fbls_single_event_search(fbls_ctx* ctx, const fbls_family* fam, const double* durations,
                             const double* rho, size_t nd, const double* ingress, size_t ns,
                             uint32_t layers, double threshold, uint32_t phase_oversample,
                             size_t capacity, fbls_row* rows, size_t* nrows,
                             fbls_event_stats* stats);

/* -- learned band proposal --------------------------------------------- */

/* A 1-D convolutional network over a coarse periodogram (L period bins x
 * in_channels duration groups): `layers` Conv1d layers of `width` channels
 * and odd `kernel`, GELU, then a per-bin sigmoid head; trained with a
 * binary cross-entropy (positive weight pos_weight) on labels from the full
 * GPU scan.  No counterpart in the reference: it proposes which period bins
 * the full-resolution scan visits.  BF16 MFMA GEMMs with FP32 accumulation,
 * FP32 master weights, AdamW (beta 0.9 / 0.999, eps 1e-8, decoupled weight
 * decay on weights only).  Tensors are channels-last: x[b][l][c]. */
typedef struct fbls_cnn fbls_cnn;
typedef struct {
  uint32_t in_channels, width, layers, kernel;
  float pos_weight;
} fbls_cnn_cfg;
int fbls_cnn_create(fbls_ctx* ctx, const fbls_cnn_cfg* cfg, uint64_t seed, fbls_cnn** out);
void fbls_cnn_destroy(fbls_cnn* net);
/* Flat FP32 parameters: per layer W[cout][k][cin] then b[cout]; then the
 * head w[width], b. */
size_t fbls_cnn_param_count(const fbls_cnn* net);
int fbls_cnn_get_params(fbls_cnn* net, float* out);
/* Sets the parameters and resets the optimiser state. */
int fbls_cnn_set_params(fbls_cnn* net, const float* in);
/* logits[b][l] of x[b][l][c] */
int fbls_cnn_forward(fbls_cnn* net, const float* x, uint32_t B, uint32_t L, float* logits);

/* Inference path of fbls_cnn_forward: 1 (default) runs each layer as one
 * fused implicit-GEMM conv + bias + GELU kernel when the shapes allow it
 * (64 channels, 8 or 64 inputs), 0 the im2col + GEMM + epilogue kernels of
 * training.  Both give bit-identical logits. */
int fbls_cnn_set_fused(fbls_cnn* net, int enable);
/* Mean loss over the B x L bins and its gradient (flat, may be NULL). */
int fbls_cnn_loss_grad(fbls_cnn* net, const float* x, const float* y, uint32_t B, uint32_t L,
                       float* loss, float* grads);
/* One AdamW step on the batch; *loss is the loss before the step. */
int fbls_cnn_train_step(fbls_cnn* net, const float* x, const float* y, uint32_t B, uint32_t L,
                        float lr, float weight_decay, float* loss);

/* Coarse periodogram (the network's input): for each star k and period i,
 * out[k][i][g] = the box SNR a / sqrt(w - w^2 / C) at the phase with the
 * largest a |a| / (w - w^2 / C) (the first such phase; this is the largest
 * SNR over phase) for a transit of duration durations[g] at nb_i =
 * clamp(ceil(2 P_i / d_0), nb_min, nb_max) phase bins (d_0 the shortest
 * duration; windows with at least 3 points and at most n - 3; -inf when
 * none).  At most 16 durations.  One launch. */
int fbls_coarse_features(fbls_ctx* ctx, const fbls_family* const* fams, size_t nfam,
                         const double* periods, size_t np, const double* durations,
                         size_t ng, uint32_t nb_min, uint32_t nb_max, double* out);

/* The network input of the same periodogram, computed on the device:
 * clip(SNR, -5, 40) / 8 as float32, 0 where no phase qualifies. */
int fbls_coarse_inputs(fbls_ctx* ctx, const fbls_family* const* fams, size_t nfam,
                       const double* periods, size_t np, const double* durations,
                       size_t ng, uint32_t nb_min, uint32_t nb_max, float* out);

/* The proposal in one call: the network inputs of fbls_coarse_inputs are
 * computed into the network's device buffer and the fused inference runs
 * on them (no copy of the inputs to the host and back); logits[k][i] for
 * star k and period bin i.  net must belong to ctx and read ng durations;
 * stars go through in batches of about 2^22 bins.  peaks (may be NULL):
 * peaks[k][i], the periodogram's largest SNR over the durations at bin i
 * (unclipped; -inf where no phase qualifies), from the same kernel. */
int fbls_propose(fbls_ctx* ctx, fbls_cnn* net, const fbls_family* const* fams, size_t nfam,
                 const double* periods, size_t np, const double* durations, size_t ng,
                 uint32_t nb_min, uint32_t nb_max, float* logits, float* peaks);

/* -- raw-mode preparation ----------------------------------------------- */

/* Preparation options of the raw path (config.rs:11-16, defaults 10, 0.4,
 * 3, 3.5, 4, 0.5). */
typedef struct {
  double bin_minutes, window_min, window_factor, window_max, clip_sigma, edge_trim;
} fbls_prep_cfg;

/* `biweight_flatten_gpu` (curve.rs:256-329, kernel kernels.rs:1049-1110):
 * Tukey biweight location in a sliding window of `window` days per
 * gap-separated segment of the sorted times, on the GPU; out[i] =
 * (1 + rel[i]) / location_i - 1.  Exact lower medians; windows of any size
 * (the reference falls back to its host loop beyond 4096 points; here the
 * kernel reads such windows from global memory). */
int fbls_biweight_flatten(fbls_ctx* ctx, const double* time, const double* rel, size_t n,
                          double window, double gap, double c, uint32_t iterations,
                          double* out);

/* `prepare_families` (curve.rs:387-425): per-group normalisation and
 * inverse-variance binning (fit_segment_noise, curve.rs:89-198), one
 * biweight-detrended family per detrending window, positive outliers
 * (flux > clip_sigma x MAD scale) and points within edge_trim windows of a
 * segment edge dropped.  err and group may be NULL.  Family k (< *nfam <=
 * nd): fam_n[k] points at fam_time/fam_flux/fam_sigma[k * n ..] (absolute
 * time), its window, the clipped count, and fam_members[k * nd + i] = 1 for
 * the durations it searches. */
int fbls_prepare_families(fbls_ctx* ctx, const double* time, const double* flux,
                          const double* err, const int64_t* group, size_t n,
                          const double* durations, size_t nd, const fbls_prep_cfg* cfg,
                          size_t* nfam, size_t* fam_n, double* fam_time, double* fam_flux,
                          double* fam_sigma, double* fam_window, size_t* fam_clipped,
                          uint8_t* fam_members);

/* `search_transits` (pipeline.rs:27-37, 70-221, without single events):
 * fbls_prepare_families on the raw flux, then the multi-pass search of
 * fbls_search_prepared over all families together (one signal search per
 * family and pass, tables concatenated, peaks selected across families,
 * every family masked with the pass's lead peak).  fbls_peak.family is the
 * family index of fbls_prepare_families.  err and group may be NULL. */
int fbls_search_transits(fbls_ctx* ctx, const double* time, const double* flux,
                         const double* err, const int64_t* group, size_t n,
                         const fbls_prep_cfg* prep, const fbls_pipeline_cfg* cfg,
                         fbls_peak* peaks, size_t peaks_cap, size_t* npeaks,
                         fbls_pass_stats* passes, size_t passes_cap, size_t* npasses,
                         fbls_search_report* report);

/* -- host helpers [host] ------------------------------------------------ */

/* [host] `period_grid` (grid.rs:5-27): count first, then fill. */
size_t fbls_period_grid_count(double minimum, double maximum, double baseline,
                              double minimum_duration, double oversample);
int fbls_period_grid(double minimum, double maximum, double baseline,
                     double minimum_duration, double oversample, double* out,
                     size_t count);

/* [host] Octave bands of one family (scan.rs:440-471, 377-381): for each
 * band, its first grid index and period count, the durations used (the
 * first nd of the sorted durations, those below lo/4), the phase bins nb of
 * the full pass and nb_coarse of the prescan (0 when there is none). */
typedef struct {
  double lo;
  size_t first, count, nd;
  uint32_t nb, nb_coarse;
} fbls_band;
int fbls_plan_bands(const double* grid, size_t count,
                    const double* sorted_durations, size_t nd,
                    uint32_t phase_oversample, uint32_t prescan_factor,
                    fbls_band* bands, size_t bands_cap, size_t* nbands);

/* [host] batches of one scan_bank call under the fold budget
 * (scan.rs:177-183); 0 when one period exceeds the budget. */
uint64_t fbls_bank_batches(size_t np, uint32_t nb, uint32_t period_batch,
                           double max_fold_mib);

/* [host] `SearchConfig::durations_days` (config.rs:177-182) */
size_t fbls_durations_days(const double* hours, size_t n, double* out_days);

/* [host] `family_state` (curve.rs:541-569) for a prepared curve:
 * fit_segment_noise without binning, prepare_ready and, when correlated is
 * nonzero, correlated_noise_factors.  Outputs sized n (t, v, w) and nd
 * (rho); *out_n receives the kept count; residual_energy (may be NULL) is
 * E = sum y v of prepare_ready (curve.rs:440). */
int fbls_family_state(const double* time, const double* flux,
                      const double* sigma, size_t n, const double* durations,
                      size_t nd, int correlated, double* t, double* v,
                      double* w, size_t* out_n, double* baseline_norm,
                      double* origin, double* rho, double* inflation,
                      size_t* segments, double* residual_energy);

/* [host] fbls_family_state for `count` stars at once, on the library's
 * preparation threads (a pool of their own at a lower CPU priority, four
 * threads short of the cores, and the calling thread), as the
 * reference's `batch` prepares its stars on worker threads before the
 * search (src/batch.rs:47-68, curve.rs:541-569 per star).  Star k reads
 * time[k], flux[k], sigma[k] of n[k] points and writes t[k], v[k], w[k]
 * (sized n[k]), rho + k nd and element k of every other output, exactly as
 * fbls_family_state of that star would; flux_offset is added to every flux
 * value first (-1 for a curve normalised to 1, rounded as the caller's own
 * flux - 1.0), 0 for flux already relative to zero.  residual_energy and
 * status (each star's code) may be NULL.  Returns FBLS_OK, or the code of
 * the first star that failed, with a message that names it; the other
 * stars are prepared all the same. */
int fbls_family_state_many(size_t count, const double* const* time,
                           const double* const* flux, const double* const* sigma,
                           const size_t* n, double flux_offset, const double* durations,
                           size_t nd, int correlated, double* const* t, double* const* v,
                           double* const* w, size_t* out_n, double* baseline_norm,
                           double* origin, double* rho, double* inflation,
                           size_t* segments, double* residual_energy, int* status);

/* [device] fbls_family_state_many and the upload of the prepared families
 * in one call: each preparation thread quantises its star and writes the
 * star's device image into pinned staging as soon as the star is prepared,
 * then the images are copied into one device allocation on the context's
 * upload stream (the work of fbls_family_state_many followed by
 * fbls_family_upload_list, without handing the prepared arrays back).
 * Outputs as fbls_family_state_many (no t, v, w; t_last, which may be
 * NULL, receives each star's last origin-relative time) plus one handle
 * per star in out; every handle or none is written.  Returns FBLS_OK or the code of
 * the first star that failed, with a message that names it. */
int fbls_family_prepare_upload_many(fbls_ctx* ctx, size_t count, const double* const* time,
                                    const double* const* flux, const double* const* sigma,
                                    const size_t* n, double flux_offset,
                                    const double* durations, size_t nd, int correlated,
                                    size_t* out_n, double* baseline_norm, double* origin,
                                    double* rho, double* inflation, size_t* segments,
                                    double* residual_energy, double* t_last, int* status,
                                    fbls_family** out);

/* [host] fixed-point representation used on the device; returns e, the
 * words standing for integers in units of 2^-e.  *block (may be NULL) = 0:
 * q = rint(x * 2^e), |q| <= 2^30 (max |x| <= 16 mean |x|); 1: block mode,
 * q = mantissa * 32 + shift for the integer mantissa * 2^shift (27-bit
 * signed mantissa, 5-bit shift).  Either way each value's error is at most
 * 2^-26 max(|x|, mean |x|) and the integers' absolute sum is below 2^53. */
int fbls_quantize(const double* x, size_t n, int32_t* q, int* block);

/* [host] seeded synthetic light curve (xorshift64 + Box-Muller as in
 * survey/numeric.rs:345-364).  See fbls_amd/synth.py for the workloads. */
typedef struct {
  uint64_t seed;
  uint32_t sectors;        /* contiguous 27.4-d sectors */
  double sector_days;      /* 27.4 */
  double gap_days;         /* mid-sector gap (1.0) */
  double cadence_days;     /* 2 min = 1/720 */
  double t0;               /* absolute start time */
  double sigma;            /* total white+red noise (relative flux) */
  double red_fraction;     /* fraction of variance in AR(1), 0 = white */
  double ar_phi;           /* AR(1) coefficient */
  double period, epoch, depth, duration; /* injected box; depth 0 = none */
  double secondary_depth;  /* eclipse at phase 0.5, 0 = none */
  double baseline_gap_days; /* extra gap inserted after half the sectors */
} fbls_synth_cfg;
/* Writes up to cap points; returns the count through *n. */
int fbls_synth_star(const fbls_synth_cfg* cfg, double* time, double* flux,
                    double* flux_err, size_t cap, size_t* n);

/* -- formats and the command-line host [host] ----------------------------- */

/* [host] `read_inputs` (main.rs:50-132) over NPZ files (npz.rs:17-113):
 * time_bjd and flux required, flux_err in all files or none, sector a
 * scalar or per observation; group numbers the (file, sector) pairs in
 * order of first appearance.  Writes min(cap, total) values of each array
 * (NULL skips one) and the total through *n; FITS input is not supported. */
int fbls_read_inputs(const char* const* paths, size_t npaths, double* time, double* flux,
                     double* err, int64_t* group, size_t cap, size_t* n, int* has_err);

/* [host] The shortest text that reads back to x, as the schema-3 report
 * prints doubles ("0.0", "1e-12", "null" when not finite).  Returns its
 * length; writes at most cap - 1 characters and a terminator. */
size_t fbls_format_f64(double x, char* buf, size_t cap);

/* [host] JSON text re-printed in the report's layout (keys sorted, two-space
 * indent); *len receives the full length. */
int fbls_json_reformat(const char* text, char* out, size_t cap, size_t* len);

/* The `fbls` command line (main.rs:137-246, batch.rs:36-109): argv[0] is
 * "search" (--input FILES [--output F] [--save-table F] [--prepared]
 * [--device N] [--pmin D] [--pmax D] [--prescan F] [--threshold S]) or
 * "batch" (--list FILE --output-dir DIR [--workers K] [--devices 0,1]
 * [--pmax D]).  Returns 0, or 1 after printing "fbls: message". */
int fbls_cli(int argc, const char* const* argv);

/* Thread-local message of the last failing call on this thread. */
const char* fbls_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

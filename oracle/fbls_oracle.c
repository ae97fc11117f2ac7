/*
 * FP64 CPU oracle of the FBLS per-family period scan.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP path, never the product.
 * Citations are path:line in the reference repository (isakruas/fbls).
 * Build with -ffp-contract=off (see oracle/Makefile).
 */
#include "fbls_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define TILE 128u

/* ---------------------------------------------------------------- config */

static int cmp_double(const void* a, const void* b) {
  double x = *(const double*)a, y = *(const double*)b;
  return (x > y) - (x < y);
}

/* config.rs:177-182: hours / 24, sorted, consecutive duplicates removed. */
size_t orc_durations_days(const double* hours, size_t n, double* out) {
  for (size_t i = 0; i < n; i++) out[i] = hours[i] / 24.0;
  qsort(out, n, sizeof(double), cmp_double);
  size_t m = 0;
  for (size_t i = 0; i < n; i++)
    if (m == 0 || out[i] != out[m - 1]) out[m++] = out[i];
  return m;
}

/* ------------------------------------------------------------------ grid */

/* grid.rs:12-13 */
size_t orc_period_grid_count(double minimum, double maximum, double baseline,
                             double minimum_duration, double oversample) {
  double log_step = log1p(minimum_duration / (oversample * baseline));
  return (size_t)ceil(log(maximum / minimum) / log_step) + 1;
}

/* grid.rs:14-27 */
/* Worker threads of the oracle's parallel loops (0: the OpenMP default). */
void orc_set_threads(int n) {
    omp_set_num_threads(n > 0 ? n : omp_get_num_procs());
}

void orc_period_grid(double minimum, double maximum, double baseline,
                     double minimum_duration, double oversample, double* out,
                     size_t count) {
  (void)baseline;
  (void)minimum_duration;
  (void)oversample;
  double a = log(minimum), b = log(maximum);
  for (size_t i = 0; i < count; i++) {
    double x = count == 1 ? a : a + (b - a) * (double)i / (double)(count - 1);
    out[i] = exp(x);
  }
  out[0] = minimum;
  out[count - 1] = maximum;
}

/* ----------------------------------------------------------------- stats */

/* stats.rs:4-11: finite values, sorted, element (len - 1) / 2. */
double orc_median_lower(const double* values, size_t n) {
  double* v = (double*)malloc((n ? n : 1) * sizeof(double));
  size_t m = 0;
  for (size_t i = 0; i < n; i++)
    if (isfinite(values[i])) v[m++] = values[i];
  double r = NAN;
  if (m) {
    qsort(v, m, sizeof(double), cmp_double);
    r = v[(m - 1) / 2];
  }
  free(v);
  return r;
}

/* stats.rs:14-18 */
double orc_mad_scale(const double* values, size_t n) {
  double m = orc_median_lower(values, n);
  double* dev = (double*)malloc((n ? n : 1) * sizeof(double));
  for (size_t i = 0; i < n; i++) dev[i] = fabs(values[i] - m);
  double r = 1.4826 * orc_median_lower(dev, n);
  free(dev);
  return r;
}

/* stable argsort by value (stats.rs:20-28 uses a stable sort) */
typedef struct {
  double x;
  size_t i;
} keyed;

static int cmp_keyed(const void* a, const void* b) {
  const keyed* p = (const keyed*)a;
  const keyed* q = (const keyed*)b;
  if (p->x < q->x) return -1;
  if (p->x > q->x) return 1;
  return (p->i > q->i) - (p->i < q->i);
}

/* ------------------------------------------------------------ noise model */

/* curve.rs:89-198 with bin_minutes = 0 and a single group (group = None). */
int orc_fit_segment_noise(const double* time, const double* flux,
                          const double* err, size_t n, double gap_days,
                          double* out_t, double* out_def, double* out_var,
                          size_t* out_n, double* origin, double* inflation,
                          size_t* segments) {
  keyed* order = (keyed*)malloc((n ? n : 1) * sizeof(keyed));
  size_t m = 0;
  for (size_t i = 0; i < n; i++) {
    double e = err ? err[i] : 1.0;
    if (isfinite(time[i]) && isfinite(flux[i]) && isfinite(e) && e > 0.0) {
      order[m].x = time[i];
      order[m].i = i;
      m++;
    }
  }
  qsort(order, m, sizeof(keyed), cmp_keyed);
  if (m < 6) {
    free(order);
    return 6; /* "insufficient usable observations" */
  }
  double* t = out_t;
  double* f = (double*)malloc(m * sizeof(double));
  double* e = (double*)malloc(m * sizeof(double));
  for (size_t k = 0; k < m; k++) {
    size_t i = order[k].i;
    t[k] = time[i];
    f[k] = flux[i];
    e[k] = err ? err[i] : 1.0;
  }
  free(order);
  double med = orc_median_lower(f, m); /* curve.rs:120-131, one label */
  if (!(med > 0.0)) {
    free(f);
    free(e);
    return 7; /* "mission flux must have a positive median" */
  }
  for (size_t k = 0; k < m; k++) {
    f[k] = f[k] / med - 1.0;
    e[k] /= med;
  }
  double org = t[0];
  for (size_t k = 0; k < m; k++) t[k] -= org;
  /* curve.rs:164-177 */
  double* deltas = (double*)malloc(m * sizeof(double));
  double* standardized = (double*)malloc(m * sizeof(double));
  size_t nd = 0;
  for (size_t i = 1; i < m; i++) {
    if (t[i] - t[i - 1] < gap_days) {
      double d = f[i] - f[i - 1];
      deltas[nd] = d;
      standardized[nd] = d / sqrt(e[i] * e[i] + e[i - 1] * e[i - 1]);
      nd++;
    }
  }
  if (nd == 0) {
    free(f);
    free(e);
    free(deltas);
    free(standardized);
    return 8; /* "no adjacent observations available for noise estimation" */
  }
  double infl = orc_mad_scale(standardized, nd);
  if (!(infl >= 1.0)) infl = 1.0; /* f64::max(1.0); NaN gives 1.0 */
  if (err) {
    for (size_t k = 0; k < m; k++) {
      double s = e[k] * infl;
      out_var[k] = s * s;
    }
  } else {
    double md = orc_median_lower(deltas, nd);
    for (size_t k = 0; k < nd; k++) deltas[k] = fabs(deltas[k] - md);
    double white = 1.4826 / sqrt(2.0) * orc_median_lower(deltas, nd);
    if (!(white >= 1e-12)) white = 1e-12;
    for (size_t k = 0; k < m; k++) out_var[k] = white * white;
  }
  size_t seg = 1;
  for (size_t i = 1; i < m; i++)
    if (t[i] - t[i - 1] > gap_days) seg++;
  for (size_t k = 0; k < m; k++) out_def[k] = -f[k];
  *out_n = m;
  *origin = org;
  *inflation = infl;
  *segments = seg;
  free(f);
  free(e);
  free(deltas);
  free(standardized);
  return 0;
}

/* curve.rs:427-450 */
int orc_prepare_ready(const double* deficit, const double* variance, size_t n,
                      double* v, double* w, double* baseline_norm,
                      double* residual_energy) {
  if (n < 6) return 6;
  double c = 0.0;
  for (size_t i = 0; i < n; i++) {
    w[i] = 1.0 / variance[i];
    c += w[i];
  }
  if (!(isfinite(c) && c > 0.0)) return 9; /* "invalid baseline precision" */
  double spy = 0.0;
  for (size_t i = 0; i < n; i++) spy += w[i] * deficit[i];
  double mean = spy / c;
  double energy = 0.0;
  for (size_t i = 0; i < n; i++) {
    double py = w[i] * deficit[i];
    v[i] = py - w[i] * mean;
  }
  for (size_t i = 0; i < n; i++) energy += deficit[i] * v[i];
  *baseline_norm = c;
  *residual_energy = energy;
  return 0;
}

/* curve.rs:456-539 */
void orc_correlated_noise_factors(const double* time, const double* residual,
                                  const double* sigma, size_t n,
                                  const double* durations, size_t nd,
                                  double* rho) {
  const double min_fill = 0.5;
  const size_t minimum_bins = 20;
  double* weights = (double*)malloc((n ? n : 1) * sizeof(double));
  for (size_t i = 0; i < n; i++) weights[i] = 1.0 / (sigma[i] * sigma[i]);
  double* steps = (double*)malloc((n ? n : 1) * sizeof(double));
  size_t ns = 0;
  for (size_t i = 1; i < n; i++) {
    double s = time[i] - time[i - 1];
    if (s > 0.0) steps[ns++] = s;
  }
  double cadence = orc_median_lower(steps, ns);
  free(steps);
  double* measured = (double*)malloc(nd * sizeof(double));
  size_t* counts = (size_t*)malloc(nd * sizeof(size_t));
  double* z = (double*)malloc((n ? n : 1) * sizeof(double));
  for (size_t di = 0; di < nd; di++) {
    double d = durations[di];
    size_t nz = 0;
    double total = 0.0, acc = 0.0;
    size_t cnt = 0;
    int64_t current = INT64_MIN;
    for (size_t i = 0; i < n; i++) {
      int64_t cell = (int64_t)floor((time[i] - time[0]) / d);
      if (cell != current) {
        if (cnt > 0 && (double)cnt >= min_fill * d / cadence)
          z[nz++] = (acc / total) * sqrt(total);
        current = cell;
        total = 0.0;
        acc = 0.0;
        cnt = 0;
      }
      total += weights[i];
      acc += weights[i] * residual[i];
      cnt++;
    }
    if (cnt > 0 && (double)cnt >= min_fill * d / cadence)
      z[nz++] = (acc / total) * sqrt(total);
    counts[di] = nz;
    if (nz >= 3) {
      double s = orc_mad_scale(z, nz);
      measured[di] = s >= 1.0 ? s : 1.0;
    } else {
      measured[di] = 1.0;
    }
  }
  free(z);
  int any = 0;
  for (size_t di = 0; di < nd; di++) any |= counts[di] >= minimum_bins;
  double* inherited = (double*)malloc(nd * sizeof(double));
  if (any) {
    for (size_t di = 0; di < nd; di++) {
      double d = durations[di];
      size_t anchor = (size_t)-1;
      double best = 0.0;
      for (size_t a = 0; a < nd; a++) {
        if (counts[a] < minimum_bins) continue;
        double dist = fabs(log(d) - log(durations[a]));
        if (anchor == (size_t)-1 || dist < best) {
          anchor = a;
          best = dist;
        }
      }
      double ratio = d / durations[anchor];
      inherited[di] = measured[anchor] * sqrt(ratio >= 1.0 ? ratio : 1.0);
    }
  } else {
    double* st = (double*)malloc((n ? n : 1) * sizeof(double));
    for (size_t i = 0; i < n; i++) st[i] = residual[i] / sigma[i];
    double scatter = orc_mad_scale(st, n);
    if (!(scatter >= 1.0)) scatter = 1.0;
    free(st);
    for (size_t di = 0; di < nd; di++) {
      double ratio = durations[di] / cadence;
      inherited[di] = scatter * sqrt(ratio >= 1.0 ? ratio : 1.0);
    }
  }
  for (size_t di = 0; di < nd; di++) {
    if (counts[di] >= minimum_bins)
      rho[di] = measured[di];
    else
      rho[di] = measured[di] >= inherited[di] ? measured[di] : inherited[di];
  }
  free(inherited);
  free(measured);
  free(counts);
  free(weights);
}

/* ------------------------------------------------------------ quantizer */

int orc_quantize(const double* x, size_t n, int32_t* q, int* block) {
  double mx = 0.0, sum = 0.0;
  for (size_t i = 0; i < n; i++) {
    double a = fabs(x[i]);
    if (a > mx) mx = a;
    sum += a;
  }
  *block = 0;
  if (!(mx > 0.0)) {
    for (size_t i = 0; i < n; i++) q[i] = 0;
    return 0;
  }
  int ex;
  if (mx <= 16.0 * (sum / (double)n)) {
    frexp(mx, &ex); /* mx = f * 2^ex, f in [0.5, 1) */
    int e = 30 - ex;
    for (size_t i = 0; i < n; i++) q[i] = (int32_t)nearbyint(ldexp(x[i], e));
    return e;
  }
  /* block mode: an integer r per value (sum |r| < 2^52 + n), kept with 26
   * significant bits: r ~ m * 2^s, 2^25 <= |m| < 2^26, word m * 32 + s */
  *block = 1;
  frexp(sum, &ex);
  int e = 52 - ex;
  for (size_t i = 0; i < n; i++) {
    double r = nearbyint(ldexp(x[i], e));
    double m = r;
    int s = 0;
    while (fabs(m) >= 67108864.0) { /* 2^26: halve the scale until it fits */
      s++;
      m = nearbyint(ldexp(r, -s));
    }
    q[i] = (int32_t)((int64_t)m * 32 + s);
  }
  return e;
}

double orc_fixed_value(int32_t word, int block) {
  if (!block) return (double)word;
  int32_t s = word & 31;
  int32_t m = (word - s) / 32; /* exact: word - s is a multiple of 32 */
  return ldexp((double)m, s);
}

/* ------------------------------------------------------------- the scan */

/* kernels.rs:314-327 */
uint32_t orc_phase_cell(double t, double scale, uint32_t nb) {
  double x = floor(t * scale);
  uint32_t v = x <= 0.0 ? 0u : (x >= 4294967295.0 ? 4294967295u : (uint32_t)x);
  return v % nb;
}

typedef struct {
  double best;
  uint64_t key;
  uint32_t di;
} lane_state;

/* kernels.rs:427-455 (rank_exact) */
static double rank_exact(lane_state* s, double a, double hp, double norm,
                         double rho, double c, int valid, uint64_t key,
                         uint32_t di, int* bad) {
  double d = hp * hp / c;
  double q = norm - d;
  if (valid && q <= 0.0) *bad = 1;
  double cand = (a / rho) / sqrt(fmax(q, 1e-300));
  double merit = (valid && q > 0.0 && cand < INFINITY) ? cand : -INFINITY;
  if (merit > s->best || (merit == s->best && key < s->key)) {
    s->best = merit;
    s->key = key;
    s->di = di;
  }
  return merit;
}

/* kernels.rs:84-113 (shape_sums) over a doubled prefix with row stride. */
static void shape_sums(const double* cv, const double* cw, size_t j, size_t k,
                       double a0, double w0, double ingress,
                       const double* factors, size_t si, size_t layers,
                       double* pa, double* php, double* pnorm) {
  if (ingress == 0.0) {
    *pa = a0;
    *php = w0;
    *pnorm = w0;
    return;
  }
  double l = (double)layers;
  double a = a0 / l;
  double hp = w0 / l;
  double norm = w0 / (l * l);
  for (size_t layer = 1; layer < layers; layer++) {
    double factor = factors[si * layers + layer];
    double wf = floor((double)k * factor + 0.5);
    size_t width = wf < 1.0 ? 1 : (size_t)wf;
    size_t start = j + (k - width) / 2;
    double dw = cw[start + width] - cw[start];
    a += (cv[start + width] - cv[start]) / l;
    hp += dw / l;
    norm += dw * ((2.0 * (double)layer + 1.0) / (l * l));
  }
  *pa = a;
  *php = hp;
  *pnorm = norm;
}

/* scan.rs:157-171 */
static void staircase_factors(const double* ingress, size_t ns, size_t layers,
                              double* factors) {
  for (size_t s = 0; s < ns; s++)
    for (size_t l = 0; l < layers; l++)
      factors[s * layers + l] =
          1.0 - 2.0 * ingress[s] *
                    (layers > 1 ? (double)l / (double)(layers - 1) : 0.0);
}

/* One period: fold (kernels.rs:499-553), doubled prefix (kernels.rs:237-309),
 * tile scan (kernels.rs:608-845) with f64 ranking, merge (kernels.rs:1113-1155).
 * Work arrays: hist 3*nb, cum 3*(2nb+1). */
static int scan_one(const double* t, const double* v, const double* w,
                    size_t n, double c, double period, const double* durations,
                    const double* rho, size_t nd, const double* ingress,
                    size_t ns, size_t layers, const double* factors,
                    double gate, uint32_t nb, double* hist, double* cum,
                    double* record, double* score, double* depth,
                    uint32_t* winner) {
  const double scale = (double)nb / period;
  const size_t stride = 2 * (size_t)nb + 1;
  double* hv = hist;
  double* hw = hist + nb;
  double* hn = hist + 2 * (size_t)nb;
  memset(hist, 0, 3 * (size_t)nb * sizeof(double));
  for (size_t i = 0; i < n; i++) {
    uint32_t b = orc_phase_cell(t[i], scale, nb);
    hv[b] += v[i];
    hw[b] += w[i];
    hn[b] += 1.0;
  }
  double* cv = cum;
  double* cw = cum + stride;
  double* cn = cum + 2 * stride;
  double* hs[3] = {hv, hw, hn};
  double* cs[3] = {cv, cw, cn};
  for (int f = 0; f < 3; f++) {
    double* h = hs[f];
    double* cf = cs[f];
    cf[0] = 0.0;
    for (size_t b = 0; b < nb; b++) cf[b + 1] = cf[b] + h[b];
    double total = cf[nb];
    for (size_t b = 0; b < nb; b++) cf[nb + b + 1] = cf[b + 1] + total;
  }
  const uint64_t nb1 = (uint64_t)nb + 1;
  const double nf = (double)n;
  const size_t tiles = ((size_t)nb + TILE - 1) / TILE;
  int bad = 0;
  double best_value = -INFINITY;
  uint64_t best_key = UINT64_MAX;
  uint32_t best_di = 0;
  lane_state lanes[TILE];
  for (size_t tile = 0; tile < tiles; tile++) {
    for (size_t l = 0; l < TILE; l++) {
      lanes[l].best = -INFINITY;
      lanes[l].key = UINT64_MAX;
      lanes[l].di = 0;
    }
    for (size_t di = 0; di < nd; di++) {
      double rk = floor(durations[di] * scale + 0.5);
      size_t raw_k = rk < 1.0 ? 1 : (size_t)rk;
      size_t k = raw_k < nb / 2 ? raw_k : nb / 2;
      double r = rho[di];
      int passing = 0;
      for (size_t l = 0; l < TILE; l++) {
        size_t j = tile * TILE + l;
        size_t jj = j < nb ? j : nb - 1;
        uint64_t key = (uint64_t)k * nb1 + jj;
        double count = cn[jj + k] - cn[jj];
        double ww = cw[jj + k] - cw[jj];
        double a = cv[jj + k] - cv[jj];
        int usable = j < nb && raw_k < nb / 2 && count >= 3.0 &&
                     count <= nf - 3.0;
        int valid = usable && a > 0.0;
        double merit = rank_exact(&lanes[l], a, ww, ww, r, c, valid, key,
                                  (uint32_t)di, &bad);
        passing |= valid && merit >= gate;
      }
      if (!passing) continue;
      for (size_t si = 1; si < ns; si++) {
        for (size_t l = 0; l < TILE; l++) {
          size_t j = tile * TILE + l;
          size_t jj = j < nb ? j : nb - 1;
          uint64_t key = ((uint64_t)si * nb1 + k) * nb1 + jj;
          double count = cn[jj + k] - cn[jj];
          double ww = cw[jj + k] - cw[jj];
          double a0 = cv[jj + k] - cv[jj];
          int usable = j < nb && raw_k < nb / 2 && count >= 3.0 &&
                       count <= nf - 3.0;
          double a, hp, norm;
          shape_sums(cv, cw, jj, k, a0, ww, ingress[si], factors, si, layers,
                     &a, &hp, &norm);
          rank_exact(&lanes[l], a, hp, norm, r, c, usable && a > 0.0, key,
                     (uint32_t)di, &bad);
        }
      }
    }
    /* block_winner: largest value, smallest key, lowest lane */
    size_t chosen = 0;
    for (size_t l = 1; l < TILE; l++) {
      const lane_state* p = &lanes[l];
      const lane_state* q = &lanes[chosen];
      if (p->best > q->best || (p->best == q->best && p->key < q->key))
        chosen = l;
    }
    /* merge_tiles: (value, smallest key); only valid winners compete, an
     * all-invalid period takes the canonical key nb1 (phase 0, width 1). */
    const lane_state* win = &lanes[chosen];
    if (win->best > -INFINITY &&
        (win->best > best_value ||
         (win->best == best_value && win->key < best_key))) {
      best_value = win->best;
      best_key = win->key;
      best_di = win->di;
    }
  }
  if (!(best_value > -INFINITY)) {
    best_key = nb1;
    best_di = 0;
  }
  /* kernels.rs:800-828: the winner recomputed in f64 */
  size_t phase = (size_t)(best_key % nb1);
  size_t width = (size_t)((best_key / nb1) % nb1);
  size_t si = (size_t)(best_key / (nb1 * nb1));
  size_t sic = si < ns - 1 ? si : ns - 1;
  double ing = ingress[sic];
  double r = rho[best_di];
  double ww = cw[phase + width] - cw[phase];
  double a00 = cv[phase + width] - cv[phase];
  double a0, hp, norm;
  shape_sums(cv, cw, phase, width, a00, ww, ing, factors, sic, layers, &a0, &hp,
             &norm);
  double a = a0 / r;
  double d = hp * hp / c;
  double q = fmax(norm - d, 1e-300);
  double value = best_value > -INFINITY ? a / sqrt(q) : -INFINITY;
  double dep = a / q * r;
  record[0] = period;
  record[1] = scale;
  record[2] = (double)phase;
  record[3] = (double)width;
  record[4] = a;
  record[5] = norm;
  record[6] = d;
  record[7] = ing;
  record[8] = r;
  *score = value;
  *depth = dep;
  if (winner) {
    winner[0] = (uint32_t)phase;
    winner[1] = (uint32_t)width;
    winner[2] = (uint32_t)si;
    winner[3] = best_di;
  }
  return bad;
}

int orc_scan_bank(const double* t, const double* v, const double* w, size_t n,
                  double c, const double* periods, size_t np,
                  const double* durations, const double* rho, size_t nd,
                  const double* ingress, size_t ns, uint32_t layers,
                  double gate, uint32_t nb, double* records, double* score,
                  double* depth, uint32_t* winner) {
  size_t nfac = ns * layers;
  double* factors = (double*)malloc((nfac ? nfac : 1) * sizeof(double));
  staircase_factors(ingress, ns, layers, factors);
  int bad = 0;
#pragma omp parallel reduction(| : bad)
  {
    double* hist = (double*)malloc(3 * (size_t)nb * sizeof(double));
    double* cum = (double*)malloc(3 * (2 * (size_t)nb + 1) * sizeof(double));
#pragma omp for schedule(dynamic, 8)
    for (long long f = 0; f < (long long)np; f++) {
      bad |= scan_one(t, v, w, n, c, periods[f], durations, rho, nd, ingress,
                      ns, layers, factors, gate, nb, hist, cum,
                      records + 9 * f, score + f, depth + f,
                      winner ? winner + 4 * f : NULL);
    }
    free(hist);
    free(cum);
  }
  free(factors);
  return bad ? 2 : 0;
}

/* scan.rs:177-183 */
uint64_t orc_bank_batches(size_t np, uint32_t nb, uint32_t period_batch,
                          double max_fold_mib) {
  if (np == 0) return 0;
  double per_period_bytes =
      (double)(3 * (4 * (uint64_t)nb + 2) * 8 + 3 * (2 * (uint64_t)nb + 1) * 8);
  uint64_t budget = (uint64_t)floor((max_fold_mib * 1048576.0) / per_period_bytes);
  if (budget == 0) return 0;
  uint64_t batch = period_batch < budget ? period_batch : budget;
  if (batch > np) batch = np;
  return (np + batch - 1) / batch;
}

static int bank_scan_rows(const double* t, const double* v, const double* w,
                          size_t n, double c, const double* periods, size_t np,
                          const double* durations, const double* rho, size_t nd,
                          const double* ingress, size_t ns,
                          const orc_bank_cfg* cfg, uint32_t nb, double* score,
                          double* rec, double* depth, orc_scan_stats* stats) {
  double per_period_bytes =
      (double)(3 * (4 * (uint64_t)nb + 2) * 8 + 3 * (2 * (uint64_t)nb + 1) * 8);
  if ((uint64_t)floor((cfg->max_fold_mib * 1048576.0) / per_period_bytes) == 0)
    return 1; /* "one folded period exceeds max_fold_mib" */
  stats->phase_bins = nb;
  int st = orc_scan_bank(t, v, w, n, c, periods, np, durations, rho, nd,
                         ingress, ns, cfg->layers, cfg->shape_gate, nb, rec,
                         score, depth, NULL);
  if (st) return st;
  stats->batches += orc_bank_batches(np, nb, cfg->period_batch, cfg->max_fold_mib);
  stats->templates_scanned += (uint64_t)np * nd * ns * nb;
  return 0;
}

/* scan.rs:363-408 */
int orc_scan_periods(const double* t, const double* v, const double* w,
                     size_t n, double c, double origin, const double* periods,
                     size_t np, const double* durations, const double* rho,
                     size_t nd, const double* ingress, size_t ns,
                     const orc_bank_cfg* cfg, orc_row* rows, size_t* nrows,
                     orc_scan_stats* stats) {
  memset(stats, 0, sizeof(*stats));
  stats->periods = np;
  *nrows = 0;
  if (np == 0 || nd == 0) return 0;
  double dmin = INFINITY;
  for (size_t i = 0; i < nd; i++) dmin = fmin(dmin, durations[i]);
  double nbf = ceil((double)cfg->phase_oversample * periods[np - 1] / dmin);
  uint32_t nb = nbf < 8.0 ? 8u : (uint32_t)nbf;
  double* survivors = (double*)malloc(np * sizeof(double));
  size_t nsurv = np;
  memcpy(survivors, periods, np * sizeof(double));
  double* score = (double*)malloc(np * sizeof(double));
  double* depth = (double*)malloc(np * sizeof(double));
  double* rec = (double*)malloc(9 * np * sizeof(double));
  int st = 0;
  if (cfg->prescan_factor > 1 && nb / cfg->prescan_factor >= 8) {
    st = bank_scan_rows(t, v, w, n, c, periods, np, durations, rho, nd,
                        ingress, ns, cfg, nb / cfg->prescan_factor, score, rec,
                        depth, stats);
    if (st) goto done;
    nsurv = 0;
    for (size_t i = 0; i < np; i++)
      if (isfinite(score[i]) &&
          score[i] >= cfg->prescan_fraction * cfg->threshold)
        survivors[nsurv++] = rec[9 * i];
    stats->prescan_survivors = nsurv;
  }
  if (nsurv) {
    st = bank_scan_rows(t, v, w, n, c, survivors, nsurv, durations, rho, nd,
                        ingress, ns, cfg, nb, score, rec, depth, stats);
    if (st) goto done;
    for (size_t i = 0; i < nsurv; i++) {
      if (!(isfinite(score[i]) && score[i] >= cfg->threshold)) continue;
      const double* r = rec + 9 * i;
      orc_row* row = &rows[(*nrows)++];
      row->period = r[0];
      row->epoch = (r[2] + r[3] / 2.0) / r[1] + origin;
      row->duration = r[3] / r[1];
      row->score = score[i];
      row->depth = depth[i];
    }
  }
done:
  free(survivors);
  free(score);
  free(depth);
  free(rec);
  return st;
}

/* scan.rs:440-471 with the reference's filter over the whole grid */
size_t orc_plan_bands(const double* grid, size_t count, const double* sd,
                      size_t nd, uint32_t phase_oversample,
                      uint32_t prescan_factor, orc_band* out, size_t cap) {
  size_t nbands = 0;
  if (count == 0 || nd == 0) return 0;
  double lo = grid[0];
  double last = grid[count - 1];
  while (lo <= last) {
    double hi = lo * 2.0;
    size_t first = (size_t)-1, sel = 0, lastsel = 0;
    for (size_t i = 0; i < count; i++) {
      if (grid[i] >= lo && grid[i] < hi) {
        if (first == (size_t)-1) first = i;
        sel++;
        lastsel = i;
      }
    }
    size_t chosen = 0;
    for (size_t i = 0; i < nd; i++)
      if (sd[i] < lo / 4.0) chosen++;
    if (sel && chosen) {
      double dmin = INFINITY;
      for (size_t i = 0; i < nd; i++)
        if (sd[i] < lo / 4.0) dmin = fmin(dmin, sd[i]);
      double nbf = ceil((double)phase_oversample * grid[lastsel] / dmin);
      uint32_t nb = nbf < 8.0 ? 8u : (uint32_t)nbf;
      if (nbands < cap) {
        orc_band* b = &out[nbands];
        b->lo = lo;
        b->first = first;
        b->count = sel;
        b->nd = chosen;
        b->nb = nb;
        b->nb_coarse =
            (prescan_factor > 1 && nb / prescan_factor >= 8) ? nb / prescan_factor : 0;
      }
      nbands++;
    }
    lo = hi;
  }
  return nbands;
}

/* scan.rs:413-473 */
int orc_signal_search(const double* t, const double* v, const double* w,
                      size_t n, double c, double origin,
                      const double* durations, const double* rho, size_t nd,
                      const double* ingress, size_t ns,
                      const orc_bank_cfg* bank, const orc_search_cfg* scfg,
                      orc_row* rows, size_t rows_cap, size_t* nrows,
                      orc_scan_stats* stats, size_t* nbands,
                      size_t* grid_count) {
  *nrows = 0;
  *nbands = 0;
  *grid_count = 0;
  if (nd == 0 || n == 0) return 0;
  keyed* ds = (keyed*)malloc(nd * sizeof(keyed));
  for (size_t i = 0; i < nd; i++) {
    ds[i].x = durations[i];
    ds[i].i = i;
  }
  qsort(ds, nd, sizeof(keyed), cmp_keyed);
  double* sd = (double*)malloc(nd * sizeof(double));
  double* sr = (double*)malloc(nd * sizeof(double));
  for (size_t i = 0; i < nd; i++) {
    sd[i] = ds[i].x;
    sr[i] = rho[ds[i].i];
  }
  free(ds);
  double baseline = t[n - 1];
  double maximum = fmin(scfg->period_max, baseline);
  double minimum = fmax(scfg->period_min, 4.0 * sd[0] * (1.0 + 1e-9));
  int st = 0;
  if (maximum <= minimum) goto out;
  size_t count = orc_period_grid_count(minimum, maximum, baseline, sd[0],
                                       scfg->period_oversample);
  double* grid = (double*)malloc(count * sizeof(double));
  orc_period_grid(minimum, maximum, baseline, sd[0], scfg->period_oversample,
                  grid, count);
  *grid_count = count;
  double lo = grid[0];
  double last = grid[count - 1];
  double* sel = (double*)malloc(count * sizeof(double));
  orc_bank_cfg b = *bank;
  while (lo <= last) {
    double hi = lo * 2.0;
    size_t ns_sel = 0;
    for (size_t i = 0; i < count; i++)
      if (grid[i] >= lo && grid[i] < hi) sel[ns_sel++] = grid[i];
    size_t nchosen = 0;
    while (nchosen < nd && sd[nchosen] < lo / 4.0) nchosen++;
    if (ns_sel && nchosen) {
      size_t got = 0;
      if (*nrows + ns_sel > rows_cap) {
        st = 4;
        break;
      }
      st = orc_scan_periods(t, v, w, n, c, origin, sel, ns_sel, sd, sr,
                            nchosen, ingress, ns, &b, rows + *nrows, &got,
                            &stats[*nbands]);
      if (st) break;
      *nrows += got;
      (*nbands)++;
    }
    lo = hi;
  }
  free(sel);
  free(grid);
out:
  free(sd);
  free(sr);
  return st;
}

/* ----------------------------------------------------------------- rng */

/* survey/numeric.rs:345-364 */
uint64_t orc_rng_new(uint64_t seed) {
  return (seed * 0x9E3779B97F4A7C15ull) | 1ull;
}

double orc_rng_uniform(uint64_t* state) {
  uint64_t x = *state;
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  *state = x;
  return (double)(x >> 11) / (double)(1ull << 53);
}

double orc_rng_normal(uint64_t* state, double mean, double sigma) {
  double u1 = orc_rng_uniform(state);
  if (!(u1 >= 1e-300)) u1 = 1e-300;
  double u2 = orc_rng_uniform(state);
  return mean + sigma * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

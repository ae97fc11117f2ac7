/* CPU oracle of the single-event scan and its selection.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of fbls_single_event_search and of
 * the single-event stage of the search (never linked into the library).
 *
 *   orc_single_event_search   isakruas/fbls src/search/events.rs:18-153 with
 *                             the kernels fold (kernels.rs:464-494),
 *                             cumulative (kernels.rs:237-309) and scan_all
 *                             (kernels.rs:849-919), sequentially
 *   orc_select_events         src/select/peaks.rs:490-550
 *
 * The fold sums the dequantised fixed-point fields in observation order
 * (exact: every partial sum is an integer multiple of a power of two below
 * 2^53), the prefix is the doubled prefix of the reference, and templates
 * are visited in (phase bin, ascending duration, shape) order; the
 * reference appends survivors through an atomic counter, so its order is
 * the hardware's. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fbls_oracle.h"

typedef struct {
  double d, r;
  size_t i;
} dur_t;

static int cmp_dur(const void* a, const void* b) {
  const dur_t* x = (const dur_t*)a;
  const dur_t* y = (const dur_t*)b;
  if (x->d != y->d) return x->d < y->d ? -1 : 1;
  return x->i < y->i ? -1 : (x->i > y->i); /* stable */
}

void orc_single_event_search(const double* t, const double* v, const double* w, size_t n,
                             double C, double origin, const double* durations,
                             const double* rho, size_t nd, const double* ingress, size_t ns,
                             size_t layers, double threshold, uint32_t phase_oversample,
                             size_t capacity, orc_row* rows, size_t* nrows,
                             orc_event_stats* stats) {
  *nrows = 0;
  memset(stats, 0, sizeof(*stats));
  dur_t* ds = (dur_t*)malloc((nd + 1) * sizeof(dur_t));
  for (size_t i = 0; i < nd; i++) {
    ds[i].d = durations[i];
    ds[i].r = rho[i];
    ds[i].i = i;
  }
  qsort(ds, nd, sizeof(dur_t), cmp_dur);
  const double baseline = t[n - 1];
  const double trial = baseline + 2.0 * ds[nd - 1].d + 1e-6;
  double nbf = ceil((double)phase_oversample * trial / ds[0].d);
  const uint32_t nb = nbf > 8.0 ? (uint32_t)nbf : 8u;
  const double scale = (double)nb / trial;
  double* factors = (double*)malloc((ns * layers + 1) * sizeof(double));
  for (size_t s = 0; s < ns; s++)
    for (size_t l = 0; l < layers; l++)
      factors[s * layers + l] =
          1.0 - 2.0 * ingress[s] * (layers > 1 ? (double)l / (double)(layers - 1) : 0.0);
  /* fold (one trial period: no wrap) */
  double* h = (double*)calloc(3 * (size_t)nb, sizeof(double));
  for (size_t i = 0; i < n; i++) {
    uint32_t b = orc_phase_cell(t[i], scale, nb);
    h[b] += v[i];
    h[nb + b] += w[i];
    h[2 * (size_t)nb + b] += 1.0;
  }
  /* doubled exclusive prefix: cum[x + 1] = sum of bins <= x, cum[nb + x + 1]
   * = cum[x + 1] + total */
  const size_t stride = 2 * (size_t)nb + 1;
  double* cum = (double*)malloc(3 * stride * sizeof(double));
  for (int f = 0; f < 3; f++) {
    double* cf = cum + f * stride;
    const double* hf = h + (size_t)f * nb;
    cf[0] = 0.0;
    for (size_t b = 0; b < nb; b++) cf[b + 1] = cf[b] + hf[b];
    for (size_t b = 0; b < nb; b++) cf[nb + b + 1] = cf[b + 1] + cf[nb];
  }
  const double* cv = cum;
  const double* cw = cum + stride;
  const double* cn = cum + 2 * stride;
  const double nf = (double)n, l = (double)layers;
  size_t count = 0;
  for (size_t j = 0; j < nb; j++) {
    for (size_t di = 0; di < nd; di++) {
      const double rk = floor(ds[di].d * scale + 0.5);
      const size_t raw_k = rk < 1.0 ? 1 : (size_t)rk;
      const size_t k = raw_k < nb / 2 ? raw_k : nb / 2;
      const double cnt = cn[j + k] - cn[j];
      if (!(raw_k < nb / 2 && cnt >= 3.0 && cnt <= nf - 3.0)) continue;
      const double a0 = cv[j + k] - cv[j];
      const double w0 = cw[j + k] - cw[j];
      for (size_t si = 0; si < ns; si++) {
        double a = a0, hp = w0, norm = w0;
        if (ingress[si] != 0.0) { /* kernels.rs:84-113 */
          a = a0 / l;
          hp = w0 / l;
          norm = w0 / (l * l);
          for (size_t layer = 1; layer < layers; layer++) {
            const double wf = floor((double)k * factors[si * layers + layer] + 0.5);
            const size_t width = wf < 1.0 ? 1 : (size_t)wf;
            const size_t start = j + (k - width) / 2;
            const double dw = cw[start + width] - cw[start];
            a += (cv[start + width] - cv[start]) / l;
            hp += dw / l;
            norm += dw * ((2.0 * (double)layer + 1.0) / (l * l));
          }
        }
        const double q = norm - hp * hp / C;
        if (!(a > 0.0 && q > 0.0)) continue;
        const double score = (a / ds[di].r) / sqrt(q);
        if (score >= threshold) {
          if (count < capacity) {
            orc_row* r = &rows[count];
            r->period = INFINITY;
            r->epoch = ((double)j + (double)k / 2.0) / scale + origin;
            r->duration = (double)k / scale;
            r->score = score;
            r->depth = a / q;
          }
          count++;
        }
      }
    }
  }
  stats->templates_scanned = (uint64_t)nb * nd * ns;
  stats->threshold_survivors = count;
  stats->overflow = count > capacity;
  stats->phase_bins = nb;
  *nrows = count > capacity ? 0 : count;
  free(ds);
  free(factors);
  free(h);
  free(cum);
}

size_t orc_select_events(const orc_row* table, size_t nrows, const double* time,
                         size_t n_time, double threshold, size_t limit,
                         uint8_t* const* masks, size_t n_masks, orc_single_event* out) {
  size_t nout = 0;
  if (nrows == 0) return 0;
  double* remaining = (double*)malloc(nrows * sizeof(double));
  for (size_t i = 0; i < nrows; i++) {
    const double s = table[i].score;
    remaining[i] = (isfinite(s) && s >= threshold) ? s : -INFINITY;
  }
  /* the selection's own masks follow the ones it was given */
  uint8_t** own = (uint8_t**)malloc((limit + 1) * sizeof(uint8_t*));
  size_t n_own = 0;
  uint8_t* inside = (uint8_t*)malloc(n_time + 1);
  size_t iterations = 0;
  while (nout < limit && iterations < 5000) {
    iterations++;
    size_t i = 0; /* max_by keeps the last of equal maxima */
    for (size_t j = 1; j < nrows; j++)
      if (remaining[j] >= remaining[i]) i = j;
    if (!isfinite(remaining[i])) break;
    const orc_row row = table[i];
    size_t points = 0;
    for (size_t j = 0; j < n_time; j++) {
      inside[j] = fabs(time[j] - row.epoch) <= row.duration / 2.0;
      points += inside[j];
    }
    for (size_t j = 0; j < nrows; j++)
      if (fabs(table[j].epoch - row.epoch) <= (table[j].duration + row.duration) / 2.0)
        remaining[j] = -INFINITY;
    int clash = points < 3;
    for (size_t m = 0; m < n_masks + n_own && !clash; m++) {
      const uint8_t* other = m < n_masks ? masks[m] : own[m - n_masks];
      size_t both = 0;
      for (size_t j = 0; j < n_time; j++) both += inside[j] && other[j];
      clash = (points ? (double)both / (double)points : 0.0) >= 0.5;
    }
    if (clash) continue;
    uint8_t* wide = (uint8_t*)malloc(n_time + 1);
    for (size_t j = 0; j < n_time; j++) wide[j] = fabs(time[j] - row.epoch) <= 0.75 * row.duration;
    own[n_own++] = wide;
    orc_single_event* e = &out[nout++];
    e->epoch_bjd = row.epoch;
    e->duration_hours = 24.0 * row.duration;
    e->snr = row.score;
    e->depth = row.depth;
    e->depth_ppm = 1e6 * row.depth;
    e->points = points;
  }
  for (size_t m = 0; m < n_own; m++) free(own[m]);
  free(own);
  free(inside);
  free(remaining);
  return nout;
}

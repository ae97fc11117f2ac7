/*
 * FP64 CPU oracle of the raw-mode preparation: per-group normalisation with
 * inverse-variance binning, the Tukey biweight detrend and the per-window
 * duration families.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the library's raw path.
 * Restates (isakruas/fbls):
 *   fit_segment_noise      src/prep/curve.rs:89-198
 *   biweight_flatten       src/prep/curve.rs:204-248
 *   segment_edges          src/prep/curve.rs:331-350
 *   detrending_windows     src/prep/curve.rs:355-377
 *   prepare_families       src/prep/curve.rs:387-425
 * Medians are the lower medians of sorted copies (stats.rs:4-11).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fbls_oracle.h"

typedef struct {
  double x;
  size_t i;
} kx;

static int cmp_kx(const void* a, const void* b) {
  const kx* p = (const kx*)a;
  const kx* q = (const kx*)b;
  if (p->x < q->x) return -1;
  if (p->x > q->x) return 1;
  return (p->i > q->i) - (p->i < q->i); /* stable */
}

typedef struct {
  int64_t bin, group;
  size_t i;
} kb;

static int cmp_kb(const void* a, const void* b) {
  const kb* p = (const kb*)a;
  const kb* q = (const kb*)b;
  if (p->bin != q->bin) return p->bin < q->bin ? -1 : 1;
  if (p->group != q->group) return p->group < q->group ? -1 : 1;
  return (p->i > q->i) - (p->i < q->i);
}

static int cmp_i64(const void* a, const void* b) {
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return x < y ? -1 : x > y;
}

/* curve.rs:89-198.  Outputs hold up to n values. */
int orc_fit_segment_noise_ex(const double* time, const double* flux, const double* err,
                             const int64_t* group, size_t n, double bin_minutes,
                             double gap, double* out_t, double* out_def, double* out_var,
                             size_t* out_n, double* origin, double* inflation,
                             size_t* segments) {
  kx* order = (kx*)malloc((n + 1) * sizeof(kx));
  size_t m = 0;
  for (size_t i = 0; i < n; i++) {
    double e = err ? err[i] : 1.0;
    if (isfinite(time[i]) && isfinite(flux[i]) && isfinite(e) && e > 0.0) {
      order[m].x = time[i];
      order[m].i = i;
      m++;
    }
  }
  qsort(order, m, sizeof(kx), cmp_kx);
  if (m < 6) {
    free(order);
    return 6;
  }
  double* t = (double*)malloc((m + 1) * 8);
  double* f = (double*)malloc((m + 1) * 8);
  double* e = (double*)malloc((m + 1) * 8);
  int64_t* sec = (int64_t*)malloc((m + 1) * 8);
  for (size_t k = 0; k < m; k++) {
    size_t i = order[k].i;
    t[k] = time[i];
    f[k] = flux[i];
    e[k] = err ? err[i] : 1.0;
    sec[k] = group ? group[i] : 0;
  }
  free(order);
  /* per-group lower median (curve.rs:120-131) */
  int64_t* labels = (int64_t*)malloc((m + 1) * 8);
  memcpy(labels, sec, m * 8);
  qsort(labels, m, 8, cmp_i64);
  double* vals = (double*)malloc((m + 1) * 8);
  int st = 0;
  for (size_t a = 0; a < m && !st;) {
    size_t b = a;
    while (b < m && labels[b] == labels[a]) b++;
    size_t nv = 0;
    for (size_t k = 0; k < m; k++)
      if (sec[k] == labels[a]) vals[nv++] = f[k];
    double med = orc_median_lower(vals, nv);
    if (!(med > 0.0)) {
      st = 7;
      break;
    }
    for (size_t k = 0; k < m; k++)
      if (sec[k] == labels[a]) {
        f[k] = f[k] / med - 1.0;
        e[k] /= med;
      }
    a = b;
  }
  free(labels);
  free(vals);
  if (st) {
    free(t);
    free(f);
    free(e);
    free(sec);
    return st;
  }
  double org = t[0];
  for (size_t k = 0; k < m; k++) t[k] -= org;
  if (bin_minutes > 0.0) { /* curve.rs:141-162 */
    double width = bin_minutes / 1440.0;
    kb* keys = (kb*)malloc((m + 1) * sizeof(kb));
    for (size_t k = 0; k < m; k++) {
      keys[k].bin = (int64_t)floor(t[k] / width);
      keys[k].group = sec[k];
      keys[k].i = k;
    }
    qsort(keys, m, sizeof(kb), cmp_kb);
    size_t* inv = (size_t*)malloc((m + 1) * sizeof(size_t));
    int64_t* usec = (int64_t*)malloc((m + 1) * 8);
    size_t u = 0;
    for (size_t k = 0; k < m; k++) {
      if (k == 0 || keys[k].bin != keys[k - 1].bin || keys[k].group != keys[k - 1].group) {
        usec[u] = keys[k].group;
        u++;
      }
      inv[keys[k].i] = u - 1;
    }
    double* den = (double*)calloc(u + 1, 8);
    double* stt = (double*)calloc(u + 1, 8);
    double* sf = (double*)calloc(u + 1, 8);
    for (size_t k = 0; k < m; k++) {
      double w = 1.0 / (e[k] * e[k]);
      den[inv[k]] += w;
      stt[inv[k]] += t[k] * w;
      sf[inv[k]] += f[k] * w;
    }
    kx* ord = (kx*)malloc((u + 1) * sizeof(kx));
    for (size_t j = 0; j < u; j++) {
      ord[j].x = stt[j] / den[j];
      ord[j].i = j;
    }
    qsort(ord, u, sizeof(kx), cmp_kx);
    for (size_t j = 0; j < u; j++) {
      size_t q = ord[j].i;
      t[j] = stt[q] / den[q];
      f[j] = sf[q] / den[q];
      e[j] = 1.0 / sqrt(den[q]);
      sec[j] = usec[q];
    }
    m = u;
    free(keys);
    free(inv);
    free(usec);
    free(den);
    free(stt);
    free(sf);
    free(ord);
  }
  double* deltas = (double*)malloc((m + 1) * 8);
  double* stdz = (double*)malloc((m + 1) * 8);
  size_t nd = 0;
  for (size_t i = 1; i < m; i++) {
    if (t[i] - t[i - 1] < gap && sec[i] == sec[i - 1]) {
      double d = f[i] - f[i - 1];
      deltas[nd] = d;
      stdz[nd] = d / sqrt(e[i] * e[i] + e[i - 1] * e[i - 1]);
      nd++;
    }
  }
  if (nd == 0) {
    st = 8;
  } else {
    double infl = orc_mad_scale(stdz, nd);
    if (!(infl >= 1.0)) infl = 1.0;
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
      if (t[i] - t[i - 1] > gap || sec[i] != sec[i - 1]) seg++;
    for (size_t k = 0; k < m; k++) {
      out_t[k] = t[k];
      out_def[k] = -f[k];
    }
    *out_n = m;
    *origin = org;
    *inflation = infl;
    *segments = seg;
  }
  free(t);
  free(f);
  free(e);
  free(sec);
  free(deltas);
  free(stdz);
  return st;
}

static size_t lower_bound(const double* v, size_t n, double x) {
  size_t lo = 0, hi = n;
  while (lo < hi) {
    size_t mid = lo + (hi - lo) / 2;
    if (v[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

static size_t upper_bound(const double* v, size_t n, double x) {
  size_t lo = 0, hi = n;
  while (lo < hi) {
    size_t mid = lo + (hi - lo) / 2;
    if (v[mid] <= x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* curve.rs:204-248 */
void orc_biweight_flatten(const double* time, const double* rel, size_t n, double window,
                          double gap, double c, size_t iterations, double* out) {
  double* flux = (double*)malloc((n + 1) * 8);
  for (size_t i = 0; i < n; i++) flux[i] = 1.0 + rel[i];
  size_t a = 0;
  for (size_t i = 1; i <= n; i++) {
    if (i < n && !(time[i] - time[i - 1] > gap)) continue;
    const size_t b = i, len = b - a;
    const double* lt = time + a;
    const double* lf = flux + a;
#pragma omp parallel for schedule(dynamic, 64)
    for (long long q = 0; q < (long long)len; q++) {
      size_t lo = lower_bound(lt, len, lt[q] - window / 2.0);
      size_t hi = upper_bound(lt, len, lt[q] + window / 2.0);
      size_t w = hi - lo;
      double* dev = (double*)malloc((w + 1) * 8);
      double location = orc_median_lower(lf + lo, w);
      for (size_t it = 0; it < iterations; it++) {
        for (size_t j = 0; j < w; j++) dev[j] = fabs(lf[lo + j] - location);
        double scale = c * orc_median_lower(dev, w);
        if (!(scale >= 1e-12)) scale = 1e-12;
        double total = 0.0, acc = 0.0;
        for (size_t j = 0; j < w; j++) {
          double v = lf[lo + j];
          double u = (v - location) / scale;
          if (fabs(u) < 1.0) {
            double wt = (1.0 - u * u) * (1.0 - u * u);
            total += wt;
            acc += wt * v;
          }
        }
        if (total > 0.0) location = acc / total;
      }
      out[a + q] = lf[q] / location - 1.0;
      free(dev);
    }
    a = b;
  }
  free(flux);
}

/* curve.rs:331-350 */
static void segment_edges(const double* t, size_t n, double gap, double* out) {
  size_t start = 0;
  for (size_t i = 1; i <= n; i++) {
    if (i == n || t[i] - t[i - 1] > gap) {
      for (size_t j = start; j < i; j++) out[j] = fmin(t[j] - t[start], t[i - 1] - t[j]);
      start = i;
    }
  }
}

static int cmp_d(const void* a, const void* b) {
  double x = *(const double*)a, y = *(const double*)b;
  return x < y ? -1 : x > y;
}

/* curve.rs:387-425.  Per family k (< *nfam <= nd): fam_n[k] points of
 * fam_time/fam_flux/fam_sigma[k * n ..], fam_window[k], fam_clipped[k],
 * fam_members[k * nd + i] = 1 when duration i belongs to the family. */
int orc_prepare_families(const double* time, const double* flux, const double* err,
                         const int64_t* group, size_t n, const double* durations, size_t nd,
                         const orc_prep_cfg* cfg, size_t* nfam, size_t* fam_n,
                         double* fam_time, double* fam_flux, double* fam_sigma,
                         double* fam_window, size_t* fam_clipped, uint8_t* fam_members) {
  *nfam = 0;
  double* t = (double*)malloc((n + 1) * 8);
  double* def = (double*)malloc((n + 1) * 8);
  double* var = (double*)malloc((n + 1) * 8);
  size_t m = 0, segs = 0;
  double origin = 0.0, infl = 0.0;
  int st = orc_fit_segment_noise_ex(time, flux, err, group, n, cfg->bin_minutes, 0.75, t,
                                    def, var, &m, &origin, &infl, &segs);
  if (st) {
    free(t);
    free(def);
    free(var);
    return st;
  }
  double* rel = (double*)malloc((m + 1) * 8);
  double* sig = (double*)malloc((m + 1) * 8);
  double* edge = (double*)malloc((m + 1) * 8);
  double* flat = (double*)malloc((m + 1) * 8);
  for (size_t i = 0; i < m; i++) {
    rel[i] = -def[i];
    sig[i] = sqrt(var[i]);
  }
  segment_edges(t, m, 0.5, edge);
  /* curve.rs:355-377 */
  double* w = (double*)malloc((nd + 1) * 8);
  double* uniq = (double*)malloc((nd + 1) * 8);
  for (size_t i = 0; i < nd; i++) {
    double x = fmin(fmax(cfg->window_factor * durations[i], cfg->window_min), cfg->window_max);
    w[i] = round(x * 1e6) / 1e6;
    uniq[i] = w[i];
  }
  qsort(uniq, nd, 8, cmp_d);
  size_t nu = 0;
  for (size_t i = 0; i < nd; i++)
    if (nu == 0 || uniq[i] != uniq[nu - 1]) uniq[nu++] = uniq[i];
  for (size_t k = 0; k < nu; k++) {
    double window = uniq[k];
    orc_biweight_flatten(t, rel, m, window, 0.5, 5.0, 10, flat);
    double scale = orc_mad_scale(flat, m);
    size_t c = 0;
    for (size_t i = 0; i < m; i++) {
      if (isfinite(flat[i]) && flat[i] <= cfg->clip_sigma * scale &&
          edge[i] >= cfg->edge_trim * window) {
        fam_time[k * n + c] = t[i] + origin;
        fam_flux[k * n + c] = flat[i];
        fam_sigma[k * n + c] = sig[i];
        c++;
      }
    }
    fam_n[k] = c;
    fam_window[k] = window;
    fam_clipped[k] = m - c;
    for (size_t i = 0; i < nd; i++) fam_members[k * nd + i] = w[i] == window;
  }
  *nfam = nu;
  free(t);
  free(def);
  free(var);
  free(rel);
  free(sig);
  free(edge);
  free(flat);
  free(w);
  free(uniq);
  return 0;
}

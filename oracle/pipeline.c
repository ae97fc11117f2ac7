/*
 * FP64 CPU oracle of the multi-pass prepared search: signal search per
 * pass, greedy peak selection with event-overlap rescoring, masking of the
 * lead peak, and the model comparison of the strongest peaks.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the library's pipeline.
 * Restates (isakruas/fbls):
 *   transit_mask / mask_curve           src/prep/curve.rs:572-603
 *   event_count, nearest_rho, positions src/select/peaks.rs:66-103
 *   rescore_rows / count_rows           src/search/kernels.rs:934-1042
 *   Rescorer::evaluate                  src/select/peaks.rs:148-240
 *   without_strongest_event             src/select/peaks.rs:243-299
 *   select_peaks                        src/select/peaks.rs:308-487
 *   search_prepared / search_families   src/pipeline.rs:41-221
 * The fields v and w are quantised exactly as the library stores them
 * (orc_quantize), so every sum here equals the device's.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fbls_oracle.h"

/* ------------------------------------------------------------ helpers */

static double rem_euclid(double a, double b) {
  double r = fmod(a, b);
  return r < 0.0 ? r + fabs(b) : r;
}

/* curve.rs:572-585 */
void orc_transit_mask(const double* time, size_t n, double period,
                      double epoch, double duration, double margin,
                      uint8_t* out) {
  for (size_t i = 0; i < n; i++) {
    double phase = rem_euclid(time[i] - epoch + period / 2.0, period) - period / 2.0;
    out[i] = fabs(phase) <= (1.0 + margin) * duration / 2.0;
  }
}

static int cmp_i64(const void* a, const void* b) {
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return x < y ? -1 : x > y;
}

/* peaks.rs:66-83 */
size_t orc_event_count(const double* time, size_t n, double period,
                       double epoch, double duration, size_t minimum_points) {
  uint8_t* inside = (uint8_t*)malloc(n ? n : 1);
  orc_transit_mask(time, n, period, epoch, duration, 0.0, inside);
  int64_t* cyc = (int64_t*)malloc((n ? n : 1) * sizeof(int64_t));
  size_t m = 0;
  for (size_t i = 0; i < n; i++)
    if (inside[i]) cyc[m++] = (int64_t)round((time[i] - epoch) / period);
  qsort(cyc, m, sizeof(int64_t), cmp_i64);
  size_t events = 0;
  for (size_t i = 0; i < m;) {
    size_t j = i;
    while (j < m && cyc[j] == cyc[i]) j++;
    if (j - i >= minimum_points) events++;
    i = j;
  }
  free(inside);
  free(cyc);
  return events;
}

/* kernels.rs:934-989 (one row, sequential) */
void orc_rescore_row(const double* time, const double* v, const double* w,
                     const uint8_t* excluded, size_t n, double origin,
                     double period, double epoch, double duration, double* out3) {
  double inverse = 1.0 / period, half = duration / 2.0;
  double a = 0.0, hp = 0.0, norm = 0.0;
  for (size_t i = 0; i < n; i++) {
    if (excluded[i]) continue;
    double x = (time[i] + origin - epoch) * inverse + 0.5;
    double phase = (x - floor(x) - 0.5) * period;
    if (fabs(phase) <= half) {
      a += v[i];
      hp += w[i];
      norm += w[i];
    }
  }
  out3[0] = a;
  out3[1] = hp;
  out3[2] = norm;
}

/* kernels.rs:992-1042 (one row, sequential) */
void orc_count_row(const double* time_all, const uint8_t* excluded, size_t n,
                   double period, double epoch, double duration,
                   uint32_t* out2) {
  double inverse = 1.0 / period, half = duration / 2.0;
  uint32_t total = 0, unshared = 0;
  for (size_t i = 0; i < n; i++) {
    double x = (time_all[i] - epoch) * inverse + 0.5;
    double phase = (x - floor(x) - 0.5) * period;
    if (fabs(phase) <= half) {
      total++;
      unshared += excluded[i] == 0;
    }
  }
  out2[0] = total;
  out2[1] = unshared;
}

/* one prepared family: the curve it was built from and its state */
typedef struct {
  double *ct, *cf, *cs; /* curve: absolute time, flux, sigma */
  size_t cn;
  double *t, *v, *w; /* state: relative time, quantised fields */
  size_t n;
  double C, origin, energy;
  double* rho;
  double* dur; /* the family's durations (days) */
  size_t* idx; /* their positions in the bank */
  size_t nd;
} fam_t;

static void fam_free(fam_t* f) {
  free(f->ct);
  free(f->cf);
  free(f->cs);
  free(f->t);
  free(f->v);
  free(f->w);
  free(f->rho);
  free(f->dur);
  free(f->idx);
  memset(f, 0, sizeof(*f));
}

/* curve.rs:541-569 plus the library's fixed-point step */
static int fam_build(const double* ct, const double* cf, const double* cs,
                     size_t cn, const double* durations, const size_t* indices,
                     size_t nd, int correlated, fam_t* f) {
  memset(f, 0, sizeof(*f));
  f->nd = nd;
  f->dur = (double*)malloc((nd + 1) * 8);
  f->idx = (size_t*)malloc((nd + 1) * sizeof(size_t));
  memcpy(f->dur, durations, nd * 8);
  memcpy(f->idx, indices, nd * sizeof(size_t));
  f->cn = cn;
  f->ct = (double*)malloc((cn + 1) * 8);
  f->cf = (double*)malloc((cn + 1) * 8);
  f->cs = (double*)malloc((cn + 1) * 8);
  memcpy(f->ct, ct, cn * 8);
  memcpy(f->cf, cf, cn * 8);
  memcpy(f->cs, cs, cn * 8);
  double* flux1 = (double*)malloc((cn + 1) * 8);
  for (size_t i = 0; i < cn; i++) flux1[i] = 1.0 + cf[i];
  double* dt = (double*)malloc((cn + 1) * 8);
  double* dd = (double*)malloc((cn + 1) * 8);
  double* dv = (double*)malloc((cn + 1) * 8);
  size_t m = 0, segs = 0;
  double infl = 0.0;
  int st = orc_fit_segment_noise(ct, flux1, cs, cn, 0.75, dt, dd, dv, &m,
                                 &f->origin, &infl, &segs);
  free(flux1);
  if (st) {
    free(dt);
    free(dd);
    free(dv);
    fam_free(f);
    return st;
  }
  f->n = m;
  f->t = dt;
  f->v = (double*)malloc((m + 1) * 8);
  f->w = (double*)malloc((m + 1) * 8);
  st = orc_prepare_ready(dd, dv, m, f->v, f->w, &f->C, &f->energy);
  if (st) {
    free(dd);
    free(dv);
    fam_free(f);
    return st;
  }
  f->rho = (double*)malloc((nd + 1) * 8);
  if (correlated) {
    double* resid = (double*)malloc((m + 1) * 8);
    double* sig = (double*)malloc((m + 1) * 8);
    for (size_t i = 0; i < m; i++) {
      resid[i] = -dd[i];
      sig[i] = sqrt(dv[i]);
    }
    orc_correlated_noise_factors(dt, resid, sig, m, durations, nd, f->rho);
    free(resid);
    free(sig);
  } else {
    for (size_t i = 0; i < nd; i++) f->rho[i] = 1.0;
  }
  free(dd);
  free(dv);
  /* fixed point as the library uploads it */
  int32_t* q = (int32_t*)malloc((m + 1) * 4);
  int block = 0;
  int e = orc_quantize(f->v, m, q, &block);
  for (size_t i = 0; i < m; i++) f->v[i] = ldexp(orc_fixed_value(q[i], block), -e);
  e = orc_quantize(f->w, m, q, &block);
  for (size_t i = 0; i < m; i++) f->w[i] = ldexp(orc_fixed_value(q[i], block), -e);
  free(q);
  return 0;
}

/* curve.rs:587-603: the family rebuilt from the curve without the masked
 * observations */
static int fam_masked(const fam_t* src, double period, double epoch,
                      double duration, double margin, int correlated, fam_t* out,
                      size_t* kept) {
  uint8_t* m = (uint8_t*)malloc(src->cn + 1);
  orc_transit_mask(src->ct, src->cn, period, epoch, duration, margin, m);
  double* t = (double*)malloc((src->cn + 1) * 8);
  double* fl = (double*)malloc((src->cn + 1) * 8);
  double* s = (double*)malloc((src->cn + 1) * 8);
  size_t k = 0;
  for (size_t i = 0; i < src->cn; i++)
    if (!m[i]) {
      t[k] = src->ct[i];
      fl[k] = src->cf[i];
      s[k] = src->cs[i];
      k++;
    }
  free(m);
  *kept = k;
  /* family_state fails below six observations (curve.rs:429-431) */
  int st = fam_build(t, fl, s, k, src->dur, src->idx, src->nd, correlated, out);
  free(t);
  free(fl);
  free(s);
  return st;
}

/* peaks.rs:85-94 */
static double nearest_rho(const double* durations, const double* rho, size_t nd,
                          double duration) {
  size_t best = 0;
  for (size_t i = 1; i < nd; i++)
    if (fabs(durations[i] - duration) < fabs(durations[best] - duration)) best = i;
  return rho[best];
}

static size_t lower_bound(const double* v, size_t n, double x) {
  size_t lo = 0, hi = n;
  while (lo < hi) {
    size_t mid = lo + (hi - lo) / 2;
    if (v[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

/* ---------------------------------------------------------- selection */

typedef struct {
  double v;
  size_t i;
} entry_t;

typedef struct {
  entry_t* a;
  size_t n, cap;
} heap_t;

/* BinaryHeap<Entry> order (peaks.rs:367-379): larger score first, then the
 * smaller row index */
static int above(entry_t x, entry_t y) { return x.v > y.v || (x.v == y.v && x.i < y.i); }

static void heap_push(heap_t* h, entry_t e) {
  if (h->n == h->cap) {
    h->cap = h->cap ? 2 * h->cap : 64;
    h->a = (entry_t*)realloc(h->a, h->cap * sizeof(entry_t));
  }
  size_t k = h->n++;
  h->a[k] = e;
  while (k > 0) {
    size_t p = (k - 1) / 2;
    if (!above(h->a[k], h->a[p])) break;
    entry_t t = h->a[k];
    h->a[k] = h->a[p];
    h->a[p] = t;
    k = p;
  }
}

static entry_t heap_pop(heap_t* h) {
  entry_t top = h->a[0];
  h->a[0] = h->a[--h->n];
  size_t k = 0;
  for (;;) {
    size_t l = 2 * k + 1, r = l + 1, b = k;
    if (l < h->n && above(h->a[l], h->a[b])) b = l;
    if (r < h->n && above(h->a[r], h->a[b])) b = r;
    if (b == k) break;
    entry_t t = h->a[k];
    h->a[k] = h->a[b];
    h->a[b] = t;
    k = b;
  }
  return top;
}

typedef struct {
  const orc_row* table;
  size_t nrows;
  const size_t* family; /* family of each row */
  fam_t* fams;
  size_t nfam;
  size_t** pos; /* per family: index of each observation in time_all */
  const double* time_all;
  size_t n_all;
  double threshold;
  uint8_t* excluded;
} sel_ctx;

/* peaks.rs:148-240 */
static void evaluate(const sel_ctx* S, const size_t* idx, size_t m, double* score,
                     double* required) {
  uint8_t** ex_fam = (uint8_t**)malloc(S->nfam * sizeof(uint8_t*));
  for (size_t k = 0; k < S->nfam; k++) {
    const fam_t* f = &S->fams[k];
    ex_fam[k] = (uint8_t*)malloc(f->n + 1);
    for (size_t i = 0; i < f->n; i++) ex_fam[k][i] = S->excluded[S->pos[k][i]];
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (long long s = 0; s < (long long)m; s++) {
    const orc_row* r = &S->table[idx[s]];
    const size_t fk = S->family[idx[s]];
    const fam_t* f = &S->fams[fk];
    uint32_t c2[2];
    orc_count_row(S->time_all, S->excluded, S->n_all, r->period, r->epoch, r->duration, c2);
    double total = (double)(c2[0] > 1 ? c2[0] : 1);
    double unshared = (double)c2[1] / total;
    required[s] = fmax(S->threshold * sqrt(unshared), 0.5 * S->threshold);
    double sums[3];
    orc_rescore_row(f->t, f->v, f->w, ex_fam[fk], f->n, f->origin, r->period, r->epoch,
                    r->duration, sums);
    double q = sums[2] - sums[1] * sums[1] / f->C;
    score[s] = (q > 0.0 && sums[0] > 0.0)
                   ? (sums[0] / nearest_rho(f->dur, f->rho, f->nd, r->duration)) /
                         sqrt(fmax(q, 1e-300))
                   : -INFINITY;
  }
  for (size_t k = 0; k < S->nfam; k++) free(ex_fam[k]);
  free(ex_fam);
}

typedef struct {
  int64_t cycle;
  double a, hp, w;
} cyc_t;

/* peaks.rs:243-299 */
static double without_strongest_event(const sel_ctx* S, const orc_row* row, size_t fk) {
  const fam_t* f = &S->fams[fk];
  const size_t* pos = S->pos[fk];
  cyc_t* c = (cyc_t*)malloc((f->n + 1) * sizeof(cyc_t));
  size_t* mem_i = (size_t*)malloc((f->n + 1) * sizeof(size_t));
  int64_t* mem_c = (int64_t*)malloc((f->n + 1) * sizeof(int64_t));
  size_t nc = 0, nm = 0;
  for (size_t i = 0; i < f->n; i++) {
    if (S->excluded[pos[i]]) continue;
    double ta = f->t[i] + f->origin;
    double phase = rem_euclid(ta - row->epoch + row->period / 2.0, row->period) - row->period / 2.0;
    if (fabs(phase) <= row->duration / 2.0) {
      int64_t cycle = (int64_t)round((ta - row->epoch) / row->period);
      size_t k = 0;
      while (k < nc && c[k].cycle != cycle) k++;
      if (k == nc) {
        c[nc].cycle = cycle;
        c[nc].a = c[nc].hp = c[nc].w = 0.0;
        nc++;
      }
      c[k].a += f->v[i];
      c[k].hp += f->w[i];
      c[k].w += f->w[i];
      mem_i[nm] = i;
      mem_c[nm] = cycle;
      nm++;
    }
  }
  double value = -INFINITY;
  if (nc) {
    double rho = nearest_rho(f->dur, f->rho, f->nd, row->duration);
    /* BTreeMap order: ascending cycle; the first strict maximum wins */
    int64_t strongest = 0;
    double best = -INFINITY;
    int64_t prev = INT64_MIN;
    for (size_t done = 0; done < nc; done++) {
      size_t k = nc;
      for (size_t j = 0; j < nc; j++)
        if (c[j].cycle > prev && (k == nc || c[j].cycle < c[k].cycle)) k = j;
      prev = c[k].cycle;
      double s = c[k].a / rho / sqrt(fmax(c[k].w - c[k].hp * c[k].hp / f->C, 1e-300));
      if (s > best) {
        best = s;
        strongest = c[k].cycle;
      }
    }
    double a = 0.0, hp = 0.0, norm = 0.0;
    for (size_t j = 0; j < nm; j++)
      if (mem_c[j] != strongest) {
        a += f->v[mem_i[j]];
        hp += f->w[mem_i[j]];
        norm += f->w[mem_i[j]];
      }
    double q = norm - hp * hp / f->C;
    value = (q > 0.0 && a > 0.0) ? a / rho / sqrt(fmax(q, 1e-300)) : -INFINITY;
  }
  free(c);
  free(mem_i);
  free(mem_c);
  return value;
}

typedef struct {
  sel_ctx* S;
  double* remaining;
  uint8_t* fresh;
  heap_t* heap;
} refresh_t;

/* peaks.rs:383-419 */
static void refresh(refresh_t* R) {
  const size_t block = 8192;
  size_t* stale = (size_t*)malloc(block * sizeof(size_t));
  size_t ns = 0;
  heap_t kept = {0, 0, 0};
  while (R->heap->n) {
    entry_t e = heap_pop(R->heap);
    if (e.v != R->remaining[e.i]) continue;
    if (R->fresh[e.i]) {
      heap_push(&kept, e);
    } else {
      stale[ns++] = e.i;
      if (ns == block) break;
    }
  }
  for (size_t k = 0; k < kept.n; k++) heap_push(R->heap, kept.a[k]);
  free(kept.a);
  if (ns) {
    double* sc = (double*)malloc(ns * 8);
    double* rq = (double*)malloc(ns * 8);
    evaluate(R->S, stale, ns, sc, rq);
    for (size_t m = 0; m < ns; m++) {
      size_t i = stale[m];
      R->remaining[i] = sc[m] >= rq[m] ? sc[m] : -INFINITY;
      R->fresh[i] = 1;
      if (isfinite(R->remaining[i])) {
        entry_t e = {R->remaining[i], i};
        heap_push(R->heap, e);
      }
    }
    free(sc);
    free(rq);
  }
  free(stale);
}

static size_t popcount_u8(const uint8_t* a, size_t n) {
  size_t c = 0;
  for (size_t i = 0; i < n; i++) c += a[i] != 0;
  return c;
}

/* peaks.rs:308-487.  masks: n_masks x n_all (grown);
 * accepted peaks appended to out (capacity cap). */
static void select_peaks(sel_ctx* S, size_t limit, size_t min_events,
                         uint8_t** masks, size_t* n_masks, size_t pass,
                         orc_peak* out, size_t* nout, size_t cap) {
  const double margin = 0.5;
  const size_t max_iterations = 20000;
  if (S->nrows == 0) return;
  double* remaining = (double*)malloc(S->nrows * 8);
  uint8_t* fresh = (uint8_t*)malloc(S->nrows);
  int any_ex = popcount_u8(S->excluded, S->n_all) > 0;
  heap_t heap = {0, 0, 0};
  for (size_t i = 0; i < S->nrows; i++) {
    double sc = S->table[i].score;
    remaining[i] = (isfinite(sc) && sc >= S->threshold) ? sc : -INFINITY;
    fresh[i] = !any_ex;
    if (isfinite(remaining[i])) {
      entry_t e = {remaining[i], i};
      heap_push(&heap, e);
    }
  }
  refresh_t R = {S, remaining, fresh, &heap};
  uint8_t* mask = (uint8_t*)malloc(S->n_all + 1);
  size_t accepted = 0, iterations = 0;
  while (accepted < limit && iterations < max_iterations) {
    iterations++;
    while (heap.n && heap.a[0].v != remaining[heap.a[0].i]) heap_pop(&heap);
    if (!heap.n) break;
    size_t i = heap.a[0].i;
    if (!fresh[i]) {
      refresh(&R);
      continue;
    }
    heap_pop(&heap);
    double isolated = remaining[i];
    remaining[i] = -INFINITY;
    const orc_row* row = &S->table[i];
    orc_transit_mask(S->time_all, S->n_all, row->period, row->epoch, row->duration, 0.0, mask);
    size_t total = popcount_u8(mask, S->n_all);
    int clash = 0;
    for (size_t k = 0; k < *n_masks && !clash; k++) {
      size_t both = 0;
      for (size_t j = 0; j < S->n_all; j++) both += mask[j] && masks[k][j];
      double ov = total ? (double)both / (double)total : 0.0;
      clash = ov >= 0.5;
    }
    if (clash) continue;
    size_t events = orc_event_count(S->time_all, S->n_all, row->period, row->epoch,
                                    row->duration, 3);
    if (events < min_events) continue;
    double leftover = without_strongest_event(S, row, S->family[i]);
    if (min_events > 1 && leftover < 0.5 * S->threshold) continue;
    uint8_t* wide = (uint8_t*)malloc(S->n_all + 1);
    orc_transit_mask(S->time_all, S->n_all, row->period, row->epoch, row->duration, margin, wide);
    for (size_t j = 0; j < S->n_all; j++) S->excluded[j] |= wide[j];
    masks[(*n_masks)++] = wide;
    if (*nout < cap) {
      orc_peak* p = &out[*nout];
      memset(p, 0, sizeof(*p));
      p->period_days = row->period;
      p->epoch_bjd = row->epoch;
      p->duration_hours = 24.0 * row->duration;
      p->snr = isolated;
      p->snr_folded = row->score;
      p->snr_without_strongest_event = leftover;
      p->depth = row->depth;
      p->depth_ppm = 1e6 * row->depth;
      p->events = events;
      p->family = S->family[i];
      p->pass = pass;
      p->row = i;
      p->planet_log_odds = p->log_bf_planet_noise = p->log_bf_binary_noise = NAN;
      p->binary_secondary_ratio = p->binary_secondary_phase = NAN;
      p->binary_odd_even_ratio = p->binary_period_days = p->ingress_fraction = NAN;
      p->evidence_epoch_bjd = p->evidence_duration_hours = NAN;
      (*nout)++;
    }
    accepted++;
    memset(fresh, 0, S->nrows);
    refresh(&R);
  }
  free(mask);
  free(remaining);
  free(fresh);
  free(heap.a);
}

/* select_peaks (peaks.rs:308-487) on a caller's table: the selection step
 * alone, for tests.  Family k: relative times, fixed-point fields and rho
 * at ft/fv/fw[foff[k] .. foff[k+1]), baseline norm fC[k], origin
 * forigin[k], durations/rho at fdur/frho[fdoff[k] .. fdoff[k+1]).
 * masks_in: n_masks_in x n_all accepted masks of earlier passes; excluded:
 * n_all flags, updated in place.  Returns the number of peaks written. */
size_t orc_select_peaks(size_t nfam, const double* ft, const double* fv, const double* fw,
                        const size_t* foff, const double* fC, const double* forigin,
                        const double* fdur, const double* frho, const size_t* fdoff,
                        const orc_row* table, const size_t* family, size_t nrows,
                        const double* time_all, size_t n_all, double threshold, size_t limit,
                        size_t min_events, const uint8_t* masks_in, size_t n_masks_in,
                        uint8_t* excluded, orc_peak* out, size_t cap) {
  fam_t* fams = (fam_t*)calloc(nfam + 1, sizeof(fam_t));
  size_t** pos = (size_t**)calloc(nfam + 1, sizeof(size_t*));
  for (size_t k = 0; k < nfam; k++) {
    fam_t* f = &fams[k];
    f->n = foff[k + 1] - foff[k];
    f->t = (double*)(ft + foff[k]);
    f->v = (double*)(fv + foff[k]);
    f->w = (double*)(fw + foff[k]);
    f->C = fC[k];
    f->origin = forigin[k];
    f->nd = fdoff[k + 1] - fdoff[k];
    f->dur = (double*)(fdur + fdoff[k]);
    f->rho = (double*)(frho + fdoff[k]);
    pos[k] = (size_t*)malloc((f->n + 1) * sizeof(size_t));
    for (size_t i = 0; i < f->n; i++) {
      size_t p = lower_bound(time_all, n_all, f->t[i] + f->origin);
      pos[k][i] = p < n_all - 1 ? p : n_all - 1;
    }
  }
  uint8_t** masks = (uint8_t**)calloc(n_masks_in + limit + 1, sizeof(uint8_t*));
  for (size_t k = 0; k < n_masks_in; k++) {
    masks[k] = (uint8_t*)malloc(n_all + 1);
    memcpy(masks[k], masks_in + k * n_all, n_all);
  }
  size_t n_masks = n_masks_in, nout = 0;
  sel_ctx S = {table, nrows, family, fams, nfam, pos, time_all, n_all, threshold, excluded};
  select_peaks(&S, limit, min_events, masks, &n_masks, 1, out, &nout, cap);
  for (size_t k = 0; k < n_masks; k++) free(masks[k]);
  free(masks);
  for (size_t k = 0; k < nfam; k++) free(pos[k]);
  free(pos);
  free(fams);
  return nout;
}

/* ------------------------------------------------------------ pipeline */

static int cmp_dbl(const void* a, const void* b) {
  double x = *(const double*)a, y = *(const double*)b;
  return x < y ? -1 : x > y;
}

/* pipeline.rs:70-221 over the families (curve k: fn[k] points at ft/ff/fs
 * + foff[k], searching the bank durations fidx[fmoff[k] .. fmoff[k + 1]]) */
static int search_families(size_t nfam, const double* ft, const double* ff, const double* fs,
                           const size_t* foff, const size_t* fn, const size_t* fidx,
                           const size_t* fmoff, const orc_pipeline_cfg* cfg, orc_peak* peaks,
                           size_t cap, size_t* npeaks, orc_pass_stats* pstats,
                           size_t* npasses, orc_events_out* events) {
  const double* dur = cfg->durations;
  int st = 0;
  fam_t* cur = (fam_t*)calloc(nfam + 1, sizeof(fam_t));
  fam_t* original = (fam_t*)calloc(nfam + 1, sizeof(fam_t));
  size_t built = 0;
  for (; built < nfam && !st; built++) {
    const size_t m = fmoff[built + 1] - fmoff[built];
    double* d = (double*)malloc((m + 1) * 8);
    for (size_t i = 0; i < m; i++) d[i] = dur[fidx[fmoff[built] + i]];
    st = fam_build(ft + foff[built], ff + foff[built], fs + foff[built], fn[built], d,
                   fidx + fmoff[built], m, cfg->correlated, &cur[built]);
    if (!st)
      st = fam_build(ft + foff[built], ff + foff[built], fs + foff[built], fn[built], d,
                     fidx + fmoff[built], m, cfg->correlated, &original[built]);
    free(d);
  }
  if (st) {
    for (size_t k = 0; k < nfam; k++) {
      fam_free(&cur[k]);
      fam_free(&original[k]);
    }
    free(cur);
    free(original);
    return st;
  }
  /* merged_time (pipeline.rs:16-24): sorted, deduplicated */
  size_t total = 0;
  for (size_t k = 0; k < nfam; k++) total += cur[k].cn;
  double* time_all = (double*)malloc((total + 1) * 8);
  size_t n_all = 0;
  for (size_t k = 0; k < nfam; k++) {
    memcpy(time_all + n_all, cur[k].ct, cur[k].cn * 8);
    n_all += cur[k].cn;
  }
  qsort(time_all, n_all, 8, cmp_dbl);
  size_t u = 0;
  for (size_t i = 0; i < n_all; i++)
    if (u == 0 || time_all[i] != time_all[u - 1]) time_all[u++] = time_all[i];
  n_all = u;
  uint8_t* excluded = (uint8_t*)calloc(n_all + 1, 1);
  uint8_t** masks = (uint8_t**)calloc(cfg->max_candidates + 1, sizeof(uint8_t*));
  size_t n_masks = 0;
  if (n_all < 6) st = 6;
  for (size_t pass = 0; !st && pass < cfg->passes; pass++) {
    size_t nrows = 0, grid_total = 0;
    size_t cap_rows = 1 << 22;
    orc_row* table = (orc_row*)malloc(cap_rows * sizeof(orc_row));
    size_t* family = (size_t*)malloc(cap_rows * sizeof(size_t));
    for (size_t k = 0; k < nfam && !st; k++) {
      size_t grid_count = 0, nr = 0, nbands = 0;
      orc_scan_stats bstats[64];
      st = orc_signal_search(cur[k].t, cur[k].v, cur[k].w, cur[k].n, cur[k].C, cur[k].origin,
                             cur[k].dur, cur[k].rho, cur[k].nd, cfg->ingress, cfg->ns, &cfg->bank,
                             &cfg->search, table + nrows, cap_rows - nrows, &nr, bstats, &nbands,
                             &grid_count);
      for (size_t i = 0; i < nr; i++) family[nrows + i] = k;
      nrows += nr;
      grid_total += grid_count;
    }
    if (st) {
      free(table);
      free(family);
      break;
    }
    orc_pass_stats* ps = &pstats[*npasses];
    memset(ps, 0, sizeof(*ps));
    ps->pass = pass + 1;
    ps->rows = nrows;
    ps->periods = grid_total;
    (*npasses)++;
    size_t** pos = (size_t**)malloc(nfam * sizeof(size_t*));
    for (size_t k = 0; k < nfam; k++) {
      pos[k] = (size_t*)malloc((cur[k].n + 1) * sizeof(size_t));
      for (size_t i = 0; i < cur[k].n; i++) {
        size_t p = lower_bound(time_all, n_all, cur[k].t[i] + cur[k].origin);
        pos[k][i] = p < n_all - 1 ? p : n_all - 1;
      }
    }
    sel_ctx S = {table, nrows, family, cur, nfam, pos, time_all, n_all, cfg->bank.threshold,
                 excluded};
    size_t before = *npeaks;
    select_peaks(&S, cfg->max_candidates - before, cfg->min_events, masks, &n_masks, pass + 1,
                 peaks, npeaks, cap);
    for (size_t k = 0; k < nfam; k++) free(pos[k]);
    free(pos);
    free(table);
    free(family);
    ps->accepted = *npeaks - before;
    if (*npeaks == before) break;
    orc_peak lead = peaks[before];
    if (pass + 1 < cfg->passes) {
      int small = 0;
      for (size_t k = 0; k < nfam && !st; k++) {
        fam_t next;
        size_t kept = 0;
        st = fam_masked(&cur[k], lead.period_days, lead.epoch_bjd, lead.duration_hours / 24.0,
                        0.5, cfg->correlated, &next, &kept);
        if (st) break;
        fam_free(&cur[k]);
        cur[k] = next;
        small |= kept < 6;
      }
      if (st || small) break;
    }
  }
  /* evidence (pipeline.rs:169-221) */
  if (!st && cfg->model_comparison && *npeaks) {
    size_t np = *npeaks;
    size_t* ord = (size_t*)malloc(np * sizeof(size_t));
    for (size_t i = 0; i < np; i++) ord[i] = i;
    for (size_t i = 1; i < np; i++) { /* stable sort by snr, descending */
      size_t x = ord[i], j = i;
      while (j > 0 && peaks[ord[j - 1]].snr < peaks[x].snr) {
        ord[j] = ord[j - 1];
        j--;
      }
      ord[j] = x;
    }
    for (size_t oi = 0; oi < np && oi < cfg->evidence_candidates; oi++) {
      orc_peak* p = &peaks[ord[oi]];
      const fam_t* base = &original[p->family];
      size_t cn = base->cn;
      double* ct = (double*)malloc((cn + 1) * 8);
      double* cf = (double*)malloc((cn + 1) * 8);
      double* cs = (double*)malloc((cn + 1) * 8);
      memcpy(ct, base->ct, cn * 8);
      memcpy(cf, base->cf, cn * 8);
      memcpy(cs, base->cs, cn * 8);
      uint8_t* m = (uint8_t*)malloc(cn + 1);
      for (size_t q = 0; q < np; q++) {
        const orc_peak* o = &peaks[q];
        if (o->row == p->row && o->pass == p->pass && o->family == p->family) continue;
        orc_transit_mask(ct, cn, o->period_days, o->epoch_bjd, o->duration_hours / 24.0, 0.5, m);
        size_t kk = 0;
        for (size_t i = 0; i < cn; i++)
          if (!m[i]) {
            ct[kk] = ct[i];
            cf[kk] = cf[i];
            cs[kk] = cs[i];
            kk++;
          }
        cn = kk;
      }
      free(m);
      if (cn >= 6) {
        fam_t s;
        st = fam_build(ct, cf, cs, cn, base->dur, base->idx, base->nd, cfg->correlated, &s);
        if (st == 0) {
          orc_evidence_row rows[2];
          size_t nr = 0;
          orc_compare_search(s.t, s.v, s.w, s.n, s.C, s.origin, s.energy, s.dur, s.rho, s.nd,
                             p->period_days, &cfg->evidence, rows, &nr);
          const orc_evidence_row* r = NULL;
          for (size_t q = 0; q < nr; q++)
            if (rows[q].selection_flags & 1) {
              r = &rows[q];
              break;
            }
          if (!r && nr) r = &rows[0];
          if (r) {
#define FIN(x) (isfinite(x) ? (x) : NAN)
            p->planet_log_odds = FIN(r->planet_log_odds);
            p->log_bf_planet_noise = FIN(r->log_bf_planet_noise);
            p->log_bf_binary_noise = FIN(r->log_bf_binary_noise);
            p->binary_secondary_ratio = FIN(r->binary_secondary_ratio);
            p->binary_secondary_phase = FIN(r->binary_secondary_phase);
            p->binary_odd_even_ratio = FIN(r->binary_odd_even_ratio);
            p->binary_period_days = FIN(r->binary_period_days);
            p->ingress_fraction = FIN(r->ingress_fraction);
#undef FIN
            p->evidence_epoch_bjd = r->epoch;
            p->evidence_duration_hours = 24.0 * r->duration;
            p->model_preferred = r->planet_log_odds >= cfg->planet_log_odds_min;
            p->binary_diagnostics_complete =
                isfinite(r->binary_secondary_ratio) && isfinite(r->binary_secondary_phase) &&
                isfinite(r->binary_odd_even_ratio) && isfinite(r->binary_period_days);
          }
          fam_free(&s);
        }
      }
      free(ct);
      free(cf);
      free(cs);
      if (st) break;
    }
    free(ord);
  }
  /* single events (pipeline.rs:222-238) of the families as the passes left them */
  if (!st && cfg->single_events && events) {
    const size_t capacity = 131072;
    const size_t limit = cfg->max_candidates;
    orc_row* rows = (orc_row*)malloc(capacity * sizeof(orc_row));
    orc_row* all = (orc_row*)malloc((nfam * limit + 1) * sizeof(orc_row));
    orc_single_event* sel = (orc_single_event*)malloc((limit + 1) * sizeof(orc_single_event));
    size_t nall = 0;
    events->nstats = 0;
    for (size_t k = 0; k < nfam; k++) {
      const fam_t* f = &cur[k];
      size_t nr = 0;
      orc_event_stats es;
      orc_single_event_search(f->t, f->v, f->w, f->n, f->C, f->origin, f->dur, f->rho, f->nd,
                              cfg->ingress, cfg->ns, cfg->bank.layers, cfg->bank.threshold,
                              cfg->bank.phase_oversample, capacity, rows, &nr, &es);
      if (events->stats && events->nstats < events->stats_cap) events->stats[events->nstats] = es;
      events->nstats++;
      size_t m = orc_select_events(rows, nr, time_all, n_all, cfg->bank.threshold, limit, masks,
                                   n_masks, sel);
      for (size_t q = 0; q < m; q++) {
        orc_row* r = &all[nall++];
        r->period = INFINITY;
        r->epoch = sel[q].epoch_bjd;
        r->duration = sel[q].duration_hours / 24.0;
        r->score = sel[q].snr;
        r->depth = sel[q].depth;
      }
    }
    for (size_t i = 1; i < nall; i++) { /* stable sort by score, descending */
      orc_row x = all[i];
      size_t j = i;
      while (j > 0 && all[j - 1].score < x.score) {
        all[j] = all[j - 1];
        j--;
      }
      all[j] = x;
    }
    size_t m = orc_select_events(all, nall, time_all, n_all, cfg->bank.threshold, limit, masks,
                                 n_masks, sel);
    events->nevents = m;
    for (size_t q = 0; q < m && q < events->events_cap; q++) events->events[q] = sel[q];
    free(rows);
    free(all);
    free(sel);
  }
  for (size_t i = 0; i < n_masks; i++) free(masks[i]);
  free(masks);
  free(excluded);
  free(time_all);
  for (size_t k = 0; k < nfam; k++) {
    fam_free(&cur[k]);
    fam_free(&original[k]);
  }
  free(cur);
  free(original);
  return st;
}

int orc_search_prepared(const double* time, const double* flux,
                        const double* err, size_t n, const orc_pipeline_cfg* cfg,
                        orc_peak* peaks, size_t cap, size_t* npeaks,
                        orc_pass_stats* pstats, size_t* npasses, orc_events_out* events) {
  *npeaks = 0;
  *npasses = 0;
  /* pipeline.rs:47-61: finite, positive error, sorted by time (stable) */
  size_t* order = (size_t*)malloc((n + 1) * sizeof(size_t));
  size_t k = 0;
  for (size_t i = 0; i < n; i++)
    if (isfinite(time[i]) && isfinite(flux[i]) && isfinite(err[i]) && err[i] > 0.0) order[k++] = i;
  for (size_t i = 1; i < k; i++) { /* stable insertion sort by time */
    size_t x = order[i], j = i;
    while (j > 0 && time[order[j - 1]] > time[x]) {
      order[j] = order[j - 1];
      j--;
    }
    order[j] = x;
  }
  double* ct = (double*)malloc((k + 1) * 8);
  double* cf = (double*)malloc((k + 1) * 8);
  double* cs = (double*)malloc((k + 1) * 8);
  for (size_t i = 0; i < k; i++) {
    ct[i] = time[order[i]];
    cf[i] = flux[order[i]];
    cs[i] = err[order[i]];
  }
  free(order);
  size_t* idx = (size_t*)malloc((cfg->nd + 1) * sizeof(size_t));
  for (size_t i = 0; i < cfg->nd; i++) idx[i] = i;
  const size_t foff[1] = {0}, fn[1] = {k}, fmoff[2] = {0, cfg->nd};
  int st = search_families(1, ct, cf, cs, foff, fn, idx, fmoff, cfg, peaks, cap, npeaks, pstats,
                           npasses, events);
  free(idx);
  free(ct);
  free(cf);
  free(cs);
  return st;
}

/* pipeline.rs:27-37: prepare_families (curve.rs:387-425) then the search */
int orc_search_transits(const double* time, const double* flux, const double* err,
                        const int64_t* group, size_t n, const orc_prep_cfg* prep,
                        const orc_pipeline_cfg* cfg, orc_peak* peaks, size_t cap,
                        size_t* npeaks, orc_pass_stats* pstats, size_t* npasses,
                        orc_events_out* events) {
  *npeaks = 0;
  *npasses = 0;
  const size_t nd = cfg->nd;
  size_t nfam = 0;
  size_t* fn = (size_t*)malloc((nd + 1) * sizeof(size_t));
  double* ft = (double*)malloc((nd * n + 1) * 8);
  double* ff = (double*)malloc((nd * n + 1) * 8);
  double* fs = (double*)malloc((nd * n + 1) * 8);
  double* fw = (double*)malloc((nd + 1) * 8);
  size_t* fc = (size_t*)malloc((nd + 1) * sizeof(size_t));
  uint8_t* mem = (uint8_t*)malloc(nd * nd + 1);
  int st = orc_prepare_families(time, flux, err, group, n, cfg->durations, nd, prep, &nfam, fn,
                                ft, ff, fs, fw, fc, mem);
  if (!st) {
    size_t* foff = (size_t*)malloc((nfam + 1) * sizeof(size_t));
    size_t* fidx = (size_t*)malloc((nfam * nd + 1) * sizeof(size_t));
    size_t* fmoff = (size_t*)malloc((nfam + 2) * sizeof(size_t));
    size_t m = 0;
    for (size_t k = 0; k < nfam; k++) {
      foff[k] = k * n;
      fmoff[k] = m;
      for (size_t i = 0; i < nd; i++)
        if (mem[k * nd + i]) fidx[m++] = i;
    }
    fmoff[nfam] = m;
    st = search_families(nfam, ft, ff, fs, foff, fn, fidx, fmoff, cfg, peaks, cap, npeaks, pstats,
                         npasses, events);
    free(foff);
    free(fidx);
    free(fmoff);
  }
  free(fn);
  free(ft);
  free(ff);
  free(fs);
  free(fw);
  free(fc);
  free(mem);
  return st;
}

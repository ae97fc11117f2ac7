/*
 * FP64 CPU oracle of the coarse periodogram that feeds the band-proposal
 * network (the library's fbls_coarse_features; no counterpart in the
 * reference).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of feature_kernel.hip.  Sums are
 * sequential FP64 sums of the fixed-point field values, which are exact, so
 * the result is bit-identical to the device's integer sums.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "fbls_oracle.h"

void orc_coarse_features(const double* t, const double* v, const double* w, size_t n,
                         double C, const double* periods, size_t np,
                         const double* durations, size_t ng, uint32_t nb_min,
                         uint32_t nb_max, double* out) {
  double d0 = INFINITY;
  for (size_t g = 0; g < ng; g++) d0 = fmin(d0, durations[g]);
  double* cv = (double*)malloc((nb_max + 1) * sizeof(double));
  double* cw = (double*)malloc((nb_max + 1) * sizeof(double));
  uint32_t* cn = (uint32_t*)malloc((nb_max + 1) * sizeof(uint32_t));
  for (size_t i = 0; i < np; i++) {
    double want = ceil(2.0 * periods[i] / d0);
    uint32_t nb = (uint32_t)fmin((double)nb_max, fmax((double)nb_min, want));
    double scale = (double)nb / periods[i];
    for (uint32_t b = 0; b <= nb; b++) {
      cv[b] = 0.0;
      cw[b] = 0.0;
      cn[b] = 0;
    }
    for (size_t o = 0; o < n; o++) {
      uint64_t cell = (uint64_t)floor(t[o] * scale);
      uint32_t b = (uint32_t)(cell % nb) + 1;
      cv[b] += v[o];
      cw[b] += w[o];
      cn[b] += 1;
    }
    for (uint32_t b = 1; b <= nb; b++) {
      cv[b] += cv[b - 1];
      cw[b] += cw[b - 1];
      cn[b] += cn[b - 1];
    }
    for (size_t g = 0; g < ng; g++) {
      double kf = floor(durations[g] * scale + 0.5);
      uint32_t k = kf < 1.0 ? 1u : (uint32_t)kf;
      if (k > nb) k = nb;
      /* the qualifying phase with the largest key a|a|/q (first on ties),
       * then its SNR */
      double best_key = -INFINITY, best = -INFINITY;
      for (uint32_t j = 0; j < nb; j++) {
        uint32_t e = j + k, cnt;
        double a, ww;
        if (e <= nb) {
          a = cv[e] - cv[j];
          ww = cw[e] - cw[j];
          cnt = cn[e] - cn[j];
        } else {
          a = cv[e - nb] + (cv[nb] - cv[j]);
          ww = cw[e - nb] + (cw[nb] - cw[j]);
          cnt = cn[e - nb] + (cn[nb] - cn[j]);
        }
        double q = ww - ww * ww / C;
        if (cnt >= 3 && cnt + 3 <= cn[nb] && q > 0.0) {
          double key = a * fabs(a) / q;
          if (key > best_key) {
            best_key = key;
            best = a / sqrt(q);
          }
        }
      }
      out[i * ng + g] = best;
    }
  }
  free(cv);
  free(cw);
  free(cn);
}

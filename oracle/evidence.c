/*
 * FP64 CPU oracle of the FBLS evidence (model comparison) at one period.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP evidence kernels.
 * Restates src/search/evidence.rs and src/search/math.rs of the reference
 * (isakruas/fbls); citations are path:line there.  Build with
 * -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fbls_oracle.h"

#define EV_TILE 64
#define PART 15
#define NORB 3
#define NPARAM 30

/* ---------------------------------------------------------------- math */

/* exp and log of the restatement: the reference evaluates them with the
 * platform libm (host) and the CUDA libdevice (kernels).  The oracle and the
 * HIP path both use the two routines below, written from the mathematics,
 * so that their evidence values agree bit for bit; orc_evidence_use_libm(1)
 * switches the oracle to glibc to show that the substitution moves no
 * result by more than FP64 rounding.
 *
 * exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split
 * into a 40-bit head (k * head exact for |k| < 4096) and a tail; e^r - 1 by
 * its Taylor series to r^13 (remainder < 4e-18 relative); 2^k by ldexp.
 * log: x = m 2^e, m in [sqrt(1/2), sqrt(2)); f = m - 1 (exact),
 * s = f / (2 + f), log(1 + f) = f - (f^2/2 - s (f^2/2 + R)) with
 * R = sum_{i=1..11} 2 s^(2i) / (2i + 1) (2 atanh(s) = 2s + s R, and
 * 2s = f - s f); truncation < 4e-20 relative. */
static int g_libm = 0;
void orc_evidence_use_libm(int on) { g_libm = on; }

static const double kLn2Hi = 0x1.62e42fefa2000p-1;
static const double kLn2Lo = 0x1.9ef35793c7673p-41;

double orc_exp(double x) {
  if (x != x) return x + x;
  if (x > 709.782712893383973096) return INFINITY;
  if (x < -745.133219101941108420) return 0.0;
  double kf = rint(x * 1.4426950408889634);
  double r = (x - kf * kLn2Hi) - kf * kLn2Lo;
  double q = 1.0 / 6227020800.0;                 /* 1/13! */
  q = q * r + 1.0 / 479001600.0;                 /* 1/12! */
  q = q * r + 1.0 / 39916800.0;
  q = q * r + 1.0 / 3628800.0;
  q = q * r + 1.0 / 362880.0;
  q = q * r + 1.0 / 40320.0;
  q = q * r + 1.0 / 5040.0;
  q = q * r + 1.0 / 720.0;
  q = q * r + 1.0 / 120.0;
  q = q * r + 1.0 / 24.0;
  q = q * r + 1.0 / 6.0;
  q = q * r + 0.5;                               /* 1/2! */
  double em1 = r + (r * r) * q;
  return ldexp(1.0 + em1, (int)kf);
}

double orc_log(double x) {
  if (x != x || x < 0.0) return NAN;
  if (x == 0.0) return -INFINITY;
  if (x == INFINITY) return x;
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654757) {
    m = m * 2.0;
    e -= 1;
  }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  double R = 2.0 / 23.0;
  R = R * z + 2.0 / 21.0;
  R = R * z + 2.0 / 19.0;
  R = R * z + 2.0 / 17.0;
  R = R * z + 2.0 / 15.0;
  R = R * z + 2.0 / 13.0;
  R = R * z + 2.0 / 11.0;
  R = R * z + 2.0 / 9.0;
  R = R * z + 2.0 / 7.0;
  R = R * z + 2.0 / 5.0;
  R = R * z + 2.0 / 3.0;
  R = R * z;
  double hfsq = 0.5 * f * f;
  double dk = (double)e;
  return dk * kLn2Hi - ((hfsq - (s * (hfsq + R) + dk * kLn2Lo)) - f);
}

#define EXP(x) (g_libm ? exp(x) : orc_exp(x))
#define LOG(x) (g_libm ? log(x) : orc_log(x))

/* W. J. Cody, "Rational Chebyshev approximations for the error function",
 * Math. Comp. 23 (1969); the coefficient table is Cody's published one
 * (as used at math.rs:8-100). */
double orc_erfc(double x) {
  static const double A[5] = {3.16112374387056560e0, 1.13864154151050156e2,
                              3.77485237685302021e2, 3.20937758913846947e3,
                              1.85777706184603153e-1};
  static const double B[4] = {2.36012909523441209e1, 2.44024637934444173e2,
                              1.28261652607737228e3, 2.84423683343917062e3};
  static const double C[9] = {5.64188496988670089e-1, 8.88314979438837594e0,
                              6.61191906371416295e1,  2.98635138197400131e2,
                              8.81952221241769090e2,  1.71204761263407058e3,
                              2.05107837782607147e3,  1.23033935479799725e3,
                              2.15311535474403846e-8};
  static const double D[8] = {1.57449261107098347e1, 1.17693950891312499e2,
                              5.37181101862009858e2, 1.62138957456669019e3,
                              3.29079923573345963e3, 4.36261909014324716e3,
                              3.43936767414372164e3, 1.23033935480374942e3};
  static const double P[6] = {3.05326634961232344e-1, 3.60344899949804439e-1,
                              1.25781726111229246e-1, 1.60837851487422766e-2,
                              6.58749161529837803e-4, 1.63153871373020978e-2};
  static const double Q[5] = {2.56852019228982242e0, 1.87295284992346725e0,
                              5.27905102951428412e-1, 6.05183413124413191e-2,
                              2.33520497626869185e-3};
  const double sqrpi = 5.6418958354775628695e-1;
  double y = fabs(x);
  double result;
  if (y <= 0.46875) {
    double ysq = y > 1.11e-16 ? y * y : 0.0;
    double xnum = A[4] * ysq, xden = ysq;
    for (int i = 0; i < 3; i++) {
      xnum = (xnum + A[i]) * ysq;
      xden = (xden + B[i]) * ysq;
    }
    return 1.0 - x * (xnum + A[3]) / (xden + B[3]);
  } else if (y <= 4.0) {
    double xnum = C[8] * y, xden = y;
    for (int i = 0; i < 7; i++) {
      xnum = (xnum + C[i]) * y;
      xden = (xden + D[i]) * y;
    }
    double r = (xnum + C[7]) / (xden + D[7]);
    double ysq = trunc(y * 16.0) / 16.0;
    double del = (y - ysq) * (y + ysq);
    result = EXP(-ysq * ysq) * EXP(-del) * r;
  } else if (y < 26.543) {
    double ysq = 1.0 / (y * y);
    double xnum = P[5] * ysq, xden = ysq;
    for (int i = 0; i < 4; i++) {
      xnum = (xnum + P[i]) * ysq;
      xden = (xden + Q[i]) * ysq;
    }
    double r = ysq * (xnum + P[4]) / (xden + Q[4]);
    r = (sqrpi - r) / y;
    double ysq2 = trunc(y * 16.0) / 16.0;
    double del = (y - ysq2) * (y + ysq2);
    result = EXP(-ysq2 * ysq2) * EXP(-del) * r;
  } else {
    result = 0.0;
  }
  return x < 0.0 ? 2.0 - result : result;
}

/* math.rs:103-110 */
double orc_log_add(double x, double y) {
  double high = fmax(x, y);
  if (high == -INFINITY) return high;
  return high + LOG(1.0 + EXP(fmin(x, y) - high));
}

/* math.rs:115-140 */
double orc_positive_amplitude_evidence(double a, double q, double tau) {
  q = fmax(q, 0.0);
  double precision = 1.0 + tau * tau * q;
  double z = tau * a / sqrt(precision);
  if (z < -26.0) {
    double zz = fmax(z * z, 1.0);
    double inv = 1.0 / zz;
    double series =
        1.0 + inv * (-1.0 + inv * (3.0 + inv * (-15.0 + inv * (105.0 +
              inv * (-945.0 + inv * (10395.0 - inv * 135135.0))))));
    return -0.5 * LOG(precision) - LOG(fmax(-z, 1.0)) - 0.22579135264472743236 +
           LOG(series);
  }
  return -0.5 * LOG(precision) + 0.5 * z * z +
         LOG(fmax(orc_erfc(-z * 0.70710678118654752440), 1e-300));
}

/* math.rs:145-177 */
double orc_subspace_upper(double a, double b, double q, double r, double cross,
                          double rho2) {
  double product = q * r;
  double det = q * r - cross * cross;
  double energy = (a * a * r + b * b * q - 2.0 * a * b * cross) / fmax(det * rho2, 1e-300);
  int safe = q > 0.0 && r > 0.0 && det > 1e-8 * product;
  double pa = fmax(a, 0.0), pb = fmax(b, 0.0);
  double boundary = 0.5 * fmax(pa * pa / fmax(q * rho2, 1e-300),
                               pb * pb / fmax(r * rho2, 1e-300));
  int interior = a * r - b * cross > 0.0 && b * q - a * cross > 0.0;
  double upper = safe ? (interior ? 0.5 * fmax(energy, 0.0) : boundary) : INFINITY;
  int only_a = r == 0.0 && b == 0.0 && cross == 0.0;
  int only_b = q == 0.0 && a == 0.0 && cross == 0.0;
  if (only_a && q > 0.0) upper = 0.5 * a * a / (q * rho2);
  if (only_b && r > 0.0) upper = 0.5 * b * b / (r * rho2);
  if (only_a && only_b) upper = 0.0;
  if (a <= 0.0 && b <= 0.0) upper = 0.0;
  return upper * 1.000001 + 1e-6;
}

/* ------------------------------------------------------------- sums */

/* evidence.rs:37-45 */
static double circular_sum(const double* cum, size_t start, size_t width,
                           size_t nb2) {
  size_t j = start % nb2;
  size_t end = j + width;
  double left = cum[j];
  double right = cum[end < nb2 ? end : nb2];
  double wrap = cum[end > nb2 ? end - nb2 : 0];
  return right - left + wrap;
}

/* evidence.rs:48-67 */
static double layered_sum(const double* cum, size_t start, size_t width,
                          const double* shapes, size_t si, size_t nb2,
                          size_t layers, int squared) {
  double value = 0.0;
  for (size_t layer = 0; layer < layers; layer++) {
    double fraction = shapes[(si * layers + layer) * 3];
    double inner_f = floor((double)width * fraction + 0.5);
    size_t inner = inner_f < 1.0 ? 1 : (size_t)inner_f;
    double coefficient = shapes[(si * layers + layer) * 3 + (squared ? 2 : 1)];
    value += coefficient * circular_sum(cum, start + (width - inner) / 2, inner, nb2);
  }
  return value;
}

static double logsumexp(const double* v, size_t n) {
  double high = -INFINITY;
  for (size_t i = 0; i < n; i++) high = fmax(high, v[i]);
  if (high == -INFINITY) return high;
  double s = 0.0;
  for (size_t i = 0; i < n; i++) s += EXP(v[i] - high);
  return high + LOG(s);
}

/* evidence.rs:665-689 */
void orc_evidence_parameters(const orc_evidence_cfg* cfg, double* p) {
  double dilution = cfg->target_flux_fraction;
  double planet_scale = cfg->planet_depth_scale;
  if (cfg->stellar_radius_solar > 0.0) {
    double x = 0.10045 * cfg->planet_radius_scale_rj / cfg->stellar_radius_solar;
    planet_scale = x * x;
  }
  p[0] = planet_scale * dilution;
  p[1] = cfg->binary_depth_scale * dilution;
  p[2] = cfg->diluted_binary_depth_scale;
  p[3] = LOG(cfg->planet_prior);
  p[4] = LOG(cfg->binary_prior);
  p[5] = LOG(1.0 - cfg->planet_prior - cfg->binary_prior);
  static const double ratios[5] = {0.0, 0.01, 0.1, 0.5, 1.0};
  static const double wp[5] = {0.75, 0.15, 0.08, 0.015, 0.005};
  static const double wb[5] = {0.10, 0.10, 0.20, 0.25, 0.35};
  for (int i = 0; i < 5; i++) {
    p[6 + 3 * i] = ratios[i];
    p[7 + 3 * i] = wp[i];
    p[8 + 3 * i] = wb[i];
  }
  static const double odd[7] = {0.25, 0.5, 0.75, 1.0, 4.0 / 3.0, 2.0, 4.0};
  for (int i = 0; i < 7; i++) p[21 + i] = odd[i];
  p[28] = 0.001;
  p[29] = 0.05;
}

/* ------------------------------------------------------------ kernels */

typedef struct {
  const double* cum; /* 3 x (nb2 + 1): v, w, count */
  size_t nb, nb2, nd, ns, layers;
  double scale, c, n, threshold;
  const double *dur, *rho, *dlm, *shapes, *priors, *params, *orbit;
} ev_ctx;

/* evidence.rs:151-273 for one tile: out[4] = {max score, lower_p, lower_b,
 * key}; left untouched (zeros) when no lane has a valid template. */
static void screen_tile(const ev_ctx* E, size_t tile, double* out) {
  double best[EV_TILE], merit_l[EV_TILE];
  double ba[EV_TILE], bq[EV_TILE], bnull[EV_TILE], bps[EV_TILE], bbs[EV_TILE];
  uint64_t bkey[EV_TILE];
  const size_t nb = E->nb, nb2 = E->nb2, stride = nb2 + 1;
  const double* cv = E->cum;
  const double* cw = E->cum + stride;
  const double* cn = E->cum + 2 * stride;
  for (size_t lane = 0; lane < EV_TILE; lane++) {
    size_t j = tile * EV_TILE + lane;
    size_t jj = j < nb ? j : nb - 1;
    double bst = -INFINITY, bm = -INFINITY;
    double best_a = 0.0, best_q = 1.0, best_null = 0.0, best_ps = 0.0, best_bs = 0.0;
    uint64_t best_key = 0;
    for (size_t di = 0; di < E->nd; di++) {
      double rk = floor(E->dur[di] * E->scale + 0.5);
      size_t raw_k = rk < 1.0 ? 1 : (size_t)rk;
      size_t k = raw_k < nb2 / 8 - 1 ? raw_k : nb2 / 8 - 1;
      double r = E->rho[di];
      double count = circular_sum(cn, jj, k, nb2) + circular_sum(cn, jj + nb, k, nb2);
      int usable = j < nb && raw_k < nb2 / 8 && count >= 3.0 && count <= E->n - 3.0;
      for (size_t si = 0; si < E->ns; si++) {
        double a = layered_sum(cv, jj, k, E->shapes, si, nb2, E->layers, 0) +
                   layered_sum(cv, jj + nb, k, E->shapes, si, nb2, E->layers, 0);
        double pe = layered_sum(cw, jj, k, E->shapes, si, nb2, E->layers, 0);
        double po = layered_sum(cw, jj + nb, k, E->shapes, si, nb2, E->layers, 0);
        double ne = layered_sum(cw, jj, k, E->shapes, si, nb2, E->layers, 1);
        double no = layered_sum(cw, jj + nb, k, E->shapes, si, nb2, E->layers, 1);
        double qe = ne - pe * pe / E->c, qo = no - po * po / E->c, qeo = -pe * po / E->c;
        double q = qe + qo + 2.0 * qeo;
        double score = a / sqrt(fmax(q * (r * r), 1e-300));
        int valid = usable && a > 0.0 && q > 0.0;
        double merit = 0.5 * score * score + E->dlm[di] + E->priors[si * 2];
        if (valid && merit > bm) {
          best_a = a / (r * r);
          best_q = q / (r * r);
          best_null = E->dlm[di];
          best_ps = E->priors[si * 2];
          best_bs = E->priors[si * 2 + 1];
          best_key = (uint64_t)(si * E->nd + di) * nb + jj;
          bm = merit;
        }
        if (valid && score > bst) bst = score;
      }
    }
    best[lane] = bst;
    merit_l[lane] = bm;
    ba[lane] = best_a;
    bq[lane] = best_q;
    bnull[lane] = best_null;
    bps[lane] = best_ps;
    bbs[lane] = best_bs;
    bkey[lane] = best_key;
  }
  double maximum = -INFINITY, bound_maximum = -INFINITY;
  size_t chosen = EV_TILE;
  for (size_t l = 0; l < EV_TILE; l++) {
    if (best[l] > maximum) maximum = best[l];
    if (merit_l[l] > bound_maximum) {
      bound_maximum = merit_l[l];
      chosen = l;
    }
  }
  if (chosen == EV_TILE) return;
  double lower_p = orc_positive_amplitude_evidence(ba[chosen], bq[chosen], E->params[0]) +
                   LOG(0.75) + bnull[chosen] + bps[chosen];
  double lower_b = orc_log_add(orc_positive_amplitude_evidence(ba[chosen], bq[chosen], E->params[1]),
                               orc_positive_amplitude_evidence(ba[chosen], bq[chosen], E->params[2])) +
                   LOG(0.025) + bnull[chosen] + bbs[chosen];
  out[0] = maximum;
  out[1] = bound_maximum > -INFINITY ? lower_p : -INFINITY;
  out[2] = bound_maximum > -INFINITY ? lower_b : -INFINITY;
  out[3] = (double)bkey[chosen];
}

typedef struct {
  double best_joint, best_signal, sum_p, sum_b;
  uint64_t joint_key, signal_key;
  double j_out[9], s_out[9];
  uint32_t omitted;
  /* per-lane values that enter tile-wide maxima */
  double upper;
} joint_lane;

/* evidence.rs:280-632 for one tile (pilot: every lane evaluates pilot_key) */
static void joint_tile(const ev_ctx* E, size_t tile, int pilot, double floor_p,
                       double floor_b, uint64_t pilot_key, double* parts) {
  const size_t nb = E->nb, nb2 = E->nb2, stride = nb2 + 1, nd = E->nd, ns = E->ns;
  const size_t layers = E->layers;
  const double* cv = E->cum;
  const double* cw = E->cum + stride;
  const double* cn = E->cum + 2 * stride;
  const double tau_p = E->params[0], tau_b = E->params[1], tau_d = E->params[2];
  if (pilot) {
    floor_p = -INFINITY;
    floor_b = -INFINITY;
  }
  joint_lane L[EV_TILE];
  size_t jl[EV_TILE], jjl[EV_TILE];
  for (size_t lane = 0; lane < EV_TILE; lane++) {
    joint_lane* s = &L[lane];
    s->best_joint = s->best_signal = -INFINITY;
    s->joint_key = s->signal_key = UINT64_MAX;
    memset(s->j_out, 0, sizeof(s->j_out));
    memset(s->s_out, 0, sizeof(s->s_out));
    s->sum_p = s->sum_b = -INFINITY;
    s->omitted = 0;
    jl[lane] = pilot ? (size_t)(pilot_key % nb) : tile * EV_TILE + lane;
    jjl[lane] = jl[lane] < nb ? jl[lane] : nb - 1;
  }
  size_t di_begin = 0, di_end = nd, si_begin = 0, si_end = ns;
  if (pilot) {
    di_begin = (size_t)((pilot_key / nb) % nd);
    di_end = di_begin + 1;
    si_begin = (size_t)(pilot_key / ((uint64_t)nb * nd));
    si_end = si_begin + 1;
  }
  /* per-lane scratch for one (di, si) */
  double ae[EV_TILE], ao[EV_TILE], pe[EV_TILE], po[EV_TILE], qe[EV_TILE], qo[EV_TILE],
      qeo[EV_TILE], ap[EV_TILE], qp[EV_TILE], snr[EV_TILE], depth[EV_TILE];
  int valid[EV_TILE];
  double lp[EV_TILE], lb[EV_TILE], comp[EV_TILE], sec_ratio[EV_TILE], sec_phase[EV_TILE],
      odd_ratio[EV_TILE], bin_mult[EV_TILE], om_upper[EV_TILE], upper[EV_TILE],
      raw_upper[EV_TILE], pa_[EV_TILE], pb_[EV_TILE], pq_[EV_TILE], pr_[EV_TILE], px_[EV_TILE];
  for (size_t di = di_begin; di < di_end; di++) {
    double rk = floor(E->dur[di] * E->scale + 0.5);
    size_t raw_k = rk < 1.0 ? 1 : (size_t)rk;
    size_t k = raw_k < nb2 / 8 - 1 ? raw_k : nb2 / 8 - 1;
    double rho2 = E->rho[di] * E->rho[di];
    double dlm = E->dlm[di];
    for (size_t si = si_begin; si < si_end; si++) {
      for (size_t lane = 0; lane < EV_TILE; lane++) {
        size_t j = jl[lane], jj = jjl[lane];
        double count = circular_sum(cn, jj, k, nb2) + circular_sum(cn, jj + nb, k, nb2);
        int usable = j < nb && raw_k < nb2 / 8 && count >= 3.0 && count <= E->n - 3.0;
        ae[lane] = layered_sum(cv, jj, k, E->shapes, si, nb2, layers, 0);
        ao[lane] = layered_sum(cv, jj + nb, k, E->shapes, si, nb2, layers, 0);
        pe[lane] = layered_sum(cw, jj, k, E->shapes, si, nb2, layers, 0);
        po[lane] = layered_sum(cw, jj + nb, k, E->shapes, si, nb2, layers, 0);
        double ne = layered_sum(cw, jj, k, E->shapes, si, nb2, layers, 1);
        double no = layered_sum(cw, jj + nb, k, E->shapes, si, nb2, layers, 1);
        qe[lane] = ne - pe[lane] * pe[lane] / E->c;
        qo[lane] = no - po[lane] * po[lane] / E->c;
        qeo[lane] = -pe[lane] * po[lane] / E->c;
        ap[lane] = ae[lane] + ao[lane];
        qp[lane] = qe[lane] + qo[lane] + 2.0 * qeo[lane];
        snr[lane] = ap[lane] / sqrt(fmax(qp[lane] * rho2, 1e-300));
        depth[lane] = ap[lane] / fmax(qp[lane], 1e-300);
        valid[lane] = usable && ap[lane] > 0.0 && qp[lane] > 0.0 && snr[lane] >= E->threshold;
      }
      if (!(raw_k < nb2 / 8)) continue;
      for (size_t lane = 0; lane < EV_TILE; lane++) {
        lp[lane] = -INFINITY;
        lb[lane] = -INFINITY;
        comp[lane] = -INFINITY;
        sec_ratio[lane] = 0.0;
        sec_phase[lane] = 0.5;
        odd_ratio[lane] = 1.0;
        bin_mult[lane] = 1.0;
        om_upper[lane] = -INFINITY;
      }
      for (size_t o = 0; o < NORB; o++) {
        const double* orb = E->orbit + (di * NORB + o) * 4;
        double phase = orb[0], duration_ratio = orb[1], logwp = orb[2], logwb = orb[3];
        int64_t offset = (int64_t)floor(phase * (double)nb + 0.5);
        double skf = floor((double)k * duration_ratio + 0.5);
        int64_t ski = skf < 1.0 ? 1 : (int64_t)skf;
        if (ski > (int64_t)nb - 1) ski = (int64_t)nb - 1;
        size_t sk = (size_t)ski;
        for (size_t lane = 0; lane < EV_TILE; lane++) {
          size_t jj = jjl[lane];
          size_t s0 = (size_t)((int64_t)jj + offset +
                               (int64_t)floor(((double)k - (double)sk) / 2.0));
          size_t s1 = s0 + nb;
          double ass = layered_sum(cv, s0, sk, E->shapes, si, nb2, layers, 0) +
                       layered_sum(cv, s1, sk, E->shapes, si, nb2, layers, 0);
          double ps = layered_sum(cw, s0, sk, E->shapes, si, nb2, layers, 0) +
                      layered_sum(cw, s1, sk, E->shapes, si, nb2, layers, 0);
          double nsum = layered_sum(cw, s0, sk, E->shapes, si, nb2, layers, 1) +
                        layered_sum(cw, s1, sk, E->shapes, si, nb2, layers, 1);
          double qs = nsum - ps * ps / E->c, qps = -(pe[lane] + po[lane]) * ps / E->c;
          pa_[lane] = ap[lane] / rho2;
          pb_[lane] = ass / rho2;
          pq_[lane] = qp[lane] / rho2;
          pr_[lane] = qs / rho2;
          px_[lane] = qps / rho2;
          double cu = 1e-6;
          double pu = orc_subspace_upper(pa_[lane], pb_[lane], pq_[lane], pr_[lane], px_[lane], 1.0) + logwp + cu;
          double bu = orc_subspace_upper(pa_[lane], pb_[lane], pq_[lane], pr_[lane], px_[lane], 1.0) + logwb + cu;
          double up = fmax(pu - floor_p, bu - floor_b) + dlm;
          if (!(logwp > -INFINITY || logwb > -INFINITY)) up = -INFINITY;
          raw_upper[lane] = logwb > -INFINITY ? bu : -INFINITY;
          upper[lane] = jl[lane] < nb ? up : -INFINITY;
        }
        double tile_max = -INFINITY;
        for (size_t l = 0; l < EV_TILE; l++)
          if (upper[l] > tile_max) tile_max = upper[l];
        for (size_t lane = 0; lane < EV_TILE; lane++) {
          if (tile_max >= 0.0) {
            for (int ri = 0; ri < 5; ri++) {
              double ratio = E->params[6 + ri * 3];
              double wp = E->params[7 + ri * 3], wb = E->params[8 + ri * 3];
              double aa = pa_[lane] + ratio * pb_[lane];
              double qq = pq_[lane] + 2.0 * ratio * px_[lane] + ratio * ratio * pr_[lane];
              double ep = orc_positive_amplitude_evidence(aa, qq, tau_p) + LOG(wp) + logwp;
              double diluted = orc_positive_amplitude_evidence(aa, qq, tau_d);
              double eb = orc_log_add(orc_positive_amplitude_evidence(aa, qq, tau_b), diluted) +
                          LOG(0.25 * wb) + logwb;
              lp[lane] = orc_log_add(lp[lane], ep);
              lb[lane] = orc_log_add(lb[lane], eb);
              if (eb > comp[lane]) {
                sec_ratio[lane] = ratio;
                sec_phase[lane] = phase;
                comp[lane] = eb;
              }
            }
          } else {
            L[lane].omitted += 1;
            if (raw_upper[lane] > om_upper[lane]) om_upper[lane] = raw_upper[lane];
          }
        }
      }
      double alias_mass = k < nb2 / 8 ? 0.0 : -INFINITY;
      for (size_t lane = 0; lane < EV_TILE; lane++) {
        pa_[lane] = ae[lane] / rho2;
        pb_[lane] = ao[lane] / rho2;
        pq_[lane] = qe[lane] / rho2;
        pr_[lane] = qo[lane] / rho2;
        px_[lane] = qeo[lane] / rho2;
        raw_upper[lane] = orc_subspace_upper(pa_[lane], pb_[lane], pq_[lane], pr_[lane], px_[lane], 1.0) +
                          1e-6 + alias_mass;
        upper[lane] = jl[lane] < nb ? raw_upper[lane] + dlm : -INFINITY;
      }
      double tile_max = -INFINITY;
      for (size_t l = 0; l < EV_TILE; l++)
        if (upper[l] > tile_max) tile_max = upper[l];
      for (size_t lane = 0; lane < EV_TILE; lane++) {
        joint_lane* s = &L[lane];
        if (tile_max >= floor_b) {
          for (int ri = 0; ri < 7; ri++) {
            double ratio = E->params[21 + ri];
            double aa = pa_[lane] + ratio * pb_[lane];
            double qq = pq_[lane] + 2.0 * ratio * px_[lane] + ratio * ratio * pr_[lane];
            double diluted = orc_positive_amplitude_evidence(aa, qq, tau_d);
            double eb = orc_log_add(orc_positive_amplitude_evidence(aa, qq, tau_b), diluted) +
                        alias_mass + LOG(0.25 / 7.0);
            lb[lane] = orc_log_add(lb[lane], eb);
            if (eb > comp[lane]) {
              sec_ratio[lane] = 0.0;
              sec_phase[lane] = 0.5;
              odd_ratio[lane] = ratio;
              bin_mult[lane] = 2.0;
              comp[lane] = eb;
            }
          }
        } else {
          s->omitted += 1;
          if (raw_upper[lane] > om_upper[lane]) om_upper[lane] = raw_upper[lane];
        }
        int certain = comp[lane] > om_upper[lane];
        if (!certain) {
          sec_ratio[lane] = NAN;
          sec_phase[lane] = NAN;
          odd_ratio[lane] = NAN;
          bin_mult[lane] = NAN;
        }
        lp[lane] += E->priors[si * 2] + dlm;
        lb[lane] += E->priors[si * 2 + 1] + dlm;
        if (jl[lane] < nb) {
          s->sum_p = orc_log_add(s->sum_p, lp[lane]);
          s->sum_b = orc_log_add(s->sum_b, lb[lane]);
        }
        double joint = lp[lane];
        uint64_t key = (uint64_t)(si * nd + di) * nb + jjl[lane];
        int bj = valid[lane] && (joint > s->best_joint || (joint == s->best_joint && key < s->joint_key));
        int bs = valid[lane] && (snr[lane] > s->best_signal || (snr[lane] == s->best_signal && key < s->signal_key));
        double vals[9] = {snr[lane], depth[lane], lp[lane], lb[lane], sec_ratio[lane],
                          sec_phase[lane], odd_ratio[lane], bin_mult[lane], joint};
        if (bj) {
          memcpy(s->j_out, vals, sizeof(vals));
          s->best_joint = joint;
          s->joint_key = key;
        }
        if (bs) {
          memcpy(s->s_out, vals, sizeof(vals));
          s->best_signal = snr[lane];
          s->signal_key = key;
        }
      }
    }
  }
  for (int sel = 0; sel < 2; sel++) {
    double maximum = -INFINITY;
    for (size_t l = 0; l < EV_TILE; l++) {
      double v = sel == 0 ? L[l].best_joint : L[l].best_signal;
      if (v > maximum) maximum = v;
    }
    uint64_t best_key = UINT64_MAX;
    size_t chosen = EV_TILE;
    for (size_t l = 0; l < EV_TILE; l++) {
      double v = sel == 0 ? L[l].best_joint : L[l].best_signal;
      uint64_t kk = sel == 0 ? L[l].joint_key : L[l].signal_key;
      if (v == maximum && kk < best_key) {
        best_key = kk;
        chosen = l;
      }
    }
    double high = -INFINITY;
    for (size_t l = 0; l < EV_TILE; l++) {
      double v = sel == 0 ? L[l].sum_p : L[l].sum_b;
      if (v > high) high = v;
    }
    double log_sum = high;
    if (high > -INFINITY) {
      double acc = 0.0;
      for (size_t l = 0; l < EV_TILE; l++) acc += EXP((sel == 0 ? L[l].sum_p : L[l].sum_b) - high);
      log_sum = high + LOG(acc);
    }
    size_t cl = chosen < EV_TILE - 1 ? chosen : EV_TILE - 1;
    const double* vals = sel == 0 ? L[cl].j_out : L[cl].s_out;
    double* out = parts + (tile * 2 + sel) * PART;
    out[0] = (double)best_key;
    out[1] = maximum;
    out[2] = sel == 0 ? vals[0] : L[cl].best_signal;
    for (int i = 0; i < 7; i++) out[3 + i] = vals[1 + i];
    out[10] = sel == 0 ? L[cl].best_joint : vals[8];
    out[13] = (double)L[cl].omitted;
    out[11 + sel] = log_sum;
  }
  /* both selections carry both sums (evidence.rs:623-631) */
  double* o0 = parts + tile * 2 * PART;
  double sp = o0[11], sb = o0[PART + 12];
  o0[12] = sb;
  o0[PART + 11] = sp;
}

/* evidence.rs:693-1003 */
int orc_compare_search_ex(const double* t, const double* v, const double* w,
                          size_t n, double baseline_norm, double origin,
                          double residual_energy, const double* durations_in,
                          const double* rho_in, size_t nd_in, double period,
                          const orc_evidence_cfg* cfg, orc_evidence_row* rows,
                          size_t* nrows, double* screen_out, size_t screen_cap,
                          double* parts_out, size_t parts_cap,
                          double* floors_out, size_t* tiles_out,
                          size_t* nb_out) {
  *nrows = 0;
  if (tiles_out) *tiles_out = 0;
  if (nb_out) *nb_out = 0;
  /* durations below a quarter period, ascending (evidence.rs:699-705) */
  double* dur = (double*)malloc((nd_in ? nd_in : 1) * sizeof(double));
  double* rho = (double*)malloc((nd_in ? nd_in : 1) * sizeof(double));
  size_t nd = 0;
  for (size_t i = 0; i < nd_in; i++) {
    if (durations_in[i] < period / 4.0) {
      size_t k = nd++;
      while (k > 0 && dur[k - 1] > durations_in[i]) {
        dur[k] = dur[k - 1];
        rho[k] = rho[k - 1];
        k--;
      }
      dur[k] = durations_in[i];
      rho[k] = rho_in[i];
    }
  }
  if (nd == 0) {
    free(dur);
    free(rho);
    return 0;
  }
  const size_t ns = cfg->ns, layers = cfg->layers;
  const double dmin = dur[0];
  double nbq = ceil((double)cfg->phase_oversample * period / dmin / 4.0);
  size_t nb = 4 * (size_t)nbq;
  if (nb < 8) nb = 8;
  const size_t nb2 = 2 * nb;
  const double scale = (double)nb / period;
  double* dlm = (double*)malloc(nd * sizeof(double));
  double dmax = -INFINITY;
  for (size_t i = 0; i < nd; i++) {
    double r2 = rho[i] * rho[i];
    dlm[i] = -0.5 * (((double)n - 1.0) * LOG(r2) + residual_energy / r2);
    dmax = fmax(dmax, dlm[i]);
  }
  for (size_t i = 0; i < nd; i++) dlm[i] -= dmax;
  /* shape priors (evidence.rs:729-738) */
  double* lgp = (double*)malloc(ns * sizeof(double));
  double* lgb = (double*)malloc(ns * sizeof(double));
  for (size_t s = 0; s < ns; s++) {
    lgp[s] = -4.0 * cfg->ingress[s];
    lgb[s] = 2.0 * cfg->ingress[s];
  }
  double lp0 = logsumexp(lgp, ns), lb0 = logsumexp(lgb, ns);
  double* priors = (double*)malloc(2 * ns * sizeof(double));
  for (size_t s = 0; s < ns; s++) {
    priors[2 * s] = lgp[s] - lp0;
    priors[2 * s + 1] = lgb[s] - lb0;
  }
  /* staircase coefficients (evidence.rs:739-748) */
  double* shapes = (double*)malloc(ns * layers * 3 * sizeof(double));
  for (size_t s = 0; s < ns; s++) {
    double previous = 0.0;
    size_t denom = (layers > 2 ? layers : 2) - 1;
    for (size_t l = 0; l < layers; l++) {
      double fraction = 1.0 - 2.0 * cfg->ingress[s] * (double)l / (double)denom;
      double inc = 1.0 / (double)layers;
      double* e = shapes + (s * layers + l) * 3;
      e[0] = fraction;
      e[1] = inc;
      e[2] = inc * (2.0 * previous + inc);
      previous += inc;
    }
  }
  double params[NPARAM];
  orc_evidence_parameters(cfg, params);
  /* orbit bank (evidence.rs:751-789) */
  static const double phases[3] = {0.25, 0.5, 0.75};
  const double logw[3] = {LOG(0.1), LOG(0.8), LOG(0.1)};
  double* lpb = (double*)malloc(nd * NORB * sizeof(double));
  double* lbb = (double*)malloc(nd * NORB * sizeof(double));
  size_t valid_count = 0;
  for (size_t di = 0; di < nd; di++) {
    double k = fmax(floor(dur[di] * scale + 0.5), 1.0);
    if (k < (double)(nb / 4)) valid_count++;
    for (size_t o = 0; o < NORB; o++) {
      double ks = fmax(floor(k + 0.5), 1.0);
      double shift = floor(phases[o] * (double)nb + 0.5);
      double start = shift + floor((k - ks) / 2.0);
      int ok = k < (double)(nb / 4) && ks < (double)(nb / 2) && start >= k &&
               start + ks <= (double)nb;
      lpb[di * NORB + o] = ok ? logw[o] : -INFINITY;
      lbb[di * NORB + o] = ok ? logw[o] : -INFINITY;
    }
  }
  double cnt = LOG((double)valid_count);
  double norm_p = cnt - logsumexp(lpb, nd * NORB);
  double norm_b = cnt - logsumexp(lbb, nd * NORB);
  double* orbit = (double*)malloc(nd * NORB * 4 * sizeof(double));
  for (size_t di = 0; di < nd; di++)
    for (size_t o = 0; o < NORB; o++) {
      double* e = orbit + (di * NORB + o) * 4;
      e[0] = phases[o];
      e[1] = 1.0;
      e[2] = lpb[di * NORB + o] + norm_p;
      e[3] = lbb[di * NORB + o] + norm_b;
    }
  /* fold_two + prefix_fields (evidence.rs:75-144) */
  double* cum = (double*)calloc(3 * (nb2 + 1), sizeof(double));
  double* hist = (double*)calloc(3 * nb2, sizeof(double));
  for (size_t i = 0; i < n; i++) {
    int64_t cell = (int64_t)floor(t[i] * scale);
    int64_t m = cell % (int64_t)nb2;
    if (m < 0) m += (int64_t)nb2;
    hist[m] += v[i];
    hist[nb2 + m] += w[i];
    hist[2 * nb2 + m] += 1.0;
  }
  for (int f = 0; f < 3; f++) {
    double* c = cum + f * (nb2 + 1);
    c[0] = 0.0;
    for (size_t b = 0; b < nb2; b++) c[b + 1] = c[b] + hist[f * nb2 + b];
  }
  ev_ctx E;
  E.cum = cum;
  E.nb = nb;
  E.nb2 = nb2;
  E.nd = nd;
  E.ns = ns;
  E.layers = layers;
  E.scale = scale;
  E.c = baseline_norm;
  E.n = (double)n;
  E.threshold = cfg->threshold;
  E.dur = dur;
  E.rho = rho;
  E.dlm = dlm;
  E.shapes = shapes;
  E.priors = priors;
  E.params = params;
  E.orbit = orbit;
  const size_t tiles = (nb + EV_TILE - 1) / EV_TILE;
  if (tiles_out) *tiles_out = tiles;
  if (nb_out) *nb_out = nb;
  double* screen = (double*)calloc(tiles * 4, sizeof(double));
  for (size_t tile = 0; tile < tiles; tile++) screen_tile(&E, tile, screen + tile * 4);
  double maximum = -INFINITY;
  for (size_t tile = 0; tile < tiles; tile++) maximum = fmax(maximum, screen[tile * 4]);
  int status = 0;
  double* parts = NULL;
  if (screen_out)
    memcpy(screen_out, screen, (tiles * 4 < screen_cap ? tiles * 4 : screen_cap) * sizeof(double));
  if (!(maximum >= cfg->threshold * (1.0 - 1e-12))) goto done;
  {
    /* Iterator::max_by keeps the last of equal maxima */
    size_t winner = 0;
    for (size_t tile = 1; tile < tiles; tile++)
      if (!(screen[tile * 4 + 1] < screen[winner * 4 + 1])) winner = tile;
    double pilot_key = screen[winner * 4 + 3];
    double margin = cfg->evidence_rtol > 0.0
                        ? LOG(cfg->evidence_rtol) - LOG((double)((NORB + 1) * nb * nd * ns))
                        : -INFINITY;
    parts = (double*)calloc(tiles * 2 * PART, sizeof(double));
    joint_tile(&E, 0, 1, -INFINITY, -INFINITY, (uint64_t)pilot_key, parts);
    double floor_p = parts[11] + margin, floor_b = parts[12] + margin;
    memset(parts, 0, tiles * 2 * PART * sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
    for (long long tile = 0; tile < (long long)tiles; tile++)
      joint_tile(&E, (size_t)tile, 0, floor_p, floor_b, (uint64_t)pilot_key, parts);
    if (parts_out)
      memcpy(parts_out, parts,
             (tiles * 2 * PART < parts_cap ? tiles * 2 * PART : parts_cap) * sizeof(double));
    if (floors_out) {
      floors_out[0] = floor_p;
      floors_out[1] = floor_b;
      floors_out[2] = pilot_key;
    }
    /* host merge (evidence.rs:936-1002) */
    double* dm = (double*)malloc(nd * sizeof(double));
    size_t nin = 0;
    for (size_t i = 0; i < nd; i++)
      if (fmax(floor(dur[i] * scale + 0.5), 1.0) < (double)(nb / 4)) dm[nin++] = dlm[i];
    double log_mass = LOG((double)nb) + logsumexp(dm, nin);
    free(dm);
    orc_evidence_row r2[2];
    double* col = (double*)malloc(tiles * sizeof(double));
    for (int sel = 0; sel < 2; sel++) {
#define VAL(tt, ii) parts[((tt) * 2 + sel) * PART + (ii)]
      double mx = -INFINITY;
      for (size_t tt = 0; tt < tiles; tt++) mx = fmax(mx, VAL(tt, 1));
      double key = INFINITY;
      for (size_t tt = 0; tt < tiles; tt++)
        if (VAL(tt, 1) == mx) key = fmin(key, VAL(tt, 0));
      size_t tile = 0;
      for (size_t tt = 0; tt < tiles; tt++)
        if (VAL(tt, 1) == mx && VAL(tt, 0) == key) {
          tile = tt;
          break;
        }
      for (size_t tt = 0; tt < tiles; tt++) col[tt] = VAL(tt, 11);
      double lp = logsumexp(col, tiles) - log_mass;
      for (size_t tt = 0; tt < tiles; tt++) col[tt] = VAL(tt, 12);
      double lb = logsumexp(col, tiles) - log_mass;
      double joint = params[3] + lp - orc_log_add(params[5], params[4] + lb);
      uint64_t decoded = mx > -INFINITY ? (uint64_t)key : 0;
      size_t j = (size_t)(decoded % nb);
      size_t di = (size_t)((decoded / nb) % nd);
      size_t si = (size_t)(decoded / ((uint64_t)nb * nd));
      double width = fmax(floor(dur[di] * scale + 0.5), 1.0);
      orc_evidence_row* r = &r2[sel];
      r->period = period;
      r->epoch = ((double)j + width / 2.0) / scale + origin;
      r->duration = width / scale;
      r->snr = mx > -INFINITY ? VAL(tile, 2) : -INFINITY;
      r->depth = VAL(tile, 3);
      r->planet_log_odds = joint;
      r->log_bf_planet_noise = lp;
      r->log_bf_binary_noise = lb;
      r->binary_secondary_ratio = VAL(tile, 6);
      r->binary_secondary_phase = VAL(tile, 7);
      r->binary_odd_even_ratio = VAL(tile, 8);
      r->binary_period_days = VAL(tile, 9) * period;
      r->ingress_fraction = cfg->ingress[si < ns - 1 ? si : ns - 1];
      r->selection_flags = sel == 0 ? 1u : 2u;
      r->phase = (uint32_t)j;
      r->duration_index = (uint32_t)di;
      r->shape = (uint32_t)si;
#undef VAL
    }
    free(col);
    int hit0 = isfinite(r2[0].snr), hit1 = isfinite(r2[1].snr);
    int same = hit0 && hit1 && r2[0].period == r2[1].period && r2[0].epoch == r2[1].epoch &&
               r2[0].duration == r2[1].duration && r2[0].snr == r2[1].snr &&
               r2[0].depth == r2[1].depth && r2[0].ingress_fraction == r2[1].ingress_fraction;
    if (hit0) {
      rows[*nrows] = r2[0];
      if (same) rows[*nrows].selection_flags = 3;
      (*nrows)++;
    }
    if (hit1 && !same) rows[(*nrows)++] = r2[1];
  }
done:
  free(parts);
  free(screen);
  free(cum);
  free(hist);
  free(orbit);
  free(lpb);
  free(lbb);
  free(shapes);
  free(priors);
  free(lgp);
  free(lgb);
  free(dlm);
  free(dur);
  free(rho);
  return status;
}

int orc_compare_search(const double* t, const double* v, const double* w,
                       size_t n, double baseline_norm, double origin,
                       double residual_energy, const double* durations,
                       const double* rho, size_t nd, double period,
                       const orc_evidence_cfg* cfg, orc_evidence_row* rows,
                       size_t* nrows) {
  return orc_compare_search_ex(t, v, w, n, baseline_norm, origin,
                               residual_energy, durations, rho, nd, period, cfg,
                               rows, nrows, NULL, 0, NULL, 0, NULL, NULL, NULL);
}

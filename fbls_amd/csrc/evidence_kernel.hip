// Joint planet / binary / noise evidence at one trial period (gfx950).
//
// Seven kernels on one stream, no host round trip in between:
//   ev_zero       clears the bin sums and the screen
//   ev_fold       two-period fold: bin = floor(t * scale) mod 2 nb, exact
//                 fixed-point sums of v and w (int64 atomics) and counts,
//                 with runs of equal bins aggregated in registers
//   ev_prefix     exclusive prefix per field (one workgroup each), exact
//   ev_screen     four waves per 64-phase tile, each over a quarter of the
//                 (duration, shape) templates: best SNR and the evidence
//                 lower bounds of the best-merit template
//   ev_pilot      threshold check of the screen, the pilot template (winner
//                 tile by lower bound, last of equals) with its scenarios
//                 spread over the lanes, and the pruning floors
//   ev_templates  one wave per (tile, template): the template's evidence at
//                 the tile's 64 phases over the secondary-eclipse and odd/even
//                 scenarios, pruned per tile by subspace upper bounds against
//                 the floors; per (template, phase) results to HBM
//   ev_reduce     one wave per tile: log-sums over the templates, winners by
//                 joint evidence and by SNR, tile log-sums, the tile parts
//
// The work of a tile of the reference (one 64-thread block looping over
// every template, evidence.rs:280-632) is spread over nd x ns waves, so a
// period with a few dozen tiles still fills the chip; what the reference
// keeps in registers between templates (the running log-sums and winners)
// goes through compact per-(template, phase) records read back in the
// reference's template order.
//
// Reference semantics restated (isakruas/fbls src/search/evidence.rs):
//   circular_sum / layered_sum   evidence.rs:37-67
//   fold_two / prefix_fields     evidence.rs:75-144
//   comparison_screen            evidence.rs:151-273
//   joint_tiles                  evidence.rs:280-632
//   screen threshold, pilot key, floors   evidence.rs:847-896
// Every floating-point expression keeps the reference's operand order (built
// with -ffp-contract=off) and the transcendental functions are fmath.h's, so
// the values are the same doubles the host restatement computes.  Sums whose
// order matters (the log-sums over templates, the tile log-sum-exp) are taken
// sequentially in the reference's order; maxima and winners are exact and
// order-free.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "evidence_kernel.h"
#include "fixed_point.h"
#include "fmath.h"

namespace fbls {

namespace {

constexpr double kInf = __builtin_inf();
constexpr uint64_t kKeyMax = ~0ull;
constexpr uint32_t kRun = 8;  // consecutive observations per fold thread

__device__ __forceinline__ double wave_max(double x) {
  for (int m = 32; m >= 1; m >>= 1) x = fmax(x, __shfl_xor(x, m, 64));
  return x;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t x) {
  for (int m = 32; m >= 1; m >>= 1) {
    const uint64_t y = __shfl_xor(x, m, 64);
    x = y < x ? y : x;
  }
  return x;
}

// lowest lane whose predicate holds, or 64
__device__ __forceinline__ uint32_t first_lane(bool p) {
  const uint64_t m = __ballot(p);
  return m ? (uint32_t)__builtin_ctzll(m) : 64u;
}

// evidence.rs:37-45
__device__ __forceinline__ double circular_sum(const double* cum, uint64_t start,
                                               uint64_t width, uint64_t nb2) {
  const uint64_t j = start % nb2;
  const uint64_t end = j + width;
  const double left = cum[j];
  const double right = cum[end < nb2 ? end : nb2];
  const double wrap = cum[end > nb2 ? end - nb2 : 0];
  return right - left + wrap;
}

// evidence.rs:48-67
__device__ __forceinline__ double layered_sum(const double* cum, uint64_t start,
                                              uint64_t width, const double* shapes,
                                              uint32_t si, uint64_t nb2,
                                              uint32_t layers, int which) {
  double value = 0.0;
  for (uint32_t layer = 0; layer < layers; layer++) {
    const double* e = shapes + (si * layers + layer) * 3;
    const double inner_f = floor((double)width * e[0] + 0.5);
    const uint64_t inner = inner_f < 1.0 ? 1 : (uint64_t)inner_f;
    value += e[which] * circular_sum(cum, start + (width - inner) / 2, inner, nb2);
  }
  return value;
}

__device__ __forceinline__ uint64_t template_bins(double duration, double scale) {
  const double rk = floor(duration * scale + 0.5);
  return rk < 1.0 ? 1 : (uint64_t)rk;
}

// ------------------------------------------------------------------ fold

__global__ __launch_bounds__(256) void ev_zero(EvArgs A) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t i = i0; i < 2 * A.nb2; i += stride) A.hist[i] = 0ull;
  for (uint32_t i = i0; i < A.nb2; i += stride) A.hcount[i] = 0u;
  for (uint32_t i = i0; i < 4 * A.tiles; i += stride) A.screen[i] = 0.0;
}

__device__ __forceinline__ void fold_flush(const EvArgs& A, uint32_t bin,
                                           long long sv, long long sw,
                                           unsigned int cnt) {
  atomicAdd(A.hist + bin, (unsigned long long)sv);
  atomicAdd(A.hist + A.nb2 + bin, (unsigned long long)sw);
  atomicAdd(A.hcount + bin, cnt);
}

// evidence.rs:75-95: bin = floor(t * scale) rem_euclid 2 nb (t >= 0)
__global__ __launch_bounds__(256) void ev_fold(EvArgs A) {
  const uint32_t i0 = (blockIdx.x * blockDim.x + threadIdx.x) * kRun;
  if (i0 >= A.n) return;
  const uint32_t i1 = min(i0 + kRun, A.n);
  uint32_t cur = ~0u;
  long long sv = 0, sw = 0;
  unsigned int cnt = 0;
  for (uint32_t i = i0; i < i1; i++) {
    const uint64_t cell = (uint64_t)floor(A.t[i] * A.scale);
    const uint32_t bin = (uint32_t)(cell % A.nb2);
    if (bin != cur) {
      if (cnt) fold_flush(A, cur, sv, sw, cnt);
      cur = bin;
      sv = sw = 0;
      cnt = 0;
    }
    sv += fixed_ivalue(A.vw[2 * i], A.fx & 1u);
    sw += fixed_ivalue(A.vw[2 * i + 1], A.fx & 2u);
    cnt++;
  }
  if (cnt) fold_flush(A, cur, sv, sw, cnt);
}

// evidence.rs:99-144: one workgroup of 1024 threads per field; each thread
// owns a contiguous run of bins.  The integer prefix is exact and converts
// to the double the sequential FP64 prefix of the same values holds.
__global__ __launch_bounds__(1024) void ev_prefix(EvArgs A) {
  __shared__ long long wave_tot[16];
  const uint32_t field = blockIdx.x;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t nb2 = A.nb2;
  const uint32_t per = (nb2 + 1023) / 1024;
  const uint32_t b0 = min(tid * per, nb2), b1 = min(b0 + per, nb2);
  auto value = [&](uint32_t b) -> long long {
    if (field == 2) return (long long)A.hcount[b];
    return (long long)A.hist[field * nb2 + b];
  };
  long long local = 0;
  for (uint32_t b = b0; b < b1; b++) local += value(b);
  // inclusive scan over the wave, then over the waves
  long long inc = local;
  for (int d = 1; d < 64; d <<= 1) {
    const long long y = __shfl_up(inc, d, 64);
    if (lane >= (uint32_t)d) inc += y;
  }
  if (lane == 63) wave_tot[wave] = inc;
  __syncthreads();
  long long base = inc - local;
  for (uint32_t w = 0; w < wave; w++) base += wave_tot[w];
  const double s = field == 0 ? A.sv : field == 1 ? A.sw : 1.0;
  double* row = A.cum + (size_t)field * (nb2 + 1);
  if (tid == 0) row[0] = 0.0;
  long long run = base;
  for (uint32_t b = b0; b < b1; b++) {
    run += value(b);
    row[b + 1] = (double)run * s;
  }
}

// -------------------------------------------------------- one template

// Largest of the values the lanes hold, NaN ignored, -inf when there is none:
// the maximum a sequential `if v > m` scan of the lanes finds.
__device__ __forceinline__ double lanes_max(double x) {
  return wave_max(x == x ? x : -kInf);
}

// The odd/even-window sums of one template at phase jj (evidence.rs:211-218,
// 360-372), in the reference's operand order.
struct TemplateSums {
  double ae, ao, pe, po, qe, qo, qeo, ap, qp;
  bool counted;  // 3 <= in-window count <= n - 3
};

__device__ __forceinline__ TemplateSums template_sums(const EvArgs& A, uint64_t jj,
                                                      uint64_t k, uint32_t si) {
  const uint64_t nb2 = A.nb2, nb = A.nb;
  const double* cv = A.cum;
  const double* cw = A.cum + (nb2 + 1);
  const double* cn = A.cum + 2 * (nb2 + 1);
  const double c = A.C;
  TemplateSums s;
  const double count = circular_sum(cn, jj, k, nb2) + circular_sum(cn, jj + nb, k, nb2);
  s.counted = count >= 3.0 && count <= (double)A.n - 3.0;
  s.ae = layered_sum(cv, jj, k, A.shapes, si, nb2, A.layers, 1);
  s.ao = layered_sum(cv, jj + nb, k, A.shapes, si, nb2, A.layers, 1);
  s.pe = layered_sum(cw, jj, k, A.shapes, si, nb2, A.layers, 1);
  s.po = layered_sum(cw, jj + nb, k, A.shapes, si, nb2, A.layers, 1);
  const double ne = layered_sum(cw, jj, k, A.shapes, si, nb2, A.layers, 2);
  const double no = layered_sum(cw, jj + nb, k, A.shapes, si, nb2, A.layers, 2);
  s.qe = ne - s.pe * s.pe / c;
  s.qo = no - s.po * s.po / c;
  s.qeo = -s.pe * s.po / c;
  s.ap = s.ae + s.ao;
  s.qp = s.qe + s.qo + 2.0 * s.qeo;
  return s;
}

// The secondary-eclipse subspace of orbit o (evidence.rs:381-415): the
// whitened two-template statistics and the bounds used for pruning.
struct Orbit {
  double pa, pb, pq, pr, px;  // primary / secondary numerators, precisions, cross
  double upper;               // pruning bound against the floors
  double raw_upper;           // binary bound (for the "certain" test)
  double phase, logwp, logwb;
};

__device__ __forceinline__ Orbit orbit_subspace(const EvArgs& A, const TemplateSums& b,
                                                uint64_t jj, uint64_t k, uint32_t si,
                                                uint32_t di, uint32_t o, double rho2,
                                                double dlm, double floor_p, double floor_b) {
  const uint64_t nb2 = A.nb2, nb = A.nb;
  const double* cv = A.cum;
  const double* cw = A.cum + (nb2 + 1);
  const double c = A.C;
  const double* orb = A.orbit + (di * kEvOrbits + o) * 4;
  Orbit r;
  r.phase = orb[0];
  r.logwp = orb[2];
  r.logwb = orb[3];
  const int64_t offset = (int64_t)floor(r.phase * (double)nb + 0.5);
  const double skf = floor((double)k * orb[1] + 0.5);
  int64_t ski = skf < 1.0 ? 1 : (int64_t)skf;
  if (ski > (int64_t)nb - 1) ski = (int64_t)nb - 1;
  const uint64_t sk = (uint64_t)ski;
  const uint64_t s0 =
      (uint64_t)((int64_t)jj + offset + (int64_t)floor(((double)k - (double)sk) / 2.0));
  const uint64_t s1 = s0 + nb;
  const double ass = layered_sum(cv, s0, sk, A.shapes, si, nb2, A.layers, 1) +
                     layered_sum(cv, s1, sk, A.shapes, si, nb2, A.layers, 1);
  const double ps = layered_sum(cw, s0, sk, A.shapes, si, nb2, A.layers, 1) +
                    layered_sum(cw, s1, sk, A.shapes, si, nb2, A.layers, 1);
  const double nsum = layered_sum(cw, s0, sk, A.shapes, si, nb2, A.layers, 2) +
                      layered_sum(cw, s1, sk, A.shapes, si, nb2, A.layers, 2);
  const double qs = nsum - ps * ps / c, qps = -(b.pe + b.po) * ps / c;
  r.pa = b.ap / rho2;
  r.pb = ass / rho2;
  r.pq = b.qp / rho2;
  r.pr = qs / rho2;
  r.px = qps / rho2;
  const double cu = 1e-6;
  const double pu = fm::subspace_upper(r.pa, r.pb, r.pq, r.pr, r.px, 1.0) + r.logwp + cu;
  const double bu = fm::subspace_upper(r.pa, r.pb, r.pq, r.pr, r.px, 1.0) + r.logwb + cu;
  r.upper = fmax(pu - floor_p, bu - floor_b) + dlm;
  if (!(r.logwp > -kInf || r.logwb > -kInf)) r.upper = -kInf;
  r.raw_upper = r.logwb > -kInf ? bu : -kInf;
  return r;
}

// Planet and binary evidence of secondary ratio index ri (evidence.rs:426-441).
__device__ __forceinline__ void secondary_scenario(const double* P, const Orbit& r, int ri,
                                                   double* ep, double* eb, double* ratio) {
  *ratio = P[6 + ri * 3];
  const double wp = P[7 + ri * 3], wb = P[8 + ri * 3];
  const double aa = r.pa + *ratio * r.pb;
  const double qq = r.pq + 2.0 * *ratio * r.px + *ratio * *ratio * r.pr;
  *ep = fm::positive_amplitude_evidence(aa, qq, P[0]) + fm::log(wp) + r.logwp;
  const double diluted = fm::positive_amplitude_evidence(aa, qq, P[2]);
  *eb = fm::log_add(fm::positive_amplitude_evidence(aa, qq, P[1]), diluted) +
        fm::log(0.25 * wb) + r.logwb;
}

// The odd/even (alternating primaries at 2P) subspace (evidence.rs:458-470).
struct OddEven {
  double ba, bb, bq, br, bx, raw_upper, upper, alias_mass;
};

__device__ __forceinline__ OddEven odd_even(const TemplateSums& b, uint64_t k, uint64_t nb2,
                                            double rho2, double dlm) {
  OddEven e;
  e.ba = b.ae / rho2;
  e.bb = b.ao / rho2;
  e.bq = b.qe / rho2;
  e.br = b.qo / rho2;
  e.bx = b.qeo / rho2;
  e.alias_mass = k < nb2 / 8 ? 0.0 : -kInf;
  e.raw_upper = fm::subspace_upper(e.ba, e.bb, e.bq, e.br, e.bx, 1.0) + 1e-6 + e.alias_mass;
  e.upper = e.raw_upper + dlm;
  return e;
}

// Binary evidence of odd/even ratio index ri (evidence.rs:479-485).
__device__ __forceinline__ double odd_even_scenario(const double* P, const OddEven& e, int ri,
                                                   double* ratio) {
  *ratio = P[21 + ri];
  const double aa = e.ba + *ratio * e.bb;
  const double qq = e.bq + 2.0 * *ratio * e.bx + *ratio * *ratio * e.br;
  const double diluted = fm::positive_amplitude_evidence(aa, qq, P[2]);
  return fm::log_add(fm::positive_amplitude_evidence(aa, qq, P[1]), diluted) + e.alias_mass +
         fm::log(0.25 / 7.0);
}

__device__ __forceinline__ uint64_t clamped_k(uint64_t raw_k, uint64_t nb2) {
  return raw_k < nb2 / 8 - 1 ? raw_k : nb2 / 8 - 1;
}

// ---------------------------------------------------------------- screen

// Amplitude screen (evidence.rs:151-273): per 64-phase tile the largest SNR
// of a valid template, and the planet / binary evidence lower bounds of the
// template of largest merit (first in (duration, shape) order, then lowest
// phase).  Four waves share a tile: wave w takes templates w, w + 4, ...;
// each keeps its first best merit, the waves' bests merge by (merit, smaller
// template), and the chosen lane recomputes its template's sums.
constexpr uint32_t kScreenWaves = 4;

__global__ __launch_bounds__(64 * kScreenWaves) void ev_screen(EvArgs A) {
  __shared__ double s_best[kScreenWaves][kEvTile], s_merit[kScreenWaves][kEvTile];
  __shared__ uint32_t s_t[kScreenWaves][kEvTile];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t tile = blockIdx.x;
  const uint64_t nb2 = A.nb2, nb = A.nb;
  const uint64_t j = (uint64_t)tile * kEvTile + lane;
  const uint64_t jj = j < nb ? j : nb - 1;
  const uint32_t T = A.nd * A.ns;
  double best = -kInf, merit_best = -kInf;
  uint32_t t_best = T;
  for (uint32_t t = wave; t < T; t += kScreenWaves) {
    const uint32_t di = t / A.ns, si = t % A.ns;
    const uint64_t raw_k = template_bins(A.dur[di], A.scale);
    const uint64_t k = clamped_k(raw_k, nb2);
    const double r = A.rho[di];
    const TemplateSums b = template_sums(A, jj, k, si);
    const bool usable = j < nb && raw_k < nb2 / 8 && b.counted;
    const double score = b.ap / sqrt(fmax(b.qp * (r * r), 1e-300));
    const bool valid = usable && b.ap > 0.0 && b.qp > 0.0;
    const double merit = 0.5 * score * score + A.dlm[di] + A.priors[si * 2];
    if (valid && merit > merit_best) {
      merit_best = merit;
      t_best = t;
    }
    if (valid && score > best) best = score;
  }
  s_best[wave][lane] = best;
  s_merit[wave][lane] = merit_best;
  s_t[wave][lane] = t_best;
  __syncthreads();
  if (wave != 0) return;
  for (uint32_t w = 1; w < kScreenWaves; w++) {
    const double bw = s_best[w][lane], mw = s_merit[w][lane];
    const uint32_t tw = s_t[w][lane];
    if (bw > best) best = bw;
    if (mw > merit_best || (mw == merit_best && mw > -kInf && tw < t_best)) {
      merit_best = mw;
      t_best = tw;
    }
  }
  const double maximum = lanes_max(best);
  const double bound_maximum = lanes_max(merit_best);
  const uint32_t chosen = first_lane(merit_best > -kInf && merit_best == bound_maximum);
  if (lane != chosen) return;
  const uint32_t di = t_best / A.ns, si = t_best % A.ns;
  const double r = A.rho[di];
  const TemplateSums b = template_sums(A, jj, clamped_k(template_bins(A.dur[di], A.scale), nb2), si);
  const double best_a = b.ap / (r * r), best_q = b.qp / (r * r);
  const double* P = A.params;
  const double lower_p = fm::positive_amplitude_evidence(best_a, best_q, P[0]) + fm::log(0.75) +
                         A.dlm[di] + A.priors[si * 2];
  const double lower_b = fm::log_add(fm::positive_amplitude_evidence(best_a, best_q, P[1]),
                                     fm::positive_amplitude_evidence(best_a, best_q, P[2])) +
                         fm::log(0.025) + A.dlm[di] + A.priors[si * 2 + 1];
  double* out = A.screen + (size_t)tile * 4;
  out[0] = maximum;
  out[1] = lower_p;
  out[2] = lower_b;
  out[3] = (double)((uint64_t)(si * A.nd + di) * nb + jj);
}

// ----------------------------------------------------------------- pilot

// evidence.rs:847-896 on the device: the screen threshold, the pilot
// template (tile of largest lower bound, the last of equals), and the
// pruning floors from the pilot's evidence.  The pilot has no floors, so
// every scenario of its one template is evaluated: lanes 0-14 take the
// (secondary phase, ratio) scenarios and lanes 15-21 the odd/even ratios in
// parallel, and lane 0 accumulates them in the reference's order.  The
// reference evaluates the pilot in all 64 lanes of a tile, so its tile
// log-sum is that of 64 equal terms.
__global__ __launch_bounds__(64) void ev_pilot(EvArgs A) {
  const uint32_t lane = threadIdx.x;
  double m = -kInf, lo = -kInf;
  for (uint32_t t = lane; t < A.tiles; t += kEvTile) {
    m = fmax(m, A.screen[(size_t)t * 4]);
    lo = fmax(lo, A.screen[(size_t)t * 4 + 1]);
  }
  const double maximum = wave_max(m);
  if (!(maximum >= A.screen_threshold)) {
    if (lane == 0) A.floors[3] = 0.0;
    return;
  }
  // Iterator::max_by keeps the last of equal maxima
  const double top = wave_max(lo);
  int64_t last = -1;
  for (uint32_t t = lane; t < A.tiles; t += kEvTile)
    if (A.screen[(size_t)t * 4 + 1] == top) last = (int64_t)t;
  for (int s = 32; s >= 1; s >>= 1) {
    const int64_t y = __shfl_xor(last, s, 64);
    last = y > last ? y : last;
  }
  const uint32_t winner = last < 0 ? A.tiles - 1 : (uint32_t)last;
  const double pilot_key = A.screen[(size_t)winner * 4 + 3];
  const uint64_t key = (uint64_t)pilot_key;
  const uint64_t nb2 = A.nb2, nb = A.nb;
  const uint64_t jj = key % nb;
  const uint32_t di = (uint32_t)((key / nb) % A.nd);
  const uint32_t si = (uint32_t)(key / (nb * A.nd));
  const uint64_t raw_k = template_bins(A.dur[di], A.scale);
  double sum_p = -kInf, sum_b = -kInf;
  if (raw_k < nb2 / 8) {  // uniform: a pilot beyond the bank contributes nothing
    const uint64_t k = clamped_k(raw_k, nb2);
    const double rho2 = A.rho[di] * A.rho[di];
    const double dlm = A.dlm[di];
    const TemplateSums b = template_sums(A, jj, k, si);
    double e_p = -kInf, e_b = -kInf;
    bool take = false;
    if (lane < 15) {
      const Orbit r = orbit_subspace(A, b, jj, k, si, di, lane / 5, rho2, dlm, -kInf, -kInf);
      const double u = r.upper == r.upper ? r.upper : -kInf;
      take = u >= 0.0;
      if (take) {
        double ratio;
        secondary_scenario(A.params, r, (int)(lane % 5), &e_p, &e_b, &ratio);
      }
    } else if (lane < 22) {
      const OddEven e = odd_even(b, k, nb2, rho2, dlm);
      const double u = e.upper == e.upper ? e.upper : -kInf;
      take = u >= -kInf;
      if (take) {
        double ratio;
        e_b = odd_even_scenario(A.params, e, (int)(lane - 15), &ratio);
      }
    }
    double lp = -kInf, lb = -kInf;
    for (uint32_t l = 0; l < 22; l++) {
      const double xp = __shfl(e_p, (int)l, 64), xb = __shfl(e_b, (int)l, 64);
      const int tk = __shfl((int)take, (int)l, 64);
      if (!tk) continue;
      if (l < 15) lp = fm::log_add(lp, xp);
      lb = fm::log_add(lb, xb);
    }
    lp += A.priors[si * 2] + dlm;
    lb += A.priors[si * 2 + 1] + dlm;
    sum_p = fm::log_add(-kInf, lp);
    sum_b = fm::log_add(-kInf, lb);
  }
  if (lane != 0) return;
  double sums[2] = {sum_p, sum_b};
  for (int sel = 0; sel < 2; sel++) {
    const double high = sums[sel] > -kInf ? sums[sel] : -kInf;
    double log_sum = high;
    if (high > -kInf) {
      const double term = fm::exp(sums[sel] - high);  // the same for all 64 lanes
      double acc = 0.0;
      for (uint32_t l = 0; l < kEvTile; l++) acc += term;
      log_sum = high + fm::log(acc);
    }
    A.floors[sel] = log_sum + A.margin;
  }
  A.floors[2] = pilot_key;
  A.floors[3] = 1.0;
}

// ------------------------------------------------------------- templates

// One wave per (64-phase tile, template): the evidence of the template at
// each of the tile's phases over every secondary-eclipse and odd/even
// scenario (evidence.rs:343-533).  A scenario group is evaluated for the
// whole tile or skipped for the whole tile, by the tile's largest pruning
// bound against the floors (the reference's block-wide decision).  Writes the
// template's planet and binary log evidence, SNR (when valid), depth and
// winning-scenario code per phase, and the evaluated groups per tile.
__global__ __launch_bounds__(64) void ev_templates(EvArgs A) {
  if (A.floors[3] == 0.0) return;
  const uint32_t lane = threadIdx.x, tile = blockIdx.x, t = blockIdx.y;
  const uint32_t di = t / A.ns, si = t % A.ns;
  const uint64_t nb2 = A.nb2, nb = A.nb;
  const uint64_t raw_k = template_bins(A.dur[di], A.scale);
  if (!(raw_k < nb2 / 8)) return;  // outside the bank: no evidence, never read
  const uint64_t k = clamped_k(raw_k, nb2);
  const uint64_t j = (uint64_t)tile * kEvTile + lane;
  const uint64_t jj = j < nb ? j : nb - 1;
  const double rho2 = A.rho[di] * A.rho[di];
  const double dlm = A.dlm[di];
  const double floor_p = A.floors[0], floor_b = A.floors[1];
  const double* P = A.params;
  const TemplateSums b = template_sums(A, jj, k, si);
  const bool usable = j < nb && b.counted;
  const double snr = b.ap / sqrt(fmax(b.qp * rho2, 1e-300));
  const double depth = b.ap / fmax(b.qp, 1e-300);
  const bool valid = usable && b.ap > 0.0 && b.qp > 0.0 && snr >= A.threshold;
  double lp = -kInf, lb = -kInf, component = -kInf, omitted_upper = -kInf;
  uint32_t code = 0, kept = 0;
  for (uint32_t o = 0; o < kEvOrbits; o++) {
    const Orbit r = orbit_subspace(A, b, jj, k, si, di, o, rho2, dlm, floor_p, floor_b);
    if (lanes_max(j < nb ? r.upper : -kInf) >= 0.0) {
      kept |= 1u << o;
      for (int ri = 0; ri < 5; ri++) {
        double ep, eb, ratio;
        secondary_scenario(P, r, ri, &ep, &eb, &ratio);
        lp = fm::log_add(lp, ep);
        lb = fm::log_add(lb, eb);
        if (eb > component) {
          component = eb;
          code = 1 + o * 5 + (uint32_t)ri;
        }
      }
    } else if (r.raw_upper > omitted_upper) {
      omitted_upper = r.raw_upper;
    }
  }
  const OddEven e = odd_even(b, k, nb2, rho2, dlm);
  if (lanes_max(j < nb ? e.upper : -kInf) >= floor_b) {
    kept |= 8u;
    for (int ri = 0; ri < 7; ri++) {
      double ratio;
      const double eb = odd_even_scenario(P, e, ri, &ratio);
      lb = fm::log_add(lb, eb);
      if (eb > component) {
        component = eb;
        code = 16 + (uint32_t)ri;
      }
    }
  } else if (e.raw_upper > omitted_upper) {
    omitted_upper = e.raw_upper;
  }
  if (!(component > omitted_upper)) code |= 0x80u;  // binary diagnostics not certain
  lp += A.priors[si * 2] + dlm;
  lb += A.priors[si * 2 + 1] + dlm;
  const size_t row = (size_t)A.tiles * kEvTile;
  const size_t at = (size_t)t * row + j;
  A.t_lp[at] = lp;
  A.t_lb[at] = lb;
  A.t_snr[at] = valid ? snr : -kInf;
  A.t_depth[at] = depth;
  A.t_code[at] = (uint8_t)code;
  if (lane == 0) A.t_kept[(size_t)tile * (A.nd * A.ns) + t] = kept;
}

// ---------------------------------------------------------------- reduce

// Per tile (one wave, lane = phase): the planet and binary log-sums over the
// templates in (duration, shape) order, the best template by joint evidence
// and by SNR (ties: smaller key), the tile's winners (largest value, smallest
// key, lowest lane), the tile log-sums taken in lane order, and the 15 part
// values of each selection (evidence.rs:520-631).
__device__ __forceinline__ void decode_scenario(const EvArgs& A, uint32_t di, uint32_t code,
                                                double* v) {
  const double* P = A.params;
  double ratio = 0.0, phase = 0.5, odd = 1.0, multiple = 1.0;
  const uint32_t c = code & 0x7fu;
  if (c >= 1 && c <= 15) {
    const uint32_t o = (c - 1) / 5, ri = (c - 1) % 5;
    ratio = P[6 + ri * 3];
    phase = A.orbit[(di * kEvOrbits + o) * 4];
  } else if (c >= 16) {
    odd = P[21 + (c - 16)];
    multiple = 2.0;
  }
  if (code & 0x80u) ratio = phase = odd = multiple = __builtin_nan("");
  v[0] = ratio;
  v[1] = phase;
  v[2] = odd;
  v[3] = multiple;
}

__global__ __launch_bounds__(64) void ev_reduce(EvArgs A) {
  __shared__ double s_sum[kEvTile];
  __shared__ double s_log;
  if (A.floors[3] == 0.0) return;
  const uint32_t lane = threadIdx.x, tile = blockIdx.x;
  const uint64_t nb2 = A.nb2, nb = A.nb;
  const uint32_t nd = A.nd, ns = A.ns, T = nd * ns;
  const uint64_t j = (uint64_t)tile * kEvTile + lane;
  const uint64_t jj = j < nb ? j : nb - 1;
  const size_t row = (size_t)A.tiles * kEvTile;
  double sum_p = -kInf, sum_b = -kInf, best_joint = -kInf, best_signal = -kInf;
  uint64_t joint_key = kKeyMax, signal_key = kKeyMax;
  uint32_t joint_t = T, signal_t = T, omitted = 0;
  for (uint32_t t = 0; t < T; t++) {
    const uint32_t di = t / ns, si = t % ns;
    if (!(template_bins(A.dur[di], A.scale) < nb2 / 8)) continue;
    omitted += 4u - (uint32_t)__builtin_popcount(A.t_kept[(size_t)tile * T + t]);
    const size_t at = (size_t)t * row + j;
    const double lp = A.t_lp[at], lb = A.t_lb[at], snr = A.t_snr[at];
    if (j < nb) {
      sum_p = fm::log_add(sum_p, lp);
      sum_b = fm::log_add(sum_b, lb);
    }
    const bool valid = snr > -kInf;
    const uint64_t key = (uint64_t)(si * nd + di) * nb + jj;
    if (valid && (lp > best_joint || (lp == best_joint && key < joint_key))) {
      best_joint = lp;
      joint_key = key;
      joint_t = t;
    }
    if (valid && (snr > best_signal || (snr == best_signal && key < signal_key))) {
      best_signal = snr;
      signal_key = key;
      signal_t = t;
    }
  }
  double log_sums[2];
  for (int sel = 0; sel < 2; sel++) {
    const double v = sel == 0 ? best_joint : best_signal;
    const uint64_t kk = sel == 0 ? joint_key : signal_key;
    const double s = sel == 0 ? sum_p : sum_b;
    const double maximum = lanes_max(v);
    const uint64_t best_key = wave_min_u64(v == maximum ? kk : kKeyMax);
    uint32_t chosen = first_lane(v == maximum && kk == best_key && kk < kKeyMax);
    if (chosen > kEvTile - 1) chosen = kEvTile - 1;
    const double high = lanes_max(s);
    // each lane's term in parallel, summed by one lane in lane order
    s_sum[lane] = high > -kInf ? fm::exp(s - high) : 0.0;
    __syncthreads();
    if (lane == 0) {
      double log_sum = high;
      if (high > -kInf) {
        double acc = 0.0;
        for (uint32_t l = 0; l < kEvTile; l++) acc += s_sum[l];
        log_sum = high + fm::log(acc);
      }
      s_log = log_sum;
    }
    __syncthreads();
    log_sums[sel] = s_log;
    __syncthreads();
    if (lane == chosen) {
      // the chosen lane's winner of this selection; zeros when it has none
      double vals[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      const uint32_t tw = sel == 0 ? joint_t : signal_t;
      if (tw < T) {
        const size_t at = (size_t)tw * row + j;
        vals[0] = A.t_snr[at];
        vals[1] = A.t_depth[at];
        vals[2] = A.t_lp[at];
        vals[3] = A.t_lb[at];
        decode_scenario(A, tw / ns, A.t_code[at], vals + 4);
        vals[8] = A.t_lp[at];
      }
      double* out = A.parts + ((size_t)tile * 2 + sel) * kEvPart;
      out[0] = (double)best_key;
      out[1] = maximum;
      out[2] = sel == 0 ? vals[0] : best_signal;
      for (int i = 0; i < 7; i++) out[3 + i] = vals[1 + i];
      out[10] = sel == 0 ? best_joint : vals[8];
      out[13] = (double)omitted;
      out[14] = 0.0;
    }
  }
  if (lane == 0) {
    // both selections carry both sums (evidence.rs:623-631)
    double* out = A.parts + (size_t)tile * 2 * kEvPart;
    out[11] = log_sums[0];
    out[12] = log_sums[1];
    out[kEvPart + 11] = log_sums[0];
    out[kEvPart + 12] = log_sums[1];
  }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

size_t evidence_scratch_bytes(uint32_t nb2, uint32_t tiles, uint32_t templates) {
  const size_t cells = (size_t)templates * tiles * kEvTile;
  size_t b = align256((size_t)2 * nb2 * 8);       // hist
  b += align256((size_t)nb2 * 4);                 // hcount
  b += align256((size_t)3 * (nb2 + 1) * 8);       // cum
  b += align256((size_t)tiles * 4 * 8);           // screen
  b += align256((size_t)tiles * 2 * kEvPart * 8); // parts
  b += align256(4 * 8);                           // floors
  b += 4 * align256(cells * 8);                   // t_lp, t_lb, t_snr, t_depth
  b += align256(cells);                           // t_code
  b += align256((size_t)tiles * templates * 4);   // t_kept
  return b;
}

void evidence_bind_scratch(EvArgs* a, void* base) {
  const size_t cells = (size_t)a->nd * a->ns * a->tiles * kEvTile;
  unsigned char* p = (unsigned char*)base;
  a->hist = (unsigned long long*)p;
  p += align256((size_t)2 * a->nb2 * 8);
  a->hcount = (unsigned int*)p;
  p += align256((size_t)a->nb2 * 4);
  a->cum = (double*)p;
  p += align256((size_t)3 * (a->nb2 + 1) * 8);
  a->screen = (double*)p;
  p += align256((size_t)a->tiles * 4 * 8);
  a->parts = (double*)p;
  p += align256((size_t)a->tiles * 2 * kEvPart * 8);
  a->floors = (double*)p;
  p += align256(4 * 8);
  a->t_lp = (double*)p;
  p += align256(cells * 8);
  a->t_lb = (double*)p;
  p += align256(cells * 8);
  a->t_snr = (double*)p;
  p += align256(cells * 8);
  a->t_depth = (double*)p;
  p += align256(cells * 8);
  a->t_code = (uint8_t*)p;
  p += align256(cells);
  a->t_kept = (uint32_t*)p;
}

hipError_t evidence_launch(const EvArgs& a, hipStream_t s) {
  const uint32_t zero_n = 2 * a.nb2 > 4 * a.tiles ? 2 * a.nb2 : 4 * a.tiles;
  const uint32_t zero_blocks = min((zero_n + 255) / 256, 1024u);
  hipLaunchKernelGGL(ev_zero, dim3(zero_blocks), dim3(256), 0, s, a);
  const uint32_t fold_threads = (a.n + kRun - 1) / kRun;
  hipLaunchKernelGGL(ev_fold, dim3((fold_threads + 255) / 256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(ev_prefix, dim3(3), dim3(1024), 0, s, a);
  hipLaunchKernelGGL(ev_screen, dim3(a.tiles), dim3(64 * kScreenWaves), 0, s, a);
  hipLaunchKernelGGL(ev_pilot, dim3(1), dim3(64), 0, s, a);
  hipLaunchKernelGGL(ev_templates, dim3(a.tiles, a.nd * a.ns), dim3(64), 0, s, a);
  hipLaunchKernelGGL(ev_reduce, dim3(a.tiles), dim3(64), 0, s, a);
  return hipGetLastError();
}

}  // namespace fbls

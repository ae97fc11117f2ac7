// Joint planet / binary / noise evidence at one trial period (gfx950).
//
// Six kernels on one stream, no host round trip in between:
//   ev_zero    clears the bin sums and the screen
//   ev_fold    two-period fold: bin = floor(t * scale) mod 2 nb, exact
//              fixed-point sums of v and w (int64 atomics) and counts,
//              with runs of equal bins aggregated in registers
//   ev_prefix  exclusive prefix per field (one workgroup each), exact
//   ev_screen  one wave per 64-bin tile: best SNR, and the planet / binary
//              evidence lower bounds of the best-merit template
//   ev_pilot   threshold check of the screen, the pilot template (winner
//              tile by lower bound, last of equals), its joint evidence and
//              the pruning floors
//   ev_joint   one wave per tile: evidence of every (phase, duration, shape)
//              with secondary-eclipse and odd/even scenarios, pruned by
//              subspace upper bounds against the floors; tile winners by
//              joint evidence and by SNR, tile log-sums
//
// Reference semantics restated (isakruas/fbls src/search/evidence.rs):
//   circular_sum / layered_sum   evidence.rs:37-67
//   fold_two / prefix_fields     evidence.rs:75-144
//   comparison_screen            evidence.rs:151-273
//   joint_tiles                  evidence.rs:280-632
//   screen threshold, pilot key, floors   evidence.rs:847-896
// A tile of the reference is a 64-thread block with shared-memory
// reductions; here it is exactly one wavefront and the reductions are
// cross-lane shuffles.  Every floating-point expression keeps the
// reference's operand order (built with -ffp-contract=off) and the
// transcendental functions are fmath.h's, so the values are the same
// doubles the host restatement computes.  Sums whose order matters (tile
// log-sum-exp) are taken sequentially in lane order, as the reference does.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "evidence_kernel.h"
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
    sv += A.vw[2 * i];
    sw += A.vw[2 * i + 1];
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

// ---------------------------------------------------------------- screen

// evidence.rs:151-273
__global__ __launch_bounds__(64) void ev_screen(EvArgs A) {
  const uint32_t lane = threadIdx.x;
  const uint32_t tile = blockIdx.x;
  const uint64_t nb2 = A.nb2, nb = A.nb;
  const uint64_t j = (uint64_t)tile * kEvTile + lane;
  const uint64_t jj = j < nb ? j : nb - 1;
  const double* cv = A.cum;
  const double* cw = A.cum + (nb2 + 1);
  const double* cn = A.cum + 2 * (nb2 + 1);
  const double n = (double)A.n, c = A.C;
  double best = -kInf, bound_merit = -kInf;
  double best_a = 0.0, best_q = 1.0, best_null = 0.0, best_ps = 0.0, best_bs = 0.0;
  uint64_t best_key = 0;
  for (uint32_t di = 0; di < A.nd; di++) {
    const uint64_t raw_k = template_bins(A.dur[di], A.scale);
    const uint64_t k = raw_k < nb2 / 8 - 1 ? raw_k : nb2 / 8 - 1;
    const double r = A.rho[di];
    const double count = circular_sum(cn, jj, k, nb2) + circular_sum(cn, jj + nb, k, nb2);
    const bool usable = j < nb && raw_k < nb2 / 8 && count >= 3.0 && count <= n - 3.0;
    for (uint32_t si = 0; si < A.ns; si++) {
      const double a = layered_sum(cv, jj, k, A.shapes, si, nb2, A.layers, 1) +
                       layered_sum(cv, jj + nb, k, A.shapes, si, nb2, A.layers, 1);
      const double pe = layered_sum(cw, jj, k, A.shapes, si, nb2, A.layers, 1);
      const double po = layered_sum(cw, jj + nb, k, A.shapes, si, nb2, A.layers, 1);
      const double ne = layered_sum(cw, jj, k, A.shapes, si, nb2, A.layers, 2);
      const double no = layered_sum(cw, jj + nb, k, A.shapes, si, nb2, A.layers, 2);
      const double qe = ne - pe * pe / c, qo = no - po * po / c, qeo = -pe * po / c;
      const double q = qe + qo + 2.0 * qeo;
      const double score = a / sqrt(fmax(q * (r * r), 1e-300));
      const bool valid = usable && a > 0.0 && q > 0.0;
      const double merit = 0.5 * score * score + A.dlm[di] + A.priors[si * 2];
      if (valid && merit > bound_merit) {
        best_a = a / (r * r);
        best_q = q / (r * r);
        best_null = A.dlm[di];
        best_ps = A.priors[si * 2];
        best_bs = A.priors[si * 2 + 1];
        best_key = (uint64_t)(si * A.nd + di) * nb + jj;
        bound_merit = merit;
      }
      if (valid && score > best) best = score;
    }
  }
  const double maximum = wave_max(best);
  const double bound_maximum = wave_max(bound_merit);
  // first lane holding the largest merit (none when every merit is -inf)
  const uint32_t chosen = first_lane(bound_merit > -kInf && bound_merit == bound_maximum);
  if (lane == chosen) {
    const double* P = A.params;
    const double lower_p = fm::positive_amplitude_evidence(best_a, best_q, P[0]) +
                           fm::log(0.75) + best_null + best_ps;
    const double lower_b = fm::log_add(fm::positive_amplitude_evidence(best_a, best_q, P[1]),
                                       fm::positive_amplitude_evidence(best_a, best_q, P[2])) +
                           fm::log(0.025) + best_null + best_bs;
    double* out = A.screen + (size_t)tile * 4;
    out[0] = maximum;
    out[1] = lower_p;
    out[2] = lower_b;
    out[3] = (double)best_key;
  }
}

// ----------------------------------------------------------------- joint

struct JointLds {
  double j_out[9][kEvTile];
  double s_out[9][kEvTile];
  double sums[kEvTile];
};

// evidence.rs:280-632 for one tile (one wave).  In pilot mode every lane
// evaluates the template pilot_key and the floors are -inf.  Returns the
// tile's log-sums {planet, binary}; writes the parts when `parts` is set.
__device__ void joint_tile(const EvArgs& A, uint32_t tile, bool pilot,
                           double floor_p, double floor_b, uint64_t pilot_key,
                           double* parts, JointLds& L, double* sp_out,
                           double* sb_out) {
  const uint32_t lane = threadIdx.x;
  const uint64_t nb2 = A.nb2, nb = A.nb;
  const uint32_t nd = A.nd, ns = A.ns, layers = A.layers;
  const double* cv = A.cum;
  const double* cw = A.cum + (nb2 + 1);
  const double* cn = A.cum + 2 * (nb2 + 1);
  const double* P = A.params;
  const double c = A.C, n = (double)A.n;
  const double tau_p = P[0], tau_b = P[1], tau_d = P[2];
  const uint64_t j = pilot ? pilot_key % nb : (uint64_t)tile * kEvTile + lane;
  const uint64_t jj = j < nb ? j : nb - 1;
  if (pilot) {
    floor_p = -kInf;
    floor_b = -kInf;
  }
  double best_joint = -kInf, best_signal = -kInf;
  uint64_t joint_key = kKeyMax, signal_key = kKeyMax;
  for (int i = 0; i < 9; i++) {
    L.j_out[i][lane] = 0.0;
    L.s_out[i][lane] = 0.0;
  }
  double sum_p = -kInf, sum_b = -kInf;
  uint32_t omitted = 0;
  uint32_t di_begin = 0, di_end = nd, si_begin = 0, si_end = ns;
  if (pilot) {
    di_begin = (uint32_t)((pilot_key / nb) % nd);
    di_end = di_begin + 1;
    si_begin = (uint32_t)(pilot_key / (nb * nd));
    si_end = si_begin + 1;
  }
  for (uint32_t di = di_begin; di < di_end; di++) {
    const uint64_t raw_k = template_bins(A.dur[di], A.scale);
    const uint64_t k = raw_k < nb2 / 8 - 1 ? raw_k : nb2 / 8 - 1;
    const double rho2 = A.rho[di] * A.rho[di];
    const double dlm = A.dlm[di];
    const double count = circular_sum(cn, jj, k, nb2) + circular_sum(cn, jj + nb, k, nb2);
    const bool usable = j < nb && raw_k < nb2 / 8 && count >= 3.0 && count <= n - 3.0;
    if (!(raw_k < nb2 / 8)) continue;  // uniform over the wave
    for (uint32_t si = si_begin; si < si_end; si++) {
      const double ae = layered_sum(cv, jj, k, A.shapes, si, nb2, layers, 1);
      const double ao = layered_sum(cv, jj + nb, k, A.shapes, si, nb2, layers, 1);
      const double pe = layered_sum(cw, jj, k, A.shapes, si, nb2, layers, 1);
      const double po = layered_sum(cw, jj + nb, k, A.shapes, si, nb2, layers, 1);
      const double ne = layered_sum(cw, jj, k, A.shapes, si, nb2, layers, 2);
      const double no = layered_sum(cw, jj + nb, k, A.shapes, si, nb2, layers, 2);
      const double qe = ne - pe * pe / c, qo = no - po * po / c, qeo = -pe * po / c;
      const double ap = ae + ao, qp = qe + qo + 2.0 * qeo;
      const double snr = ap / sqrt(fmax(qp * rho2, 1e-300));
      const double depth = ap / fmax(qp, 1e-300);
      const bool valid = usable && ap > 0.0 && qp > 0.0 && snr >= A.threshold;
      double lp = -kInf, lb = -kInf, component_best = -kInf;
      double sec_ratio = 0.0, sec_phase = 0.5, odd_ratio = 1.0, binary_multiple = 1.0;
      double omitted_upper = -kInf;
      for (uint32_t o = 0; o < kEvOrbits; o++) {
        const double* orb = A.orbit + (di * kEvOrbits + o) * 4;
        const double phase = orb[0], duration_ratio = orb[1];
        const double logwp = orb[2], logwb = orb[3];
        const int64_t offset = (int64_t)floor(phase * (double)nb + 0.5);
        const double skf = floor((double)k * duration_ratio + 0.5);
        int64_t ski = skf < 1.0 ? 1 : (int64_t)skf;
        if (ski > (int64_t)nb - 1) ski = (int64_t)nb - 1;
        const uint64_t sk = (uint64_t)ski;
        const uint64_t s0 = (uint64_t)((int64_t)jj + offset +
                                       (int64_t)floor(((double)k - (double)sk) / 2.0));
        const uint64_t s1 = s0 + nb;
        const double ass = layered_sum(cv, s0, sk, A.shapes, si, nb2, layers, 1) +
                           layered_sum(cv, s1, sk, A.shapes, si, nb2, layers, 1);
        const double ps = layered_sum(cw, s0, sk, A.shapes, si, nb2, layers, 1) +
                          layered_sum(cw, s1, sk, A.shapes, si, nb2, layers, 1);
        const double nsum = layered_sum(cw, s0, sk, A.shapes, si, nb2, layers, 2) +
                            layered_sum(cw, s1, sk, A.shapes, si, nb2, layers, 2);
        const double qs = nsum - ps * ps / c, qps = -(pe + po) * ps / c;
        const double pa = ap / rho2, pb = ass / rho2, pq = qp / rho2, pr = qs / rho2,
                     px = qps / rho2;
        const double cu = 1e-6;
        const double pu = fm::subspace_upper(pa, pb, pq, pr, px, 1.0) + logwp + cu;
        const double bu = fm::subspace_upper(pa, pb, pq, pr, px, 1.0) + logwb + cu;
        double upper = fmax(pu - floor_p, bu - floor_b) + dlm;
        if (!(logwp > -kInf || logwb > -kInf)) upper = -kInf;
        const double raw_upper = logwb > -kInf ? bu : -kInf;
        const double tile_max = wave_max(j < nb ? upper : -kInf);
        if (tile_max >= 0.0) {
          for (int ri = 0; ri < 5; ri++) {
            const double ratio = P[6 + ri * 3];
            const double wp = P[7 + ri * 3], wb = P[8 + ri * 3];
            const double aa = pa + ratio * pb;
            const double qq = pq + 2.0 * ratio * px + ratio * ratio * pr;
            const double ep = fm::positive_amplitude_evidence(aa, qq, tau_p) + fm::log(wp) + logwp;
            const double diluted = fm::positive_amplitude_evidence(aa, qq, tau_d);
            const double eb = fm::log_add(fm::positive_amplitude_evidence(aa, qq, tau_b), diluted) +
                              fm::log(0.25 * wb) + logwb;
            lp = fm::log_add(lp, ep);
            lb = fm::log_add(lb, eb);
            if (eb > component_best) {
              sec_ratio = ratio;
              sec_phase = phase;
              component_best = eb;
            }
          }
        } else {
          omitted += 1;
          if (raw_upper > omitted_upper) omitted_upper = raw_upper;
        }
      }
      // alternating primaries at twice the period (odd/even windows)
      const double ba = ae / rho2, bb = ao / rho2, bq = qe / rho2, br = qo / rho2,
                   bx = qeo / rho2;
      const double alias_mass = k < nb2 / 8 ? 0.0 : -kInf;
      const double raw_upper = fm::subspace_upper(ba, bb, bq, br, bx, 1.0) + 1e-6 + alias_mass;
      const double upper = raw_upper + dlm;
      const double tile_max = wave_max(j < nb ? upper : -kInf);
      if (tile_max >= floor_b) {
        for (int ri = 0; ri < 7; ri++) {
          const double ratio = P[21 + ri];
          const double aa = ba + ratio * bb;
          const double qq = bq + 2.0 * ratio * bx + ratio * ratio * br;
          const double diluted = fm::positive_amplitude_evidence(aa, qq, tau_d);
          const double eb = fm::log_add(fm::positive_amplitude_evidence(aa, qq, tau_b), diluted) +
                            alias_mass + fm::log(0.25 / 7.0);
          lb = fm::log_add(lb, eb);
          if (eb > component_best) {
            sec_ratio = 0.0;
            sec_phase = 0.5;
            odd_ratio = ratio;
            binary_multiple = 2.0;
            component_best = eb;
          }
        }
      } else {
        omitted += 1;
        if (raw_upper > omitted_upper) omitted_upper = raw_upper;
      }
      if (!(component_best > omitted_upper)) {
        sec_ratio = __builtin_nan("");
        sec_phase = __builtin_nan("");
        odd_ratio = __builtin_nan("");
        binary_multiple = __builtin_nan("");
      }
      lp += A.priors[si * 2] + dlm;
      lb += A.priors[si * 2 + 1] + dlm;
      if (j < nb) {
        sum_p = fm::log_add(sum_p, lp);
        sum_b = fm::log_add(sum_b, lb);
      }
      const double joint = lp;
      const uint64_t key = (uint64_t)(si * nd + di) * nb + jj;
      const bool bj = valid && (joint > best_joint || (joint == best_joint && key < joint_key));
      const bool bs = valid && (snr > best_signal || (snr == best_signal && key < signal_key));
      if (bj || bs) {
        const double vals[9] = {snr, depth, lp, lb, sec_ratio, sec_phase,
                                odd_ratio, binary_multiple, joint};
        for (int i = 0; i < 9; i++) {
          if (bj) L.j_out[i][lane] = vals[i];
          if (bs) L.s_out[i][lane] = vals[i];
        }
      }
      if (bj) {
        best_joint = joint;
        joint_key = key;
      }
      if (bs) {
        best_signal = snr;
        signal_key = key;
      }
    }
  }
  double log_sums[2];
  for (int sel = 0; sel < 2; sel++) {
    const double v = sel == 0 ? best_joint : best_signal;
    const uint64_t kk = sel == 0 ? joint_key : signal_key;
    const double s = sel == 0 ? sum_p : sum_b;
    const double maximum = wave_max(v);
    const uint64_t best_key = wave_min_u64(v == maximum ? kk : kKeyMax);
    uint32_t chosen = first_lane(v == maximum && kk == best_key && kk < kKeyMax);
    if (chosen > kEvTile - 1) chosen = kEvTile - 1;
    const double high = wave_max(s);
    L.sums[lane] = s;
    __syncthreads();
    double log_sum = high;
    if (high > -kInf) {
      double acc = 0.0;
      for (uint32_t l = 0; l < kEvTile; l++) acc += fm::exp(L.sums[l] - high);
      log_sum = high + fm::log(acc);
    }
    __syncthreads();
    log_sums[sel] = log_sum;
    if (parts && lane == chosen) {
      double* out = parts + ((size_t)tile * 2 + sel) * kEvPart;
      const double (*vals)[kEvTile] = sel == 0 ? L.j_out : L.s_out;
      out[0] = (double)best_key;
      out[1] = maximum;
      out[2] = sel == 0 ? vals[0][lane] : best_signal;
      for (int i = 0; i < 7; i++) out[3 + i] = vals[1 + i][lane];
      out[10] = sel == 0 ? best_joint : vals[8][lane];
      out[13] = (double)omitted;
      out[14] = 0.0;
    }
  }
  if (parts && lane == 0) {
    // both selections carry both sums (evidence.rs:623-631)
    double* out = parts + (size_t)tile * 2 * kEvPart;
    out[11] = log_sums[0];
    out[12] = log_sums[1];
    out[kEvPart + 11] = log_sums[0];
    out[kEvPart + 12] = log_sums[1];
  }
  *sp_out = log_sums[0];
  *sb_out = log_sums[1];
}

// evidence.rs:847-896 on the device: screen threshold, pilot template,
// floors.  One wave.
__global__ __launch_bounds__(64) void ev_pilot(EvArgs A) {
  __shared__ JointLds L;
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
  double sp, sb;
  joint_tile(A, 0, true, -kInf, -kInf, (uint64_t)pilot_key, nullptr, L, &sp, &sb);
  if (lane == 0) {
    A.floors[0] = sp + A.margin;
    A.floors[1] = sb + A.margin;
    A.floors[2] = pilot_key;
    A.floors[3] = 1.0;
  }
}

__global__ __launch_bounds__(64) void ev_joint(EvArgs A) {
  __shared__ JointLds L;
  if (A.floors[3] == 0.0) return;
  double sp, sb;
  joint_tile(A, blockIdx.x, false, A.floors[0], A.floors[1], (uint64_t)A.floors[2],
             A.parts, L, &sp, &sb);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

size_t evidence_scratch_bytes(uint32_t nb2, uint32_t tiles) {
  size_t b = align256((size_t)2 * nb2 * 8);       // hist
  b += align256((size_t)nb2 * 4);                 // hcount
  b += align256((size_t)3 * (nb2 + 1) * 8);       // cum
  b += align256((size_t)tiles * 4 * 8);           // screen
  b += align256((size_t)tiles * 2 * kEvPart * 8); // parts
  b += align256(4 * 8);                           // floors
  return b;
}

void evidence_bind_scratch(EvArgs* a, void* base) {
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
}

hipError_t evidence_launch(const EvArgs& a, hipStream_t s) {
  const uint32_t zero_n = 2 * a.nb2 > 4 * a.tiles ? 2 * a.nb2 : 4 * a.tiles;
  const uint32_t zero_blocks = min((zero_n + 255) / 256, 1024u);
  hipLaunchKernelGGL(ev_zero, dim3(zero_blocks), dim3(256), 0, s, a);
  const uint32_t fold_threads = (a.n + kRun - 1) / kRun;
  hipLaunchKernelGGL(ev_fold, dim3((fold_threads + 255) / 256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(ev_prefix, dim3(3), dim3(1024), 0, s, a);
  hipLaunchKernelGGL(ev_screen, dim3(a.tiles), dim3(64), 0, s, a);
  hipLaunchKernelGGL(ev_pilot, dim3(1), dim3(64), 0, s, a);
  hipLaunchKernelGGL(ev_joint, dim3(a.tiles), dim3(64), 0, s, a);
  return hipGetLastError();
}

}  // namespace fbls

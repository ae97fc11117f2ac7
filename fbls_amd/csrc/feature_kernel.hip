// Coarse periodogram of the band-proposal network (gfx950): for every star
// of a batch and every period of a shared log-spaced list, the largest box
// SNR over phase for each of a few representative durations.
//
// One workgroup (512 threads) per star and group of up to 16 consecutive
// periods whose histograms fit the LDS budget (each lane's observations are
// loaded once for the whole group): the star's fixed-point fields are folded
// into an LDS histogram of nb bins per period (nb = clamp(ceil(2 P / d_0),
// nb_min, nb_max), d_0 the shortest duration) with exact integer sums held
// in doubles (fold.h: runs of consecutive observations in one bin summed
// from per-lane prefix sums), turned into an exclusive prefix by the waves in
// parallel.  Then one wave per (period, duration d_g) task (width k =
// max(1, floor(d_g nb / P + 0.5)) bins, circular): the phases j that hold at
// least three points (and at most n - 3) with a positive precision q = w -
// w^2 / C qualify; the value is a / sqrt(q) at the qualifying phase with the
// largest key a |a| / q (a monotone function of a / sqrt(q); the smallest
// phase on equal keys), -inf when none qualifies.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "feature_kernel.h"
#include "fixed_point.h"
#include "fold.h"

namespace fbls {

namespace {

constexpr uint32_t kBlock = 512;
constexpr double kInf = __builtin_inf();

// floor(t * scale) mod nb, plus one (slot b + 1 accumulates bin b)
__device__ __forceinline__ uint32_t slot_of(double t, double scale, uint32_t nb) {
  const double cell = floor(t * scale);
  if (cell < 4294967296.0) return (uint32_t)cell % nb + 1;
  return (uint32_t)((uint64_t)cell % nb) + 1;
}

// template width in bins for duration D: max(1, floor(D * scale + 0.5)), <= nb
__device__ __forceinline__ uint32_t width_of(double D, double scale, uint32_t nb) {
  const double kf = floor(D * scale + 0.5);
  uint32_t k = kf < 1.0 ? 1u : (uint32_t)kf;
  return k > nb ? nb : k;
}

template <bool UNI>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
void coarse_features(FeatureArgs A) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  __shared__ double red_m[kFeatureMaxGroup][kFeatureMaxDurations];
  const uint32_t gi = blockIdx.x % A.ngroups, star = blockIdx.x / A.ngroups;
  const uint32_t p0 = A.gstart[gi], gcount = A.gstart[gi + 1] - p0;
  const FeatureStar S = A.stars[star];
  const uint32_t slots = A.gstride[gi];               // slots per period of this group
  const uint32_t pstride = (slots * 20u + 15u) & ~15u; // {v, w} pairs, then counts
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  {
    uint4* h4 = (uint4*)lds;
    for (uint32_t i = tid; i < gcount * pstride / 16; i += kBlock) h4[i] = make_uint4(0, 0, 0, 0);
  }
  // lane g holds period p0 + g's scale, phase bins and float reciprocal (a
  // memory read inside the period loop would wait on the counter the LDS
  // atomics share and drain them)
  double my_scale = 0.0;
  uint32_t my_nb = 1;
  float my_inv = 0.0f;
  if (lane < gcount) {
    my_nb = A.nb[p0 + lane];
    my_scale = (double)my_nb / A.periods[p0 + lane];
    my_inv = fold::inv_nb_low(my_nb);
  }
  const uint64_t my_bits = (uint64_t)__double_as_longlong(my_scale);
  const int my_lo = (int)(uint32_t)my_bits, my_hi = (int)(uint32_t)(my_bits >> 32);
  __syncthreads();
  // fold: each lane loads 8 consecutive observations once for the whole
  // group and folds them into every period's histogram (fold.h)
  const uint32_t nblk = S.n_pad / fold::kObsPerWave;
  for (uint32_t blk = wave; blk < nblk; blk += kBlock / 64) {
    const uint32_t base = blk * fold::kObsPerWave + lane * fold::kObsPerLane;
    double tt[fold::kObsPerLane], pv[fold::kObsPerLane], pw[fold::kObsPerLane];
    fold::load_obs(S.t, S.vw, S.fx, base, tt, pv, pw);
#pragma unroll
    for (uint32_t r = 1; r < fold::kObsPerLane; r++) {  // inclusive prefix, exact
      pv[r] += pv[r - 1];
      if (!UNI) pw[r] += pw[r - 1];
    }
    const bool full = blk * fold::kObsPerWave + fold::kObsPerWave <= S.n;
    for (uint32_t g = 0; g < gcount; g++) {
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane(my_lo, (int)g);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane(my_hi, (int)g);
      const double scale = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
      const uint32_t nb = (uint32_t)__builtin_amdgcn_readlane((int)my_nb, (int)g);
      const float inv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_inv), (int)g));
      fold::HistAt h;
      h.vo = g * pstride;
      h.no = h.vo + 16u * slots;
      if (A.narrow) {
        uint32_t u[fold::kObsPerLane];
#pragma unroll
        for (uint32_t r = 0; r < fold::kObsPerLane; r++) u[r] = fold::phase_cell_u(tt[r], scale);
        if (!full)
          fold::fold_lane<false, true, false, UNI, false>(lds, h, u, pv, pw, nb, inv, base, S.n);
        else if (S.sorted && __all(u[fold::kObsPerLane - 1] - u[0] < min(fold::kOver, nb)))
          fold::fold_lane<false, true, true, UNI, true>(lds, h, u, pv, pw, nb, inv, base, S.n);
        else
          fold::fold_lane<false, true, true, UNI, false>(lds, h, u, pv, pw, nb, inv, base, S.n);
      } else {
        // cells beyond 2^24: one update per observation, bins by floor and
        // modulo in 64 bits
#pragma unroll
        for (uint32_t r = 0; r < fold::kObsPerLane; r++) {
          if (base + r >= S.n) continue;
          const uint32_t b = slot_of(tt[r], scale, nb);
          const double v = r ? pv[r] - pv[r - 1] : pv[0];
          const double w = UNI ? 0.0 : (r ? pw[r] - pw[r - 1] : pw[0]);
          fold::add_run<UNI>(lds, h.vo + 16u * b, h.no + 4u * b, v, w, 1u);
        }
      }
    }
  }
  __syncthreads();
  // overflow slots back into the first bins, equal weights rebuilt, then the
  // exclusive prefix of every period (slot x = sum of bins < x)
  for (uint32_t g = 0; g < gcount; g++) {
    const uint32_t nb = A.nb[p0 + g];
    const uint32_t wrap = min(fold::kOver, nb);
    double2* vw = (double2*)(lds + (size_t)g * pstride);
    uint32_t* n = (uint32_t*)(vw + slots);
    for (uint32_t x = tid; x < wrap; x += kBlock) {
      const double2 o = vw[nb + 1 + x], e = vw[1 + x];
      vw[1 + x] = make_double2(e.x + o.x, e.y + o.y);
      n[1 + x] += n[nb + 1 + x];
    }
  }
  __syncthreads();
  if (UNI) {
    const double w0 = (double)S.uniform_w;
    for (uint32_t g = 0; g < gcount; g++) {
      const uint32_t nb = A.nb[p0 + g];
      double2* vw = (double2*)(lds + (size_t)g * pstride);
      const uint32_t* n = (const uint32_t*)(vw + slots);
      for (uint32_t b = tid; b < nb; b += kBlock) vw[b + 1].y = (double)n[b + 1] * w0;
    }
    __syncthreads();
  }
  for (uint32_t task = wave; task < 2 * gcount; task += kBlock / 64) {
    const uint32_t g = task >> 1;
    double2* vw = (double2*)(lds + (size_t)g * pstride);
    if (task & 1)
      fold::wave_prefix_n((uint32_t*)(vw + slots), A.nb[p0 + g], lane);
    else
      fold::wave_prefix_vw(vw, A.nb[p0 + g], lane, 1.0, 1.0);
  }
  __syncthreads();
  // best SNR over phase per (period, duration): one wave per (period,
  // duration) task.  Each lane keeps the largest key a|a| / q (monotone in
  // a / sqrt(q)) over its phases j = lane, lane + 64, ... (ascending, so its
  // phase is the smallest one of its key) with the sums a, q it came from;
  // the wave's winner is the largest key, the smallest phase on equal keys,
  // and its SNR a / sqrt(q).
  const uint32_t ntask = gcount * A.ng;
  for (uint32_t task = wave; task < ntask; task += kBlock / 64) {
    const uint32_t g = task / A.ng, d = task - g * A.ng;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane(my_lo, (int)g);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane(my_hi, (int)g);
    const double scale = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
    const uint32_t nb = (uint32_t)__builtin_amdgcn_readlane((int)my_nb, (int)g);
    const uint32_t k = width_of(A.durations[d], scale, nb);
    const double2* vw = (const double2*)(lds + (size_t)g * pstride);
    const uint32_t* gn = (const uint32_t*)(vw + slots);
    const uint32_t tn = gn[nb];
    const double tv = vw[nb].x, tw = vw[nb].y;
    double best = -kInf, ba = 0.0, bq = 1.0;
    uint32_t bj = 0xffffffffu;
    for (uint32_t j = lane; j < nb; j += 64) {
      const uint32_t e = j + k;
      double a, w;
      uint32_t c;
      if (e <= nb) {
        a = (vw[e].x - vw[j].x) * S.sv;
        w = (vw[e].y - vw[j].y) * S.sw;
        c = gn[e] - gn[j];
      } else {  // exact integers: the order of the sums does not matter
        a = (vw[e - nb].x + (tv - vw[j].x)) * S.sv;
        w = (vw[e - nb].y + (tw - vw[j].y)) * S.sw;
        c = gn[e - nb] + (tn - gn[j]);
      }
      const double qq = w - w * w / S.C;
      if (c >= 3 && c + 3 <= tn && qq > 0.0) {
        const double key = a * fabs(a) / qq;
        if (key > best) {
          best = key;
          bj = j;
          ba = a;
          bq = qq;
        }
      }
    }
    double mx = best;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    double m = -kInf;
    if (mx > -kInf) {
      uint64_t at_max = __ballot(best == mx);
      if (__popcll(at_max) > 1) {  // equal keys: the smallest phase
        uint32_t jm = best == mx ? bj : 0xffffffffu;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) jm = min(jm, (uint32_t)__shfl_xor((int)jm, off, 64));
        at_max = __ballot(best == mx && bj == jm);
      }
      const int wl = __builtin_ctzll(at_max);
      m = __shfl(ba, wl, 64) / sqrt(__shfl(bq, wl, 64));
    }
    if (lane == 0) {
      const size_t at = ((size_t)star * A.np + p0 + g) * A.ng + d;
      if (A.out) A.out[at] = m;
      if (A.inputs) {  // the network input: clip(SNR, -5, 40) / 8, 0 when none
        const double x = m > -kInf ? m : 0.0;
        A.inputs[at] = (float)(fmin(fmax(x, -5.0), 40.0) / 8.0);
      }
      red_m[g][d] = m;
    }
  }
  if (A.peaks) {  // the largest SNR over the durations (unclipped)
    __syncthreads();
    if (tid < gcount) {
      double mx = -kInf;
      for (uint32_t d = 0; d < A.ng; d++) mx = fmax(mx, red_m[tid][d]);
      A.peaks[(size_t)star * A.np + p0 + tid] = mx > -kInf ? (float)mx : -INFINITY;
    }
  }
}

}  // namespace

size_t feature_lds_bytes(uint32_t nb_max, uint32_t G) {
  return (size_t)G * (((nb_max + 1 + fold::kOver) * 20 + 15) & ~(size_t)15);
}

uint32_t feature_slots(uint32_t nb_max) { return nb_max + 1 + fold::kOver; }

hipError_t feature_launch(const FeatureArgs& a, uint32_t nstars, hipStream_t s) {
  const size_t lds = a.lds_bytes;
  if (a.ng > kFeatureMaxDurations || a.G > kFeatureMaxGroup || a.ngroups == 0)
    return hipErrorInvalidValue;
  const void* fn = a.uniform ? (const void*)coarse_features<true>
                            : (const void*)coarse_features<false>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  if (a.uniform)
    hipLaunchKernelGGL((coarse_features<true>), dim3(nstars * a.ngroups), dim3(kBlock), lds, s, a);
  else
    hipLaunchKernelGGL((coarse_features<false>), dim3(nstars * a.ngroups), dim3(kBlock), lds, s, a);
  return hipGetLastError();
}

}  // namespace fbls

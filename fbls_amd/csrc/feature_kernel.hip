// Coarse periodogram of the band-proposal network (gfx950): for every star
// of a batch and every period of a shared log-spaced list, the largest box
// SNR over phase for each of a few representative durations.
//
// One workgroup (256 threads) per star and group of up to 4 consecutive
// periods whose histograms fit 24 KiB of LDS (each lane's observations are
// loaded once for the whole group): the star's fixed-point
// fields are folded into an LDS histogram of nb bins (nb = clamp(ceil(2 P /
// d_0), nb_min, nb_max), d_0 the shortest duration) with exact integer sums
// held in doubles (runs of consecutive observations in one bin are summed in
// registers first; the bins of a lane's 8 ascending times follow from the
// first by the differences of floor(t * scale)), turned into an exclusive
// prefix by the waves in parallel.  For every duration d_g (width k =
// max(1, floor(d_g nb / P + 0.5)) bins, circular) the phases j that hold at
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

// Sums of the circular window [j, j + k) from the exclusive prefix ({v, w}
// pairs and counts); true when it holds 3 .. n - 3 points and a positive
// precision q = w - w^2 / C.
__device__ __forceinline__ bool window(const double2* vw, const uint32_t* gn, uint32_t nb,
                                       uint32_t tn, uint32_t j, uint32_t k,
                                       const FeatureStar& S, double* pa, double* pq) {
  const uint32_t e = j + k;
  double a, w;
  uint32_t c;
  if (e <= nb) {
    a = (vw[e].x - vw[j].x) * S.sv;
    w = (vw[e].y - vw[j].y) * S.sw;
    c = gn[e] - gn[j];
  } else {
    a = (vw[e - nb].x + (vw[nb].x - vw[j].x)) * S.sv;
    w = (vw[e - nb].y + (vw[nb].y - vw[j].y)) * S.sw;
    c = gn[e - nb] + (gn[nb] - gn[j]);
  }
  const double q = w - w * w / S.C;
  *pa = a;
  *pq = q;
  return c >= 3 && c + 3 <= tn && q > 0.0;
}

template <uint32_t NG, bool UNI>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
void coarse_features(FeatureArgs A) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  __shared__ double red_key[kBlock / 64][NG];
  __shared__ uint32_t red_j[kBlock / 64][NG];
  __shared__ double red_m[NG];
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
  // best SNR over phase per (period, duration): the phase with the largest
  // key a|a| / q (monotone in a / sqrt(q); the smallest phase on ties),
  // found per thread, then by wave and workgroup reductions; its SNR
  // a / sqrt(q) is computed once at the end
  for (uint32_t g = 0; g < gcount; g++) {
    const uint32_t nb = A.nb[p0 + g];
    const double scale = (double)nb / A.periods[p0 + g];
    const double2* vw = (const double2*)(lds + (size_t)g * pstride);
    const uint32_t* gn = (const uint32_t*)(vw + slots);
    const uint32_t tn = gn[nb];
    const double tv = vw[nb].x, tw = vw[nb].y;
    uint32_t kd[NG];
    double best[NG];
    uint32_t bj[NG];
#pragma unroll
    for (uint32_t d = 0; d < NG; d++) {
      kd[d] = d < A.ng ? width_of(A.durations[d], scale, nb) : 1u;
      best[d] = -kInf;
      bj[d] = 0xffffffffu;
    }
    for (uint32_t j = tid; j < nb; j += kBlock) {
      const double vj = vw[j].x, wj = vw[j].y;
      const uint32_t nj = gn[j];
#pragma unroll
      for (uint32_t d = 0; d < NG; d++) {
        if (d < A.ng) {
          const uint32_t e = j + kd[d];
          double a, w;
          uint32_t c;
          if (e <= nb) {
            a = (vw[e].x - vj) * S.sv;
            w = (vw[e].y - wj) * S.sw;
            c = gn[e] - nj;
          } else {  // exact integers: the order of the sums does not matter
            a = (vw[e - nb].x + (tv - vj)) * S.sv;
            w = (vw[e - nb].y + (tw - wj)) * S.sw;
            c = gn[e - nb] + (tn - nj);
          }
          const double qq = w - w * w / S.C;
          if (c >= 3 && c + 3 <= tn && qq > 0.0) {
            const double key = a * fabs(a) / qq;
            if (key > best[d]) {
              best[d] = key;
              bj[d] = j;
            }
          }
        }
      }
    }
#pragma unroll
    for (uint32_t d = 0; d < NG; d++) {
      double bk = best[d];
      uint32_t b = bj[d];
      for (int off = 32; off >= 1; off >>= 1) {
        const double ok = __shfl_xor(bk, off, 64);
        const uint32_t oj = __shfl_xor(b, off, 64);
        if (ok > bk || (ok == bk && oj < b)) {
          bk = ok;
          b = oj;
        }
      }
      if (lane == 0) {
        red_key[wave][d] = bk;
        red_j[wave][d] = b;
      }
    }
    __syncthreads();
    if (tid < NG) {
      const uint32_t d = tid;
      double m = -kInf;
      if (d < A.ng) {
        double bk = red_key[0][d];
        uint32_t b = red_j[0][d];
        for (uint32_t w = 1; w < kBlock / 64; w++) {
          const double ok = red_key[w][d];
          const uint32_t oj = red_j[w][d];
          if (ok > bk || (ok == bk && oj < b)) {
            bk = ok;
            b = oj;
          }
        }
        if (b != 0xffffffffu) {
          double a, qq;
          window(vw, gn, nb, tn, b, width_of(A.durations[d], scale, nb), S, &a, &qq);
          m = a / sqrt(qq);
        }
        const size_t at = ((size_t)star * A.np + p0 + g) * A.ng + d;
        if (A.out) A.out[at] = m;
        if (A.inputs) {  // the network input: clip(SNR, -5, 40) / 8, 0 when none
          const double x = m > -kInf ? m : 0.0;
          A.inputs[at] = (float)(fmin(fmax(x, -5.0), 40.0) / 8.0);
        }
      }
      red_m[d] = m;
    }
    __syncthreads();
    if (A.peaks && tid == 0) {  // the largest SNR over the durations (unclipped)
      double mx = -kInf;
      for (uint32_t d = 0; d < A.ng; d++) mx = fmax(mx, red_m[d]);
      A.peaks[(size_t)star * A.np + p0 + g] = mx > -kInf ? (float)mx : -INFINITY;
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
  if (a.ng > 16 || a.G > kFeatureMaxGroup || a.ngroups == 0) return hipErrorInvalidValue;
  const void* fn;
  void (*launch)(const FeatureArgs&, uint32_t, size_t, hipStream_t);
  if (a.ng <= 8) {
    if (a.uniform) {
      fn = (const void*)coarse_features<8, true>;
      launch = [](const FeatureArgs& x, uint32_t g, size_t l, hipStream_t st) {
        hipLaunchKernelGGL((coarse_features<8, true>), dim3(g), dim3(kBlock), l, st, x);
      };
    } else {
      fn = (const void*)coarse_features<8, false>;
      launch = [](const FeatureArgs& x, uint32_t g, size_t l, hipStream_t st) {
        hipLaunchKernelGGL((coarse_features<8, false>), dim3(g), dim3(kBlock), l, st, x);
      };
    }
  } else {
    if (a.uniform) {
      fn = (const void*)coarse_features<16, true>;
      launch = [](const FeatureArgs& x, uint32_t g, size_t l, hipStream_t st) {
        hipLaunchKernelGGL((coarse_features<16, true>), dim3(g), dim3(kBlock), l, st, x);
      };
    } else {
      fn = (const void*)coarse_features<16, false>;
      launch = [](const FeatureArgs& x, uint32_t g, size_t l, hipStream_t st) {
        hipLaunchKernelGGL((coarse_features<16, false>), dim3(g), dim3(kBlock), l, st, x);
      };
    }
  }
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  launch(a, nstars * a.ngroups, lds, s);
  return hipGetLastError();
}

}  // namespace fbls

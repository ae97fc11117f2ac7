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

namespace fbls {

namespace {

constexpr uint32_t kBlock = 256;
constexpr double kInf = __builtin_inf();

// floor(t * scale) mod nb, plus one (slot b + 1 accumulates bin b)
__device__ __forceinline__ uint32_t slot_of(double t, double scale, uint32_t nb) {
  const double cell = floor(t * scale);
  if (cell < 4294967296.0) return (uint32_t)cell % nb + 1;
  return (uint32_t)((uint64_t)cell % nb) + 1;
}

// floor(t * scale) mod nb for a 32-bit cell u < 2^24: the quotient from a
// float reciprocal biased low by ~2^-22 is floor(u / nb) or one less, so one
// conditional subtraction (as an unsigned min) finishes.
__device__ __forceinline__ uint32_t bin_of(uint32_t u, uint32_t nb, float inv_lo) {
  const uint32_t q = (uint32_t)((float)u * inv_lo);
  const uint32_t r = u - __umul24(q, nb);
  return min(r, r - nb);
}

// The sums are integers below 2^53 held in doubles, so every addition is
// exact and the LDS double atomics are order-independent.
__device__ __forceinline__ void flush(double* hv, double* hw, uint32_t* hn, uint32_t b,
                                      double v, double w, uint32_t c) {
  __hip_atomic_fetch_add(hv + b, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  __hip_atomic_fetch_add(hw + b, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  __hip_atomic_fetch_add(hn + b, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Inclusive scan of slots 0..nb of one array by one wave: lane l owns a
// contiguous block of slots; block sums, a wave scan of them, then the
// running sums written back (slot 0 holds zero, so slot x ends as the sum
// of bins < x).
template <class T>
__device__ void wave_inclusive(T* a, uint32_t len, uint32_t lane) {
  const uint32_t per = (len + 63) / 64;
  const uint32_t b0 = min(lane * per, len), b1 = min(b0 + per, len);
  T sum = 0;
  for (uint32_t x = b0; x < b1; x++) sum += a[x];
  T inc = sum;
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const T y = __shfl_up(inc, off, 64);
    if (lane >= off) inc += y;
  }
  T run = inc - sum;
  for (uint32_t x = b0; x < b1; x++) {
    run += a[x];
    a[x] = run;
  }
}

// template width in bins for duration D: max(1, floor(D * scale + 0.5)), <= nb
__device__ __forceinline__ uint32_t width_of(double D, double scale, uint32_t nb) {
  const double kf = floor(D * scale + 0.5);
  uint32_t k = kf < 1.0 ? 1u : (uint32_t)kf;
  return k > nb ? nb : k;
}

// Sums of the circular window [j, j + k) from the exclusive prefix; true
// when it holds 3 .. n - 3 points and a positive precision q = w - w^2 / C.
__device__ __forceinline__ bool window(const double* gv, const double* gw, const uint32_t* gn,
                                       uint32_t nb, uint32_t tn, uint32_t j, uint32_t k,
                                       const FeatureStar& S, double* pa, double* pq) {
  const uint32_t e = j + k;
  double a, w;
  uint32_t c;
  if (e <= nb) {
    a = (gv[e] - gv[j]) * S.sv;
    w = (gw[e] - gw[j]) * S.sw;
    c = gn[e] - gn[j];
  } else {
    a = (gv[e - nb] + (gv[nb] - gv[j])) * S.sv;
    w = (gw[e - nb] + (gw[nb] - gw[j])) * S.sw;
    c = gn[e - nb] + (gn[nb] - gn[j]);
  }
  const double q = w - w * w / S.C;
  *pa = a;
  *pq = q;
  return c >= 3 && c + 3 <= tn && q > 0.0;
}

__device__ __forceinline__ double wave_max(double x) {
  for (int off = 32; off >= 1; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
  return x;
}

template <uint32_t NG>
__global__ __launch_bounds__(kBlock) void coarse_features(FeatureArgs A) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  const uint32_t gi = blockIdx.x % A.ngroups, star = blockIdx.x / A.ngroups;
  const uint32_t p0 = A.gstart[gi], gcount = A.gstart[gi + 1] - p0;
  const FeatureStar S = A.stars[star];
  const uint32_t stride = A.gstride[gi];  // slots per period of this group
  double* hv = (double*)lds;              // [gcount][stride]
  double* hw = hv + (size_t)gcount * stride;
  uint32_t* hn = (uint32_t*)(hw + (size_t)gcount * stride);
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (uint32_t b = tid; b < gcount * stride; b += kBlock) {
    hv[b] = 0;
    hw[b] = 0;
    hn[b] = 0;
  }
  __syncthreads();
  // fold: each lane loads 8 consecutive observations once per step and, for
  // every period of the group, sums runs of equal bins in registers (times
  // are sorted, so neighbours share bins) before one LDS update per run.
  // The group's phase-bin counts and scales sit in registers (G <= 4): a
  // memory read inside the period loop would wait on the counter the LDS
  // atomics share and drain them.
  uint32_t nbg[4];
  double scg[4];
  float invg[4];
#pragma unroll
  for (uint32_t g = 0; g < 4; g++) {
    const uint32_t p = p0 + (g < gcount ? g : 0);
    nbg[g] = A.nb[p];
    scg[g] = (double)nbg[g] / A.periods[p];
    invg[g] = (float)((1.0 / (double)nbg[g]) * (1.0 - 0x1p-22));
  }
  for (uint32_t base = 0; base < S.n; base += kBlock * 8) {
    const uint32_t i0 = base + tid * 8;
    if (i0 >= S.n) continue;
    const uint32_t cnt = min(8u, S.n - i0);
    double tt[8], vv[8], ww[8];
#pragma unroll
    for (uint32_t r = 0; r < 8; r++) {
      const uint32_t i = r < cnt ? i0 + r : i0;
      tt[r] = S.t[i];
      vv[r] = fixed_dvalue(S.vw[2 * (size_t)i], S.fx & 1u);
      ww[r] = fixed_dvalue(S.vw[2 * (size_t)i + 1], S.fx & 2u);
    }
#pragma unroll
    for (uint32_t g = 0; g < 4; g++) {
      if (g >= gcount) break;
      const uint32_t nb = nbg[g];
      const double scale = scg[g];
      double* gv = hv + (size_t)g * stride;
      double* gw = hw + (size_t)g * stride;
      uint32_t* gn = hn + (size_t)g * stride;
      if (A.narrow && cnt == 8) {
        // 32-bit cells; when the times ascend and the lane's cells span
        // fewer than nb, each bin is the previous one advanced by the
        // difference of the cells, wrapped once
        uint32_t u[8];
#pragma unroll
        for (uint32_t r = 0; r < 8; r++) u[r] = (uint32_t)(tt[r] * scale);
        const bool incr = S.sorted && __all(u[7] - u[0] < nb);
        uint32_t cb = bin_of(u[0], nb, invg[g]);
        double av = vv[0], aw = ww[0];
        uint32_t ac = 1;
#pragma unroll
        for (uint32_t r = 1; r < 8; r++) {
          uint32_t b;
          bool change;
          if (incr) {
            const uint32_t dlt = u[r] - u[r - 1];
            change = dlt != 0;
            b = cb + dlt;
            b = min(b, b - nb);
          } else {
            b = bin_of(u[r], nb, invg[g]);
            change = b != cb;
          }
          if (change) flush(gv, gw, gn, cb + 1, av, aw, ac);
          const double sv = av + vv[r], sw = aw + ww[r];
          av = change ? vv[r] : sv;
          aw = change ? ww[r] : sw;
          ac = change ? 1u : ac + 1u;
          cb = b;
        }
        flush(gv, gw, gn, cb + 1, av, aw, ac);
      } else {
        // the star's last partial chunk, or cells beyond 2^24: one update
        // per observation, bins by floor and modulo
        for (uint32_t r = 0; r < cnt; r++)
          flush(gv, gw, gn, slot_of(tt[r], scale, nb), vv[r], ww[r], 1u);
      }
    }
  }
  __syncthreads();
  // exclusive prefix of every (period, field): slot x = sum of bins < x
  for (uint32_t task = wave; task < 3 * gcount; task += kBlock / 64) {
    const uint32_t g = task / 3, f = task % 3;
    const uint32_t len = A.nb[p0 + g] + 1;
    if (f == 0) wave_inclusive(hv + (size_t)g * stride, len, lane);
    if (f == 1) wave_inclusive(hw + (size_t)g * stride, len, lane);
    if (f == 2) wave_inclusive(hn + (size_t)g * stride, len, lane);
  }
  __syncthreads();
  // best SNR over phase per (period, duration): the phase with the largest
  // key a|a| / q (monotone in a / sqrt(q); the smallest phase on ties) is
  // found per thread (each thread reads its phase's prefix once for every
  // duration), then by wave and workgroup reductions; its SNR a / sqrt(q) is
  // computed once at the end
  __shared__ double red_key[kBlock / 64][4 * NG];
  __shared__ uint32_t red_j[kBlock / 64][4 * NG];
  for (uint32_t g = 0; g < gcount; g++) {
    const uint32_t nb = A.nb[p0 + g];
    const double scale = (double)nb / A.periods[p0 + g];
    const double* gv = hv + (size_t)g * stride;
    const double* gw = hw + (size_t)g * stride;
    const uint32_t* gn = hn + (size_t)g * stride;
    const uint32_t tn = gn[nb];
    const double tv = gv[nb], tw = gw[nb];
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
      const double vj = gv[j], wj = gw[j];
      const uint32_t nj = gn[j];
#pragma unroll
      for (uint32_t d = 0; d < NG; d++) {
        if (d < A.ng) {
          const uint32_t e = j + kd[d];
          double a, w;
          uint32_t c;
          if (e <= nb) {
            a = (gv[e] - vj) * S.sv;
            w = (gw[e] - wj) * S.sw;
            c = gn[e] - nj;
          } else {  // exact integers: the order of the sums does not matter
            a = (gv[e - nb] + (tv - vj)) * S.sv;
            w = (gw[e - nb] + (tw - wj)) * S.sw;
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
      if (lane == 0 && d < A.ng) {
        red_key[wave][g * NG + d] = bk;
        red_j[wave][g * NG + d] = b;
      }
    }
  }
  __syncthreads();
  double m = -kInf;
  const bool mine = tid < gcount * NG && tid % NG < A.ng;
  if (mine) {
    const uint32_t g = tid / NG, d = tid % NG;
    double bk = red_key[0][tid];
    uint32_t b = red_j[0][tid];
    for (uint32_t w = 1; w < kBlock / 64; w++) {
      const double ok = red_key[w][tid];
      const uint32_t oj = red_j[w][tid];
      if (ok > bk || (ok == bk && oj < b)) {
        bk = ok;
        b = oj;
      }
    }
    if (b != 0xffffffffu) {
      const uint32_t nb = A.nb[p0 + g];
      const double scale = (double)nb / A.periods[p0 + g];
      double a, qq;
      window(hv + (size_t)g * stride, hw + (size_t)g * stride, hn + (size_t)g * stride, nb,
             hn[(size_t)g * stride + nb], b, width_of(A.durations[d], scale, nb), S, &a, &qq);
      m = a / sqrt(qq);
    }
    const size_t at = ((size_t)star * A.np + p0 + g) * A.ng + d;
    if (A.out) A.out[at] = m;
    if (A.inputs) {  // the network input: clip(SNR, -5, 40) / 8, 0 when none
      const double x = m > -kInf ? m : 0.0;
      A.inputs[at] = (float)(fmin(fmax(x, -5.0), 40.0) / 8.0);
    }
  }
  if (A.peaks) {  // per period, the largest SNR over the durations (unclipped)
    __syncthreads();  // every read of red_key above is done
    if (tid < gcount * NG) red_key[0][tid] = mine ? m : -kInf;
    __syncthreads();
    if (tid < gcount) {
      double mx = -kInf;
      for (uint32_t d = 0; d < A.ng; d++) mx = fmax(mx, red_key[0][tid * NG + d]);
      A.peaks[(size_t)star * A.np + p0 + tid] = mx > -kInf ? (float)mx : -INFINITY;
    }
  }
}

}  // namespace

size_t feature_lds_bytes(uint32_t nb_max, uint32_t G) {
  return (size_t)G * (nb_max + 1) * 20 + 16;
}

hipError_t feature_launch(const FeatureArgs& a, uint32_t nstars, hipStream_t s) {
  const size_t lds = a.lds_bytes;
  if (a.ng > 16 || a.G > 4 || a.ngroups == 0) return hipErrorInvalidValue;
  const void* fn = a.ng <= 8 ? (const void*)coarse_features<8> : (const void*)coarse_features<16>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const uint32_t groups = a.ngroups;
  if (a.ng <= 8)
    hipLaunchKernelGGL(coarse_features<8>, dim3(nstars * groups), dim3(kBlock), lds, s, a);
  else
    hipLaunchKernelGGL(coarse_features<16>, dim3(nstars * groups), dim3(kBlock), lds, s, a);
  return hipGetLastError();
}

}  // namespace fbls

// Coarse periodogram of the band-proposal network (gfx950): for every star
// of a batch and every period of a shared log-spaced list, the largest box
// SNR over phase for each of a few representative durations.
//
// One workgroup (256 threads) per (star, period): the star's fixed-point
// fields are folded into an LDS histogram of nb bins (nb = clamp(ceil(2 P /
// d_0), nb_min, nb_max), d_0 the shortest duration) with exact 64-bit
// integer sums, turned into an exclusive prefix, and every phase j of every
// duration d_g (width k = max(1, floor(d_g nb / P + 0.5)) bins, circular) is
// scored as a / sqrt(w - w^2 / C) when it holds at least three points and a
// positive precision.  The value is -inf when no phase qualifies.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "feature_kernel.h"

namespace fbls {

namespace {

constexpr uint32_t kBlock = 256;
constexpr double kInf = __builtin_inf();

__global__ __launch_bounds__(kBlock) void coarse_features(FeatureArgs A) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  const uint32_t star = blockIdx.x / A.np, pi = blockIdx.x % A.np;
  const FeatureStar S = A.stars[star];
  const double P = A.periods[pi];
  const uint32_t nb = A.nb[pi];
  const double scale = (double)nb / P;
  long long* hv = (long long*)lds;   // [nb + 1]
  long long* hw = hv + (nb + 1);     // [nb + 1]
  uint32_t* hn = (uint32_t*)(hw + (nb + 1));  // [nb + 1]
  __shared__ double red[kBlock];
  for (uint32_t b = threadIdx.x; b <= nb; b += kBlock) {
    hv[b] = 0;
    hw[b] = 0;
    hn[b] = 0;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < S.n; i += kBlock) {
    const uint64_t cell = (uint64_t)floor(S.t[i] * scale);
    const uint32_t b = (uint32_t)(cell % nb) + 1;  // slot b + 1 holds bin b
    atomicAdd((unsigned long long*)&hv[b], (unsigned long long)(long long)S.vw[2 * i]);
    atomicAdd((unsigned long long*)&hw[b], (unsigned long long)(long long)S.vw[2 * i + 1]);
    atomicAdd(&hn[b], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // exclusive prefix: slot x = sum of bins < x
    for (uint32_t b = 1; b <= nb; b++) {
      hv[b] += hv[b - 1];
      hw[b] += hw[b - 1];
      hn[b] += hn[b - 1];
    }
  }
  __syncthreads();
  const uint32_t tn = hn[nb];
  for (uint32_t g = 0; g < A.ng; g++) {
    const double kf = floor(A.durations[g] * scale + 0.5);
    uint32_t k = kf < 1.0 ? 1u : (uint32_t)kf;
    if (k > nb) k = nb;
    double best = -kInf;
    for (uint32_t j = threadIdx.x; j < nb; j += kBlock) {
      const uint32_t e = j + k;
      double a, w;
      uint32_t cnt;
      if (e <= nb) {
        a = (double)(hv[e] - hv[j]) * S.sv;
        w = (double)(hw[e] - hw[j]) * S.sw;
        cnt = hn[e] - hn[j];
      } else {
        a = (double)(hv[e - nb] + (hv[nb] - hv[j])) * S.sv;
        w = (double)(hw[e - nb] + (hw[nb] - hw[j])) * S.sw;
        cnt = hn[e - nb] + (hn[nb] - hn[j]);
      }
      const double q = w - w * w / S.C;
      if (cnt >= 3 && cnt + 3 <= tn && q > 0.0) {
        const double snr = a / sqrt(q);
        best = snr > best ? snr : best;
      }
    }
    red[threadIdx.x] = best;
    __syncthreads();
    for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) A.out[((size_t)star * A.np + pi) * A.ng + g] = red[0];
    __syncthreads();
  }
}

}  // namespace

size_t feature_lds_bytes(uint32_t nb_max) { return (size_t)(nb_max + 1) * 20 + 16; }

hipError_t feature_launch(const FeatureArgs& a, uint32_t nstars, uint32_t nb_max,
                          hipStream_t s) {
  const size_t lds = feature_lds_bytes(nb_max);
  hipError_t e = hipFuncSetAttribute((const void*)coarse_features,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(coarse_features, dim3(nstars * a.np), dim3(kBlock), lds, s, a);
  return hipGetLastError();
}

}  // namespace fbls

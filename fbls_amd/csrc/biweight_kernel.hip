// Tukey biweight location in a sliding time window, one thread per point
// (gfx950).  Restates the reference's biweight_flatten (isakruas/fbls
// src/prep/curve.rs:204-248, kernel src/search/kernels.rs:1049-1110):
// location = lower median of the window, then `iterations` times
// scale = max(c * lower median |v - location|, 1e-12) and
// location = sum w v / sum w over |u| < 1, u = (v - location) / scale,
// w = (1 - u^2)^2; output v_i / location - 1.
//
// A workgroup owns up to 128 consecutive points and stages the union of
// their windows (<= 4096 values) in LDS; a larger union is read from global
// memory instead.  Lower medians are exact: the k-th smallest value (k =
// (W - 1) / 2) is found by bisection over order-preserving 64-bit keys of
// the doubles (64 counting passes), which selects the same element a sort
// would.  Sums run in window order, as the reference's loops do.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "biweight_kernel.h"

namespace fbls {

namespace {

__device__ __forceinline__ uint64_t key_of(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double value_of(uint64_t k) {
  const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)u);
}

// k-th smallest of v[a..e) (or of |v - center| when absolute)
__device__ double kth(const double* v, uint32_t a, uint32_t e, uint32_t k, double center,
                      bool absolute) {
  uint64_t prefix = 0;
  for (int b = 63; b >= 0; b--) {
    const uint64_t cand = prefix | (1ull << b);
    uint32_t below = 0;
    for (uint32_t j = a; j < e; j++) {
      const double x = absolute ? fabs(v[j] - center) : v[j];
      below += key_of(x) < cand;
    }
    if (below <= k) prefix = cand;
  }
  return value_of(prefix);
}

__global__ __launch_bounds__(kBiweightBlock) void biweight(BiweightArgs A) {
  __shared__ double win[kBiweightWindow];
  const uint32_t lane = threadIdx.x;
  const uint32_t* blk = A.blocks + 4 * blockIdx.x;
  const uint32_t first = blk[0], count = blk[1], ulo = blk[2], uhi = blk[3];
  const bool staged = uhi - ulo <= kBiweightWindow;
  if (staged)
    for (uint32_t q = lane; q < uhi - ulo; q += kBiweightBlock) win[q] = A.flux[ulo + q];
  __syncthreads();
  if (lane >= count) return;
  const double* v = staged ? win : A.flux + ulo;
  const uint32_t p = first + lane;
  const uint32_t a = A.lo[p] - ulo, e = A.hi[p] - ulo;
  const uint32_t k = (e - a - 1) / 2;
  double location = kth(v, a, e, k, 0.0, false);
  for (uint32_t it = 0; it < A.iterations; it++) {
    const double loc = location;
    double scale = A.c * kth(v, a, e, k, loc, true);
    scale = scale >= 1e-12 ? scale : 1e-12;
    double total = 0.0, acc = 0.0;
    for (uint32_t j = a; j < e; j++) {
      const double x = v[j];
      const double u = (x - loc) / scale;
      if (fabs(u) < 1.0) {
        const double w = (1.0 - u * u) * (1.0 - u * u);
        total += w;
        acc += w * x;
      }
    }
    if (total > 0.0) location = acc / total;
  }
  A.out[p] = v[p - ulo] / location - 1.0;
}

}  // namespace

hipError_t biweight_launch(const BiweightArgs& a, uint32_t nblocks, hipStream_t s) {
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(biweight, dim3(nblocks), dim3(kBiweightBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace fbls

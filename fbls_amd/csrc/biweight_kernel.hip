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

// The staged form: the union of the workgroup's windows (<= 4096 values) is
// sorted once in LDS (bitonic sort of order-preserving keys with their time
// indices); each point's lower median is then the (k + 1)-th member of its
// window in sorted order, and the lower median of |v - location| the
// (k + 1)-th step of a merge walking outward from the location through the
// sorted members (both sides ascend in |v - location|).  Every median is the
// same element the counting selection finds; the weighted sums still run in
// window order.
__global__ __launch_bounds__(kBiweightBlock) void biweight_sorted(BiweightArgs A) {
  __shared__ double win[kBiweightWindow];
  __shared__ uint64_t key[kBiweightWindow];
  __shared__ uint16_t idx[kBiweightWindow];
  const uint32_t lane = threadIdx.x;
  const uint32_t* blk = A.blocks + 4 * blockIdx.x;
  const uint32_t first = blk[0], count = blk[1], ulo = blk[2], uhi = blk[3];
  const uint32_t U = uhi - ulo;
  uint32_t P2 = 1;
  while (P2 < U) P2 <<= 1;
  for (uint32_t q = lane; q < P2; q += kBiweightBlock) {
    if (q < U) {
      const double x = A.flux[ulo + q];
      win[q] = x;
      key[q] = key_of(x);
      idx[q] = (uint16_t)q;
    } else {
      key[q] = ~0ull;
      idx[q] = 0xffffu;
    }
  }
  __syncthreads();
  for (uint32_t size = 2; size <= P2; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t i = lane; i < P2 / 2; i += kBiweightBlock) {
        const uint32_t lo = 2 * stride * (i / stride) + (i % stride), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t kl = key[lo], kh = key[hi];
        if ((kl > kh) == up) {
          key[lo] = kh;
          key[hi] = kl;
          const uint16_t t = idx[lo];
          idx[lo] = idx[hi];
          idx[hi] = t;
        }
      }
      __syncthreads();
    }
  }
  if (lane >= count) return;
  const uint32_t p = first + lane;
  const uint32_t a = A.lo[p] - ulo, e = A.hi[p] - ulo;
  const uint32_t k = (e - a - 1) / 2;
  auto member = [&](uint32_t s) { return idx[s] >= a && idx[s] < e; };
  // lower median: the (k + 1)-th member in sorted order
  double location = 0.0;
  for (uint32_t s = 0, c = 0; s < U; s++)
    if (member(s) && c++ == k) {
      location = value_of(key[s]);
      break;
    }
  for (uint32_t it = 0; it < A.iterations; it++) {
    const double loc = location;
    // first sorted position whose value is >= loc
    const uint64_t kl = key_of(loc);
    uint32_t lo = 0, hi = U;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if (key[mid] < kl)
        lo = mid + 1;
      else
        hi = mid;
    }
    // merge outward: the (k + 1)-th smallest |v - loc| among members
    int32_t L = (int32_t)lo - 1;
    uint32_t R = lo, c = 0;
    double mad = 0.0;
    for (;;) {
      while (L >= 0 && !member((uint32_t)L)) L--;
      while (R < U && !member(R)) R++;
      const double dl = L >= 0 ? fabs(value_of(key[L]) - loc) : __builtin_inf();
      const double dr = R < U ? fabs(value_of(key[R]) - loc) : __builtin_inf();
      const double d = dl <= dr ? dl : dr;
      if (c++ == k) {
        mad = d;
        break;
      }
      if (dl <= dr)
        L--;
      else
        R++;
    }
    double scale = A.c * mad;
    scale = scale >= 1e-12 ? scale : 1e-12;
    double total = 0.0, acc = 0.0;
    for (uint32_t j = a; j < e; j++) {
      const double x = win[j];
      const double u = (x - loc) / scale;
      if (fabs(u) < 1.0) {
        const double w = (1.0 - u * u) * (1.0 - u * u);
        total += w;
        acc += w * x;
      }
    }
    if (total > 0.0) location = acc / total;
  }
  A.out[p] = win[p - ulo] / location - 1.0;
}

}  // namespace

hipError_t biweight_launch(const BiweightArgs& a, uint32_t nblocks, hipStream_t s) {
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(biweight, dim3(nblocks), dim3(kBiweightBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t biweight_sorted_launch(const BiweightArgs& a, uint32_t nblocks, hipStream_t s) {
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(biweight_sorted, dim3(nblocks), dim3(kBiweightBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace fbls

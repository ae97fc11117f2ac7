// Rescoring kernels of the greedy peak selection (gfx950).
//
// Reference semantics restated (isakruas/fbls src/search/kernels.rs):
//   rescore_rows  kernels.rs:934-989  box sums (a, hp, norm) per row over the
//                 family's observations in transit and not excluded
//   count_rows    kernels.rs:992-1042 in-transit observations per row over
//                 the merged time vector, and those not excluded
// One workgroup of 256 threads per row, observations strided over the
// threads.  The fields are the family's 32-bit fixed-point values, so the
// sums are exact int64 and independent of the reduction order; hp and norm
// are the same sum of w for a box template.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "select_kernel.h"
#include "fixed_point.h"

namespace fbls {

namespace {

constexpr uint32_t kBlock = 256;

__device__ __forceinline__ bool in_transit(double t, double epoch, double inverse,
                                           double period, double half) {
  const double x = (t - epoch) * inverse + 0.5;
  const double phase = (x - floor(x) - 0.5) * period;
  return fabs(phase) <= half;
}

template <typename T>
__device__ __forceinline__ T block_sum(T x, T* red) {
  for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, 64);
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = x;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0)
    for (uint32_t w = 0; w < kBlock / 64; w++) s += red[w];
  return s;
}

__global__ __launch_bounds__(kBlock) void rescore_rows(RescoreArgs A) {
  __shared__ long long red[kBlock / 64];
  const uint32_t r = blockIdx.x;
  const double period = A.rows[3 * r], epoch = A.rows[3 * r + 1], duration = A.rows[3 * r + 2];
  const double inverse = 1.0 / period, half = duration / 2.0;
  long long a = 0, w = 0;
  for (uint32_t i = threadIdx.x; i < A.n; i += kBlock) {
    if (A.excluded[i]) continue;
    // (time + origin - epoch) as the reference forms it
    if (in_transit(A.t[i] + A.origin, epoch, inverse, period, half)) {
      a += fixed_ivalue(A.vw[2 * i], A.fx & 1u);
      w += fixed_ivalue(A.vw[2 * i + 1], A.fx & 2u);
    }
  }
  const long long sa = block_sum(a, red);
  __syncthreads();
  const long long sw = block_sum(w, red);
  if (threadIdx.x == 0) {
    A.sums[2 * r] = sa;
    A.sums[2 * r + 1] = sw;
  }
}

__global__ __launch_bounds__(kBlock) void count_rows(CountArgs A) {
  __shared__ uint32_t red[kBlock / 64];
  const uint32_t r = blockIdx.x;
  const double period = A.rows[3 * r], epoch = A.rows[3 * r + 1], duration = A.rows[3 * r + 2];
  const double inverse = 1.0 / period, half = duration / 2.0;
  uint32_t total = 0, unshared = 0;
  for (uint32_t i = threadIdx.x; i < A.n; i += kBlock) {
    if (in_transit(A.time_all[i], epoch, inverse, period, half)) {
      total++;
      unshared += A.excluded[i] == 0;
    }
  }
  const uint32_t st = block_sum(total, red);
  __syncthreads();
  const uint32_t su = block_sum(unshared, red);
  if (threadIdx.x == 0) {
    A.counts[2 * r] = st;
    A.counts[2 * r + 1] = su;
  }
}

}  // namespace

hipError_t rescore_launch(const RescoreArgs& a, uint32_t nrows, hipStream_t s) {
  if (nrows == 0) return hipSuccess;
  hipLaunchKernelGGL(rescore_rows, dim3(nrows), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t count_launch(const CountArgs& a, uint32_t nrows, hipStream_t s) {
  if (nrows == 0) return hipSuccess;
  hipLaunchKernelGGL(count_rows, dim3(nrows), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace fbls

// Device-side band pipeline (band_kernel.h): prescan survivors, the full
// scan's segment tables and the reported rows, computed where the scores
// are, so that the full scan of the survivors follows the first scan in the
// same stream without a round trip through the host.
//
// Reference semantics restated here (isakruas/fbls):
//   survivors     src/search/scan.rs:390-395 (finite score >= fraction * threshold,
//                 in period order; scales nb / P as in scan.rs:204)
//   rows          src/search/scan.rs:397-405 (finite score >= threshold)
//   row decode    src/search/scan.rs:339-349 (epoch (phase + width / 2) / scale + origin,
//                 duration width / scale)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "band_kernel.h"

namespace fbls {

namespace {

constexpr uint32_t kBlock = 1024;
constexpr uint32_t kPer = 4;  // consecutive items per thread in a compaction step

// Exclusive scan of x over the workgroup (BLOCK threads, wave64); *total
// receives the sum.  wsum: BLOCK / 64 words of LDS.  Ends at a barrier, so
// wsum may be reused at once.
template <uint32_t BLOCK>
__device__ uint32_t block_excl_scan(uint32_t x, uint32_t* wsum, uint32_t* total) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t incl = x;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(incl, off, 64);
    if (lane >= (uint32_t)off) incl += y;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < BLOCK / 64; w++) {
    const uint32_t s = wsum[w];
    before += w < wave ? s : 0u;
    all += s;
  }
  *total = all;
  __syncthreads();
  return before + incl - x;
}

__device__ __forceinline__ bool kept(double s, double cut) {
  return __builtin_isfinite(s) && s >= cut;
}

__global__ __launch_bounds__(kBlock) void survivors_kernel(BandArgs a) {
  __shared__ uint32_t wsum[kBlock / 64];
  const BandDesc B = a.bands[blockIdx.x];
  const uint32_t tid = threadIdx.x;
  if (!B.prescan) {
    if (tid == 0) a.counts[2 * blockIdx.x] = 0;
    return;
  }
  uint32_t base = 0;
  for (uint32_t c0 = 0; c0 < B.np; c0 += kBlock * kPer) {
    const uint32_t i0 = c0 + tid * kPer;
    bool f[kPer];
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) {
      f[j] = i0 + j < B.np && kept(B.score1[i0 + j], a.cut);
      mine += f[j] ? 1u : 0u;
    }
    uint32_t total = 0;
    uint32_t at = base + block_excl_scan<kBlock>(mine, wsum, &total);
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) {
      if (f[j]) {
        const double P = B.periods[i0 + j];
        B.surv_p[at] = P;
        B.surv_s[at] = B.nb / P;
        at++;
      }
    }
    base += total;
  }
  if (tid == 0) a.counts[2 * blockIdx.x] = base;
}

__global__ __launch_bounds__(kBlock) void plan_kernel(BandArgs a) {
  __shared__ uint32_t wsum[kBlock / 64];
  const uint32_t tid = threadIdx.x;
  for (uint32_t b = tid; b < a.nbands; b += kBlock) {
    const BandDesc& B = a.bands[b];
    if (B.prescan) a.fine[B.fine_table].segs[B.fine_seg].np = a.counts[2 * b];
  }
  __syncthreads();
  for (uint32_t t = 0; t < 2; t++) {
    const FineTable& T = a.fine[t];
    if (T.nseg == 0) continue;
    uint32_t periods = 0;
    for (uint32_t s = tid; s < T.nseg; s += kBlock) periods += T.segs[s].np;
    uint32_t total = 0;
    (void)block_excl_scan<kBlock>(periods, wsum, &total);
    const uint32_t cap = total / 1024 > 1 ? total / 1024 : 1u;
    uint32_t base = 0;
    for (uint32_t s0 = 0; s0 < T.nseg; s0 += kBlock) {
      const uint32_t s = s0 + tid;
      uint32_t groups = 0, G = 1;
      if (s < T.nseg) {
        G = T.segs[s].G;
        if (T.cap_by_total && G > cap) G = cap;
        groups = (T.segs[s].np + G - 1) / G;
      }
      uint32_t sum = 0;
      const uint32_t at = base + block_excl_scan<kBlock>(groups, wsum, &sum);
      if (s < T.nseg) {
        T.segs[s].G = G;
        T.segs[s].group0 = at;
      }
      base += sum;
    }
    if (tid == 0) {
      *T.vgroups = base;
      *T.ticket = 0;
    }
  }
}

__global__ __launch_bounds__(kBlock) void rows_kernel(BandArgs a) {
  __shared__ uint32_t wsum[kBlock / 64];
  const BandDesc B = a.bands[blockIdx.x];
  const uint32_t tid = threadIdx.x;
  const double* rec = B.prescan ? B.rec2 : B.rec1;
  const double* sc = B.prescan ? B.score2 : B.score1;
  const double* dp = B.prescan ? B.depth2 : B.depth1;
  const uint32_t m = B.prescan ? a.counts[2 * blockIdx.x] : B.np;
  uint32_t base = 0;
  for (uint32_t c0 = 0; c0 < m; c0 += kBlock * kPer) {
    const uint32_t i0 = c0 + tid * kPer;
    bool f[kPer];
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) {
      f[j] = i0 + j < m && kept(sc[i0 + j], a.threshold);
      mine += f[j] ? 1u : 0u;
    }
    uint32_t total = 0;
    uint32_t at = base + block_excl_scan<kBlock>(mine, wsum, &total);
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) {
      if (f[j]) {
        const uint32_t k = i0 + j;
        const double* q = rec + 9 * (size_t)k;
        RowOut r;
        r.period = q[0];
        r.epoch = (q[2] + q[3] / 2.0) / q[1] + B.origin;
        r.duration = q[3] / q[1];
        r.score = sc[k];
        r.depth = dp[k];
        B.rows[at++] = r;
      }
    }
    base += total;
  }
  if (tid == 0) a.counts[2 * blockIdx.x + 1] = base;
}

__global__ __launch_bounds__(kBlock) void rows_prefix_kernel(BandArgs a) {
  __shared__ uint32_t wsum[kBlock / 64];
  const uint32_t tid = threadIdx.x;
  uint32_t base = 0;
  for (uint32_t b0 = 0; b0 < a.nbands; b0 += kBlock) {
    const uint32_t b = b0 + tid;
    const uint32_t n = b < a.nbands ? a.counts[2 * b + 1] : 0u;
    uint32_t sum = 0;
    const uint32_t at = base + block_excl_scan<kBlock>(n, wsum, &sum);
    if (b < a.nbands) a.row_off[b] = at;
    base += sum;
  }
  if (tid == 0) a.counts[2 * a.nbands] = base;
}

__global__ __launch_bounds__(256) void rows_gather_kernel(BandArgs a) {
  const uint32_t b = blockIdx.x;
  const uint32_t n = a.counts[2 * b + 1];
  const RowOut* src = a.bands[b].rows;
  RowOut* dst = a.out + a.row_off[b];
  for (uint32_t i = threadIdx.x; i < n; i += 256) dst[i] = src[i];
}

}  // namespace

hipError_t band_survivors(const BandArgs& a, hipStream_t stream) {
  if (a.nbands == 0) return hipSuccess;
  hipLaunchKernelGGL(survivors_kernel, dim3(a.nbands), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

hipError_t band_plan(const BandArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL(plan_kernel, dim3(1), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

hipError_t band_rows(const BandArgs& a, hipStream_t stream) {
  if (a.nbands == 0) return hipSuccess;
  hipLaunchKernelGGL(rows_kernel, dim3(a.nbands), dim3(kBlock), 0, stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(rows_prefix_kernel, dim3(1), dim3(kBlock), 0, stream, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(rows_gather_kernel, dim3(a.nbands), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace fbls

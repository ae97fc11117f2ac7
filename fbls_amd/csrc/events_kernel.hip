// Single-event scan for gfx950 (isakruas/fbls src/search/events.rs:18-153,
// kernels scan_all and read_counter, src/search/kernels.rs:849-930): every
// template above threshold at one trial period longer than the baseline.
//
// The trial period exceeds the span of the times, so the phase bin
// floor(t * scale) never wraps and is non-decreasing along the sorted
// times.  The per-bin histogram and its exclusive prefix therefore need no
// atomics: one workgroup gives each thread a contiguous range of bins, the
// thread walks the observations of its range once (the first one found by
// binary search) writing the range-local prefix, and a workgroup scan of the
// range totals supplies the offsets.  Sums are exact 64-bit integers of the
// fixed-point fields, converted to doubles by a power-of-two scale, so the
// prefix equals the sequential double sums of the dequantised fields.  The
// count prefix is the observation index itself.
//
// The scan kernel gives one thread to each phase bin j; for every duration
// (k = min(max(1, floor(D * scale + 0.5)), nb / 2), usable when raw_k <
// nb / 2 and 3 <= N <= n - 3) and shape it forms the staircase sums over the
// doubled prefix (cum[nb + x] = cum[x] + total), q = norm - hp^2 / C and the
// score (a / rho) / sqrt(q), and appends (j, k, score, a / q) with its key
// when the score reaches the threshold.  The host orders the records by key,
// which is the loop order of a sequential scan.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "events_kernel.h"
#include "fixed_point.h"

namespace fbls {

namespace {

__device__ __forceinline__ uint32_t cell(double t, double scale) {
  return (uint32_t)__builtin_floor(t * scale);
}

__global__ __launch_bounds__(kEventPrefixBlock) void event_prefix(EventPrefixArgs A) {
  __shared__ long long wave_v[kEventPrefixBlock / 64], wave_w[kEventPrefixBlock / 64];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t chunk = (A.nb + kEventPrefixBlock - 1) / kEventPrefixBlock;
  const uint32_t x0 = min(tid * chunk, A.nb), x1 = min(x0 + chunk, A.nb);
  // first observation whose bin is >= x0
  uint32_t lo = 0, hi = A.n;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (cell(A.t[mid], A.scale) < x0)
      lo = mid + 1;
    else
      hi = mid;
  }
  uint32_t i = lo;
  long long* lv = (long long*)A.cv;
  long long* lw = (long long*)A.cw;
  long long av = 0, aw = 0;
  for (uint32_t x = x0; x < x1; x++) {
    lv[x] = av;
    lw[x] = aw;
    A.cn[x] = i;
    while (i < A.n && cell(A.t[i], A.scale) == x) {
      av += fixed_ivalue(A.vw[2 * (size_t)i], A.fx & 1u);
      aw += fixed_ivalue(A.vw[2 * (size_t)i + 1], A.fx & 2u);
      i++;
    }
  }
  // workgroup exclusive scan of the range totals
  long long sv = av, sw = aw;
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const long long yv = __shfl_up(sv, off, 64);
    const long long yw = __shfl_up(sw, off, 64);
    if (lane >= off) {
      sv += yv;
      sw += yw;
    }
  }
  if (lane == 63) {
    wave_v[wave] = sv;
    wave_w[wave] = sw;
  }
  __syncthreads();
  long long bv = 0, bw = 0, tv = 0, tw = 0;
  for (uint32_t q = 0; q < kEventPrefixBlock / 64; q++) {
    if (q < wave) {
      bv += wave_v[q];
      bw += wave_w[q];
    }
    tv += wave_v[q];
    tw += wave_w[q];
  }
  const long long ov = bv + sv - av, ow = bw + sw - aw;
  for (uint32_t x = x0; x < x1; x++) {
    const long long pv = lv[x], pw = lw[x];
    A.cv[x] = (double)(ov + pv) * A.sv;
    A.cw[x] = (double)(ow + pw) * A.sw;
  }
  if (tid == 0) {
    A.cv[A.nb] = (double)tv * A.sv;
    A.cw[A.nb] = (double)tw * A.sw;
    A.cn[A.nb] = A.n;
    *A.counter = 0;
  }
}

__global__ __launch_bounds__(kEventScanBlock) void event_scan(EventScanArgs A) {
  const uint32_t j = blockIdx.x * kEventScanBlock + threadIdx.x;
  if (j >= A.nb) return;
  const uint32_t nb = A.nb;
  const double tv = A.cv[nb], tw = A.cw[nb];
  const uint32_t tn = A.cn[nb];
  auto V = [&](uint32_t x) { return x <= nb ? A.cv[x] : A.cv[x - nb] + tv; };
  auto W = [&](uint32_t x) { return x <= nb ? A.cw[x] : A.cw[x - nb] + tw; };
  auto N = [&](uint32_t x) { return x <= nb ? A.cn[x] : A.cn[x - nb] + tn; };
  const double nf = (double)A.n;
  const double l = (double)A.layers;
  for (uint32_t di = 0; di < A.nd; di++) {
    const double rk = __builtin_floor(A.durations[di] * A.scale + 0.5);
    const uint32_t raw_k = rk < 1.0 ? 1u : (uint32_t)rk;
    const uint32_t k = raw_k < nb / 2 ? raw_k : nb / 2;
    const double count = (double)(N(j + k) - N(j));
    if (!(raw_k < nb / 2 && count >= 3.0 && count <= nf - 3.0)) continue;
    const double a0 = V(j + k) - V(j);
    const double w0 = W(j + k) - W(j);
    const double r = A.rho[di];
    for (uint32_t si = 0; si < A.ns; si++) {
      double a = a0, hp = w0, norm = w0;
      if (A.ingress[si] != 0.0) {
        a = a0 / l;
        hp = w0 / l;
        norm = w0 / (l * l);
        for (uint32_t layer = 1; layer < A.layers; layer++) {
          const double factor = A.factors[si * A.layers + layer];
          const double wf = __builtin_floor((double)k * factor + 0.5);
          const uint32_t width = wf < 1.0 ? 1u : (uint32_t)wf;
          const uint32_t start = j + (k - width) / 2;
          const double dw = W(start + width) - W(start);
          a += (V(start + width) - V(start)) / l;
          hp += dw / l;
          norm += dw * ((2.0 * (double)layer + 1.0) / (l * l));
        }
      }
      const double q = norm - hp * hp / A.C;
      if (!(a > 0.0 && q > 0.0)) continue;
      const double score = (a / r) / __builtin_sqrt(q);
      if (score >= A.threshold) {
        const uint32_t slot = atomicAdd(A.counter, 1u);
        if (slot < A.capacity) {
          double* rec = A.records + 4 * (size_t)slot;
          rec[0] = (double)j;
          rec[1] = (double)k;
          rec[2] = score;
          rec[3] = a / q;
          A.keys[slot] = ((uint64_t)j * A.nd + di) * A.ns + si;
        }
      }
    }
  }
}

}  // namespace

hipError_t event_prefix_launch(const EventPrefixArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(event_prefix, dim3(1), dim3(kEventPrefixBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t event_scan_launch(const EventScanArgs& a, hipStream_t s) {
  const uint32_t blocks = (a.nb + kEventScanBlock - 1) / kEventScanBlock;
  hipLaunchKernelGGL(event_scan, dim3(blocks), dim3(kEventScanBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace fbls

// fbls_scan_bank_noalloc: the scan of one bank (the reference's scan_bank,
// src/search/scan.rs:144-357) with no device or pinned-host allocation.
// Every device buffer the scan needs beyond the caller's family -- the
// launch table, the periods and their scales, the counters and the staged
// outputs -- lives in `scratch`, a fixed array of this library's code object
// on each device.  The kernels are the ones fbls_scan_bank launches, with the
// same geometry, so the results are bit-identical.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>

#include "../../include/fbls.h"
#include "scan_host.h"
#include "scan_kernel.h"

using namespace fbls::host;

namespace {

constexpr size_t kScratchBytes = size_t(8) << 20;  // per device, held until exit
constexpr size_t kMaxPeriods = 65536;                // periods per call
constexpr int kMaxDevices = 64;

// In host code this name is only a stand-in for the device array: the
// array's device address comes from hipGetSymbolAddress (noalloc_setup).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The per-call launch table, staged in one host-to-device copy.  The
// counters come first so that the copy clears them.
struct Table {
  uint32_t counters[4];  // nonpositive-precision flag, exact tile rescans
  double ingress[8];
  double factors[8 * 16];
  double dur[32];
  double rho[32];
  fbls::Segment seg;
};

constexpr size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }
// periods | scales | records (9) | score | depth: 13 doubles, winner 4 u32
constexpr size_t kPerPeriod = 13 * 8 + 16;
static_assert(round256(sizeof(Table)) + 6 * 256 + kMaxPeriods * kPerPeriod <= kScratchBytes,
              "scratch holds kMaxPeriods periods");

struct DeviceScratch {
  std::once_flag once;
  void* ptr = nullptr;  // the array's address on this device
  hipError_t err = hipErrorNotInitialized;
  std::mutex busy;      // one call at a time per device: they share the array
};
DeviceScratch g_dev[kMaxDevices];

// scales[k] = nb / periods[k] (scan.rs:204); IEEE division, as the host's.
__global__ void stage_scales(const double* __restrict__ periods, double* __restrict__ scales,
                             uint32_t np, double nb) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) scales[i] = nb / periods[i];
}

}  // namespace

void fbls::host::noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  DeviceScratch& d = g_dev[device];
  std::call_once(d.once, [&d]() {
    void* p = nullptr;
    d.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    d.ptr = d.err == hipSuccess ? p : nullptr;
  });
}

extern "C" int fbls_scan_bank_noalloc(fbls_ctx* c, const fbls_family* f,
                                      const double* periods, size_t np,
                                      const fbls_bank* b, uint32_t nb, double* records,
                                      double* score, double* depth, uint32_t* winner,
                                      fbls_scan_stats* stats) {
  return guard([&]() -> int {
    if (!c || !f || (!periods && np) || !records || !score || !depth)
      return fail(FBLS_E_ARG, "null argument");
    int st = check_bank(b);
    if (st) return st;
    const uint64_t budget = bank_batch(nb, b->period_batch, b->max_fold_mib);
    if (budget == 0) return fail(FBLS_E_FOLD_BUDGET, "one folded period exceeds max_fold_mib");
    if (np > kMaxPeriods)
      return fail(FBLS_E_ARG, "at most %zu periods per call without allocation, got %zu",
                  kMaxPeriods, np);
    if (c->device < 0 || c->device >= kMaxDevices)
      return fail(FBLS_E_ARG, "device %d has no scratch slot", c->device);
    DeviceScratch& ds = g_dev[c->device];
    if (!ds.ptr)
      return fail(FBLS_E_HIP, "scratch address unavailable: %s", hipGetErrorString(ds.err));
    Work x;
    x.fam = f;
    x.periods = periods;
    x.np = np;
    x.nb = nb;
    x.nd = (uint32_t)b->nd;
    x.dur = b->durations;
    x.rho = b->rho;
    x.records = records;
    x.score = score;
    x.depth = depth;
    x.winner = winner;
    const Shared sh = shared_of(b);
    double secs = 0.0;
    if (np) {
      const bool split = needs_split(c, x, sh);
      SegShape shape;
      size_t lds = 0;
      uint32_t groups = 0;
      if ((st = segment_shapes(c, &x, 1, sh, split, &shape, &lds, &groups))) return st;
      double smax = 0.0;
      if ((st = check_periods(x, &smax))) return st;

      std::lock_guard<std::mutex> hold(ds.busy);
      HIP_TRY(hipSetDevice(c->device));
      unsigned char* base = (unsigned char*)ds.ptr;
      size_t at = round256(sizeof(Table));
      auto take = [&](size_t bytes) {
        unsigned char* p = base + at;
        at += round256(bytes);
        return p;
      };
      double* d_periods = (double*)take(np * 8);
      double* d_scales = (double*)take(np * 8);
      double* d_records = (double*)take(np * 72);
      double* d_score = (double*)take(np * 8);
      double* d_depth = (double*)take(np * 8);
      uint32_t* d_winner = (uint32_t*)take(np * 16);
      // device addresses of the table's fields (never dereferenced here)
      unsigned int* d_counters = (unsigned int*)(base + offsetof(Table, counters));
      const double* d_ingress = (const double*)(base + offsetof(Table, ingress));
      const double* d_factors = (const double*)(base + offsetof(Table, factors));
      const double* d_dur = (const double*)(base + offsetof(Table, dur));
      const double* d_rho = (const double*)(base + offsetof(Table, rho));
      const fbls::Segment* d_seg = (const fbls::Segment*)(base + offsetof(Table, seg));

      Table tab;
      std::memset(&tab, 0, sizeof(tab));
      for (uint32_t s = 0; s < sh.ns; s++) tab.ingress[s] = sh.ingress[s];
      bank_factors(sh, tab.factors);
      for (uint32_t k = 0; k < x.nd; k++) {
        tab.dur[k] = x.dur[k];
        tab.rho[k] = x.rho[k];
      }
      segment_fields(c, x, shape, split, smax, &tab.seg);
      tab.seg.periods = d_periods;
      tab.seg.scales = d_scales;
      tab.seg.dur = d_dur;
      tab.seg.rho = d_rho;
      tab.seg.records = d_records;
      tab.seg.score = d_score;
      tab.seg.depth = d_depth;
      tab.seg.winner = winner ? d_winner : nullptr;
      tab.seg.group0 = 0;
      tab.seg.index = 0;

      fbls::LaunchArgs A;
      std::memset(&A, 0, sizeof(A));
      A.segs = d_seg;
      A.nseg = 1;
      A.ns = sh.ns;
      A.layers = sh.layers;
      A.run_fold = c->run_fold;
      A.ingress = d_ingress;
      A.factors = d_factors;
      A.gate = sh.gate;
      A.counters = d_counters;

      // host arrays are read and written in place (pageable copies); the
      // stream is drained before the function returns, also on an error
      uint32_t counters[4] = {0, 0, 0, 0};
      auto enqueue = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(base, &tab, sizeof(tab), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(d_periods, periods, np * 8, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(stage_scales, dim3((uint32_t)((np + 255) / 256)), dim3(256), 0,
                           c->stream, d_periods, d_scales, (uint32_t)np, (double)nb);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(c->ev0, c->stream));
        HIP_TRY(fbls::scan_launch(A, groups, launch_block(c, split, lds), lds, split,
                                  c->stream));
        HIP_TRY(hipEventRecord(c->ev1, c->stream));
        HIP_TRY(hipMemcpyAsync(records, d_records, np * 72, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(score, d_score, np * 8, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(depth, d_depth, np * 8, hipMemcpyDeviceToHost, c->stream));
        if (winner)
          HIP_TRY(hipMemcpyAsync(winner, d_winner, np * 16, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(counters, d_counters, sizeof(counters), hipMemcpyDeviceToHost,
                               c->stream));
        return FBLS_OK;
      };
      st = enqueue();
      const hipError_t drained = hipStreamSynchronize(c->stream);
      if (st) return st;
      if (drained != hipSuccess)
        return fail(FBLS_E_HIP, "hipStreamSynchronize failed: %s", hipGetErrorString(drained));
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      secs = ms * 1e-3;
      c->totals.launches += 1;
      c->totals.seconds += secs;
      c->totals.periods += np;
      c->totals.fold_pairs += (uint64_t)np * f->n;
      if (counters[0])
        return fail(FBLS_E_NONPOSITIVE_PRECISION,
                    "nonpositive template precision; inspect noise model");
      x.retries = counters[1];
    }
    if (stats) account(stats, x, b, budget, secs);
    return FBLS_OK;
  });
}

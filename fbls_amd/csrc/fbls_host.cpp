// Host side of the MI355X FBLS hot path: C ABI (include/fbls.h), device
// context, family upload, and the scan drivers that restate the reference's
// src/search/scan.rs:144-473 around the fused HIP kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fbls.h"
#include "prep.h"
#include "scan_kernel.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                     \
  do {                                                                    \
    hipError_t _e = (expr);                                               \
    if (_e != hipSuccess)                                                 \
      return fail(FBLS_E_HIP, "%s failed: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

double now_s() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Growable device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return FBLS_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes, 4096);
    if (hipMalloc(&ptr, want) != hipSuccess)
      return fail(FBLS_E_OOM, "device allocation of %zu bytes failed", want);
    cap = want;
    return FBLS_OK;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};

constexpr size_t kMaxLds = 160 * 1024;  // gfx950 LDS per CU
constexpr uint32_t kObsAlign = 512;     // observations per wave block

}  // namespace

struct fbls_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t lds_override = 0, block_override = 0;
  size_t max_lds = kMaxLds;
  DevBuf periods, scales, records, score, depth, winner, counters, bank;
};

struct fbls_family {
  fbls_ctx* ctx = nullptr;
  DevBuf t, vw;
  uint32_t n = 0, n_pad = 0;
  double C = 0.0, origin = 0.0, t_max = 0.0, t_last = 0.0;
  int ev = 0, ew = 0;
  std::vector<int32_t> vq, wq;
};

extern "C" const char* fbls_last_error(void) { return g_last_error.c_str(); }

// ---------------------------------------------------------------- context

extern "C" int fbls_ctx_create(int device, fbls_ctx** out) {
  if (!out) return fail(FBLS_E_ARG, "null output pointer");
  *out = nullptr;
  int count = 0;
  HIP_TRY(hipGetDeviceCount(&count));
  if (device < 0 || device >= count)
    return fail(FBLS_E_ARG, "device %d out of range (%d devices)", device, count);
  HIP_TRY(hipSetDevice(device));
  fbls_ctx* c = new fbls_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess &&
      prop.maxSharedMemoryPerMultiProcessor > 0)
    c->max_lds = std::min<size_t>(kMaxLds, prop.maxSharedMemoryPerMultiProcessor);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess ||
      hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return fail(FBLS_E_HIP, "stream/event creation failed");
  }
  hipError_t e = fbls::scan_configure(c->max_lds);
  if (e != hipSuccess) {
    fbls_ctx_destroy(c);
    return fail(FBLS_E_HIP, "kernel LDS configuration failed: %s",
                hipGetErrorString(e));
  }
  *out = c;
  return FBLS_OK;
}

extern "C" void fbls_ctx_destroy(fbls_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (DevBuf* b : {&c->periods, &c->scales, &c->records, &c->score,
                    &c->depth, &c->winner, &c->counters, &c->bank})
    b->release();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" int fbls_mem_info(fbls_ctx* c, size_t* free_bytes, size_t* total) {
  if (!c || !free_bytes || !total) return fail(FBLS_E_ARG, "null argument");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemGetInfo(free_bytes, total));
  return FBLS_OK;
}

extern "C" int fbls_ctx_set_launch(fbls_ctx* c, uint32_t lds_bytes,
                                   uint32_t block) {
  if (!c) return fail(FBLS_E_ARG, "null context");
  if (block != 0 && block != 256 && block != 512)
    return fail(FBLS_E_ARG, "block must be 0, 256 or 512");
  if (lds_bytes > c->max_lds)
    return fail(FBLS_E_ARG, "lds_bytes %u exceeds %zu", lds_bytes, c->max_lds);
  c->lds_override = lds_bytes;
  c->block_override = block;
  return FBLS_OK;
}

// ----------------------------------------------------------------- family

extern "C" int fbls_quantize(const double* x, size_t n, int32_t* q) {
  return fbls::quantize(x, n, q);
}

extern "C" int fbls_family_upload(fbls_ctx* c, const double* t,
                                  const double* v, const double* w, size_t n,
                                  double baseline_norm, double origin,
                                  fbls_family** out) {
  if (!c || !t || !v || !w || !out) return fail(FBLS_E_ARG, "null argument");
  *out = nullptr;
  if (n < 6) return fail(FBLS_E_DATA, "at least six observations are required");
  if (n > (1u << 22))
    return fail(FBLS_E_ARG, "at most 2^22 observations per family (exact sums)");
  if (!(std::isfinite(baseline_norm) && baseline_norm > 0.0))
    return fail(FBLS_E_DATA, "invalid baseline precision");
  double tmax = 0.0;
  for (size_t i = 0; i < n; i++) {
    if (!(t[i] >= 0.0) || !std::isfinite(t[i]))
      return fail(FBLS_E_DATA, "times must be finite and origin-relative (>= 0)");
    if (!std::isfinite(v[i]) || !std::isfinite(w[i]))
      return fail(FBLS_E_DATA, "fields must be finite");
    tmax = std::max(tmax, t[i]);
  }
  HIP_TRY(hipSetDevice(c->device));
  fbls_family* f = new fbls_family();
  f->ctx = c;
  f->n = (uint32_t)n;
  f->n_pad = (uint32_t)((n + kObsAlign - 1) / kObsAlign * kObsAlign);
  f->C = baseline_norm;
  f->origin = origin;
  f->t_max = tmax;
  f->t_last = t[n - 1];
  f->vq.resize(n);
  f->wq.resize(n);
  f->ev = fbls::quantize(v, n, f->vq.data());
  f->ew = fbls::quantize(w, n, f->wq.data());
  std::vector<double> tp(f->n_pad, 0.0);
  std::vector<int32_t> vw(2 * (size_t)f->n_pad, 0);
  std::memcpy(tp.data(), t, n * sizeof(double));
  for (size_t i = 0; i < n; i++) {
    vw[2 * i] = f->vq[i];
    vw[2 * i + 1] = f->wq[i];
  }
  int st = f->t.ensure(tp.size() * sizeof(double));
  if (!st) st = f->vw.ensure(vw.size() * sizeof(int32_t));
  if (st) {
    fbls_family_free(f);
    return st;
  }
  if (hipMemcpy(f->t.ptr, tp.data(), tp.size() * sizeof(double),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(f->vw.ptr, vw.data(), vw.size() * sizeof(int32_t),
                hipMemcpyHostToDevice) != hipSuccess) {
    fbls_family_free(f);
    return fail(FBLS_E_HIP, "family upload failed");
  }
  *out = f;
  return FBLS_OK;
}

extern "C" void fbls_family_free(fbls_family* f) {
  if (!f) return;
  if (f->ctx) (void)hipSetDevice(f->ctx->device);
  f->t.release();
  f->vw.release();
  delete f;
}

extern "C" int fbls_family_fields(const fbls_family* f, double* v_out,
                                  double* w_out, int* ev, int* ew) {
  if (!f) return fail(FBLS_E_ARG, "null family");
  for (uint32_t i = 0; i < f->n; i++) {
    if (v_out) v_out[i] = std::ldexp((double)f->vq[i], -f->ev);
    if (w_out) w_out[i] = std::ldexp((double)f->wq[i], -f->ew);
  }
  if (ev) *ev = f->ev;
  if (ew) *ew = f->ew;
  return FBLS_OK;
}

// ------------------------------------------------------------- scan_bank

namespace {

int check_bank(const fbls_bank* b) {
  if (!b) return fail(FBLS_E_ARG, "null bank");
  if (b->nd == 0 || b->nd > 32)
    return fail(FBLS_E_ARG, "1..32 durations per band are supported, got %zu", b->nd);
  if (b->ns == 0 || b->ns > 8) return fail(FBLS_E_ARG, "1..8 shapes supported");
  if (b->layers < 1 || b->layers > 16) return fail(FBLS_E_ARG, "1..16 layers supported");
  for (size_t i = 0; i < b->ns; i++)
    if (!(b->ingress[i] >= 0.0 && b->ingress[i] < 0.5))
      return fail(FBLS_E_ARG, "ingress fractions must be in [0, 0.5)");
  for (size_t i = 0; i < b->nd; i++)
    if (!(b->durations[i] > 0.0 && std::isfinite(b->durations[i])) ||
        !(b->rho[i] > 0.0 && std::isfinite(b->rho[i])))
      return fail(FBLS_E_ARG, "durations and rho must be positive and finite");
  return FBLS_OK;
}

// scan.rs:177-183: periods per batch under the fold budget (0 = too small).
uint64_t bank_batch(uint32_t nb, uint32_t period_batch, double max_fold_mib) {
  double per_period =
      (double)(3 * (4 * (uint64_t)nb + 2) * 8 + 3 * (2 * (uint64_t)nb + 1) * 8);
  return (uint64_t)std::floor((max_fold_mib * 1048576.0) / per_period);
}

struct Launch {
  uint32_t G, block;
  size_t lds;
  uint32_t stride;
};

int plan_launch(fbls_ctx* c, uint32_t nb, size_t np, const fbls_bank* b,
                Launch* L) {
  uint32_t nd = (uint32_t)b->nd, ns = (uint32_t)b->ns, layers = b->layers;
  size_t one = fbls::scan_lds_bytes(nb, 1, nd, ns, layers, &L->stride);
  if (one > c->max_lds)
    return fail(FBLS_E_ARG,
                "%u phase bins need %zu bytes of LDS per period (limit %zu)",
                nb, one, c->max_lds);
  size_t budget = c->lds_override ? c->lds_override : 64 * 1024;
  if (budget < one) budget = one;
  uint32_t G = (uint32_t)std::max<size_t>(1, budget / one);
  while (G > 1 && fbls::scan_lds_bytes(nb, G, nd, ns, layers, nullptr) > budget) G--;
  G = std::min<uint32_t>(G, 16);
  // keep at least ~2 workgroups per CU in flight when there is enough work
  uint32_t want_groups = 512;
  uint32_t cap = (uint32_t)std::max<size_t>(1, np / want_groups);
  G = std::max<uint32_t>(1, std::min(G, cap));
  L->G = G;
  L->lds = fbls::scan_lds_bytes(nb, G, nd, ns, layers, nullptr);
  L->block = c->block_override ? c->block_override : 512;
  return FBLS_OK;
}

// One scan of `np` periods at `nb` bins: the work of one scan_bank call.
int run_scan(fbls_ctx* c, const fbls_family* f, const double* periods,
             size_t np, const fbls_bank* b, uint32_t nb, double* records,
             double* score, double* depth, uint32_t* winner,
             fbls_scan_stats* stats) {
  if (np == 0) return FBLS_OK;
  if (nb < 2 || nb >= 65536)
    return fail(FBLS_E_ARG, "phase bins must be in [2, 65535], got %u", nb);
  if (np > 0xffffffffu) return fail(FBLS_E_ARG, "too many periods");
  HIP_TRY(hipSetDevice(c->device));
  Launch L;
  int st = plan_launch(c, nb, np, b, &L);
  if (st) return st;
  const uint32_t nd = (uint32_t)b->nd, ns = (uint32_t)b->ns;
  std::vector<double> scales(np);
  double smax = 0.0;
  for (size_t i = 0; i < np; i++) {
    if (!(periods[i] > 0.0) || !std::isfinite(periods[i]))
      return fail(FBLS_E_ARG, "periods must be positive and finite");
    scales[i] = (double)nb / periods[i];  // scan.rs:204
    smax = std::max(smax, scales[i]);
  }
  // bank arrays: dur | rho | ingress | factors (scan.rs:157-171)
  std::vector<double> bank(2 * nd + ns + ns * b->layers);
  for (uint32_t i = 0; i < nd; i++) {
    bank[i] = b->durations[i];
    bank[nd + i] = b->rho[i];
  }
  for (uint32_t s = 0; s < ns; s++) {
    bank[2 * nd + s] = b->ingress[s];
    for (uint32_t l = 0; l < b->layers; l++)
      bank[2 * nd + ns + s * b->layers + l] =
          1.0 - 2.0 * b->ingress[s] *
                    (b->layers > 1 ? (double)l / (double)(b->layers - 1) : 0.0);
  }
  if ((st = c->periods.ensure(np * 8)) || (st = c->scales.ensure(np * 8)) ||
      (st = c->records.ensure(np * 9 * 8)) || (st = c->score.ensure(np * 8)) ||
      (st = c->depth.ensure(np * 8)) ||
      (st = c->winner.ensure(winner ? np * 16 : 16)) ||
      (st = c->counters.ensure(64)) || (st = c->bank.ensure(bank.size() * 8)))
    return st;
  HIP_TRY(hipMemcpyAsync(c->periods.ptr, periods, np * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->scales.ptr, scales.data(), np * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->bank.ptr, bank.data(), bank.size() * 8, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(c->counters.ptr, 0, 64, c->stream));
  fbls::ScanArgs A;
  std::memset(&A, 0, sizeof(A));
  A.t = (const double*)f->t.ptr;
  A.vw = (const int32_t*)f->vw.ptr;
  A.n = f->n;
  A.n_pad = f->n_pad;
  A.sv = std::ldexp(1.0, -f->ev);
  A.sw = std::ldexp(1.0, -f->ew);
  A.C = f->C;
  A.invC = 1.0 / f->C;
  A.periods = (const double*)c->periods.ptr;
  A.scales = (const double*)c->scales.ptr;
  A.np = (uint32_t)np;
  A.nb = nb;
  A.G = L.G;
  A.nd = nd;
  A.ns = ns;
  A.layers = b->layers;
  const double* db = (const double*)c->bank.ptr;
  A.dur = db;
  A.rho = db + nd;
  A.ingress = db + 2 * nd;
  A.factors = db + 2 * nd + ns;
  A.gate = b->shape_gate;
  A.wide_cells = (f->t_max * smax >= 16777215.0) ? 1u : 0u;
  if (f->t_max * smax >= 4294967295.0)
    return fail(FBLS_E_ARG, "t * nb / P exceeds 2^32");
  A.period_stride = L.stride;
  A.records = (double*)c->records.ptr;
  A.score = (double*)c->score.ptr;
  A.depth = (double*)c->depth.ptr;
  A.winner = winner ? (uint32_t*)c->winner.ptr : nullptr;
  A.counters = (unsigned int*)c->counters.ptr;
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  HIP_TRY(fbls::scan_launch(A, L.block, L.lds, c->stream));
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  unsigned int counters[2];
  HIP_TRY(hipMemcpyAsync(records, c->records.ptr, np * 9 * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(score, c->score.ptr, np * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(depth, c->depth.ptr, np * 8, hipMemcpyDeviceToHost, c->stream));
  if (winner)
    HIP_TRY(hipMemcpyAsync(winner, c->winner.ptr, np * 16, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(counters, c->counters.ptr, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  if (counters[0])
    return fail(FBLS_E_NONPOSITIVE_PRECISION,
                "nonpositive template precision; inspect noise model");
  if (stats) {
    stats->exact_tile_retries += counters[1];
    stats->seconds_device += ms * 1e-3;
    stats->seconds_scan += ms * 1e-3;
  }
  return FBLS_OK;
}

int scan_bank_impl(fbls_ctx* c, const fbls_family* f, const double* periods,
                   size_t np, const fbls_bank* b, uint32_t nb, double* records,
                   double* score, double* depth, uint32_t* winner,
                   fbls_scan_stats* stats) {
  uint64_t budget = bank_batch(nb, b->period_batch, b->max_fold_mib);
  if (budget == 0) return fail(FBLS_E_FOLD_BUDGET, "one folded period exceeds max_fold_mib");
  int st = run_scan(c, f, periods, np, b, nb, records, score, depth, winner, stats);
  if (st) return st;
  if (stats) {
    stats->phase_bins = nb;
    uint64_t batch = std::min<uint64_t>(std::min<uint64_t>(b->period_batch, budget), np);
    if (batch) stats->batches += (np + batch - 1) / batch;
    stats->templates_scanned += (uint64_t)np * b->nd * b->ns * nb;  // scan.rs:352
  }
  return FBLS_OK;
}

}  // namespace

extern "C" int fbls_scan_bank(fbls_ctx* c, const fbls_family* f,
                              const double* periods, size_t np,
                              const fbls_bank* b, uint32_t nb, double* records,
                              double* score, double* depth, uint32_t* winner,
                              fbls_scan_stats* stats) {
  if (!c || !f || (!periods && np) || !records || !score || !depth)
    return fail(FBLS_E_ARG, "null argument");
  int st = check_bank(b);
  if (st) return st;
  return scan_bank_impl(c, f, periods, np, b, nb, records, score, depth, winner, stats);
}

// scan.rs:363-408
extern "C" int fbls_scan_periods(fbls_ctx* c, const fbls_family* f,
                                 const double* periods, size_t np,
                                 const fbls_bank* b, fbls_row* rows,
                                 size_t* nrows, fbls_scan_stats* stats) {
  if (!c || !f || !nrows || (np && (!periods || !rows)))
    return fail(FBLS_E_ARG, "null argument");
  double t0 = now_s();
  fbls_scan_stats local;
  std::memset(&local, 0, sizeof(local));
  local.periods = np;
  *nrows = 0;
  if (np == 0 || b == nullptr || b->nd == 0) {
    if (stats) *stats = local;
    return FBLS_OK;
  }
  int st = check_bank(b);
  if (st) return st;
  double dmin = INFINITY;
  for (size_t i = 0; i < b->nd; i++) dmin = std::min(dmin, b->durations[i]);
  double nbf = std::ceil((double)b->phase_oversample * periods[np - 1] / dmin);
  uint32_t nb = nbf < 8.0 ? 8u : (nbf > 4e9 ? 0xffffffffu : (uint32_t)nbf);
  std::vector<double> survivors(periods, periods + np);
  std::vector<double> rec(np * 9), score(np), depth(np);
  if (b->prescan_factor > 1 && nb / b->prescan_factor >= 8) {
    st = scan_bank_impl(c, f, periods, np, b, nb / b->prescan_factor, rec.data(),
                        score.data(), depth.data(), nullptr, &local);
    if (st) return st;
    survivors.clear();
    const double cut = b->prescan_fraction * b->threshold;
    for (size_t i = 0; i < np; i++)
      if (std::isfinite(score[i]) && score[i] >= cut) survivors.push_back(rec[9 * i]);
    local.prescan_survivors = survivors.size();
  }
  if (!survivors.empty()) {
    size_t ns = survivors.size();
    st = scan_bank_impl(c, f, survivors.data(), ns, b, nb, rec.data(),
                        score.data(), depth.data(), nullptr, &local);
    if (st) return st;
    for (size_t i = 0; i < ns; i++) {
      if (!(std::isfinite(score[i]) && score[i] >= b->threshold)) continue;
      const double* r = &rec[9 * i];
      fbls_row& row = rows[(*nrows)++];
      row.period = r[0];
      row.epoch = (r[2] + r[3] / 2.0) / r[1] + f->origin;  // scan.rs:345
      row.duration = r[3] / r[1];
      row.score = score[i];
      row.depth = depth[i];
    }
  }
  local.seconds = now_s() - t0;
  if (stats) *stats = local;
  return FBLS_OK;
}

// ---------------------------------------------------------- signal_search

extern "C" size_t fbls_period_grid_count(double minimum, double maximum,
                                         double baseline, double dmin,
                                         double oversample) {
  return fbls::period_grid_count(minimum, maximum, baseline, dmin, oversample);
}

extern "C" int fbls_period_grid(double minimum, double maximum,
                                double baseline, double dmin,
                                double oversample, double* out, size_t count) {
  if (!out || count == 0) return fail(FBLS_E_ARG, "empty grid");
  fbls::period_grid(minimum, maximum, out, count);
  return FBLS_OK;
}

extern "C" size_t fbls_durations_days(const double* hours, size_t n,
                                      double* out) {
  return fbls::durations_days(hours, n, out);
}

// scan.rs:413-473
extern "C" int fbls_signal_search(fbls_ctx* c, const fbls_family* f,
                                  const double* durations, const double* rho,
                                  size_t nd, const fbls_bank* bank,
                                  const fbls_search_cfg* cfg, fbls_row* rows,
                                  size_t rows_cap, size_t* nrows,
                                  fbls_scan_stats* stats, size_t stats_cap,
                                  size_t* nbands, double* grid_out,
                                  size_t grid_cap, size_t* grid_count) {
  if (!c || !f || !bank || !cfg || !nrows || !nbands || !grid_count)
    return fail(FBLS_E_ARG, "null argument");
  *nrows = 0;
  *nbands = 0;
  *grid_count = 0;
  if (nd == 0) return FBLS_OK;
  std::vector<std::pair<double, double>> ds(nd);
  for (size_t i = 0; i < nd; i++) ds[i] = {durations[i], rho[i]};
  std::stable_sort(ds.begin(), ds.end(),
                   [](const std::pair<double, double>& a,
                      const std::pair<double, double>& b) { return a.first < b.first; });
  const double baseline = f->t_last;  // last relative time (scan.rs:425)
  const double maximum = std::min(cfg->period_max, baseline);
  const double minimum = std::max(cfg->period_min, 4.0 * ds[0].first * (1.0 + 1e-9));
  if (maximum <= minimum) return FBLS_OK;
  size_t count = fbls::period_grid_count(minimum, maximum, baseline, ds[0].first,
                                         cfg->period_oversample);
  std::vector<double> grid(count);
  fbls::period_grid(minimum, maximum, grid.data(), count);
  *grid_count = count;
  if (grid_out) {
    if (grid_cap < count) return fail(FBLS_E_ARG, "grid capacity %zu < %zu", grid_cap, count);
    std::memcpy(grid_out, grid.data(), count * sizeof(double));
  }
  if (rows_cap < count) return fail(FBLS_E_ARG, "rows capacity %zu < grid %zu", rows_cap, count);
  // Gpu::fold_budget_mib (scan.rs:93-98)
  fbls_bank b = *bank;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess)
    b.max_fold_mib = std::min(b.max_fold_mib, (double)fr / 1048576.0 / 2.0);
  std::vector<double> sd(nd), sr(nd);
  for (size_t i = 0; i < nd; i++) {
    sd[i] = ds[i].first;
    sr[i] = ds[i].second;
  }
  double lo = grid[0];
  const double last = grid[count - 1];
  std::vector<double> sel;
  while (lo <= last) {
    const double hi = lo * 2.0;
    sel.clear();
    for (double p : grid)
      if (p >= lo && p < hi) sel.push_back(p);
    size_t chosen = 0;
    while (chosen < nd && sd[chosen] < lo / 4.0) chosen++;
    if (!sel.empty() && chosen) {
      if (stats && *nbands >= stats_cap)
        return fail(FBLS_E_ARG, "band statistics capacity %zu exceeded", stats_cap);
      b.durations = sd.data();
      b.rho = sr.data();
      b.nd = chosen;
      size_t got = 0;
      fbls_scan_stats bs;
      int st = fbls_scan_periods(c, f, sel.data(), sel.size(), &b, rows + *nrows,
                                 &got, &bs);
      if (st) return st;
      *nrows += got;
      if (stats) stats[*nbands] = bs;
      (*nbands)++;
    }
    lo = hi;
  }
  return FBLS_OK;
}

// ------------------------------------------------------------ host helpers

extern "C" int fbls_family_state(const double* time, const double* flux,
                                 const double* sigma, size_t n,
                                 const double* durations, size_t nd,
                                 int correlated, double* t, double* v,
                                 double* w, size_t* out_n,
                                 double* baseline_norm, double* origin,
                                 double* rho, double* inflation,
                                 size_t* segments) {
  std::string msg;
  int st = fbls::family_state(time, flux, sigma, n, durations, nd, correlated,
                              t, v, w, out_n, baseline_norm, origin, rho,
                              inflation, segments, &msg);
  if (st) return fail(st, "%s", msg.c_str());
  return FBLS_OK;
}

extern "C" int fbls_synth_star(const fbls_synth_cfg* cfg, double* time,
                               double* flux, double* flux_err, size_t cap,
                               size_t* n) {
  std::string msg;
  int st = fbls::synth_star(cfg, time, flux, flux_err, cap, n, &msg);
  if (st) return fail(st, "%s", msg.c_str());
  return FBLS_OK;
}

// Host side of the single-event scan (isakruas/fbls src/search/events.rs:18-153):
// trial period, bin count, staircase factors, the two kernel launches of
// events_kernel.hip and the decoding of the records into rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/fbls.h"
#include "events_kernel.h"
#include "host_internal.h"

using namespace fbls::host;

extern "C" int fbls_single_event_search(fbls_ctx* c, const fbls_family* fam,
                                        const double* durations, const double* rho, size_t nd,
                                        const double* ingress, size_t ns, uint32_t layers,
                                        double threshold, uint32_t phase_oversample,
                                        size_t capacity, fbls_row* rows, size_t* nrows,
                                        fbls_event_stats* stats) {
  return fbls::host::guard([&]() -> int {
    if (!c || !fam || !durations || !rho || !ingress || !nrows || (capacity && !rows))
      return fail(FBLS_E_ARG, "null argument");
    *nrows = 0;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (nd == 0 || ns == 0 || layers == 0) return fail(FBLS_E_ARG, "empty template bank");
    if (phase_oversample == 0) return fail(FBLS_E_ARG, "phase_oversample must be >= 1");
    if (capacity > 0xffffffffu) return fail(FBLS_E_ARG, "capacity above 2^32 - 1");
    if (!fam->sorted) return fail(FBLS_E_ARG, "the single-event scan needs ascending times");
    if (fam->n == 0) return fail(FBLS_E_DATA, "empty family");
    for (size_t i = 0; i < nd; i++)
      if (!(durations[i] > 0.0) || !std::isfinite(durations[i]))
        return fail(FBLS_E_ARG, "durations must be positive and finite");
    // events.rs:24-35: durations ascending with their rho
    std::vector<size_t> ord(nd);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](size_t a, size_t b) { return durations[a] < durations[b]; });
    std::vector<double> params;
    for (size_t i : ord) params.push_back(durations[i]);
    for (size_t i : ord) params.push_back(rho[i]);
    const double d0 = params[0], dmax = params[nd - 1];
    const double baseline = fam->t_last;
    const double trial = baseline + 2.0 * dmax + 1e-6;
    const double nbf = std::ceil((double)phase_oversample * trial / d0);
    if (!(nbf < 2147483648.0)) return fail(FBLS_E_ARG, "single-event bin count above 2^31");
    const uint32_t nb = std::max<uint32_t>(8u, (uint32_t)nbf);
    const double scale = (double)nb / trial;
    if (!(std::floor(fam->t_max * scale) < (double)nb))
      return fail(FBLS_E_ARG, "single-event trial period shorter than the baseline");
    // scan.rs:157-171 staircase factors
    for (size_t s = 0; s < ns; s++) params.push_back(ingress[s]);
    for (size_t s = 0; s < ns; s++)
      for (uint32_t l = 0; l < layers; l++)
        params.push_back(1.0 - 2.0 * ingress[s] *
                                   (layers > 1 ? (double)l / (double)(layers - 1) : 0.0));
    const size_t nb1 = (size_t)nb + 1;
    const size_t o_cw = nb1 * 8, o_cn = 2 * nb1 * 8, o_par = ((2 * nb1 * 8 + nb1 * 4) + 255) & ~(size_t)255;
    const size_t o_rec = 256, o_key = o_rec + ((capacity * 32 + 255) & ~(size_t)255);
    HIP_TRY(hipSetDevice(c->device));
    if (int st = c->dev_in.ensure(o_par + params.size() * 8)) return st;
    if (int st = c->dev_out.ensure(o_key + capacity * 8 + 8)) return st;
    unsigned char* din = (unsigned char*)c->dev_in.ptr;
    unsigned char* dout = (unsigned char*)c->dev_out.ptr;
    hipStream_t s = c->stream;
    HIP_TRY(hipMemcpyAsync(din + o_par, params.data(), params.size() * 8, hipMemcpyHostToDevice, s));
    fbls::EventPrefixArgs P;
    P.t = fam->t_dev;
    P.vw = fam->vw_dev;
    P.fx = fam->fx;
    P.n = fam->n;
    P.nb = nb;
    P.scale = scale;
    P.sv = std::ldexp(1.0, -fam->ev);
    P.sw = std::ldexp(1.0, -fam->ew);
    P.cv = (double*)din;
    P.cw = (double*)(din + o_cw);
    P.cn = (uint32_t*)(din + o_cn);
    P.counter = (uint32_t*)dout;
    HIP_TRY(fbls::event_prefix_launch(P, s));
    const double* dpar = (const double*)(din + o_par);
    fbls::EventScanArgs A;
    A.cv = P.cv;
    A.cw = P.cw;
    A.cn = P.cn;
    A.durations = dpar;
    A.rho = dpar + nd;
    A.ingress = dpar + 2 * nd;
    A.factors = dpar + 2 * nd + ns;
    A.scale = scale;
    A.C = fam->C;
    A.threshold = threshold;
    A.n = fam->n;
    A.nb = nb;
    A.nd = (uint32_t)nd;
    A.ns = (uint32_t)ns;
    A.layers = layers;
    A.capacity = (uint32_t)capacity;
    A.counter = P.counter;
    A.records = (double*)(dout + o_rec);
    A.keys = (uint64_t*)(dout + o_key);
    HIP_TRY(fbls::event_scan_launch(A, s));
    uint32_t count = 0;
    HIP_TRY(hipMemcpyAsync(&count, P.counter, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (stats) {
      stats->templates_scanned = (uint64_t)nb * nd * ns;
      stats->threshold_survivors = count;
      stats->overflow = count > capacity;
      stats->phase_bins = nb;
    }
    if (count > capacity || count == 0) return FBLS_OK;
    std::vector<double> rec((size_t)count * 4);
    std::vector<uint64_t> keys(count);
    HIP_TRY(hipMemcpyAsync(rec.data(), A.records, rec.size() * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(keys.data(), A.keys, keys.size() * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<uint32_t> idx(count);
    std::iota(idx.begin(), idx.end(), 0u);
    std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    // events.rs:137-151
    for (uint32_t m = 0; m < count; m++) {
      const double* r = &rec[4 * (size_t)idx[m]];
      fbls_row& o = rows[m];
      o.period = INFINITY;
      o.epoch = (r[0] + r[1] / 2.0) / scale + fam->origin;
      o.duration = r[1] / scale;
      o.score = r[2];
      o.depth = r[3];
    }
    *nrows = count;
    return FBLS_OK;
  });
}

// Host side of the coarse periodogram that feeds the band-proposal network
// (feature_kernel.hip): one launch for every (star, period) of a batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/fbls.h"
#include "feature_kernel.h"
#include "host_internal.h"

using namespace fbls::host;

namespace {

// The periodogram into host `out` (doubles) or host `inputs` (network
// inputs), or, with device_inputs, the network inputs left in that device
// buffer (no copy back).  peaks (host, may be null): per (star, period) the
// largest unclipped SNR over the durations, copied back on the stream (the
// caller synchronizes it when the inputs stay on the device).
int coarse(fbls_ctx* c, const fbls_family* const* fams, size_t nfam, const double* periods,
           size_t np, const double* durations, size_t ng, uint32_t nb_min, uint32_t nb_max,
           double* out, float* inputs, float* device_inputs = nullptr, float* peaks = nullptr) {
  if (!c || (nfam && !fams) || !periods || !durations || (!out && !inputs && !device_inputs))
    return fail(FBLS_E_ARG, "null argument");
  if (np == 0 || ng == 0 || nfam == 0) return FBLS_OK;
  if (ng > 16) return fail(FBLS_E_ARG, "at most 16 durations");
  if (nb_min < 2 || nb_max < nb_min || nb_max > 4096)
    return fail(FBLS_E_ARG, "phase bins must satisfy 2 <= nb_min <= nb_max <= 4096");
  if ((uint64_t)nfam * np > 0x7fffffffull) return fail(FBLS_E_ARG, "too many (star, period) pairs");
  double d0 = INFINITY;
  for (size_t g = 0; g < ng; g++) {
    if (!(durations[g] > 0.0 && std::isfinite(durations[g])))
      return fail(FBLS_E_ARG, "durations must be positive and finite");
    d0 = std::min(d0, durations[g]);
  }
  std::vector<uint32_t> nb(np);
  for (size_t i = 0; i < np; i++) {
    if (!(periods[i] > 0.0 && std::isfinite(periods[i])))
      return fail(FBLS_E_ARG, "periods must be positive and finite");
    const double want = std::ceil(2.0 * periods[i] / d0);
    nb[i] = (uint32_t)std::min<double>(nb_max, std::max<double>(nb_min, want));
  }
  const double pmin = *std::min_element(periods, periods + np);
  double cell_max = 0.0;  // largest phase cell t * nb / P of the launch
  bool uniform = true;    // every star has equal fixed-point weights
  std::vector<fbls::FeatureStar> stars(nfam);
  for (size_t k = 0; k < nfam; k++) {
    const fbls_family* f = fams[k];
    if (!f) return fail(FBLS_E_ARG, "null family %zu", k);
    if (f->t_max * ((double)nb_max / pmin) > 9.0e15) return fail(FBLS_E_ARG, "phase cells exceed 2^53");
    cell_max = std::max(cell_max, f->t_max * ((double)nb_max / pmin));
    stars[k].t = f->t_dev;
    stars[k].vw = f->vw_dev;
    stars[k].fx = f->fx;
    stars[k].n = f->n;
    stars[k].n_pad = f->n_pad;
    stars[k].sorted = f->sorted ? 1u : 0u;
    stars[k].uniform_w = f->w_uniform;
    stars[k].pad_ = 0;
    uniform = uniform && f->w_uniform > 0;
    stars[k].sv = std::ldexp(1.0, -f->ev);
    stars[k].sw = std::ldexp(1.0, -f->ew);
    stars[k].C = f->C;
  }
  const size_t b_stars = nfam * sizeof(fbls::FeatureStar);
  const size_t o_per = (b_stars + 255) & ~(size_t)255;
  const size_t o_nb = o_per + ((np * 8 + 255) & ~(size_t)255);
  const size_t o_dur = o_nb + ((np * 4 + 255) & ~(size_t)255);
  // groups of up to 16 consecutive periods whose histograms (20 B per slot,
  // each sized by the group's largest nb) fit half the CU's LDS (two
  // 512-thread workgroups per CU): a star's observations are loaded once
  // per group and folded into every period of it
  const size_t budget = c->max_lds / 2 - 4096;  // the kernel's static LDS aside
  std::vector<uint32_t> gstart{0}, gstride;
  size_t lds_bytes = 0;
  uint32_t gmax = 1;
  for (size_t p = 0; p < np;) {
    uint32_t G = 1, mx = nb[p];
    while (G < fbls::kFeatureMaxGroup && p + G < np) {
      const uint32_t m2 = std::max(mx, nb[p + G]);
      if (fbls::feature_lds_bytes(m2, G + 1) > budget) break;
      mx = m2;
      G++;
    }
    gstride.push_back(fbls::feature_slots(mx));
    lds_bytes = std::max(lds_bytes, fbls::feature_lds_bytes(mx, G));
    gmax = std::max(gmax, G);
    p += G;
    gstart.push_back((uint32_t)p);
  }
  const size_t ngrp = gstride.size();
  const size_t o_gs = o_dur + ((ng * 8 + 255) & ~(size_t)255);
  const size_t o_gw = o_gs + (((ngrp + 1) * 4 + 255) & ~(size_t)255);
  const size_t o_out = o_gw + ((ngrp * 4 + 255) & ~(size_t)255);
  const size_t out_bytes = device_inputs ? 0 : nfam * np * ng * (out ? 8 : 4);
  const size_t o_peaks = o_out + ((out_bytes + 255) & ~(size_t)255);
  const size_t peak_bytes = peaks ? nfam * np * 4 : 0;
  HIP_TRY(hipSetDevice(c->device));
  if (int st = c->dev_in.ensure(o_peaks + peak_bytes)) return st;
  unsigned char* d = (unsigned char*)c->dev_in.ptr;
  hipStream_t s = c->stream;
  HIP_TRY(hipMemcpyAsync(d, stars.data(), b_stars, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d + o_per, periods, np * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d + o_nb, nb.data(), np * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d + o_dur, durations, ng * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d + o_gs, gstart.data(), (ngrp + 1) * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d + o_gw, gstride.data(), ngrp * 4, hipMemcpyHostToDevice, s));
  fbls::FeatureArgs A;
  A.stars = (const fbls::FeatureStar*)d;
  A.periods = (const double*)(d + o_per);
  A.nb = (const uint32_t*)(d + o_nb);
  A.np = (uint32_t)np;
  A.ng = (uint32_t)ng;
  A.durations = (const double*)(d + o_dur);
  A.out = out ? (double*)(d + o_out) : nullptr;
  A.inputs = out ? nullptr : (device_inputs ? device_inputs : (float*)(d + o_out));
  A.peaks = peaks ? (float*)(d + o_peaks) : nullptr;
  A.nb_max = *std::max_element(nb.begin(), nb.end());
  A.G = gmax;
  A.ngroups = (uint32_t)ngrp;
  A.gstart = (const uint32_t*)(d + o_gs);
  A.gstride = (const uint32_t*)(d + o_gw);
  A.lds_bytes = (uint32_t)lds_bytes;
  A.narrow = cell_max < 16777215.0 ? 1u : 0u;  // (nb <= nb_max, P >= pmin)
  A.uniform = uniform ? 1u : 0u;
  HIP_TRY(fbls::feature_launch(A, (uint32_t)nfam, s));
  if (peaks) HIP_TRY(hipMemcpyAsync(peaks, d + o_peaks, peak_bytes, hipMemcpyDeviceToHost, s));
  if (device_inputs) return FBLS_OK;  // stream-ordered before the network
  HIP_TRY(hipMemcpyAsync(out ? (void*)out : (void*)inputs, d + o_out, out_bytes,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return FBLS_OK;
}

}  // namespace

extern "C" int fbls_coarse_features(fbls_ctx* c, const fbls_family* const* fams, size_t nfam,
                                    const double* periods, size_t np, const double* durations,
                                    size_t ng, uint32_t nb_min, uint32_t nb_max, double* out) {
  return fbls::host::guard([&]() -> int {
    if (!out) return fail(FBLS_E_ARG, "null argument");
    return coarse(c, fams, nfam, periods, np, durations, ng, nb_min, nb_max, out, nullptr);
  });
}

extern "C" int fbls_coarse_inputs(fbls_ctx* c, const fbls_family* const* fams, size_t nfam,
                                  const double* periods, size_t np, const double* durations,
                                  size_t ng, uint32_t nb_min, uint32_t nb_max, float* out) {
  return fbls::host::guard([&]() -> int {
    if (!out) return fail(FBLS_E_ARG, "null argument");
    return coarse(c, fams, nfam, periods, np, durations, ng, nb_min, nb_max, nullptr, out);
  });
}

extern "C" int fbls_propose(fbls_ctx* c, fbls_cnn* net, const fbls_family* const* fams,
                            size_t nfam, const double* periods, size_t np,
                            const double* durations, size_t ng, uint32_t nb_min,
                            uint32_t nb_max, float* logits, float* peaks) {
  return fbls::host::guard([&]() -> int {
    if (!c || !net || !logits) return fail(FBLS_E_ARG, "null argument");
    if (nfam == 0 || np == 0) return FBLS_OK;
    if (fbls::host::cnn_context(net) != c)
      return fail(FBLS_E_ARG, "the network belongs to another context");
    if (fbls::host::cnn_in_channels(net) != ng)
      return fail(FBLS_E_ARG, "the network reads %u durations, not %zu",
                  fbls::host::cnn_in_channels(net), ng);
    // in batches of stars whose activations stay within a bounded footprint
    const size_t per = std::max<size_t>(1, ((size_t)1 << 22) / np);
    for (size_t s0 = 0; s0 < nfam; s0 += per) {
      const size_t m = std::min(per, nfam - s0);
      float* staged = nullptr;
      if (int st = fbls::host::cnn_stage_input(net, (uint32_t)m, (uint32_t)np, &staged)) return st;
      if (int st = coarse(c, fams + s0, m, periods, np, durations, ng, nb_min, nb_max, nullptr,
                          nullptr, staged, peaks ? peaks + s0 * np : nullptr))
        return st;
      if (int st = fbls::host::cnn_run_staged(net, (uint32_t)m, (uint32_t)np, logits + s0 * np))
        return st;
      HIP_TRY(hipStreamSynchronize(c->stream));  // the peaks' copy, on the same stream
    }
    return FBLS_OK;
  });
}

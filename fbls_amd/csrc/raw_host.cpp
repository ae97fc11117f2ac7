// Host side of the raw-mode preparation: the biweight detrend on the GPU
// (biweight_kernel.hip) and the per-window duration families
// (isakruas/fbls src/prep/curve.rs:89-198, 204-328, 387-425).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fbls.h"
#include "biweight_kernel.h"
#include "host_internal.h"
#include "prep.h"

using namespace fbls::host;

namespace fbls::host {

// curve.rs:266-310: window bounds per point within its gap-separated
// segment, then workgroups of up to 128 points whose windows' union is
// staged (or, beyond 4096 values, read from global memory).
int biweight_gpu(fbls_ctx* c, const std::vector<double>& t, const std::vector<double>& rel,
                double window, double gap, double cst, int iterations, std::vector<double>* out) {
  const size_t n = t.size();
  out->assign(n, 0.0);
  if (n == 0) return FBLS_OK;
  if (n > 0xffffffffu) return fail(FBLS_E_ARG, "too many points");
  std::vector<double> flux(n);
  for (size_t i = 0; i < n; i++) flux[i] = 1.0 + rel[i];
  std::vector<uint32_t> lo(n), hi(n);
  size_t start = 0;
  for (size_t i = 1; i <= n; i++) {
    if (i == n || t[i] - t[i - 1] > gap) {
      const double* seg = t.data() + start;
      const size_t len = i - start;
      for (size_t j = start; j < i; j++) {
        lo[j] = (uint32_t)(start + (std::lower_bound(seg, seg + len, t[j] - window / 2.0) - seg));
        hi[j] = (uint32_t)(start + (std::upper_bound(seg, seg + len, t[j] + window / 2.0) - seg));
      }
      start = i;
    }
  }
  // workgroups whose windows' union fits the staging size go to the sorted
  // kernel (listed first), the others to the counting kernel
  std::vector<uint32_t> staged, wide;
  for (size_t first = 0; first < n;) {
    size_t count = 1;
    while (count < fbls::kBiweightBlock && first + count < n &&
           hi[first + count] - lo[first] <= fbls::kBiweightWindow)
      count++;
    const uint32_t ulo = lo[first], uhi = hi[first + count - 1];
    std::vector<uint32_t>& list = uhi - ulo <= fbls::kBiweightWindow ? staged : wide;
    list.insert(list.end(), {(uint32_t)first, (uint32_t)count, ulo, uhi});
    first += count;
  }
  const size_t n_staged = staged.size() / 4;
  std::vector<uint32_t> blocks = staged;
  blocks.insert(blocks.end(), wide.begin(), wide.end());
  const size_t nb = blocks.size() / 4;
  const size_t b_flux = n * 8, b_idx = n * 4, b_blk = blocks.size() * 4;
  const size_t o_lo = (b_flux + 255) & ~(size_t)255;
  const size_t o_hi = o_lo + ((b_idx + 255) & ~(size_t)255);
  const size_t o_blk = o_hi + ((b_idx + 255) & ~(size_t)255);
  const size_t o_out = o_blk + ((b_blk + 255) & ~(size_t)255);
  HIP_TRY(hipSetDevice(c->device));
  if (int st = c->dev_in.ensure(o_out + n * 8)) return st;
  unsigned char* d = (unsigned char*)c->dev_in.ptr;
  hipStream_t s = c->stream;
  HIP_TRY(hipMemcpyAsync(d, flux.data(), b_flux, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d + o_lo, lo.data(), b_idx, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d + o_hi, hi.data(), b_idx, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d + o_blk, blocks.data(), b_blk, hipMemcpyHostToDevice, s));
  fbls::BiweightArgs A;
  A.flux = (const double*)d;
  A.lo = (const uint32_t*)(d + o_lo);
  A.hi = (const uint32_t*)(d + o_hi);
  A.blocks = (const uint32_t*)(d + o_blk);
  A.c = cst;
  A.iterations = (uint32_t)iterations;
  A.out = (double*)(d + o_out);
  HIP_TRY(fbls::biweight_sorted_launch(A, (uint32_t)n_staged, s));
  A.blocks = (const uint32_t*)(d + o_blk) + 4 * n_staged;
  HIP_TRY(fbls::biweight_launch(A, (uint32_t)(nb - n_staged), s));
  HIP_TRY(hipMemcpyAsync(out->data(), d + o_out, n * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return FBLS_OK;
}

}  // namespace fbls::host

extern "C" int fbls_biweight_flatten(fbls_ctx* c, const double* time, const double* rel,
                                     size_t n, double window, double gap, double cst,
                                     uint32_t iterations, double* out) {
  return fbls::host::guard([&]() -> int {
    if (!c || (n && (!time || !rel || !out))) return fail(FBLS_E_ARG, "null argument");
    if (!(window >= 0.0) || !std::isfinite(window))
      return fail(FBLS_E_ARG, "window must be finite and non-negative");
    for (size_t i = 1; i < n; i++)
      if (!(time[i] >= time[i - 1])) return fail(FBLS_E_ARG, "times must be sorted and finite");
    std::vector<double> t(time, time + n), r(rel, rel + n), o;
    if (int st = biweight_gpu(c, t, r, window, gap, cst, (int)iterations, &o)) return st;
    if (n) std::memcpy(out, o.data(), n * 8);
    return FBLS_OK;
  });
}

extern "C" int fbls_prepare_families(fbls_ctx* c, const double* time, const double* flux,
                                     const double* err, const int64_t* group, size_t n,
                                     const double* durations, size_t nd,
                                     const fbls_prep_cfg* cfg, size_t* nfam, size_t* fam_n,
                                     double* fam_time, double* fam_flux, double* fam_sigma,
                                     double* fam_window, size_t* fam_clipped,
                                     uint8_t* fam_members) {
  return fbls::host::guard([&]() -> int {
    if (!c || !cfg || !nfam || (n && (!time || !flux)) || (nd && !durations) || !fam_n ||
        !fam_time || !fam_flux || !fam_sigma || !fam_window || !fam_clipped || !fam_members)
      return fail(FBLS_E_ARG, "null argument");
    *nfam = 0;
    std::vector<fbls::Family> fams;
    std::string msg;
    auto flatten = [c](const std::vector<double>& t, const std::vector<double>& rel, double w,
                       double gap, double cst, int it, std::vector<double>* out) {
      return biweight_gpu(c, t, rel, w, gap, cst, it, out);
    };
    int st = fbls::prepare_families(time, flux, err, group, n, durations, nd, cfg, flatten,
                                    &fams, &msg);
    if (st) return msg.empty() ? st : fail(st, "%s", msg.c_str());
    for (size_t k = 0; k < fams.size(); k++) {
      const fbls::Family& f = fams[k];
      const size_t m = f.time.size();
      fam_n[k] = m;
      std::memcpy(fam_time + k * n, f.time.data(), m * 8);
      std::memcpy(fam_flux + k * n, f.flux.data(), m * 8);
      std::memcpy(fam_sigma + k * n, f.sigma.data(), m * 8);
      fam_window[k] = f.window;
      fam_clipped[k] = f.clipped;
      for (size_t i = 0; i < nd; i++) fam_members[k * nd + i] = 0;
      for (size_t i : f.indices) fam_members[k * nd + i] = 1;
    }
    *nfam = fams.size();
    return FBLS_OK;
  });
}

// Host driver of the evidence (model comparison) path: restates the host
// side of compare_search (isakruas/fbls src/search/evidence.rs:693-1003) --
// bank selection, two-period bin count, duration log masses, shape priors,
// staircase coefficients, orbit bank, pruning margin, and the final merge of
// tile parts into evidence rows -- around the kernels of evidence_kernel.hip.
// One pinned upload of the tables, seven kernels on the context's stream, one
// copy back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/fbls.h"
#include "evidence_kernel.h"
#include "fmath.h"
#include "host_internal.h"

using namespace fbls::host;
namespace fm = fbls::fm;

namespace {

constexpr double kNegInf = -__builtin_inf();

// evidence.rs:655-663 (sequential sum, as Iterator::sum)
double logsumexp(const double* v, size_t n) {
  double high = kNegInf;
  for (size_t i = 0; i < n; i++) high = std::fmax(high, v[i]);
  if (high == kNegInf) return high;
  double s = 0.0;
  for (size_t i = 0; i < n; i++) s += fm::exp(v[i] - high);
  return high + fm::log(s);
}

int check_cfg(const fbls_evidence_cfg* c) {
  if (!c) return fail(FBLS_E_ARG, "null evidence configuration");
  if (!c->ingress || c->ns == 0 || c->ns > 8) return fail(FBLS_E_ARG, "1..8 shapes supported");
  for (size_t i = 0; i < c->ns; i++)
    if (!(c->ingress[i] >= 0.0 && c->ingress[i] < 0.5))
      return fail(FBLS_E_ARG, "ingress fractions must be in [0, 0.5)");
  if (c->layers < 1 || c->layers > 16) return fail(FBLS_E_ARG, "1..16 layers supported");
  if (c->phase_oversample < 1) return fail(FBLS_E_ARG, "phase_oversample must be >= 1");
  return FBLS_OK;
}

}  // namespace

// evidence.rs:665-689
extern "C" int fbls_evidence_parameters(const fbls_evidence_cfg* cfg, double* p) {
  return fbls::host::guard([&]() -> int {
    if (!cfg || !p) return fail(FBLS_E_ARG, "null argument");
    const double dilution = cfg->target_flux_fraction;
    double planet_scale = cfg->planet_depth_scale;
    if (cfg->stellar_radius_solar > 0.0) {
      const double x = 0.10045 * cfg->planet_radius_scale_rj / cfg->stellar_radius_solar;
      planet_scale = x * x;
    }
    p[0] = planet_scale * dilution;
    p[1] = cfg->binary_depth_scale * dilution;
    p[2] = cfg->diluted_binary_depth_scale;
    p[3] = fm::log(cfg->planet_prior);
    p[4] = fm::log(cfg->binary_prior);
    p[5] = fm::log(1.0 - cfg->planet_prior - cfg->binary_prior);
    const double ratios[5] = {0.0, 0.01, 0.1, 0.5, 1.0};
    const double wp[5] = {0.75, 0.15, 0.08, 0.015, 0.005};
    const double wb[5] = {0.10, 0.10, 0.20, 0.25, 0.35};
    for (int i = 0; i < 5; i++) {
      p[6 + 3 * i] = ratios[i];
      p[7 + 3 * i] = wp[i];
      p[8 + 3 * i] = wb[i];
    }
    const double odd[7] = {0.25, 0.5, 0.75, 1.0, 4.0 / 3.0, 2.0, 4.0};
    for (int i = 0; i < 7; i++) p[21 + i] = odd[i];
    p[28] = 0.001;
    p[29] = 0.05;
    return FBLS_OK;
  });
}

extern "C" int fbls_compare(fbls_ctx* c, const fbls_family* f, double period,
                            const double* durations_in, const double* rho_in,
                            size_t nd_in, double residual_energy,
                            const fbls_evidence_cfg* cfg, fbls_evidence_row* rows,
                            size_t* nrows, fbls_evidence_debug* debug) {
  return fbls::host::guard([&]() -> int {
    if (!c || !f || !rows || !nrows || (nd_in && (!durations_in || !rho_in)))
      return fail(FBLS_E_ARG, "null argument");
    *nrows = 0;
    if (int st = check_cfg(cfg)) return st;
    if (!(period > 0.0 && std::isfinite(period)))
      return fail(FBLS_E_ARG, "period must be positive and finite");
    if (!std::isfinite(residual_energy)) return fail(FBLS_E_ARG, "residual energy must be finite");
    if (debug) {
      debug->tiles = 0;
      debug->nb = 0;
      for (double& x : debug->floors) x = 0.0;
    }
    // durations below a quarter period, ascending, stable (evidence.rs:699-710)
    std::vector<std::pair<double, double>> ds;
    for (size_t i = 0; i < nd_in; i++) {
      if (!(durations_in[i] > 0.0 && std::isfinite(durations_in[i])) ||
          !(rho_in[i] > 0.0 && std::isfinite(rho_in[i])))
        return fail(FBLS_E_ARG, "durations and rho must be positive and finite");
      if (durations_in[i] < period / 4.0) ds.emplace_back(durations_in[i], rho_in[i]);
    }
    std::stable_sort(ds.begin(), ds.end(),
                     [](const std::pair<double, double>& a, const std::pair<double, double>& b) {
                       return a.first < b.first;
                     });
    if (ds.empty()) return FBLS_OK;
    const size_t nd = ds.size();
    if (nd > 64) return fail(FBLS_E_ARG, "at most 64 durations below period/4");
    std::vector<double> dur(nd), rho(nd);
    for (size_t i = 0; i < nd; i++) {
      dur[i] = ds[i].first;
      rho[i] = ds[i].second;
    }
    const size_t ns = cfg->ns, layers = cfg->layers;
    const uint32_t n = f->n;
    const double dmin = dur[0];
    const double nbq = std::ceil((double)cfg->phase_oversample * period / dmin / 4.0);
    if (!(nbq < 1e8)) return fail(FBLS_E_ARG, "period / duration ratio too large");
    const size_t nb = std::max<size_t>(8, 4 * (size_t)nbq);
    const size_t nb2 = 2 * nb;
    if (nb2 > (1u << 26)) return fail(FBLS_E_ARG, "two-period bins (%zu) exceed 2^26", nb2);
    const double scale = (double)nb / period;
    if (!(f->t_max * scale < 9.0e15)) return fail(FBLS_E_ARG, "phase cells exceed 2^53");
    const size_t tiles = (nb + fbls::kEvTile - 1) / fbls::kEvTile;

    // duration log masses (evidence.rs:719-726)
    std::vector<double> dlm(nd);
    double dmax = kNegInf;
    for (size_t i = 0; i < nd; i++) {
      const double r2 = rho[i] * rho[i];
      dlm[i] = -0.5 * (((double)n - 1.0) * fm::log(r2) + residual_energy / r2);
      dmax = std::fmax(dmax, dlm[i]);
    }
    for (size_t i = 0; i < nd; i++) dlm[i] -= dmax;
    // shape priors (evidence.rs:728-738)
    std::vector<double> lgp(ns), lgb(ns), priors(2 * ns);
    for (size_t s = 0; s < ns; s++) {
      lgp[s] = -4.0 * cfg->ingress[s];
      lgb[s] = 2.0 * cfg->ingress[s];
    }
    const double lp0 = logsumexp(lgp.data(), ns), lb0 = logsumexp(lgb.data(), ns);
    for (size_t s = 0; s < ns; s++) {
      priors[2 * s] = lgp[s] - lp0;
      priors[2 * s + 1] = lgb[s] - lb0;
    }
    // staircase coefficients (evidence.rs:739-748)
    std::vector<double> shapes(ns * layers * 3);
    for (size_t s = 0; s < ns; s++) {
      double previous = 0.0;
      const size_t denom = std::max<size_t>(layers, 2) - 1;
      for (size_t l = 0; l < layers; l++) {
        const double fraction = 1.0 - 2.0 * cfg->ingress[s] * (double)l / (double)denom;
        const double inc = 1.0 / (double)layers;
        double* e = &shapes[(s * layers + l) * 3];
        e[0] = fraction;
        e[1] = inc;
        e[2] = inc * (2.0 * previous + inc);
        previous += inc;
      }
    }
    double params[fbls::kEvParams];
    fbls_evidence_parameters(cfg, params);
    // orbit bank (evidence.rs:750-789)
    const double phases[3] = {0.25, 0.5, 0.75};
    const double logw[3] = {fm::log(0.1), fm::log(0.8), fm::log(0.1)};
    std::vector<double> lpb(nd * 3, kNegInf), lbb(nd * 3, kNegInf);
    size_t valid_count = 0;
    for (size_t di = 0; di < nd; di++) {
      const double k = std::fmax(std::floor(dur[di] * scale + 0.5), 1.0);
      if (k < (double)(nb / 4)) valid_count++;
      for (size_t o = 0; o < 3; o++) {
        const double ks = std::fmax(std::floor(k + 0.5), 1.0);
        const double shift = std::floor(phases[o] * (double)nb + 0.5);
        const double start = shift + std::floor((k - ks) / 2.0);
        const bool ok = k < (double)(nb / 4) && ks < (double)(nb / 2) && start >= k &&
                        start + ks <= (double)nb;
        if (ok) {
          lpb[di * 3 + o] = logw[o];
          lbb[di * 3 + o] = logw[o];
        }
      }
    }
    const double cnt = fm::log((double)valid_count);
    const double norm_p = cnt - logsumexp(lpb.data(), nd * 3);
    const double norm_b = cnt - logsumexp(lbb.data(), nd * 3);
    std::vector<double> orbit(nd * 3 * 4);
    for (size_t di = 0; di < nd; di++)
      for (size_t o = 0; o < 3; o++) {
        double* e = &orbit[(di * 3 + o) * 4];
        e[0] = phases[o];
        e[1] = 1.0;
        e[2] = lpb[di * 3 + o] + norm_p;
        e[3] = lbb[di * 3 + o] + norm_b;
      }
    const double margin =
        cfg->evidence_rtol > 0.0
            ? fm::log(cfg->evidence_rtol) - fm::log((double)((3 + 1) * nb * nd * ns))
            : kNegInf;

    // one staging buffer: tables in, floors + parts (+ screen) out
    const size_t n_tab = 4 * nd + ns * layers * 3 + 2 * ns + fbls::kEvParams + nd * 12;
    const size_t tab_bytes = n_tab * 8;
    const size_t scratch =
        fbls::evidence_scratch_bytes((uint32_t)nb2, (uint32_t)tiles, (uint32_t)(nd * ns));
    const size_t tab_off = (scratch + 255) & ~(size_t)255;
    if (int st = c->dev_ev.ensure(tab_off + tab_bytes)) return st;
    const size_t out_doubles = 4 + tiles * 2 * fbls::kEvPart + tiles * 4;
    if (int st = c->pin_in.ensure(tab_bytes)) return st;
    if (int st = c->pin_out.ensure(out_doubles * 8)) return st;
    double* tab = (double*)c->pin_in.ptr;
    size_t at = 0;
    auto put = [&](const double* src, size_t cnt_) {
      std::memcpy(tab + at, src, cnt_ * 8);
      at += cnt_;
    };
    const size_t o_dur = at;
    put(dur.data(), nd);
    const size_t o_rho = at;
    put(rho.data(), nd);
    const size_t o_dlm = at;
    put(dlm.data(), nd);
    const size_t o_shapes = at;
    put(shapes.data(), shapes.size());
    const size_t o_priors = at;
    put(priors.data(), priors.size());
    const size_t o_params = at;
    put(params, fbls::kEvParams);
    const size_t o_orbit = at;
    put(orbit.data(), orbit.size());

    unsigned char* dbase = (unsigned char*)c->dev_ev.ptr;
    const double* dtab = (const double*)(dbase + tab_off);
    fbls::EvArgs A{};
    A.t = f->t_dev;
    A.vw = f->vw_dev;
    A.fx = f->fx;
    A.n = n;
    A.nb = (uint32_t)nb;
    A.nb2 = (uint32_t)nb2;
    A.nd = (uint32_t)nd;
    A.ns = (uint32_t)ns;
    A.layers = (uint32_t)layers;
    A.tiles = (uint32_t)tiles;
    A.sv = std::ldexp(1.0, -f->ev);
    A.sw = std::ldexp(1.0, -f->ew);
    A.scale = scale;
    A.C = f->C;
    A.threshold = cfg->threshold;
    A.screen_threshold = cfg->threshold * (1.0 - 1e-12);
    A.margin = margin;
    A.dur = dtab + o_dur;
    A.rho = dtab + o_rho;
    A.dlm = dtab + o_dlm;
    A.shapes = dtab + o_shapes;
    A.priors = dtab + o_priors;
    A.params = dtab + o_params;
    A.orbit = dtab + o_orbit;
    fbls::evidence_bind_scratch(&A, dbase);

    HIP_TRY(hipSetDevice(c->device));
    double* hout = (double*)c->pin_out.ptr;
    // The call's whole device sequence: tables in, seven kernels, results out.
    auto enqueue = [&]() -> hipError_t {
      hipError_t e;
      if ((e = hipMemcpyAsync((void*)dtab, tab, tab_bytes, hipMemcpyHostToDevice, c->stream)))
        return e;
      if ((e = fbls::evidence_launch(A, c->stream))) return e;
      if ((e = hipMemcpyAsync(hout, A.floors, 4 * 8, hipMemcpyDeviceToHost, c->stream))) return e;
      if ((e = hipMemcpyAsync(hout + 4, A.parts, tiles * 2 * fbls::kEvPart * 8,
                              hipMemcpyDeviceToHost, c->stream)))
        return e;
      if (debug && debug->screen)
        return hipMemcpyAsync(hout + 4 + tiles * 2 * fbls::kEvPart, A.screen, tiles * 4 * 8,
                              hipMemcpyDeviceToHost, c->stream);
      return hipSuccess;
    };
    if (c->use_graphs) {
      // captured per call and updated into the cached executable graph (one
      // submission instead of nine); a new topology instantiates afresh
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const hipError_t ee = enqueue();
      const hipError_t ce = hipStreamEndCapture(c->stream, &g);
      if (ee != hipSuccess || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(FBLS_E_HIP, "evidence graph capture failed: %s",
                    hipGetErrorString(ee != hipSuccess ? ee : ce));
      }
      bool ready = false;
      if (c->ev_exec) {
        hipGraphExecUpdateResult r;
        hipGraphNode_t bad = nullptr;
        ready = hipGraphExecUpdate(c->ev_exec, g, &bad, &r) == hipSuccess &&
                r == hipGraphExecUpdateSuccess;
        if (!ready) {
          (void)hipGetLastError();
          (void)hipGraphExecDestroy(c->ev_exec);
          c->ev_exec = nullptr;
        }
      }
      hipError_t ie = hipSuccess;
      if (!ready) ie = hipGraphInstantiate(&c->ev_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        c->ev_exec = nullptr;
        return fail(FBLS_E_HIP, "evidence graph instantiation failed: %s", hipGetErrorString(ie));
      }
      HIP_TRY(hipGraphLaunch(c->ev_exec, c->stream));
      c->graph_launches++;
    } else {
      HIP_TRY(enqueue());
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double* floors = hout;
    const double* parts = hout + 4;
    const bool passed = floors[3] != 0.0;
    if (debug) {
      debug->tiles = tiles;
      debug->nb = nb;
      for (int i = 0; i < 4; i++) debug->floors[i] = passed ? floors[i] : 0.0;
      if (debug->screen)
        std::memcpy(debug->screen, hout + 4 + tiles * 2 * fbls::kEvPart,
                    std::min(debug->screen_cap, tiles * 4) * 8);
      if (debug->parts && passed)
        std::memcpy(debug->parts, parts,
                    std::min(debug->parts_cap, tiles * 2 * fbls::kEvPart) * 8);
    }
    if (!passed) return FBLS_OK;

    // merge of the tile parts (evidence.rs:936-1002)
    std::vector<double> dm;
    for (size_t i = 0; i < nd; i++)
      if (std::fmax(std::floor(dur[i] * scale + 0.5), 1.0) < (double)(nb / 4)) dm.push_back(dlm[i]);
    const double log_mass = fm::log((double)nb) + logsumexp(dm.data(), dm.size());
    fbls_evidence_row r2[2];
    std::vector<double> col(tiles);
    for (int sel = 0; sel < 2; sel++) {
      auto value = [&](size_t t, size_t i) { return parts[(t * 2 + sel) * fbls::kEvPart + i]; };
      double mx = kNegInf;
      for (size_t t = 0; t < tiles; t++) mx = std::fmax(mx, value(t, 1));
      double key = __builtin_inf();
      for (size_t t = 0; t < tiles; t++)
        if (value(t, 1) == mx) key = std::fmin(key, value(t, 0));
      size_t tile = 0;
      for (size_t t = 0; t < tiles; t++)
        if (value(t, 1) == mx && value(t, 0) == key) {
          tile = t;
          break;
        }
      for (size_t t = 0; t < tiles; t++) col[t] = value(t, 11);
      const double lp = logsumexp(col.data(), tiles) - log_mass;
      for (size_t t = 0; t < tiles; t++) col[t] = value(t, 12);
      const double lb = logsumexp(col.data(), tiles) - log_mass;
      const double joint = params[3] + lp - fm::log_add(params[5], params[4] + lb);
      const uint64_t decoded = mx > kNegInf ? (uint64_t)key : 0;
      const size_t j = (size_t)(decoded % nb);
      const size_t di = (size_t)((decoded / nb) % nd);
      const size_t si = (size_t)(decoded / ((uint64_t)nb * nd));
      const double width = std::fmax(std::floor(dur[di] * scale + 0.5), 1.0);
      fbls_evidence_row& r = r2[sel];
      r.period = period;
      r.epoch = ((double)j + width / 2.0) / scale + f->origin;
      r.duration = width / scale;
      r.snr = mx > kNegInf ? value(tile, 2) : kNegInf;
      r.depth = value(tile, 3);
      r.planet_log_odds = joint;
      r.log_bf_planet_noise = lp;
      r.log_bf_binary_noise = lb;
      r.binary_secondary_ratio = value(tile, 6);
      r.binary_secondary_phase = value(tile, 7);
      r.binary_odd_even_ratio = value(tile, 8);
      r.binary_period_days = value(tile, 9) * period;
      r.ingress_fraction = cfg->ingress[si < ns - 1 ? si : ns - 1];
      r.selection_flags = sel == 0 ? 1u : 2u;
      r.phase = (uint32_t)j;
      r.duration_index = (uint32_t)di;
      r.shape = (uint32_t)si;
    }
    const bool hit0 = std::isfinite(r2[0].snr), hit1 = std::isfinite(r2[1].snr);
    const bool same = hit0 && hit1 && r2[0].period == r2[1].period &&
                      r2[0].epoch == r2[1].epoch && r2[0].duration == r2[1].duration &&
                      r2[0].snr == r2[1].snr && r2[0].depth == r2[1].depth &&
                      r2[0].ingress_fraction == r2[1].ingress_fraction;
    if (hit0) {
      rows[*nrows] = r2[0];
      if (same) rows[*nrows].selection_flags = 3;
      (*nrows)++;
    }
    if (hit1 && !same) rows[(*nrows)++] = r2[1];
    return FBLS_OK;
  });
}

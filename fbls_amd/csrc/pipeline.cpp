// Multi-pass search (isakruas/fbls src/pipeline.rs:27-221 without single
// events): per pass the signal search of every family on the GPU, the
// greedy peak selection with event-overlap rescoring over all families
// (src/select/peaks.rs), masking of the lead peak in every family and
// rebuilt family states; then the model comparison of the strongest peaks.
// Prepared input is one family with every duration; raw input is one family
// per detrending window (prepare_families).  Rescoring runs on the GPU
// (select_kernel.hip) over the family's fixed-point fields, so every sum is
// exact; the host-side statistics use the same fixed-point values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <queue>
#include <string>
#include <vector>

#include "../../include/fbls.h"
#include "host_internal.h"
#include "prep.h"
#include "select_kernel.h"

using namespace fbls::host;

namespace {

double rem_euclid(double a, double b) {
  const double r = std::fmod(a, b);
  return r < 0.0 ? r + std::fabs(b) : r;
}

// curve.rs:572-585
void transit_mask(const std::vector<double>& time, double period, double epoch,
                  double duration, double margin, std::vector<uint8_t>& out) {
  out.resize(time.size());
  for (size_t i = 0; i < time.size(); i++) {
    const double phase = rem_euclid(time[i] - epoch + period / 2.0, period) - period / 2.0;
    out[i] = std::fabs(phase) <= (1.0 + margin) * duration / 2.0;
  }
}

// peaks.rs:66-83
size_t event_count(const std::vector<double>& time, double period, double epoch,
                   double duration, size_t minimum_points) {
  std::vector<uint8_t> inside;
  transit_mask(time, period, epoch, duration, 0.0, inside);
  std::map<int64_t, size_t> counts;
  for (size_t i = 0; i < time.size(); i++)
    if (inside[i]) counts[(int64_t)std::round((time[i] - epoch) / period)]++;
  size_t events = 0;
  for (const auto& kv : counts) events += kv.second >= minimum_points;
  return events;
}

struct Curve {
  std::vector<double> t, f, s;  // absolute time, flux, sigma
};

// A family: its curve, its state and the device copy of the state.
struct Fam {
  Curve curve;
  std::vector<double> dur;           // the family's durations (days)
  std::vector<size_t> indices;       // their positions in the bank
  std::vector<double> t, v, w, rho;  // relative time, fixed-point fields
  double C = 0.0, origin = 0.0, energy = 0.0, inflation = 0.0;
  size_t segments = 0;
  fbls_family* dev = nullptr;
  Fam() = default;
  Fam(const Fam&) = delete;
  Fam& operator=(const Fam&) = delete;
  ~Fam() { fbls_family_free(dev); }
};

// curve.rs:541-569 then the upload; fails below six observations
int build_fam(fbls_ctx* c, Curve curve, const std::vector<double>& durs,
              const std::vector<size_t>& indices, int correlated, Fam* out) {
  const double* dur = durs.data();
  const size_t nd = durs.size();
  const size_t n = curve.t.size();
  std::vector<double> t(n + 1), v(n + 1), w(n + 1), rho(nd);
  size_t m = 0, segs = 0;
  double C = 0.0, origin = 0.0, infl = 0.0, energy = 0.0;
  std::string msg;
  int st = fbls::family_state(curve.t.data(), curve.f.data(), curve.s.data(), n, dur, nd,
                              correlated, t.data(), v.data(), w.data(), &m, &C, &origin,
                              rho.data(), &infl, &segs, &energy, &msg);
  if (st) return fail(st, "%s", msg.c_str());
  fbls_family* dev = nullptr;
  st = fbls_family_upload(c, t.data(), v.data(), w.data(), m, C, origin, &dev);
  if (st) return st;
  out->curve = std::move(curve);
  out->dur = durs;
  out->indices = indices;
  out->t.assign(t.begin(), t.begin() + m);
  out->v.resize(m);
  out->w.resize(m);
  for (size_t i = 0; i < m; i++) {
    out->v[i] = std::ldexp(fbls::fixed_value(dev->vq[i], dev->fx & 1), -dev->ev);
    out->w[i] = std::ldexp(fbls::fixed_value(dev->wq[i], dev->fx & 2), -dev->ew);
  }
  out->rho = rho;
  out->C = C;
  out->origin = origin;
  out->energy = energy;
  out->inflation = infl;
  out->segments = segs;
  fbls_family_free(out->dev);
  out->dev = dev;
  return FBLS_OK;
}

Curve masked(const Curve& c, double period, double epoch, double duration, double margin) {
  std::vector<uint8_t> m;
  transit_mask(c.t, period, epoch, duration, margin, m);
  Curve out;
  for (size_t i = 0; i < c.t.size(); i++)
    if (!m[i]) {
      out.t.push_back(c.t[i]);
      out.f.push_back(c.f[i]);
      out.s.push_back(c.s[i]);
    }
  return out;
}

// peaks.rs:85-94
double nearest_rho(const double* dur, const std::vector<double>& rho, size_t nd, double d) {
  size_t best = 0;
  for (size_t i = 1; i < nd; i++)
    if (std::fabs(dur[i] - d) < std::fabs(dur[best] - d)) best = i;
  return rho[best];
}

struct Entry {
  double v;
  size_t i;
};
// BinaryHeap<Entry> order (peaks.rs:367-379): larger score, then smaller row
struct Below {
  bool operator()(const Entry& x, const Entry& y) const {
    return y.v > x.v || (y.v == x.v && y.i < x.i);
  }
};
using Heap = std::priority_queue<Entry, std::vector<Entry>, Below>;

// Selection state of one pass over every family.
struct Selector {
  fbls_ctx* c;
  const fbls_row* table;
  size_t nrows;
  std::vector<const Fam*> fams;
  const std::vector<size_t>* family;  // family of each row
  const std::vector<double>* time_all;
  std::vector<uint8_t>* excluded;
  std::vector<std::vector<size_t>> pos;  // per family: index in time_all
  double threshold;
  const double* d_time_all = nullptr;  // merged times on the device (borrowed)
  DevBuf d_ex_all, d_ex_fam, d_rows, d_out;
  ~Selector() {
    d_ex_all.release();
    d_ex_fam.release();
    d_rows.release();
    d_out.release();
  }

  // Rescorer::evaluate (peaks.rs:148-240) on the GPU: the counts over the
  // merged times for every row, then the box sums per family
  int evaluate(const std::vector<size_t>& idx, std::vector<double>& score,
               std::vector<double>& required) {
    const size_t m = idx.size();
    const size_t n_all = time_all->size();
    std::vector<double> rows(3 * m);
    for (size_t s = 0; s < m; s++) {
      const fbls_row& r = table[idx[s]];
      rows[3 * s] = r.period;
      rows[3 * s + 1] = r.epoch;
      rows[3 * s + 2] = r.duration;
    }
    int st;
    if ((st = d_ex_all.ensure(n_all + 1)) || (st = d_rows.ensure(rows.size() * 8)) ||
        (st = d_out.ensure(m * 24)))
      return st;
    HIP_TRY(hipMemcpyAsync(d_ex_all.ptr, excluded->data(), n_all, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d_rows.ptr, rows.data(), rows.size() * 8, hipMemcpyHostToDevice,
                           c->stream));
    fbls::CountArgs ca;
    ca.time_all = d_time_all;
    ca.excluded = (const uint8_t*)d_ex_all.ptr;
    ca.n = (uint32_t)n_all;
    ca.rows = (const double*)d_rows.ptr;
    ca.counts = (uint32_t*)((unsigned char*)d_out.ptr + m * 16);
    HIP_TRY(fbls::count_launch(ca, (uint32_t)m, c->stream));
    std::vector<unsigned char> out(m * 24);
    HIP_TRY(hipMemcpyAsync(out.data() + m * 16, (unsigned char*)d_out.ptr + m * 16, m * 8,
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const uint32_t* counts = (const uint32_t*)(out.data() + m * 16);
    score.assign(m, -INFINITY);
    required.resize(m);
    for (size_t s = 0; s < m; s++) {
      const double total = (double)std::max<uint32_t>(counts[2 * s], 1);
      const double unshared = (double)counts[2 * s + 1] / total;
      required[s] = std::fmax(threshold * std::sqrt(unshared), 0.5 * threshold);
    }
    for (size_t k = 0; k < fams.size(); k++) {
      std::vector<size_t> slots;
      for (size_t s = 0; s < m; s++)
        if ((*family)[idx[s]] == k) slots.push_back(s);
      if (slots.empty()) continue;
      const Fam* fam = fams[k];
      const size_t n = fam->t.size(), q = slots.size();
      std::vector<uint8_t> ex_fam(n + 1, 0);
      for (size_t i = 0; i < n; i++) ex_fam[i] = (*excluded)[pos[k][i]];
      std::vector<double> frows(3 * q);
      for (size_t j = 0; j < q; j++)
        for (int r = 0; r < 3; r++) frows[3 * j + r] = rows[3 * slots[j] + r];
      if ((st = d_ex_fam.ensure(n + 1))) return st;
      HIP_TRY(hipMemcpyAsync(d_ex_fam.ptr, ex_fam.data(), n, hipMemcpyHostToDevice, c->stream));
      HIP_TRY(hipMemcpyAsync(d_rows.ptr, frows.data(), q * 24, hipMemcpyHostToDevice, c->stream));
      fbls::RescoreArgs ra;
      ra.t = fam->dev->t_dev;
      ra.vw = fam->dev->vw_dev;
      ra.fx = fam->dev->fx;
      ra.excluded = (const uint8_t*)d_ex_fam.ptr;
      ra.n = (uint32_t)n;
      ra.origin = fam->origin;
      ra.rows = (const double*)d_rows.ptr;
      ra.sums = (long long*)d_out.ptr;
      HIP_TRY(fbls::rescore_launch(ra, (uint32_t)q, c->stream));
      HIP_TRY(hipMemcpyAsync(out.data(), d_out.ptr, q * 16, hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      const long long* sums = (const long long*)out.data();
      for (size_t j = 0; j < q; j++) {
        const size_t s = slots[j];
        const double a = std::ldexp((double)sums[2 * j], -fam->dev->ev);
        const double hp = std::ldexp((double)sums[2 * j + 1], -fam->dev->ew);
        const double norm = hp;
        const double qq = norm - hp * hp / fam->C;
        score[s] = (qq > 0.0 && a > 0.0)
                       ? (a / nearest_rho(fam->dur.data(), fam->rho, fam->dur.size(),
                                          table[idx[s]].duration)) /
                             std::sqrt(std::fmax(qq, 1e-300))
                       : -INFINITY;
      }
    }
    return FBLS_OK;
  }

  // peaks.rs:243-299
  double without_strongest_event(const fbls_row& row, size_t k) const {
    struct Acc {
      double a = 0.0, hp = 0.0, w = 0.0;
    };
    const Fam* fam = fams[k];
    std::map<int64_t, Acc> sums;
    std::vector<std::pair<size_t, int64_t>> members;
    for (size_t i = 0; i < fam->t.size(); i++) {
      if ((*excluded)[pos[k][i]]) continue;
      const double ta = fam->t[i] + fam->origin;
      const double phase =
          rem_euclid(ta - row.epoch + row.period / 2.0, row.period) - row.period / 2.0;
      if (std::fabs(phase) <= row.duration / 2.0) {
        const int64_t cycle = (int64_t)std::round((ta - row.epoch) / row.period);
        Acc& e = sums[cycle];
        e.a += fam->v[i];
        e.hp += fam->w[i];
        e.w += fam->w[i];
        members.emplace_back(i, cycle);
      }
    }
    if (sums.empty()) return -INFINITY;
    const double rho = nearest_rho(fam->dur.data(), fam->rho, fam->dur.size(), row.duration);
    int64_t strongest = 0;
    double best = -INFINITY;
    for (const auto& kv : sums) {
      const Acc& e = kv.second;
      const double s = e.a / rho / std::sqrt(std::fmax(e.w - e.hp * e.hp / fam->C, 1e-300));
      if (s > best) {
        best = s;
        strongest = kv.first;
      }
    }
    double a = 0.0, hp = 0.0, norm = 0.0;
    for (const auto& m : members)
      if (m.second != strongest) {
        a += fam->v[m.first];
        hp += fam->w[m.first];
        norm += fam->w[m.first];
      }
    const double q = norm - hp * hp / fam->C;
    return (q > 0.0 && a > 0.0) ? a / rho / std::sqrt(std::fmax(q, 1e-300)) : -INFINITY;
  }
};

// peaks.rs:383-419
int refresh(Selector& S, std::vector<double>& remaining, std::vector<uint8_t>& fresh,
            Heap& heap) {
  const size_t block = 8192;
  std::vector<size_t> stale;
  std::vector<Entry> kept;
  while (!heap.empty()) {
    const Entry e = heap.top();
    heap.pop();
    if (e.v != remaining[e.i]) continue;
    if (fresh[e.i]) {
      kept.push_back(e);
    } else {
      stale.push_back(e.i);
      if (stale.size() == block) break;
    }
  }
  for (const Entry& e : kept) heap.push(e);
  if (stale.empty()) return FBLS_OK;
  std::vector<double> score, required;
  if (int st = S.evaluate(stale, score, required)) return st;
  for (size_t m = 0; m < stale.size(); m++) {
    const size_t i = stale[m];
    remaining[i] = score[m] >= required[m] ? score[m] : -INFINITY;
    fresh[i] = 1;
    if (std::isfinite(remaining[i])) heap.push(Entry{remaining[i], i});
  }
  return FBLS_OK;
}

void empty_evidence(fbls_peak* p) {
  const double nan = std::nan("");
  p->planet_log_odds = p->log_bf_planet_noise = p->log_bf_binary_noise = nan;
  p->binary_secondary_ratio = p->binary_secondary_phase = nan;
  p->binary_odd_even_ratio = p->binary_period_days = p->ingress_fraction = nan;
  p->evidence_epoch_bjd = p->evidence_duration_hours = nan;
  p->model_preferred = 0;
  p->binary_diagnostics_complete = 0;
}

// peaks.rs:308-487
int select_peaks(Selector& S, size_t limit, size_t min_events,
                 std::vector<std::vector<uint8_t>>& masks, size_t pass,
                 std::vector<fbls_peak>& accepted) {
  const double margin = 0.5;
  const size_t max_iterations = 20000;
  if (S.nrows == 0) return FBLS_OK;
  const std::vector<double>& time_all = *S.time_all;
  std::vector<uint8_t>& excluded = *S.excluded;
  const bool any_ex = std::any_of(excluded.begin(), excluded.end(), [](uint8_t x) { return x; });
  std::vector<double> remaining(S.nrows);
  std::vector<uint8_t> fresh(S.nrows, any_ex ? 0 : 1);
  Heap heap;
  for (size_t i = 0; i < S.nrows; i++) {
    const double sc = S.table[i].score;
    remaining[i] = (std::isfinite(sc) && sc >= S.threshold) ? sc : -INFINITY;
    if (std::isfinite(remaining[i])) heap.push(Entry{remaining[i], i});
  }
  std::vector<uint8_t> mask, wide;
  size_t done = 0, iterations = 0;
  while (done < limit && iterations < max_iterations) {
    iterations++;
    while (!heap.empty() && heap.top().v != remaining[heap.top().i]) heap.pop();
    if (heap.empty()) break;
    const size_t i = heap.top().i;
    if (!fresh[i]) {
      if (int st = refresh(S, remaining, fresh, heap)) return st;
      continue;
    }
    heap.pop();
    const double isolated = remaining[i];
    remaining[i] = -INFINITY;
    const fbls_row& row = S.table[i];
    transit_mask(time_all, row.period, row.epoch, row.duration, 0.0, mask);
    size_t total = 0;
    for (uint8_t x : mask) total += x;
    bool clash = false;
    for (const auto& other : masks) {
      size_t both = 0;
      for (size_t j = 0; j < mask.size(); j++) both += mask[j] && other[j];
      if ((total ? (double)both / (double)total : 0.0) >= 0.5) {
        clash = true;
        break;
      }
    }
    if (clash) continue;
    const size_t events = event_count(time_all, row.period, row.epoch, row.duration, 3);
    if (events < min_events) continue;
    const size_t k = (*S.family)[i];
    const double leftover = S.without_strongest_event(row, k);
    if (min_events > 1 && leftover < 0.5 * S.threshold) continue;
    transit_mask(time_all, row.period, row.epoch, row.duration, margin, wide);
    for (size_t j = 0; j < excluded.size(); j++) excluded[j] |= wide[j];
    masks.push_back(wide);
    fbls_peak p;
    std::memset(&p, 0, sizeof(p));
    p.period_days = row.period;
    p.epoch_bjd = row.epoch;
    p.duration_hours = 24.0 * row.duration;
    p.snr = isolated;
    p.snr_folded = row.score;
    p.snr_without_strongest_event = leftover;
    p.depth = row.depth;
    p.depth_ppm = 1e6 * row.depth;
    p.events = events;
    p.family = k;
    p.pass = pass;
    p.row = i;
    empty_evidence(&p);
    accepted.push_back(p);
    done++;
    std::fill(fresh.begin(), fresh.end(), 0);
    if (int st = refresh(S, remaining, fresh, heap)) return st;
  }
  return FBLS_OK;
}

// peaks.rs:490-550
std::vector<fbls_single_event> select_events(const std::vector<fbls_row>& table,
                                             const std::vector<double>& time, double threshold,
                                             size_t limit,
                                             const std::vector<std::vector<uint8_t>>& taken) {
  std::vector<fbls_single_event> output;
  std::vector<std::vector<uint8_t>> masks = taken;
  if (table.empty()) return output;
  std::vector<double> remaining(table.size());
  for (size_t i = 0; i < table.size(); i++) {
    const double sc = table[i].score;
    remaining[i] = (std::isfinite(sc) && sc >= threshold) ? sc : -INFINITY;
  }
  std::vector<uint8_t> inside(time.size());
  size_t iterations = 0;
  while (output.size() < limit && iterations < 5000) {
    iterations++;
    size_t i = 0;  // Iterator::max_by: the last of equal maxima
    for (size_t j = 1; j < table.size(); j++)
      if (remaining[j] >= remaining[i]) i = j;
    if (!std::isfinite(remaining[i])) break;
    const fbls_row row = table[i];
    size_t points = 0;
    for (size_t j = 0; j < time.size(); j++) {
      inside[j] = std::fabs(time[j] - row.epoch) <= row.duration / 2.0;
      points += inside[j];
    }
    for (size_t j = 0; j < table.size(); j++)
      if (std::fabs(table[j].epoch - row.epoch) <= (table[j].duration + row.duration) / 2.0)
        remaining[j] = -INFINITY;
    bool clash = points < 3;
    for (size_t m = 0; m < masks.size() && !clash; m++) {
      size_t both = 0;
      for (size_t j = 0; j < time.size(); j++) both += inside[j] && masks[m][j];
      clash = (points ? (double)both / (double)points : 0.0) >= 0.5;
    }
    if (clash) continue;
    std::vector<uint8_t> wide(time.size());
    for (size_t j = 0; j < time.size(); j++)
      wide[j] = std::fabs(time[j] - row.epoch) <= 0.75 * row.duration;
    masks.push_back(std::move(wide));
    fbls_single_event e;
    e.epoch_bjd = row.epoch;
    e.duration_hours = 24.0 * row.duration;
    e.snr = row.score;
    e.depth = row.depth;
    e.depth_ppm = 1e6 * row.depth;
    e.points = points;
    output.push_back(e);
  }
  return output;
}

int check_pipeline(const fbls_pipeline_cfg* cfg) {
  if (!cfg || !cfg->durations || cfg->nd == 0) return fail(FBLS_E_ARG, "no durations");
  if (cfg->passes < 1) return fail(FBLS_E_ARG, "passes must be >= 1");
  if (!cfg->bank.ingress || cfg->bank.ns == 0) return fail(FBLS_E_ARG, "no shapes");
  const fbls_search_cfg& s = cfg->search;
  if (!(std::isfinite(s.period_oversample) && s.period_oversample > 0.0))
    return fail(FBLS_E_ARG, "period_oversample must be finite and positive");
  if (!(std::isfinite(s.period_min) && std::isfinite(s.period_max) && 0.0 < s.period_min &&
        s.period_min < s.period_max))
    return fail(FBLS_E_ARG, "require 0 < period_min < period_max");
  return FBLS_OK;
}

// One family handed to the search: its curve, bank durations, detrending
// window (NaN for prepared input) and clipped count.
struct FamInput {
  Curve curve;
  std::vector<size_t> indices;
  double window = std::nan("");
  size_t clipped = 0;
};

template <class T>
void put(T* arr, size_t cap, size_t* count, const T& x) {
  if (arr && *count < cap) arr[*count] = x;
  (*count)++;
}

// pipeline.rs:70-289 over the families
int search_families(fbls_ctx* c, std::vector<FamInput> input, const fbls_pipeline_cfg* cfg,
                    double start, fbls_peak* peaks, size_t peaks_cap, size_t* npeaks,
                    fbls_pass_stats* passes, size_t passes_cap, size_t* npasses,
                    fbls_search_report* report) {
  const size_t nfam = input.size();
  std::vector<std::unique_ptr<Fam>> cur(nfam);
  std::vector<Curve> original(nfam);
  std::vector<double> time_all;
  for (size_t k = 0; k < nfam; k++) {
    std::vector<double> durs;
    for (size_t i : input[k].indices) {
      if (i >= cfg->nd) return fail(FBLS_E_ARG, "duration index %zu out of range", i);
      durs.push_back(cfg->durations[i]);
    }
    if (durs.empty()) return fail(FBLS_E_ARG, "family %zu has no durations", k);
    cur[k] = std::make_unique<Fam>();
    if (int st = build_fam(c, std::move(input[k].curve), durs, input[k].indices, cfg->correlated,
                           cur[k].get()))
      return st;
    original[k] = cur[k]->curve;
    time_all.insert(time_all.end(), cur[k]->curve.t.begin(), cur[k]->curve.t.end());
  }
  // merged_time (pipeline.rs:16-24)
  std::sort(time_all.begin(), time_all.end());
  time_all.erase(std::unique(time_all.begin(), time_all.end()), time_all.end());
  if (time_all.size() < 6) return fail(FBLS_E_DATA, "at least six usable observations are required");
  fbls_search_report dummy;
  std::memset(&dummy, 0, sizeof(dummy));
  fbls_search_report& R = report ? *report : dummy;
  R.nevents = R.nevent_stats = R.period_table_rows = R.nfamilies = R.nbands = 0;
  R.observations = time_all.size();
  R.baseline_days = time_all.back() - time_all.front();
  R.nfam = nfam;
  R.periods_tested = 0;
  R.seconds_preparation = now_s() - start;
  R.seconds_scan = R.seconds_select = R.seconds_evidence = R.seconds_single_events = 0.0;
  const double search_start = now_s();
  std::vector<uint8_t> excluded(time_all.size(), 0);
  std::vector<std::vector<uint8_t>> masks;
  std::vector<fbls_peak> found;
  struct Guard {
    DevBuf b;
    ~Guard() { b.release(); }
  } g;
  DevBuf& d_time_all = g.b;
  if (int st = d_time_all.ensure(time_all.size() * 8)) return st;
  HIP_TRY(hipMemcpy(d_time_all.ptr, time_all.data(), time_all.size() * 8, hipMemcpyHostToDevice));
  for (size_t pass = 0; pass < cfg->passes; pass++) {
    const double t0 = now_s();
    // signal_search of every family (scan.rs:413-473), tables concatenated
    std::vector<fbls_row> table;
    std::vector<size_t> family;
    size_t grid_total = 0;
    for (size_t k = 0; k < nfam; k++) {
      const Fam& f = *cur[k];
      const size_t nd = f.dur.size();
      const double d0 = *std::min_element(f.dur.begin(), f.dur.end());
      const double hi = std::min(cfg->search.period_max, f.dev->t_last);
      const double lo = std::max(cfg->search.period_min, 4.0 * d0 * (1.0 + 1e-9));
      size_t cap = 1;
      if (hi > lo)
        cap = std::max<size_t>(1, fbls_period_grid_count(lo, hi, f.dev->t_last, d0,
                                                         cfg->search.period_oversample));
      const size_t base = table.size();
      table.resize(base + cap);
      std::vector<fbls_scan_stats> bstats(64);
      size_t nrows = 0, nbands = 0, grid_count = 0;
      fbls_bank bank = cfg->bank;
      bank.durations = f.dur.data();
      bank.rho = f.rho.data();
      bank.nd = nd;
      int st = fbls_signal_search(c, f.dev, f.dur.data(), f.rho.data(), nd, &bank, &cfg->search,
                                  table.data() + base, cap, &nrows, bstats.data(), bstats.size(),
                                  &nbands, nullptr, 0, &grid_count);
      if (st) return st;
      table.resize(base + nrows);
      family.resize(base + nrows, k);
      grid_total += grid_count;
      if (pass == 0) R.periods_tested += grid_count;
      fbls_family_pass fp;
      std::memset(&fp, 0, sizeof(fp));
      fp.pass = pass + 1;
      fp.family = k;
      fp.window_days = input[k].window;
      fp.observations = f.curve.t.size();
      fp.clipped = input[k].clipped;
      fp.error_inflation = f.inflation;
      fp.segments = f.segments;
      fp.periods = grid_count;
      fp.nd = nd;
      for (size_t i = 0; i < nd && i < 64; i++) {
        fp.durations_hours[i] = f.dur[i] * 24.0;
        fp.rho[i] = f.rho[i];
      }
      fp.band_first = R.nbands;
      fp.nbands = std::min(nbands, bstats.size());
      for (size_t b = 0; b < fp.nbands; b++) put(R.bands, R.bands_cap, &R.nbands, bstats[b]);
      put(R.families, R.families_cap, &R.nfamilies, fp);
    }
    const double t1 = now_s();
    R.seconds_scan += t1 - t0;
    if (pass == 0)
      for (size_t i = 0; i < table.size(); i++) {
        put(R.period_table, R.period_table_cap, &R.period_table_rows, table[i]);
        if (R.period_table_family && i < R.period_table_cap) R.period_table_family[i] = family[i];
      }
    Selector S;
    S.c = c;
    S.table = table.data();
    S.nrows = table.size();
    S.family = &family;
    S.time_all = &time_all;
    S.excluded = &excluded;
    S.threshold = cfg->bank.threshold;
    S.fams.resize(nfam);
    S.pos.resize(nfam);
    for (size_t k = 0; k < nfam; k++) {
      const Fam& f = *cur[k];
      S.fams[k] = &f;
      S.pos[k].resize(f.t.size());
      for (size_t i = 0; i < f.t.size(); i++) {
        const size_t p =
            std::lower_bound(time_all.begin(), time_all.end(), f.t[i] + f.origin) -
            time_all.begin();
        S.pos[k][i] = std::min(p, time_all.size() - 1);
      }
    }
    S.d_time_all = (const double*)d_time_all.ptr;
    const size_t before = found.size();
    int st = select_peaks(S, cfg->max_candidates - before, cfg->min_events, masks, pass + 1, found);
    if (st) return st;
    R.seconds_select += now_s() - t1;
    if (*npasses < passes_cap && passes) {
      fbls_pass_stats& ps = passes[*npasses];
      ps.pass = pass + 1;
      ps.rows = table.size();
      ps.periods = grid_total;
      ps.accepted = found.size() - before;
      ps.seconds_scan = t1 - t0;
      ps.seconds_select = now_s() - t1;
    }
    (*npasses)++;
    if (found.size() == before) break;
    const fbls_peak lead = found[before];
    if (pass + 1 < cfg->passes) {
      bool small = false;
      for (size_t k = 0; k < nfam; k++) {
        Curve next = masked(cur[k]->curve, lead.period_days, lead.epoch_bjd,
                            lead.duration_hours / 24.0, 0.5);
        auto fam = std::make_unique<Fam>();
        if ((st = build_fam(c, std::move(next), cur[k]->dur, cur[k]->indices, cfg->correlated,
                            fam.get())))
          return st;
        cur[k] = std::move(fam);
        small |= cur[k]->curve.t.size() < 6;
      }
      if (small) break;
    }
  }
  // model comparison of the strongest peaks (pipeline.rs:169-221)
  const double t_evidence = now_s();
  if (cfg->model_comparison && !found.empty()) {
    std::vector<size_t> ord(found.size());
    for (size_t i = 0; i < ord.size(); i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(),
                     [&](size_t a, size_t b) { return found[a].snr > found[b].snr; });
    for (size_t oi = 0; oi < ord.size() && oi < cfg->evidence_candidates; oi++) {
      fbls_peak& p = found[ord[oi]];
      const Fam& base = *cur[p.family];
      Curve cv = original[p.family];
      for (const fbls_peak& o : found) {
        if (o.row == p.row && o.pass == p.pass && o.family == p.family) continue;
        cv = masked(cv, o.period_days, o.epoch_bjd, o.duration_hours / 24.0, 0.5);
      }
      if (cv.t.size() < 6) continue;
      Fam s;
      if (int st = build_fam(c, std::move(cv), base.dur, base.indices, cfg->correlated, &s))
        return st;
      fbls_evidence_row rows[2];
      size_t nr = 0;
      if (int st = fbls_compare(c, s.dev, p.period_days, s.dur.data(), s.rho.data(),
                                s.dur.size(), s.energy, &cfg->evidence, rows, &nr, nullptr))
        return st;
      const fbls_evidence_row* r = nullptr;
      for (size_t q = 0; q < nr && !r; q++)
        if (rows[q].selection_flags & 1) r = &rows[q];
      if (!r && nr) r = &rows[0];
      if (!r) continue;
      auto fin = [](double x) { return std::isfinite(x) ? x : std::nan(""); };
      p.planet_log_odds = fin(r->planet_log_odds);
      p.log_bf_planet_noise = fin(r->log_bf_planet_noise);
      p.log_bf_binary_noise = fin(r->log_bf_binary_noise);
      p.binary_secondary_ratio = fin(r->binary_secondary_ratio);
      p.binary_secondary_phase = fin(r->binary_secondary_phase);
      p.binary_odd_even_ratio = fin(r->binary_odd_even_ratio);
      p.binary_period_days = fin(r->binary_period_days);
      p.ingress_fraction = fin(r->ingress_fraction);
      p.evidence_epoch_bjd = r->epoch;
      p.evidence_duration_hours = 24.0 * r->duration;
      p.model_preferred = r->planet_log_odds >= cfg->planet_log_odds_min;
      p.binary_diagnostics_complete =
          std::isfinite(r->binary_secondary_ratio) && std::isfinite(r->binary_secondary_phase) &&
          std::isfinite(r->binary_odd_even_ratio) && std::isfinite(r->binary_period_days);
    }
    R.seconds_evidence = now_s() - t_evidence;
  }
  // single events (pipeline.rs:222-238) over the families as the passes left them
  const double t_events = now_s();
  if (cfg->single_events) {
    const size_t capacity = 131072;
    std::vector<fbls_row> all;
    std::vector<fbls_row> rows(capacity);
    for (size_t k = 0; k < nfam; k++) {
      const Fam& f = *cur[k];
      size_t nr = 0;
      fbls_event_stats es;
      if (int st = fbls_single_event_search(
              c, f.dev, f.dur.data(), f.rho.data(), f.dur.size(), cfg->bank.ingress,
              cfg->bank.ns, cfg->bank.layers, cfg->bank.threshold, cfg->bank.phase_oversample,
              capacity, rows.data(), &nr, &es))
        return st;
      put(R.event_stats, R.event_stats_cap, &R.nevent_stats, es);
      std::vector<fbls_row> table(rows.begin(), rows.begin() + nr);
      for (const fbls_single_event& e :
           select_events(table, time_all, cfg->bank.threshold, cfg->max_candidates, masks)) {
        fbls_row r;
        r.period = INFINITY;
        r.epoch = e.epoch_bjd;
        r.duration = e.duration_hours / 24.0;
        r.score = e.snr;
        r.depth = e.depth;
        all.push_back(r);
      }
    }
    std::stable_sort(all.begin(), all.end(),
                     [](const fbls_row& a, const fbls_row& b) { return a.score > b.score; });
    for (const fbls_single_event& e :
         select_events(all, time_all, cfg->bank.threshold, cfg->max_candidates, masks))
      put(R.events, R.events_cap, &R.nevents, e);
  }
  R.seconds_single_events = now_s() - t_events;
  R.seconds_search = now_s() - search_start;
  *npeaks = found.size();
  if (peaks) std::memcpy(peaks, found.data(), std::min(found.size(), peaks_cap) * sizeof(fbls_peak));
  if (found.size() > peaks_cap)
    return fail(FBLS_E_ARG, "peaks capacity %zu < %zu", peaks_cap, found.size());
  return FBLS_OK;
}

}  // namespace

extern "C" int fbls_search_prepared(fbls_ctx* c, const double* time, const double* flux,
                                    const double* err, size_t n,
                                    const fbls_pipeline_cfg* cfg, fbls_peak* peaks,
                                    size_t peaks_cap, size_t* npeaks,
                                    fbls_pass_stats* passes, size_t passes_cap,
                                    size_t* npasses, fbls_search_report* report) {
  return fbls::host::guard([&]() -> int {
    const double start = now_s();
    if (!c || (n && (!time || !flux || !err)) || !npeaks || !npasses)
      return fail(FBLS_E_ARG, "null argument");
    *npeaks = 0;
    *npasses = 0;
    if (int st = check_pipeline(cfg)) return st;
    // pipeline.rs:47-61
    std::vector<size_t> order;
    for (size_t i = 0; i < n; i++)
      if (std::isfinite(time[i]) && std::isfinite(flux[i]) && std::isfinite(err[i]) && err[i] > 0.0)
        order.push_back(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return time[a] < time[b]; });
    Curve curve;
    for (size_t i : order) {
      curve.t.push_back(time[i]);
      curve.f.push_back(flux[i]);
      curve.s.push_back(err[i]);
    }
    HIP_TRY(hipSetDevice(c->device));
    std::vector<size_t> all(cfg->nd);
    for (size_t i = 0; i < cfg->nd; i++) all[i] = i;
    std::vector<FamInput> input(1);
    input[0].curve = std::move(curve);
    input[0].indices = std::move(all);
    return search_families(c, std::move(input), cfg, start, peaks, peaks_cap, npeaks, passes,
                           passes_cap, npasses, report);
  });
}

extern "C" int fbls_search_transits(fbls_ctx* c, const double* time, const double* flux,
                                    const double* err, const int64_t* group, size_t n,
                                    const fbls_prep_cfg* prep, const fbls_pipeline_cfg* cfg,
                                    fbls_peak* peaks, size_t peaks_cap, size_t* npeaks,
                                    fbls_pass_stats* passes, size_t passes_cap,
                                    size_t* npasses, fbls_search_report* report) {
  return fbls::host::guard([&]() -> int {
    const double start = now_s();
    if (!c || !prep || (n && (!time || !flux)) || !npeaks || !npasses)
      return fail(FBLS_E_ARG, "null argument");
    *npeaks = 0;
    *npasses = 0;
    if (int st = check_pipeline(cfg)) return st;
    HIP_TRY(hipSetDevice(c->device));
    std::vector<fbls::Family> fams;
    std::string msg;
    auto flatten = [c](const std::vector<double>& t, const std::vector<double>& rel, double w,
                       double gap, double cst, int it, std::vector<double>* out) {
      return biweight_gpu(c, t, rel, w, gap, cst, it, out);
    };
    int st = fbls::prepare_families(time, flux, err, group, n, cfg->durations, cfg->nd, prep,
                                    flatten, &fams, &msg);
    if (st) return msg.empty() ? st : fail(st, "%s", msg.c_str());
    std::vector<FamInput> input;
    for (fbls::Family& f : fams) {
      FamInput in;
      in.curve.t = std::move(f.time);
      in.curve.f = std::move(f.flux);
      in.curve.s = std::move(f.sigma);
      in.indices = std::move(f.indices);
      in.window = f.window;
      in.clipped = f.clipped;
      input.push_back(std::move(in));
    }
    return search_families(c, std::move(input), cfg, start, peaks, peaks_cap, npeaks, passes,
                           passes_cap, npasses, report);
  });
}

// The command-line host above the C ABI: `fbls search` (isakruas/fbls
// src/main.rs:137-229) and `fbls batch` (src/batch.rs:1-109) with NPZ input,
// the schema-3 JSON report (src/pipeline.rs:239-289) and the period-table
// CSV.  The fbls executable is a thin main around fbls_cli.
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include <sys/stat.h>

#include "../../include/fbls.h"
#include "config.h"
#include "host_internal.h"
#include "json.h"
#include "npz.h"
#include "prep.h"

namespace fbls {
namespace {

using json::Value;

struct Failure {
  std::string message;
};

[[noreturn]] void raise(const std::string& m) { throw Failure{m}; }

void check(int st) {
  if (st) raise(fbls_last_error());
}

Value peak_json(const fbls_peak& p) {
  Value v = Value::object();
  v["period_days"] = Value::number(p.period_days);
  v["epoch_bjd"] = Value::number(p.epoch_bjd);
  v["duration_hours"] = Value::number(p.duration_hours);
  v["snr"] = Value::number(p.snr);
  v["snr_folded"] = Value::number(p.snr_folded);
  v["snr_without_strongest_event"] = Value::number(p.snr_without_strongest_event);
  v["depth"] = Value::number(p.depth);
  v["depth_ppm"] = Value::number(p.depth_ppm);
  v["events"] = Value::unsigned_(p.events);
  v["family"] = Value::unsigned_(p.family);
  v["pass"] = Value::unsigned_(p.pass);
  v["planet_log_odds"] = Value::optional(p.planet_log_odds);
  v["log_bf_planet_noise"] = Value::optional(p.log_bf_planet_noise);
  v["log_bf_binary_noise"] = Value::optional(p.log_bf_binary_noise);
  v["binary_secondary_ratio"] = Value::optional(p.binary_secondary_ratio);
  v["binary_secondary_phase"] = Value::optional(p.binary_secondary_phase);
  v["binary_odd_even_ratio"] = Value::optional(p.binary_odd_even_ratio);
  v["binary_period_days"] = Value::optional(p.binary_period_days);
  v["ingress_fraction"] = Value::optional(p.ingress_fraction);
  v["evidence_epoch_bjd"] = Value::optional(p.evidence_epoch_bjd);
  v["evidence_duration_hours"] = Value::optional(p.evidence_duration_hours);
  v["model_preferred"] = Value::boolean(p.model_preferred);
  v["binary_diagnostics_complete"] = Value::boolean(p.binary_diagnostics_complete);
  return v;
}

Value row_json(const fbls_row& r) {
  Value v = Value::object();
  v["period"] = Value::number(r.period);
  v["epoch"] = Value::number(r.epoch);
  v["duration"] = Value::number(r.duration);
  v["score"] = Value::number(r.score);
  v["depth"] = Value::number(r.depth);
  return v;
}

Value band_json(const fbls_scan_stats& s) {
  Value v = Value::object();
  v["periods"] = Value::unsigned_(s.periods);
  v["prescan_survivors"] = Value::unsigned_(s.prescan_survivors);
  v["phase_bins"] = Value::unsigned_(s.phase_bins);
  v["batches"] = Value::unsigned_(s.batches);
  v["templates_scanned"] = Value::unsigned_(s.templates_scanned);
  v["exact_tile_retries"] = Value::unsigned_(s.exact_tile_retries);
  v["seconds"] = Value::number(s.seconds);
  v["seconds_fold"] = Value::number(s.seconds_fold);
  v["seconds_cumulative"] = Value::number(s.seconds_cumulative);
  v["seconds_scan"] = Value::number(s.seconds_scan);
  return v;
}

// pipeline.rs:241-259: model-preferred peaks first, by planet evidence
double model_key(const fbls_peak& p) {
  double e = p.model_preferred ? p.log_bf_planet_noise : NAN;
  if (std::isnan(e)) e = p.planet_log_odds;
  return std::isnan(e) ? INFINITY : -e;
}

// An upper bound on the rows of the first pass: every family's grid is at
// most the grid of the shortest duration over the whole time span.
size_t table_bound(const SearchConfig& cfg, const io::Observations& obs, bool prepared) {
  if (obs.time.empty()) return 0;
  double t0 = INFINITY, t1 = -INFINITY;
  for (double t : obs.time)
    if (std::isfinite(t)) {
      t0 = std::min(t0, t);
      t1 = std::max(t1, t);
    }
  if (!(t1 > t0)) return 0;
  const std::vector<double> d = cfg.durations_days();
  const double span = t1 - t0;
  const double lo = std::max(cfg.period_min, 4.0 * d[0] * (1.0 + 1e-9));
  const double hi = std::min(cfg.period_max, span);
  if (!(hi > lo)) return 0;
  const size_t count = fbls_period_grid_count(lo, hi, span, d[0], cfg.period_oversample);
  const size_t fams =
      prepared ? 1
               : detrending_windows(d.data(), d.size(), cfg.window_days, cfg.window_factor,
                                    cfg.window_max_days)
                     .size();
  return count * std::max<size_t>(fams, 1);
}

// search_transits / search_prepared and the report (pipeline.rs:239-289)
Value run_search(fbls_ctx* ctx, const SearchConfig& cfg, const io::Observations& obs,
                 bool prepared, int device) {
  CabiConfig cc(cfg);
  const size_t n = obs.time.size();
  if (prepared && !obs.has_err) raise("prepared input requires flux_err");
  const size_t nd = cc.durations.size();
  std::vector<fbls_peak> peaks(std::max<size_t>(cfg.max_candidates, 1));
  std::vector<fbls_pass_stats> passes(std::max<size_t>(cfg.passes, 1));
  const size_t table_cap = table_bound(cfg, obs, prepared);
  std::vector<fbls_single_event> events(std::max<size_t>(cfg.max_candidates, 1));
  std::vector<fbls_event_stats> estats(nd);
  std::vector<fbls_row> table(std::max<size_t>(table_cap, 1));
  std::vector<uint64_t> table_family(table.size());
  std::vector<fbls_family_pass> fams(nd * passes.size());
  std::vector<fbls_scan_stats> bands(fams.size() * 64);
  fbls_search_report R;
  std::memset(&R, 0, sizeof(R));
  R.events = events.data();
  R.events_cap = events.size();
  R.event_stats = estats.data();
  R.event_stats_cap = estats.size();
  R.period_table = table.data();
  R.period_table_cap = table_cap;
  R.period_table_family = table_family.data();
  R.families = fams.data();
  R.families_cap = fams.size();
  R.bands = bands.data();
  R.bands_cap = bands.size();
  size_t npeaks = 0, npasses = 0;
  if (prepared) {
    check(fbls_search_prepared(ctx, obs.time.data(), obs.flux.data(), obs.err.data(), n,
                               &cc.pipeline, peaks.data(), peaks.size(), &npeaks,
                               passes.data(), passes.size(), &npasses, &R));
  } else {
    check(fbls_search_transits(ctx, obs.time.data(), obs.flux.data(),
                               obs.has_err ? obs.err.data() : nullptr, obs.group.data(), n,
                               &cc.prep, &cc.pipeline, peaks.data(), peaks.size(), &npeaks,
                               passes.data(), passes.size(), &npasses, &R));
  }
  if (R.period_table_rows > R.period_table_cap || R.nfamilies > R.families_cap ||
      R.nbands > R.bands_cap || R.nevent_stats > R.event_stats_cap)
    raise("report capacity exceeded");
  peaks.resize(std::min(npeaks, peaks.size()));
  // pass statistics with their families (pipeline.rs:112-143)
  Value pass_list = Value::array();
  for (size_t p = 0; p < std::min(npasses, passes.size()); p++) {
    Value ps = Value::object();
    ps["pass"] = Value::unsigned_(passes[p].pass);
    ps["rows"] = Value::unsigned_(passes[p].rows);
    ps["accepted"] = Value::unsigned_(passes[p].accepted);
    Value fl = Value::array();
    for (size_t k = 0; k < R.nfamilies; k++) {
      const fbls_family_pass& f = fams[k];
      if (f.pass != passes[p].pass) continue;
      Value fv = Value::object();
      fv["family"] = Value::unsigned_(f.family);
      fv["window_days"] = Value::number(f.window_days);
      fv["observations"] = Value::unsigned_(f.observations);
      fv["clipped"] = Value::unsigned_(f.clipped);
      fv["error_inflation"] = Value::number(f.error_inflation);
      fv["segments"] = Value::unsigned_(f.segments);
      const size_t m = std::min<size_t>(f.nd, 64);
      fv["durations_hours"] = json::numbers(f.durations_hours, m);
      fv["periods"] = Value::unsigned_(f.periods);
      fv["rho"] = json::numbers(f.rho, m);
      Value bl = Value::array();
      for (size_t b = 0; b < f.nbands; b++) bl.push(band_json(bands[f.band_first + b]));
      fv["bands"] = bl;
      fl.push(fv);
    }
    ps["families"] = fl;
    pass_list.push(ps);
  }
  Value event_scan = Value::object();
  for (size_t k = 0; k < R.nevent_stats; k++) {
    Value e = Value::object();
    e["templates_scanned"] = Value::unsigned_(estats[k].templates_scanned);
    e["threshold_survivors"] = Value::unsigned_(estats[k].threshold_survivors);
    e["overflow"] = Value::boolean(estats[k].overflow);
    event_scan[std::to_string(k)] = e;
  }
  std::vector<fbls_peak> by_signal = peaks;
  std::stable_sort(by_signal.begin(), by_signal.end(),
                   [](const fbls_peak& a, const fbls_peak& b) { return a.snr > b.snr; });
  std::vector<fbls_peak> by_model = peaks;
  std::stable_sort(by_model.begin(), by_model.end(), [](const fbls_peak& a, const fbls_peak& b) {
    const int ga = a.model_preferred ? 0 : 1, gb = b.model_preferred ? 0 : 1;
    if (ga != gb) return ga < gb;
    return model_key(a) < model_key(b);
  });
  if (!cfg.model_comparison) by_model = by_signal;
  const size_t lim = cfg.max_candidates;
  Value cands = Value::array(), sig = Value::array(), evs = Value::array(), tab = Value::array();
  for (size_t i = 0; i < by_model.size() && i < lim; i++) cands.push(peak_json(by_model[i]));
  for (size_t i = 0; i < by_signal.size() && i < lim; i++) sig.push(peak_json(by_signal[i]));
  for (size_t i = 0; i < R.nevents && i < lim; i++) {
    const fbls_single_event& e = events[i];
    Value v = Value::object();
    v["epoch_bjd"] = Value::number(e.epoch_bjd);
    v["duration_hours"] = Value::number(e.duration_hours);
    v["snr"] = Value::number(e.snr);
    v["depth"] = Value::number(e.depth);
    v["depth_ppm"] = Value::number(e.depth_ppm);
    v["points"] = Value::unsigned_(e.points);
    evs.push(v);
  }
  for (size_t i = 0; i < R.period_table_rows; i++) tab.push(row_json(table[i]));
  Value st = Value::object();
  st["device"] = Value::string("hip:" + std::to_string(device));
  st["observations"] = Value::unsigned_(R.observations);
  st["baseline_days"] = Value::number(R.baseline_days);
  st["periodic_period_max_days"] = Value::number(std::fmin(cfg.period_max, R.baseline_days));
  st["periods_tested"] = Value::unsigned_(R.periods_tested);
  st["qualifying_periods"] = Value::unsigned_(R.period_table_rows);
  st["resolved_peaks"] = Value::unsigned_(peaks.size());
  st["single_events"] = Value::unsigned_(std::min<size_t>(R.nevents, lim));
  st["seconds_preparation"] = Value::number(R.seconds_preparation);
  st["seconds_search"] = Value::number(R.seconds_search);
  Value ev = Value::object();
  ev["seconds_search"] = Value::number(R.seconds_evidence);
  st["evidence"] = ev;
  st["seconds_scan"] = Value::number(R.seconds_scan);
  st["seconds_select"] = Value::number(R.seconds_select);
  st["seconds_single_events"] = Value::number(R.seconds_single_events);
  st["seconds_total"] = Value::number(R.seconds_preparation + R.seconds_search);
  st["passes"] = pass_list;
  st["single_event_scan"] = event_scan;
  Value rep = Value::object();
  rep["schema_version"] = Value::unsigned_(3);
  rep["software_version"] = Value::string("0.1.0");
  rep["config"] = cfg.to_json();
  rep["statistics"] = st;
  rep["candidates"] = cands;
  rep["signal_candidates"] = sig;
  rep["single_events"] = evs;
  rep["period_table"] = tab;
  return rep;
}

double parse_f64(const std::string& s) {
  char* end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (s.empty() || *end) raise("invalid float literal");
  return v;
}

size_t parse_usize(const std::string& s) {
  char* end = nullptr;
  const unsigned long long v = std::strtoull(s.c_str(), &end, 10);
  if (s.empty() || *end || s[0] == '-' || s[0] == '+') raise("invalid digit found in string");
  return (size_t)v;
}

// std::fs::create_dir_all
void create_dirs(const std::string& path) {
  for (size_t p = 1; p <= path.size(); p++)
    if (p == path.size() || path[p] == '/') {
      const std::string prefix = path.substr(0, p);
      struct stat sb;
      if (stat(prefix.c_str(), &sb) == 0) {
        if (!S_ISDIR(sb.st_mode)) raise(prefix + ": not a directory");
      } else if (mkdir(prefix.c_str(), 0777) != 0 && errno != EEXIST) {
        raise(prefix + ": cannot create directory");
      }
    }
}

void write_file(const std::string& path, const std::string& text) {
  std::ofstream f(path, std::ios::binary);
  if (!f) raise(path + ": cannot open for writing");
  f << text;
  if (!f) raise(path + ": write failed");
}

struct Ctx {
  fbls_ctx* c = nullptr;
  explicit Ctx(int device) { check(fbls_ctx_create(device, &c)); }
  ~Ctx() { fbls_ctx_destroy(c); }
};

// main.rs:137-229
void search_cmd(const std::vector<std::string>& args) {
  std::vector<std::string> inputs;
  std::string output = "-";
  std::string table_path;
  bool prepared = false, save_table = false;
  int device = 0;
  SearchConfig cfg;
  auto next = [&](size_t i) -> const std::string& {
    if (i + 1 >= args.size()) raise("missing value for " + args[i]);
    return args[i + 1];
  };
  for (size_t i = 0; i < args.size(); i++) {
    const std::string& a = args[i];
    if (a == "--input") {
      while (i + 1 < args.size() && args[i + 1].rfind("--", 0) != 0) inputs.push_back(args[++i]);
    } else if (a == "--output") {
      output = next(i), i++;
    } else if (a == "--save-table") {
      table_path = next(i), i++, save_table = true;
    } else if (a == "--prepared") {
      prepared = true;
    } else if (a == "--device") {
      device = (int)parse_usize(next(i)), i++;
    } else if (a == "--pmin") {
      cfg.period_min = parse_f64(next(i)), i++;
    } else if (a == "--pmax") {
      cfg.period_max = parse_f64(next(i)), i++;
    } else if (a == "--prescan") {
      cfg.prescan_factor = parse_usize(next(i)), i++;
    } else if (a == "--threshold") {
      cfg.threshold = parse_f64(next(i)), i++;
    } else {
      raise("unknown argument " + a);
    }
  }
  if (inputs.empty()) raise("--input is required");
  const std::string bad = cfg.validate();
  if (!bad.empty()) raise(bad);
  io::Observations obs;
  std::string error;
  if (!io::read_inputs(inputs, &obs, &error)) raise(error);
  Ctx ctx(device);
  Value rep = run_search(ctx.c, cfg, obs, prepared, device);
  Value files = Value::array();
  for (const std::string& p : inputs) files.push(Value::string(p));
  rep["input_files"] = files;
  rep["input_mode"] = Value::string(prepared ? "prepared" : "raw");
  if (save_table) {
    std::string text = "period_days,epoch_bjd,duration_days,snr,depth\n";
    for (const Value& r : rep["period_table"].a) {
      text += json::pretty(r.o.at("period")) + "," + json::pretty(r.o.at("epoch")) + "," +
              json::pretty(r.o.at("duration")) + "," + json::pretty(r.o.at("score")) + "," +
              json::pretty(r.o.at("depth")) + "\n";
    }
    write_file(table_path, text);
  }
  rep.o.erase("period_table");
  const std::string text = json::pretty(rep);
  if (output == "-") {
    std::fwrite(text.data(), 1, text.size(), stdout);
    std::fputc('\n', stdout);
    std::fflush(stdout);
  } else {
    write_file(output, text);
  }
}

struct Job {
  std::string name;
  bool has_pmax = false;
  double pmax = 0.0;
  std::vector<std::string> inputs;
};

// batch.rs:16-34: `name pmax file1 file2 ...`, pmax may be `-`
std::vector<Job> parse_list(const std::string& path) {
  std::ifstream f(path);
  if (!f) raise(path + ": No such file or directory (os error 2)");
  std::vector<Job> jobs;
  std::string line;
  while (std::getline(f, line)) {
    std::vector<std::string> fields;
    size_t p = 0;
    while (p < line.size()) {
      while (p < line.size() && std::isspace((unsigned char)line[p])) p++;
      size_t q = p;
      while (q < line.size() && !std::isspace((unsigned char)line[q])) q++;
      if (q > p) fields.push_back(line.substr(p, q - p));
      p = q;
    }
    if (fields.size() < 3 || fields[0][0] == '#') continue;
    Job j;
    j.name = fields[0];
    if (fields[1] != "-") {
      j.has_pmax = true;
      j.pmax = parse_f64(fields[1]);
    }
    j.inputs.assign(fields.begin() + 2, fields.end());
    jobs.push_back(std::move(j));
  }
  return jobs;
}

// batch.rs:36-109
void batch_cmd(const std::vector<std::string>& args) {
  auto get = [&](const std::string& flag, std::string* out) {
    for (size_t i = 0; i + 1 < args.size(); i++)
      if (args[i] == flag) {
        *out = args[i + 1];
        return true;
      }
    return false;
  };
  std::string list, output_dir, w, d, p;
  if (!get("--list", &list)) raise("--list FILE is required");
  if (!get("--output-dir", &output_dir)) raise("--output-dir DIR is required");
  const size_t workers = get("--workers", &w) ? parse_usize(w) : 2;
  std::vector<int> devices;
  if (get("--devices", &d)) {
    size_t s = 0;
    while (s <= d.size()) {
      size_t c = d.find(',', s);
      if (c == std::string::npos) c = d.size();
      devices.push_back((int)parse_usize(d.substr(s, c - s)));
      s = c + 1;
    }
  } else {
    devices.push_back(0);
  }
  const double pmax_default = get("--pmax", &p) ? parse_f64(p) : 20.0;
  create_dirs(output_dir);
  std::vector<Job> jobs = parse_list(list);
  const size_t total = jobs.size();
  std::mutex lock;
  size_t done = 0;
  const auto start = std::chrono::steady_clock::now();
  std::vector<std::thread> threads;
  std::vector<std::string> thread_errors;
  std::string spawn_error;
  // batch.rs:76-101: a star's error goes to <name>.error and the batch goes on
  auto job_error = [&](const std::string& name, const std::string& msg) {
    std::lock_guard<std::mutex> g(lock);
    std::fprintf(stderr, "%s: %s\n", name.c_str(), msg.c_str());
    std::ofstream(output_dir + "/" + name + ".error") << msg;
  };
  // a worker that cannot be started leaves the list to the ones that were
  // (every thread is joined before an error is raised)
  try {
    for (int device : devices)
      for (size_t k = 0; k < workers; k++)
      threads.emplace_back([&, device]() {
        try {
          Ctx ctx(device);
          // a stream capture in one thread fails allocations in the others
          if (workers * devices.size() > 1) check(fbls_ctx_set_graphs(ctx.c, 0, nullptr));
          for (;;) {
            Job job;
            {
              std::lock_guard<std::mutex> g(lock);
              if (jobs.empty()) return;
              job = std::move(jobs.back());
              jobs.pop_back();
            }
            SearchConfig cfg;
            cfg.period_max = job.has_pmax ? job.pmax : pmax_default;
            try {
              io::Observations obs;
              std::string error;
              if (!io::read_inputs(job.inputs, &obs, &error)) raise(error);
              Value rep = run_search(ctx.c, cfg, obs, false, device);
              Value files = Value::array();
              for (const std::string& f : job.inputs) files.push(Value::string(f));
              rep["input_files"] = files;
              rep["input_mode"] = Value::string("raw");
              write_file(output_dir + "/" + job.name + ".json", json::pretty(rep));
              const double secs = rep["statistics"]["seconds_total"].f;
              std::lock_guard<std::mutex> g(lock);
              done++;
              std::fprintf(stderr, "[%zu/%zu] %s hip:%d %.1fs\n", done, total, job.name.c_str(),
                           device, secs);
            } catch (const Failure& e) {
              job_error(job.name, e.message);
            } catch (const std::exception& e) {
              // e.g. std::bad_alloc from an archive that declares a huge array
              job_error(job.name, e.what());
            }
          }
        } catch (const Failure& e) {
          std::lock_guard<std::mutex> g(lock);
          thread_errors.push_back(e.message);
        } catch (const std::exception& e) {
          std::lock_guard<std::mutex> g(lock);
          thread_errors.push_back(e.what());
        }
      });
  } catch (const std::exception& e) {
    spawn_error = e.what();
  }
  for (std::thread& t : threads) t.join();
  if (!thread_errors.empty()) raise(thread_errors[0]);
  if (!spawn_error.empty()) {
    if (threads.empty()) raise("cannot start a batch worker: " + spawn_error);
    std::fprintf(stderr, "fbls: started %zu of %zu workers: %s\n", threads.size(),
                 workers * devices.size(), spawn_error.c_str());
  }
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
  std::fprintf(stderr, "%zu stars in %.0f s\n", total, secs);
}

}  // namespace
}  // namespace fbls

extern "C" int fbls_cli(int argc, const char* const* argv) {
  std::vector<std::string> args;
  for (int i = 1; i < argc; i++) args.emplace_back(argv[i]);
  const std::string cmd = argc > 0 && argv[0] ? argv[0] : "";
  try {
    if (cmd == "search") {
      fbls::search_cmd(args);
    } else if (cmd == "batch") {
      fbls::batch_cmd(args);
    } else {
      fbls::raise("usage: fbls search | batch");
    }
  } catch (const fbls::Failure& e) {
    std::fprintf(stderr, "fbls: %s\n", e.message.c_str());
    return 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "fbls: %s\n", e.what());
    return 1;
  }
  return 0;
}

extern "C" size_t fbls_format_f64(double x, char* buf, size_t cap) {
  try {
    const std::string s = fbls::json::format_double(x);
    if (buf && cap) {
      const size_t m = std::min(cap - 1, s.size());
      std::memcpy(buf, s.data(), m);
      buf[m] = 0;
    }
    return s.size();
  } catch (const std::exception&) {  // no exception crosses the C boundary
    if (buf && cap) buf[0] = 0;
    return 0;
  }
}

extern "C" int fbls_json_reformat(const char* text, char* out, size_t cap, size_t* len) {
  return fbls::host::guard([&]() -> int {
    if (!text || !len) return fbls::host::fail(FBLS_E_ARG, "null argument");
    fbls::json::Value v;
    std::string err;
    if (!fbls::json::parse(text, &v, &err)) return fbls::host::fail(FBLS_E_DATA, "%s", err.c_str());
    const std::string s = fbls::json::pretty(v);
    *len = s.size();
    if (out && cap) {
      const size_t m = std::min(cap - 1, s.size());
      std::memcpy(out, s.data(), m);
      out[m] = 0;
    }
    return 0;
  });
}

extern "C" int fbls_read_inputs(const char* const* paths, size_t npaths, double* time,
                                double* flux, double* err, int64_t* group, size_t cap,
                                size_t* n, int* has_err) {
  return fbls::host::guard([&]() -> int {
    if ((npaths && !paths) || !n || !has_err) return fbls::host::fail(FBLS_E_ARG, "null argument");
    std::vector<std::string> ps;
    for (size_t i = 0; i < npaths; i++) ps.emplace_back(paths[i]);
    fbls::io::Observations obs;
    std::string error;
    if (!fbls::io::read_inputs(ps, &obs, &error))
      return fbls::host::fail(FBLS_E_DATA, "%s", error.c_str());
    *n = obs.time.size();
    *has_err = obs.has_err;
    const size_t m = std::min(cap, obs.time.size());
    if (time) std::memcpy(time, obs.time.data(), m * 8);
    if (flux) std::memcpy(flux, obs.flux.data(), m * 8);
    if (err && obs.has_err) std::memcpy(err, obs.err.data(), m * 8);
    if (group) std::memcpy(group, obs.group.data(), m * 8);
    return 0;
  });
}

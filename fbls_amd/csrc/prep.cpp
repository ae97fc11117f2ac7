// Host-side input preparation of the FBLS hot path (see prep.h).
// Floating-point operations keep the reference's order and rounding; the
// library is compiled with -ffp-contract=off so nothing is fused.
#include "prep.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <map>
#include <vector>

namespace fbls {

Fixed quantize(const double* x, size_t n, int32_t* q) {
  Fixed f;
  double m = 0.0, sum = 0.0;
  for (size_t i = 0; i < n; i++) {
    m = std::max(m, std::fabs(x[i]));
    sum += std::fabs(x[i]);
  }
  if (!(m > 0.0)) {
    for (size_t i = 0; i < n; i++) q[i] = 0;
    return f;
  }
  int ex = 0;
  f.block = m > 16.0 * (sum / (double)n);
  if (!f.block) {
    std::frexp(m, &ex);  // m = g * 2^ex, g in [0.5, 1)
    f.e = 30 - ex;
    if (f.e >= -1022 && f.e <= 1023) {
      // x * 2^e with 2^e a normal double is ldexp(x, e) exactly (one
      // correctly rounded product, no libm call per value)
      const double p2 = std::ldexp(1.0, f.e);
      for (size_t i = 0; i < n; i++) q[i] = (int32_t)std::nearbyint(x[i] * p2);
    } else {
      for (size_t i = 0; i < n; i++) q[i] = (int32_t)std::nearbyint(std::ldexp(x[i], f.e));
    }
    return f;
  }
  // sum |x| * 2^e < 2^52; each value rounds to an integer r (|r| < 2^52),
  // then to 26 significant bits when |r| >= 2^26 (mantissa in [2^25, 2^26))
  std::frexp(sum, &ex);
  f.e = 52 - ex;
  const bool normal = f.e >= -1022 && f.e <= 1023;
  const double p2 = normal ? std::ldexp(1.0, f.e) : 0.0;
  for (size_t i = 0; i < n; i++) {
    const double r = std::nearbyint(normal ? x[i] * p2 : std::ldexp(x[i], f.e));
    int64_t mant = (int64_t)r;
    int shift = 0;
    if (std::fabs(r) >= 67108864.0) {  // 2^26
      int k = 0;
      std::frexp(std::fabs(r), &k);  // |r| = g * 2^k, k >= 27
      shift = k - 26;
      mant = (int64_t)std::nearbyint(std::ldexp(r, -shift));
      if (mant == 67108864 || mant == -67108864) {  // rounded up to 2^26
        mant /= 2;
        shift++;
      }
    }
    q[i] = (int32_t)(mant * 32 + shift);
  }
  return f;
}

size_t period_grid_count(double minimum, double maximum, double baseline,
                         double minimum_duration, double oversample) {
  // grid.rs:12-13
  const double log_step = std::log1p(minimum_duration / (oversample * baseline));
  const double steps = std::ceil(std::log(maximum / minimum) / log_step);
  // not a number (invalid arguments): no grid; beyond 2^62: saturated (the
  // scan entry points refuse grids above 2^31 periods)
  if (!(steps >= 0.0)) return 0;
  if (steps >= 4611686018427387904.0) return (size_t)1 << 62;
  return (size_t)steps + 1;
}

void period_grid(double minimum, double maximum, double* out, size_t count) {
  period_grid_range(minimum, maximum, out, count, 0, count);
}

void period_grid_range(double minimum, double maximum, double* out, size_t count, size_t i0,
                       size_t i1) {
  // grid.rs:14-26
  const double a = std::log(minimum), b = std::log(maximum);
  for (size_t i = i0; i < i1; i++)
    out[i] = std::exp(count == 1 ? a : a + (b - a) * (double)i / (double)(count - 1));
  if (i0 == 0 && i1 > 0) out[0] = minimum;
  if (i0 < count && i1 == count) out[count - 1] = maximum;
}

GridFormula::GridFormula(double mn, double mx, size_t n)
    : minimum(mn), maximum(mx), a(std::log(mn)), b(std::log(mx)), count(n) {}

double GridFormula::operator()(size_t i) const {
  // period_grid: the endpoints are exact, the rest exp of the interpolation
  if (i + 1 == count) return maximum;
  if (i == 0) return minimum;
  return std::exp(a + (b - a) * (double)i / (double)(count - 1));
}

size_t durations_days(const double* hours, size_t n, double* out) {
  std::vector<double> d(hours, hours + n);
  for (double& x : d) x /= 24.0;
  std::sort(d.begin(), d.end());
  d.erase(std::unique(d.begin(), d.end()), d.end());
  std::copy(d.begin(), d.end(), out);
  return d.size();
}

std::vector<fbls_band> plan_bands(const double* grid, size_t count,
                                  const double* sd, size_t nd,
                                  uint32_t phase_oversample,
                                  uint32_t prescan_factor) {
  std::vector<fbls_band> out;
  if (count == 0 || nd == 0) return out;
  double lo = grid[0];
  const double last = grid[count - 1];
  size_t i = 0;
  while (lo <= last) {
    const double hi = lo * 2.0;
    // the grid ascends, so the band's periods [lo, hi) are contiguous
    while (i < count && grid[i] < lo) i++;
    size_t j = i;
    while (j < count && grid[j] < hi) j++;
    size_t chosen = 0;
    while (chosen < nd && sd[chosen] < lo / 4.0) chosen++;
    if (j > i && chosen) {
      fbls_band b;
      b.lo = lo;
      b.first = i;
      b.count = j - i;
      b.nd = chosen;
      // scan.rs:377-379; the durations ascend so the minimum is sd[0]
      const double nbf = std::ceil((double)phase_oversample * grid[j - 1] / sd[0]);
      b.nb = nbf < 8.0 ? 8u : (nbf > 4e9 ? 0xffffffffu : (uint32_t)nbf);
      b.nb_coarse = (prescan_factor > 1 && b.nb / prescan_factor >= 8)
                        ? b.nb / prescan_factor
                        : 0u;
      out.push_back(b);
    }
    i = j;
    lo = hi;
  }
  return out;
}

std::vector<fbls_band> plan_bands(const GridFormula& grid, const double* sd, size_t nd,
                                  uint32_t phase_oversample, uint32_t prescan_factor) {
  // the same bands as plan_bands over the materialised grid: the grid
  // ascends, so the first index at or above a bound is a binary search
  std::vector<fbls_band> out;
  const size_t count = grid.count;
  if (count == 0 || nd == 0) return out;
  auto first_at_least = [&](size_t from, double bound) {
    size_t lo = from, hi = count;
    while (lo < hi) {
      const size_t mid = lo + (hi - lo) / 2;
      if (grid(mid) < bound)
        lo = mid + 1;
      else
        hi = mid;
    }
    return lo;
  };
  double lo = grid(0);
  const double last = grid(count - 1);
  size_t i = 0;
  while (lo <= last) {
    const double hi = lo * 2.0;
    i = first_at_least(i, lo);
    const size_t j = first_at_least(i, hi);
    size_t chosen = 0;
    while (chosen < nd && sd[chosen] < lo / 4.0) chosen++;
    if (j > i && chosen) {
      fbls_band b;
      b.lo = lo;
      b.first = i;
      b.count = j - i;
      b.nd = chosen;
      const double nbf = std::ceil((double)phase_oversample * grid(j - 1) / sd[0]);
      b.nb = nbf < 8.0 ? 8u : (nbf > 4e9 ? 0xffffffffu : (uint32_t)nbf);
      b.nb_coarse = (prescan_factor > 1 && b.nb / prescan_factor >= 8)
                        ? b.nb / prescan_factor
                        : 0u;
      out.push_back(b);
    }
    i = j;
    lo = hi;
  }
  return out;
}

namespace {

// Per-thread working buffers of the preparation.  Each call of family_state
// reuses the capacity the thread's previous calls left (a fresh allocation
// of a few hundred KiB is mapped and faulted in page by page, and returned
// to the system when freed: several times the arithmetic's cost, and
// serialised between threads); trim() frees them after an unusually large
// star.
struct Scratch {
  std::vector<double> flux1, t, f, e, deltas, standardized, resid, sig;
  std::vector<double> weights, steps, x, wr, st, vals;
  std::vector<double> med_v, med_in, mad_dev;
  std::vector<size_t> keep;
  std::vector<int64_t> sec;
  std::vector<uint32_t> gi, med_count;
  std::vector<uint16_t> med_id;
  std::vector<std::vector<double>> z;

  size_t bytes() const {
    size_t b = 0;
    for (const auto* v : {&flux1, &t, &f, &e, &deltas, &standardized, &resid, &sig, &weights,
                          &steps, &x, &wr, &st, &vals, &med_v, &med_in, &mad_dev})
      b += v->capacity() * sizeof(double);
    b += keep.capacity() * sizeof(size_t) + sec.capacity() * sizeof(int64_t);
    b += (gi.capacity() + med_count.capacity()) * sizeof(uint32_t);
    b += med_id.capacity() * sizeof(uint16_t);
    for (const auto& v : z) b += v.capacity() * sizeof(double);
    return b;
  }
  void trim(size_t limit) {
    if (bytes() > limit) *this = Scratch();
  }
};

Scratch& scratch() {
  thread_local Scratch s;
  return s;
}

// stats.rs:4-11: the lower median of the finite values, the element a sort
// would place at (m - 1) / 2, found by selection in linear time: a pass of
// counts over value buckets, then selection within the one bucket that holds
// it (a +0.0 and a -0.0 may trade places, which no caller can see: each
// median is a divisor checked positive, or subtracted before an absolute
// value)
double median_lower(const std::vector<double>& values) {
  Scratch& sc = scratch();
  size_t m = 0;
  double lo = INFINITY, hi = -INFINITY;
  for (double x : values)
    if (std::isfinite(x)) {
      m++;
      lo = std::min(lo, x);
      hi = std::max(hi, x);
    }
  if (m == 0) return NAN;
  const size_t k = (m - 1) / 2;
  const double span = hi - lo;
  const size_t nbk = std::min<size_t>(4096, m / 8);
  const double scale = span > 0.0 ? (double)nbk / span : 0.0;
  if (m >= 512 && std::isfinite(span) && scale > 0.0 && std::isfinite(scale)) {
    // value buckets: bucket(x) = floor((x - lo) * scale) is nondecreasing in
    // x, so the k-th smallest value lies in the bucket where the running
    // count passes k, at rank k minus the count of the buckets below it
    // (x - lo) * scale lies in [0, nbk] (up to rounding): a signed
    // conversion, one instruction, gives the same floor as an unsigned one
    const int64_t top = (int64_t)nbk - 1;
    auto bucket = [&](double x) {
      const int64_t b = (int64_t)((x - lo) * scale);
      return (size_t)(b < top ? b : top);
    };
    std::vector<uint32_t>& count = sc.med_count;
    count.assign(nbk, 0);
    std::vector<double>& in = sc.med_in;
    in.clear();
    if (m == values.size()) {  // every value finite: bucket ids kept (nbk < 2^16)
      std::vector<uint16_t>& id = sc.med_id;
      id.resize(m);
      for (size_t i = 0; i < m; i++) {
        const size_t bi = bucket(values[i]);
        id[i] = (uint16_t)bi;
        count[bi]++;
      }
      size_t below = 0, b = 0;
      while (below + count[b] <= k) below += count[b++];
      in.reserve(count[b]);
      for (size_t i = 0; i < m; i++)
        if (id[i] == b) in.push_back(values[i]);
      const size_t r = k - below;
      std::nth_element(in.begin(), in.begin() + (ptrdiff_t)r, in.end());
      return in[r];
    }
    for (double x : values)
      if (std::isfinite(x)) count[bucket(x)]++;
    size_t below = 0, b = 0;
    while (below + count[b] <= k) below += count[b++];
    in.reserve(count[b]);
    for (double x : values)
      if (std::isfinite(x) && bucket(x) == b) in.push_back(x);
    const size_t r = k - below;
    std::nth_element(in.begin(), in.begin() + (ptrdiff_t)r, in.end());
    return in[r];
  }
  std::vector<double>& v = sc.med_v;
  v.clear();
  v.reserve(m);
  for (double x : values)
    if (std::isfinite(x)) v.push_back(x);
  std::nth_element(v.begin(), v.begin() + (ptrdiff_t)k, v.end());
  return v[k];
}

// stats.rs:14-18
double mad_scale(const std::vector<double>& values) {
  const double m = median_lower(values);
  std::vector<double>& dev = scratch().mad_dev;
  dev.resize(values.size());
  for (size_t i = 0; i < values.size(); i++) dev[i] = std::fabs(values[i] - m);
  return 1.4826 * median_lower(dev);
}

double max1(double x) { return x >= 1.0 ? x : 1.0; }  // f64::max(x, 1.0), NaN -> 1

struct Noise {
  std::vector<double> t, deficit, variance;
  double origin = 0.0, inflation = 1.0;
  size_t segments = 0;
};

// curve.rs:89-198: per-group median normalisation, optional inverse-variance
// binning (never across groups), robust inflation from adjacent differences.
int fit_segment_noise(const double* time, const double* flux, const double* err,
                      const int64_t* group, size_t n, double bin_minutes, double gap,
                      Noise* out, std::string* msg) {
  Scratch& sc = scratch();
  std::vector<size_t>& keep = sc.keep;
  keep.clear();
  for (size_t i = 0; i < n; i++) {
    const double e = err ? err[i] : 1.0;
    if (std::isfinite(time[i]) && std::isfinite(flux[i]) && std::isfinite(e) && e > 0.0)
      keep.push_back(i);
  }
  bool ascending = true;
  for (size_t k = 1; k < keep.size() && ascending; k++)
    ascending = !(time[keep[k]] < time[keep[k - 1]]);
  if (!ascending)  // a stable sort of ascending times is the identity
    std::stable_sort(keep.begin(), keep.end(),
                     [&](size_t a, size_t b) { return time[a] < time[b]; });
  if (keep.size() < 6) {
    *msg = "insufficient usable observations";
    return FBLS_E_DATA;
  }
  size_t m = keep.size();
  std::vector<double>&t = sc.t, &f = sc.f, &e = sc.e;
  std::vector<int64_t>& sec = sc.sec;
  t.resize(m);
  f.resize(m);
  e.resize(m);
  sec.resize(m);
  for (size_t k = 0; k < m; k++) {
    t[k] = time[keep[k]];
    f[k] = flux[keep[k]];
    e[k] = err ? err[keep[k]] : 1.0;
    sec[k] = group ? group[keep[k]] : 0;
  }
  if (!group) {  // one group: its median alone
    const double med = median_lower(f);
    if (!(med > 0.0)) {
      *msg = "mission flux must have a positive median";
      return FBLS_E_DATA;
    }
    for (size_t k = 0; k < m; k++) {
      f[k] = f[k] / med - 1.0;
      e[k] /= med;
    }
  } else {
    // distinct group labels in ascending order, and each observation's group
    std::vector<int64_t> labels;
    for (size_t k = 0; k < m; k++)
      if (std::find(labels.begin(), labels.end(), sec[k]) == labels.end()) {
        labels.push_back(sec[k]);
        if (labels.size() > 64) {  // many groups: sort them all
          labels = sec;
          std::sort(labels.begin(), labels.end());
          labels.erase(std::unique(labels.begin(), labels.end()), labels.end());
          break;
        }
      }
    std::sort(labels.begin(), labels.end());
    std::vector<uint32_t>& gi = sc.gi;
    gi.resize(m);
    std::vector<size_t> gcount(labels.size(), 0);
    for (size_t k = 0; k < m; k++) {
      gi[k] = (uint32_t)(std::lower_bound(labels.begin(), labels.end(), sec[k]) - labels.begin());
      gcount[gi[k]]++;
    }
    std::vector<double> med(labels.size());
    for (size_t g = 0; g < labels.size(); g++) {
      std::vector<double>& vals = sc.vals;
      vals.clear();
      vals.reserve(gcount[g]);
      for (size_t k = 0; k < m; k++)
        if (gi[k] == g) vals.push_back(f[k]);
      med[g] = median_lower(vals);
      if (!(med[g] > 0.0)) {
        *msg = "mission flux must have a positive median";
        return FBLS_E_DATA;
      }
    }
    for (size_t k = 0; k < m; k++) {
      f[k] = f[k] / med[gi[k]] - 1.0;
      e[k] /= med[gi[k]];
    }
  }
  const double origin = t[0];
  for (double& x : t) x -= origin;
  if (bin_minutes > 0.0) {
    const double width = bin_minutes / 1440.0;
    std::map<std::pair<int64_t, int64_t>, size_t> slot;
    std::vector<std::pair<int64_t, int64_t>> keys(m);
    for (size_t k = 0; k < m; k++) {
      keys[k] = {(int64_t)std::floor(t[k] / width), sec[k]};
      slot.emplace(keys[k], 0);
    }
    size_t u = 0;
    std::vector<int64_t> usec;
    for (auto& kv : slot) {
      kv.second = u++;
      usec.push_back(kv.first.second);
    }
    std::vector<double> den(u, 0.0), st(u, 0.0), sf(u, 0.0);
    for (size_t k = 0; k < m; k++) {
      const size_t j = slot[keys[k]];
      const double w = 1.0 / (e[k] * e[k]);
      den[j] += w;
      st[j] += t[k] * w;
      sf[j] += f[k] * w;
    }
    std::vector<double> bt(u), bf(u), be(u);
    for (size_t j = 0; j < u; j++) {
      bt[j] = st[j] / den[j];
      bf[j] = sf[j] / den[j];
      be[j] = 1.0 / std::sqrt(den[j]);
    }
    std::vector<size_t> order(u);
    for (size_t j = 0; j < u; j++) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return bt[a] < bt[b]; });
    t.resize(u);
    f.resize(u);
    e.resize(u);
    sec.resize(u);
    for (size_t j = 0; j < u; j++) {
      t[j] = bt[order[j]];
      f[j] = bf[order[j]];
      e[j] = be[order[j]];
      sec[j] = usec[order[j]];
    }
    m = u;
  }
  std::vector<double>&deltas = sc.deltas, &standardized = sc.standardized;
  deltas.clear();
  standardized.clear();
  for (size_t i = 1; i < m; i++) {
    if (t[i] - t[i - 1] < gap && sec[i] == sec[i - 1]) {
      const double d = f[i] - f[i - 1];
      deltas.push_back(d);
      standardized.push_back(d / std::sqrt(e[i] * e[i] + e[i - 1] * e[i - 1]));
    }
  }
  if (deltas.empty()) {
    *msg = "no adjacent observations available for noise estimation";
    return FBLS_E_DATA;
  }
  const double inflation = max1(mad_scale(standardized));
  out->variance.resize(m);
  if (err) {
    for (size_t k = 0; k < m; k++) {
      const double s = e[k] * inflation;
      out->variance[k] = s * s;
    }
  } else {
    const double md = median_lower(deltas);
    std::vector<double> dev(deltas.size());
    for (size_t k = 0; k < deltas.size(); k++) dev[k] = std::fabs(deltas[k] - md);
    double white = 1.4826 / std::sqrt(2.0) * median_lower(dev);
    if (!(white >= 1e-12)) white = 1e-12;
    for (size_t k = 0; k < m; k++) out->variance[k] = white * white;
  }
  size_t seg = 1;
  for (size_t i = 1; i < m; i++)
    if (t[i] - t[i - 1] > gap || sec[i] != sec[i - 1]) seg++;
  out->deficit.resize(m);
  for (size_t k = 0; k < m; k++) out->deficit[k] = -f[k];
  out->t.assign(t.begin(), t.begin() + m);
  out->origin = origin;
  out->inflation = inflation;
  out->segments = seg;
  return 0;
}

int fit_segment_noise(const double* time, const double* flux, const double* err,
                      size_t n, double gap, Noise* out, std::string* msg) {
  return fit_segment_noise(time, flux, err, nullptr, n, 0.0, gap, out, msg);
}

// curve.rs:456-539
void correlated_noise_factors(const std::vector<double>& time,
                              const std::vector<double>& residual,
                              const std::vector<double>& sigma,
                              const double* durations, size_t nd, double* rho) {
  const double min_fill = 0.5;
  const size_t minimum_bins = 20;
  const size_t n = time.size();
  Scratch& sc = scratch();
  std::vector<double>& weights = sc.weights;
  weights.resize(n);
  for (size_t i = 0; i < n; i++) weights[i] = 1.0 / (sigma[i] * sigma[i]);
  std::vector<double>& steps = sc.steps;
  steps.clear();
  for (size_t i = 1; i < n; i++) {
    const double s = time[i] - time[i - 1];
    if (s > 0.0) steps.push_back(s);
  }
  const double cadence = median_lower(steps);
  std::vector<double> measured(nd);
  std::vector<size_t> counts(nd);
  // A point's bin is floor(fl(x / d)) for x = t - t0, nondecreasing in x
  // because the division is; each duration keeps the smallest x of its next
  // bin (the edge), so a point is compared with the edges, and the bin of a
  // point at or past one comes from the division.  Within a duration the
  // points are added in order from zero at each bin's first point, as in a
  // pass of that duration alone, so every bin's sums are the same doubles.
  std::vector<double>&x = sc.x, &wr = sc.wr;
  x.resize(n);
  wr.resize(n);
  for (size_t i = 0; i < n; i++) {
    x[i] = time[i] - time[0];
    wr[i] = weights[i] * residual[i];
  }
  std::vector<std::vector<double>>& z = sc.z;
  if (z.size() < nd) z.resize(nd);
  for (size_t d = 0; d < nd; d++) z[d].clear();
  std::vector<double> need(nd);
  for (size_t l = 0; l < nd; l++) need[l] = min_fill * durations[l] / cadence;
  auto flush = [&](size_t l, size_t count, double total, double acc) {
    if (count > 0 && (double)count >= need[l]) z[l].push_back((acc / total) * std::sqrt(total));
  };
  // the smallest double b with fl(b / d) >= c + 1: the first x of the bins
  // after bin c (b = (c + 1) d within a few ulps, then stepped exactly)
  auto next_edge = [](double c, double d) {
    const double target = c + 1.0;
    double e = target * d;
    while (!(e / d >= target)) e = std::nextafter(e, INFINITY);
    for (;;) {
      const double below = std::nextafter(e, -INFINITY);
      if (!(below / d >= target)) break;
      e = below;
    }
    return e;
  };
  // eight durations per pass, their sums in registers; a point below every
  // lane's edge costs one comparison and the sixteen additions
  constexpr size_t kLanes = 8;
  for (size_t d0 = 0; d0 < nd; d0 += kLanes) {
    const size_t lanes = std::min(kLanes, nd - d0);
    double total[kLanes], acc[kLanes], edge[kLanes];
    size_t first[kLanes];
    for (size_t l = 0; l < kLanes; l++) {
      total[l] = acc[l] = 0.0;
      edge[l] = l < lanes ? -INFINITY : INFINITY;
      first[l] = 0;
    }
    double low = -INFINITY;  // the smallest edge
    for (size_t i = 0; i < n; i++) {
      const double xi = x[i];
      if (xi >= low) {  // a new bin of some duration
        low = INFINITY;
        for (size_t l = 0; l < lanes; l++) {
          if (xi >= edge[l]) {
            const double d = durations[d0 + l];
            flush(d0 + l, i - first[l], total[l], acc[l]);
            edge[l] = next_edge(std::floor(xi / d), d);
            total[l] = acc[l] = 0.0;
            first[l] = i;
          }
          low = std::min(low, edge[l]);
        }
      }
      const double wi = weights[i], ri = wr[i];
      for (size_t l = 0; l < kLanes; l++) {  // lanes past the last duration are never read
        total[l] += wi;
        acc[l] += ri;
      }
    }
    for (size_t l = 0; l < lanes; l++) flush(d0 + l, n - first[l], total[l], acc[l]);
  }
  for (size_t di = 0; di < nd; di++) {
    counts[di] = z[di].size();
    measured[di] = z[di].size() >= 3 ? max1(mad_scale(z[di])) : 1.0;
  }
  bool any = false;
  for (size_t c : counts) any = any || c >= minimum_bins;
  std::vector<double> inherited(nd);
  if (any) {
    for (size_t di = 0; di < nd; di++) {
      const double d = durations[di];
      size_t anchor = nd;
      double best = 0.0;
      for (size_t a = 0; a < nd; a++) {
        if (counts[a] < minimum_bins) continue;
        const double dist = std::fabs(std::log(d) - std::log(durations[a]));
        if (anchor == nd || dist < best) {
          anchor = a;
          best = dist;
        }
      }
      const double ratio = d / durations[anchor];
      inherited[di] = measured[anchor] * std::sqrt(ratio >= 1.0 ? ratio : 1.0);
    }
  } else {
    std::vector<double>& st = sc.st;
    st.resize(n);
    for (size_t i = 0; i < n; i++) st[i] = residual[i] / sigma[i];
    const double scatter = max1(mad_scale(st));
    for (size_t di = 0; di < nd; di++) {
      const double ratio = durations[di] / cadence;
      inherited[di] = scatter * std::sqrt(ratio >= 1.0 ? ratio : 1.0);
    }
  }
  for (size_t di = 0; di < nd; di++)
    rho[di] = counts[di] >= minimum_bins
                  ? measured[di]
                  : (measured[di] >= inherited[di] ? measured[di] : inherited[di]);
}

}  // namespace

// curve.rs:331-350: distance of each point to the nearest edge of its
// gap-separated segment
static std::vector<double> segment_edges(const std::vector<double>& t, double gap) {
  const size_t n = t.size();
  std::vector<double> out(n, 0.0);
  size_t start = 0;
  for (size_t i = 1; i <= n; i++) {
    if (i == n || t[i] - t[i - 1] > gap) {
      for (size_t j = start; j < i; j++)
        out[j] = std::min(t[j] - t[start], t[i - 1] - t[j]);
      start = i;
    }
  }
  return out;
}

std::vector<std::pair<double, std::vector<size_t>>> detrending_windows(
    const double* durations, size_t nd, double minimum, double factor, double maximum) {
  // curve.rs:355-377: clamp(factor D, min, max) rounded to 1e-6 d
  std::vector<double> w(nd);
  for (size_t i = 0; i < nd; i++) {
    const double x = std::min(std::max(factor * durations[i], minimum), maximum);
    w[i] = std::round(x * 1e6) / 1e6;
  }
  std::vector<double> uniq = w;
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  std::vector<std::pair<double, std::vector<size_t>>> out;
  for (double x : uniq) {
    std::vector<size_t> idx;
    for (size_t i = 0; i < nd; i++)
      if (w[i] == x) idx.push_back(i);
    out.emplace_back(x, idx);
  }
  return out;
}

// curve.rs:387-425
int prepare_families(const double* time, const double* flux, const double* err,
                     const int64_t* group, size_t n, const double* durations, size_t nd,
                     const fbls_prep_cfg* cfg, const Flatten& flatten,
                     std::vector<Family>* out, std::string* msg) {
  Noise diag;
  int st = fit_segment_noise(time, flux, err, group, n, cfg->bin_minutes, 0.75, &diag, msg);
  if (st) return st;
  const size_t m = diag.t.size();
  std::vector<double> sigma(m), relative(m);
  for (size_t i = 0; i < m; i++) {
    sigma[i] = std::sqrt(diag.variance[i]);
    relative[i] = -diag.deficit[i];
  }
  const std::vector<double> edge = segment_edges(diag.t, 0.5);
  out->clear();
  for (auto& wi : detrending_windows(durations, nd, cfg->window_min, cfg->window_factor,
                                     cfg->window_max)) {
    const double window = wi.first;
    std::vector<double> flat;
    if ((st = flatten(diag.t, relative, window, 0.5, 5.0, 10, &flat))) return st;
    const double scale = mad_scale(flat);
    Family fam;
    fam.window = window;
    fam.indices = wi.second;
    for (size_t i = 0; i < m; i++) {
      if (std::isfinite(flat[i]) && flat[i] <= cfg->clip_sigma * scale &&
          edge[i] >= cfg->edge_trim * window) {
        fam.time.push_back(diag.t[i] + diag.origin);
        fam.flux.push_back(flat[i]);
        fam.sigma.push_back(sigma[i]);
      }
    }
    fam.clipped = m - fam.time.size();
    out->push_back(std::move(fam));
  }
  return 0;
}

int family_state(const double* time, const double* flux, const double* sigma,
                 size_t n, const double* durations, size_t nd, int correlated,
                 double* t, double* v, double* w, size_t* out_n,
                 double* baseline_norm, double* origin, double* rho,
                 double* inflation, size_t* segments, double* residual_energy,
                 std::string* msg, double flux_offset) {
  // curve.rs:547-548: flux1 = 1 + flux, fit without binning, gap 0.75 d
  // (flux + flux_offset: the flux relative to zero of a curve normalised to
  // 1 - flux_offset, rounded as the caller's own subtraction would be)
  Scratch& sc = scratch();
  std::vector<double>& flux1 = sc.flux1;
  flux1.resize(n);
  for (size_t i = 0; i < n; i++) flux1[i] = 1.0 + (flux[i] + flux_offset);
  thread_local Noise model;
  int st = fit_segment_noise(time, flux1.data(), sigma, n, 0.75, &model, msg);
  if (st) return st;
  const size_t m = model.t.size();
  // curve.rs:427-450 (prepare_ready)
  if (m < 6) {
    *msg = "at least six observations are required";
    return FBLS_E_DATA;
  }
  double c = 0.0;
  for (size_t i = 0; i < m; i++) {
    w[i] = 1.0 / model.variance[i];
    c += w[i];
  }
  if (!(std::isfinite(c) && c > 0.0)) {
    *msg = "invalid baseline precision";
    return FBLS_E_DATA;
  }
  double spy = 0.0;
  for (size_t i = 0; i < m; i++) spy += w[i] * model.deficit[i];
  const double mean = spy / c;
  for (size_t i = 0; i < m; i++) {
    const double py = w[i] * model.deficit[i];
    v[i] = py - w[i] * mean;
    t[i] = model.t[i];
  }
  // E = sum y v (curve.rs:440)
  double energy = 0.0;
  for (size_t i = 0; i < m; i++) energy += model.deficit[i] * v[i];
  *residual_energy = energy;
  *out_n = m;
  *baseline_norm = c;
  *origin = model.origin;
  *inflation = model.inflation;
  *segments = model.segments;
  if (correlated) {
    std::vector<double>&resid = sc.resid, &sig = sc.sig;
    resid.resize(m);
    sig.resize(m);
    for (size_t i = 0; i < m; i++) {
      resid[i] = -model.deficit[i];
      sig[i] = std::sqrt(model.variance[i]);
    }
    correlated_noise_factors(model.t, resid, sig, durations, nd, rho);
  } else {
    for (size_t i = 0; i < nd; i++) rho[i] = 1.0;
  }
  // an unusually large star does not keep its buffers (64 MiB per thread)
  sc.trim((size_t)64 << 20);
  if (model.t.capacity() > ((size_t)8 << 20)) model = Noise();
  return 0;
}

// ------------------------------------------------------------ synthetic

namespace {

// survey/numeric.rs:345-364
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s((seed * 0x9E3779B97F4A7C15ull) | 1ull) {}
  double uniform() {
    uint64_t x = s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    s = x;
    return (double)(x >> 11) / (double)(1ull << 53);
  }
  double normal(double mean, double sigma) {
    double u1 = uniform();
    if (!(u1 >= 1e-300)) u1 = 1e-300;
    const double u2 = uniform();
    return mean + sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

}  // namespace

// Light curve of `sectors` contiguous sectors at a fixed cadence, each with a
// gap at mid-sector, optionally a long gap after half of the sectors.  Flux is
// 1 + white noise + stationary AR(1) red noise - box transit(s); flux_err is
// the constant total sigma.  Draw order per point: white, then red innovation.
int synth_star(const fbls_synth_cfg* cfg, double* time, double* flux,
               double* flux_err, size_t cap, size_t* n, std::string* msg) {
  if (!cfg || !n) {
    *msg = "null argument";
    return FBLS_E_ARG;
  }
  if (!(cfg->cadence_days > 0.0) || !(cfg->sector_days > 0.0) || cfg->sectors == 0) {
    *msg = "invalid synthetic configuration";
    return FBLS_E_ARG;
  }
  Rng rng(cfg->seed);
  double epoch = cfg->epoch;
  if (!std::isfinite(epoch)) epoch = cfg->t0 + cfg->period * rng.uniform();
  const double var = cfg->sigma * cfg->sigma;
  const double red_var = var * cfg->red_fraction;
  const double white = std::sqrt(var - red_var);
  const double red = std::sqrt(red_var);
  const double innov = red * std::sqrt(1.0 - cfg->ar_phi * cfg->ar_phi);
  size_t m = 0;
  const size_t per = (size_t)std::floor(cfg->sector_days / cfg->cadence_days);
  for (uint32_t s = 0; s < cfg->sectors; s++) {
    double start = cfg->t0 + s * cfg->sector_days;
    if (cfg->baseline_gap_days > 0.0 && s >= (cfg->sectors + 1) / 2)
      start += cfg->baseline_gap_days;
    const double mid = start + 0.5 * cfg->sector_days;
    double ar = red > 0.0 ? rng.normal(0.0, red) : 0.0;
    for (size_t i = 0; i < per; i++) {
      const double t = start + (double)i * cfg->cadence_days;
      if (std::fabs(t - mid) < 0.5 * cfg->gap_days) continue;
      double x = white > 0.0 ? rng.normal(0.0, white) : 0.0;
      if (red > 0.0) {
        ar = cfg->ar_phi * ar + rng.normal(0.0, innov);
        x += ar;
      }
      if (cfg->depth > 0.0 && cfg->period > 0.0) {
        double ph = std::fmod(t - epoch, cfg->period);
        if (ph < 0.0) ph += cfg->period;
        if (ph > 0.5 * cfg->period) ph -= cfg->period;
        if (std::fabs(ph) <= 0.5 * cfg->duration) x -= cfg->depth;
        if (cfg->secondary_depth > 0.0) {
          double ps = std::fmod(t - epoch - 0.5 * cfg->period, cfg->period);
          if (ps < 0.0) ps += cfg->period;
          if (ps > 0.5 * cfg->period) ps -= cfg->period;
          if (std::fabs(ps) <= 0.5 * cfg->duration) x -= cfg->secondary_depth;
        }
      }
      if (m < cap) {
        if (time) time[m] = t;
        if (flux) flux[m] = 1.0 + x;
        if (flux_err) flux_err[m] = cfg->sigma;
      }
      m++;
    }
  }
  *n = m;
  if (m > cap && (time || flux || flux_err)) {
    *msg = "output capacity too small";
    return FBLS_E_ARG;
  }
  return 0;
}

}  // namespace fbls

// The preparation's lower median (a bucket pass, then selection within one
// bucket) against a full sort: duplicates, huge and subnormal values, NaN
// and infinities (ignored as in stats.rs:4-11), signed zeros, values spread
// over decades.  Built and run by tests/test_host.py.
#include "../../fbls_amd/csrc/prep.cpp"

#include <cstdio>
#include <random>

int main() {
  std::mt19937_64 g(7);
  int bad = 0;
  for (int it = 0; it < 20000; it++) {
    const size_t m = 1 + g() % 5000;
    std::vector<double> v(m);
    const int kind = it % 7;
    for (double& x : v) {
      const double u = std::uniform_real_distribution<double>(-1, 1)(g);
      switch (kind) {
        case 0: x = u; break;
        case 1: x = std::floor(u * 5); break;
        case 2: x = u * 1e300; break;
        case 3: x = (g() % 10 == 0) ? NAN : u * 1e-310; break;
        case 4: x = (g() % 3 == 0) ? INFINITY : 1.0 + u * 1e-15; break;
        case 5: x = (g() % 2) ? 0.0 : -0.0; break;
        default: x = std::exp(u * 700); break;
      }
    }
    std::vector<double> f;
    for (double x : v)
      if (std::isfinite(x)) f.push_back(x);
    double want = NAN;
    if (!f.empty()) {
      std::sort(f.begin(), f.end());
      want = f[(f.size() - 1) / 2];
    }
    const double got = fbls::median_lower(v);
    if (!(got == want || (std::isnan(got) && std::isnan(want)))) bad++;
  }
  std::printf("mismatches %d\n", bad);
  return bad != 0;
}

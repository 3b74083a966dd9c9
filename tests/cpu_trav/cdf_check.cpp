// CPU check of the cut-point CDF search (rtg_internal.h cdf_guide, rtg_device.h cdf_sample_guided):
// for random CDFs built as the library builds them (running float sums of non-negative weights,
// normalised, with runs of zero weights, one dominant entry and nearly uniform rows), the guided
// search returns the index of the full binary search (cdf_sample) for every draw: 200k random
// draws per CDF, every CDF value itself with its float neighbours, and every cut value k / 64 with
// its neighbours. The cut points are also checked against their definition counted directly (the
// number of CDF values below k / 64).
// Prints a summary line, exit 1 on failure.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "rtg_device.h"

using namespace rtg;
using namespace rtg::dev;

int main() {
  std::mt19937 rng(7);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  long checked = 0, bad = 0, badguide = 0;
  for (int t = 0; t < 400; t++) {
    const int n = 2 + (int)(rng() % 1025);  // weights
    std::vector<float> w(n);
    for (int i = 0; i < n; i++) {
      switch (t % 4) {
        case 0: w[i] = U(rng); break;                                   // uniform-ish
        case 1: w[i] = (rng() % 5 == 0) ? 0.f : U(rng) * U(rng); break;  // runs of zeros
        case 2: w[i] = (i == (int)(rng() % n)) ? 1000.f : 1e-4f * U(rng); break;  // one dominant
        default: w[i] = std::pow(U(rng), 8.f); break;                   // heavy skew
      }
    }
    // core/CDF.h: c[0] = 0, running sums, normalised by the total (float, as the host builds it)
    std::vector<float> c(n + 1, 0.f);
    for (int i = 0; i < n; i++) c[i + 1] = c[i] + w[i];
    const float sum = c[n];
    if (!(sum > 0.f)) continue;
    for (int i = 1; i <= n; i++) c[i] /= sum;
    std::vector<int32_t> g;
    cdf_guide(c.data(), n + 1, g);
    for (int k = 0; k < kCdfGuide; k++) {
      const float v = (float)k / (float)kCdfGuide;
      int below = 0;
      for (int i = 0; i <= n; i++) below += c[i] < v;  // c is non-decreasing: the lower bound
      if (g[k] != below) badguide++;
    }
    if (g[kCdfGuide] != n + 1) badguide++;
    auto probe = [&](float s) {
      if (!(s >= 0.f && s < 1.f)) return;
      checked++;
      if (cdf_sample_guided(c.data(), n + 1, s, g.data()) != cdf_sample(c.data(), n + 1, s)) bad++;
    };
    for (int q = 0; q < 200000; q++) probe(U(rng));
    for (int i = 0; i <= n; i++) {
      probe(c[i]);
      probe(std::nextafter(c[i], 0.f));
      probe(std::nextafter(c[i], 1.f));
    }
    for (int k = 0; k < kCdfGuide; k++) {
      const float v = (float)k / (float)kCdfGuide;
      probe(v);
      probe(std::nextafter(v, 0.f));
      probe(std::nextafter(v, 1.f));
    }
    probe(std::nextafter(1.f, 0.f));
  }
  printf("guided CDF search: %ld draws, %ld differ from the full search; %ld wrong cut points\n", checked, bad,
         badguide);
  const bool ok = bad == 0 && badguide == 0 && checked > 1000000;
  printf(ok ? "OK\n" : "FAIL\n");
  return ok ? 0 : 1;
}

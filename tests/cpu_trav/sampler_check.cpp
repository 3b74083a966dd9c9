// CPU check of the device sampler's stratum arithmetic (in_stratum in rtg_device.h, compiled for
// the host through device_cpu.h): for every sample count n up to 4225 and every cell, values from
// the extreme jitters (0, the largest float below 1 and its neighbours) and from random jitters
// stay inside [cell / n, (cell + 1) / n) in exact arithmetic, and the permutation of the
// stratified sampler is a bijection of [0, n). Usage: sampler_check; prints OK or the failures.
#include <cstdio>
#include <random>
#include <vector>

#include "rtg_device.h"

using namespace rtg::dev;

int main() {
  std::mt19937 rng(1);
  long bad = 0, checked = 0;
  const float jit[] = {0.f, 0.99999994f, 0.9999999f, 0.99999976f, 0.5f, 1e-8f};
  for (uint32_t n = 1; n <= 4225; n += (n < 300 ? 1 : 37)) {
    for (uint32_t c = 0; c < n; c++) {
      for (int q = 0; q < 10; q++) {
        float j = q < 6 ? jit[q] : to_unit((uint32_t)rng());
        float x = in_stratum(c, j, n);
        long double e = (long double)x * n;
        checked++;
        if (!(e >= c && e < c + 1) || !(x >= 0.f && x < 1.f)) {
          if (bad < 10) printf("n %u cell %u jitter %.9g -> x %.9g (x n = %.12Lf)\n", n, c, j, x, e);
          bad++;
        }
      }
    }
    std::vector<int> seen(n, 0);
    const uint32_t p = (uint32_t)rng();
    for (uint32_t i = 0; i < n; i++) seen[permute(i, n, p)]++;
    for (uint32_t i = 0; i < n; i++)
      if (seen[i] != 1) {
        if (bad < 10) printf("permute n %u: %u hit %d times\n", n, i, seen[i]);
        bad++;
      }
  }
  printf("checked %ld values, %ld failures\n", checked, bad);
  printf(bad ? "FAIL\n" : "OK\n");
  return bad ? 1 : 0;
}

// CPU check of the CDF guide tables (hikari_amd/csrc/hk_cdf.h): for many
// CDFs and samples, the guided index search must return the index of the
// search over the whole CDF.  With --plant, every guide entry above zero is
// moved up by one first, and the check must then find mismatches (so a wrong
// table cannot pass unnoticed).  Prints "checked N mismatches M".
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../hikari_amd/csrc/hk_cdf.h"

int main(int argc, char **argv) {
  bool plant = argc > 1 && std::strcmp(argv[1], "--plant") == 0;
  std::mt19937 gen(7);
  std::uniform_real_distribution<float> u01(0.0f, 1.0f);
  long checked = 0, bad = 0;
  const int sizes[] = {1, 2, 3, 5, 7, 64, 255, 256, 511, 512, 1000, 4097};
  for (int n : sizes) {
    for (int kind = 0; kind < 6; kind++) {
      std::vector<float> f(n);
      for (int i = 0; i < n; i++) {
        float x = u01(gen);
        switch (kind) {
          case 0: f[i] = 1.0f; break;                                        // uniform
          case 1: f[i] = -std::log(1.0f - x); break;                         // exponential weights
          case 2: f[i] = x < 0.97f ? 0.0f : 1e3f * x; break;                 // spikes, flat runs
          case 3: f[i] = 0.0f; break;                                        // zero: uniform fallback
          case 4: f[i] = std::ldexp(1.0f, int(x * 100.0f) - 50); break;      // wide dynamic range
          default: f[i] = float(i % 3 == 0) * (1.0f + x); break;             // every third nonzero
        }
      }
      std::vector<float> cdf(n + 1);
      float integral;
      hk::build_cdf(f.data(), n, cdf.data(), integral);
      std::vector<uint32_t> guide(n + 1);
      hk::build_cdf_guide(cdf.data(), n, guide.data());
      if (plant)
        for (int g = 1; g < n; g++) guide[g] += 1;
      std::vector<float> s;
      for (int k = 0; k < 20000; k++) s.push_back(u01(gen));
      for (float c : cdf) {
        s.push_back(c);
        s.push_back(std::nextafter(c, 2.0f));
        s.push_back(std::nextafter(c, -1.0f));
      }
      for (int g = 0; g <= n; g++) {  // bucket boundaries and their neighbours
        float b = float(g) / float(n);
        for (int d = -3; d <= 3; d++) {
          float v = b;
          for (int k = 0; k < (d < 0 ? -d : d); k++) v = std::nextafter(v, d < 0 ? -1.0f : 2.0f);
          s.push_back(v);
        }
      }
      const float special[] = {0.0f, -0.0f, 1e-45f, 1e-38f, 1.0f, std::nextafter(1.0f, 0.0f), 1.5f, -1.0f,
                               NAN, INFINITY, -INFINITY};
      for (float v : special) s.push_back(v);
      for (float v : s) {
        int a = hk::cdf_find_guided(cdf.data(), guide.data(), n, v);
        int b = hk::cdf_find_full(cdf.data(), n, v);
        checked++;
        if (a != b) bad++;
      }
    }
  }
  std::printf("checked %ld mismatches %ld\n", checked, bad);
  return 0;
}

// hk_cdf.h -- piecewise-constant CDFs of the emitters (area-light triangle
// choice, environment-map rows and marginal): construction
// (DiscretePDF1D, src/core/DiscretePDF.cpp:5-26, in FP32) and the index search
// of its sampling (DiscretePDF.cpp:32-78), accelerated by a guide table that
// leaves every result unchanged.  Plain C++ apart from the host/device
// qualifiers, so the same source serves the kernels, the runtime and a CPU
// check of the guide tables.
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIP__)
#define HK_CDF_HD __host__ __device__ inline
#else
#define HK_CDF_HD inline
#endif

namespace hk {

// cdf[0..n] of func[0..n-1] (uniform when func integrates to zero).
inline void build_cdf(const float *func, int n, float *cdf, float &integral) {
  cdf[0] = 0.0f;
  for (int i = 1; i < n + 1; i++) cdf[i] = cdf[i - 1] + func[i - 1] / n;
  integral = cdf[n];
  if (integral == 0.0f) {
    for (int i = 1; i < n + 1; i++) cdf[i] = float(i) / float(n);
  } else {
    for (int i = 1; i < n + 1; i++) cdf[i] /= integral;
  }
}

// First j in [lo, lo + cnt) with cdf[j] >= s, or lo + cnt.
HK_CDF_HD int cdf_lower_bound(const float *cdf, int lo, int cnt, float s) {
  while (cnt > 0) {
    int step = cnt >> 1;
    if (cdf[lo + step] < s) {
      lo += step + 1;
      cnt -= step + 1;
    } else {
      cnt = step;
    }
  }
  return lo;
}

// Guide bucket of a sample: clamp((int)(s * n), 0, n - 1) in FP32 (NaN -> 0).
HK_CDF_HD int cdf_bucket(float s, int n) {
  float prod = s * (float)n;
  if (!(prod >= 1.0f)) return 0;  // also NaN and negative s
  if (prod >= (float)n) return n - 1;
  return (int)prod;
}

// The interval index of DiscretePDF1D's sampling: the lower bound of s over
// the n + 1 entries, minus one, clamped to [0, n - 1].  guide: the table of
// build_cdf_guide; the search runs inside the range it gives the bucket.
HK_CDF_HD int cdf_find_guided(const float *cdf, const uint32_t *guide, int n, float s) {
  int g = cdf_bucket(s, n);
  int lo = (int)guide[g];
  int i = cdf_lower_bound(cdf, lo, (int)guide[g + 1] - lo, s) - 1;
  return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}

// The same index from a search of the whole CDF.
HK_CDF_HD int cdf_find_full(const float *cdf, int n, float s) {
  int i = cdf_lower_bound(cdf, 0, n + 1, s) - 1;
  return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}

// Guide table (n + 1 entries) of cdf[0..n]: guide[g] is the lower bound of
// the smallest float v with cdf_bucket(v, n) >= g, guide[n] = n + 1.  Bucket
// and lower bound are both monotone in s, so the lower bound of any s in
// bucket g lies in [guide[g], guide[g + 1]] and the guided search returns
// the full search's index.
inline void build_cdf_guide(const float *cdf, int n, uint32_t *guide) {
  guide[0] = 0;
  for (int g = 1; g < n; g++) {
    // bisection over the bit patterns of [0, 1] (ordered like the values):
    // bucket(lo) < g <= bucket(hi)
    uint32_t lo = 0x00000000u, hi = 0x3f800000u;
    while (hi - lo > 1) {
      uint32_t mid = lo + (hi - lo) / 2;
      float v;
      memcpy(&v, &mid, 4);
      if (cdf_bucket(v, n) >= g) hi = mid;
      else lo = mid;
    }
    float v;
    memcpy(&v, &hi, 4);
    guide[g] = (uint32_t)cdf_lower_bound(cdf, 0, n + 1, v);
  }
  guide[n] = (uint32_t)n + 1;
}

}  // namespace hk

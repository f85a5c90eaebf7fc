// hk_oracle.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of Hikari's `pathMIS` hot path, used as the checker for the
// HIP implementation in hikari_amd/.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it; the product never does.
//
// Every function cites the reference file:line whose behaviour it restates.
// Templated on the scalar type R: R=float reproduces the reference's FP32
// arithmetic (Eigen 3.3.5 evaluation order, no FMA contraction: build with
// -ffp-contract=off), R=double is the FP64 evaluation of the same formulas.
// Random numbers are always produced as pcg32 floats (as in the reference) and
// promoted to R.
//
// Pinning (what checks each part; tests/test_oracle.py and
// tests/test_oracle_pins.py hold the checks, and each check of the second file
// is shown to fail with a planted mistake, enum Slip below):
//   pcg32               ext/pcg32/pcg32-demo.out known-answer text, plus the
//                       reference's own pcg32-demo built under oracle/_ref.
//   vector ops / M-T    bit-exact against Eigen 3.3.5 from the reference tree
//                       (oracle/_ref/eigen_ops), closed-form ray/triangle cases,
//                       BVH == brute force on random rays.
//   Fresnel             closed forms (normal incidence, TIR, k=0 conductor);
//                       the conductor against the complex-index equations.
//   microfacet D / G1   normalisation by quadrature; the Smith identity
//                       int G1 <v.m> D dm = cos(theta_v) (GGX exact, Beckmann
//                       to its rational fit).
//   BSDFs               chi^2 of sampled directions against Pdf (the
//                       reference's Chi2Test procedure), Sample == Eval/Pdf,
//                       and per class a closed form or limit: rough conductor
//                       -> ks F (alpha -> 0) and albedo <= 1 (F = 1); rough
//                       dielectric -> F / (1-F) eta^-2; plastic and rough
//                       plastic -> ks F (kd = 0) and their diffuse albedo;
//                       coating of index 1 == its nested BSDF, absorption
//                       exp(-sigma d (1/cos_i + 1/cos_o)), albedo in closed
//                       form; rough coating -> coating (alpha -> 0).
//   RFT tables          knot values reproduce table entries; T(eta=4, a=0,
//                       cos=1) = 1 - ((4-1)/(4+1))^2; T and EvalDiffuse at
//                       alpha = 0 against 1 - F and 1 - the cosine-weighted
//                       Fresnel average.
//   film                one sample's splat against the tabulated Gaussian
//                       with half-pixel centres.
//   estimator           furnace 1/(1-a) (scene/pa5/tests/test-furnace.xml) and
//                       direct-lighting references (test-direct.xml); the
//                       environment sampler against the closed-form
//                       expectation of a plane under a constant sky; the
//                       roulette start depth (no reference counterpart)
//                       against the furnace closed form.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace hko {

// Planted mistakes.  The pinning tests show that each of their checks fails
// when the restatement carries a plausible slip: a test selects one through
// ho_set_slip, runs the check, expects it to fail, and selects kSlipNone
// again.  Every other use runs with kSlipNone.
enum Slip {
  kSlipNone = 0,
  kSlipGgxG1,              // GGX G1 with (alpha tan)^2 where alpha tan belongs
  kSlipBeckmannG1,         // Beckmann G1 fit without its a^2 numerator term
  kSlipConductorFresnel,   // conductor Fresnel with cos_i in place of 2 cos_i
  kSlipRoughConductorG,    // rough conductor Sample() without the shadowing term
  kSlipRoughDielectricEta, // rough dielectric transmission without eta^-2
  kSlipPlasticFresnel,     // plastic specular weight 1 - F in place of F
  kSlipPlasticEta2,        // plastic diffuse weight without eta^-2
  kSlipRoughPlasticEta,    // rough plastic Fresnel with 1/eta in place of eta
  kSlipCoatingSigma,       // coating absorption over the entering path only
  kSlipCoatingExit,        // coating exit Fresnel with eta in place of 1/eta
  kSlipRoughCoatingT,      // rough coating nested term with T(cos_i) twice
  kSlipRftDiffuse,         // EvalDiffuse from the table half of 1/eta
  kSlipFilmCentre,         // film footprint without the half-pixel centre offset
  kSlipRrStartWeight,      // roulette weight 1 / 0.95 also below the roulette start depth
};
inline int &slip() {
  static int s = kSlipNone;
  return s;
}

constexpr double kPi = 3.14159265358979323846;
constexpr double kInvPi = 0.31830988618379067154;

// ---------------------------------------------------------------- pcg32
// ext/pcg32/pcg32.h:38-199 (Jakob's mini PCG32, XSH-RR 64->32).
struct Pcg32 {
  uint64_t state = 0x853c49e6748fea9bULL, inc = 0xda3e39cb94b95bdbULL;
  void seed(uint64_t initstate, uint64_t initseq = 1) {  // pcg32.h:58-64
    state = 0U;
    inc = (initseq << 1u) | 1u;
    next_uint();
    state += initstate;
    next_uint();
  }
  uint32_t next_uint() {  // pcg32.h:67-73
    uint64_t old = state;
    state = old * 0x5851f42d4c957f2dULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31));
  }
  uint32_t next_uint(uint32_t bound) {  // pcg32.h:76-103
    uint32_t threshold = (~bound + 1u) % bound;
    for (;;) {
      uint32_t r = next_uint();
      if (r >= threshold) return r % bound;
    }
  }
  float next_float() {  // pcg32.h:106-115
    union { uint32_t u; float f; } x;
    x.u = (next_uint() >> 9) | 0x3f800000UL;
    return x.f - 1.0f;
  }
  template <typename It>
  void shuffle(It begin, It end) {  // pcg32.h:160-163
    for (It it = end - 1; it > begin; --it)
      std::iter_swap(it, begin + next_uint((uint32_t)(it - begin + 1)));
  }
};

// Per-path stream key shared with the HIP path (the reference seeds one stream
// per 32x32 tile, src/sampler/IndependentSampler.cpp:20-23, whose draw order is
// compiler dependent; this build keys one stream per (request, pixel, sample)).
inline uint64_t mix64(uint64_t z) {  // splitmix64 finaliser (Steele/Vigna)
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
inline void seed_path(Pcg32 &r, uint32_t req_seed, uint32_t pixel, uint32_t sample) {
  r.seed(mix64(((uint64_t)req_seed << 32) | sample), pixel);
}

// ---------------------------------------------------------------- vectors
// Eigen 3.3.5 fixed-size evaluation order (ext/eigen/Eigen/src/Core/Redux.h:88-105:
// a 3-term sum reduces as x0 + (x1 + x2); OrthoMethods.h:43-47 for cross).
template <typename R>
struct V3 {
  R x = 0, y = 0, z = 0;
  V3() {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  explicit V3(R a) : x(a), y(a), z(a) {}
  R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  R &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <typename R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename R> inline V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <typename R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
template <typename R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R> inline R sqnorm(V3<R> a) { return a.x * a.x + (a.y * a.y + a.z * a.z); }
template <typename R> inline R norm(V3<R> a) { return std::sqrt(sqnorm(a)); }
template <typename R> inline V3<R> normalized(V3<R> a) {  // Eigen Dot.h:124-134
  R z = sqnorm(a);
  if (z > R(0)) { R s = std::sqrt(z); return {a.x / s, a.y / s, a.z / s}; }
  return a;
}
template <typename R> inline bool is_zero_fuzzy(V3<R> c) {  // Eigen isZero, dummy_precision<float>
  const R p = R(1e-5f);
  return std::abs(c.x) <= p && std::abs(c.y) <= p && std::abs(c.z) <= p;
}
template <typename R> inline V3<R> vexp(V3<R> a) { return {std::exp(a.x), std::exp(a.y), std::exp(a.z)}; }
template <typename R> inline R luminance(V3<R> c) {  // src/core/Color.cpp:95-98
  return c.x * R(0.212671f) + c.y * R(0.715160f) + c.z * R(0.072169f);
}
template <typename R> inline bool color_valid(V3<R> c) {  // src/core/Color.cpp:44-55
  for (int i = 0; i < 3; i++)
    if (c[i] < R(0) || !std::isfinite(c[i])) return false;
  return true;
}
template <typename R> inline R signum(R v) { return std::copysign(R(1), v); }  // Common.hpp:675-682
template <typename R> inline R clampr(R v, R lo, R hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------- frames
// src/core/Common.cpp:591-604
template <typename R>
inline void coordinate_system(V3<R> a, V3<R> &b, V3<R> &c) {
  if (std::abs(a.x) > std::abs(a.y)) {
    R inv = R(1) / std::sqrt(a.x * a.x + a.z * a.z);
    c = V3<R>(a.z * inv, R(0), -a.x * inv);
  } else {
    R inv = R(1) / std::sqrt(a.y * a.y + a.z * a.z);
    c = V3<R>(R(0), a.z * inv, -a.y * inv);
  }
  b = cross(c, a);
}
// src/core/Frame.cpp
template <typename R>
struct Frame {
  V3<R> s, t, n;
  Frame() {}
  explicit Frame(V3<R> nn) : n(nn) { coordinate_system(nn, s, t); }          // Frame.cpp:27-30
  Frame(V3<R> nn, V3<R> dpdu) : n(nn) {                                      // Frame.cpp:9-14
    s = normalized(dpdu - n * dot(n, dpdu));
    t = cross(n, s);
  }
  V3<R> to_local(V3<R> v) const { return {dot(v, s), dot(v, t), dot(v, n)}; }  // Frame.cpp:32-35
  V3<R> to_world(V3<R> v) const { return (s * v.x + t * v.y) + n * v.z; }      // Frame.cpp:37-40
};
template <typename R> inline R cos_theta(V3<R> v) { return v.z; }
template <typename R> inline R sin_theta2(V3<R> v) { return R(1) - v.z * v.z; }
template <typename R> inline R sin_theta(V3<R> v) {  // Frame.cpp:38-46
  R tmp = sin_theta2(v);
  return tmp <= R(0) ? R(0) : std::sqrt(tmp);
}
template <typename R> inline R sin_phi(V3<R> v) {  // Frame.cpp:68-76
  R st = sin_theta(v);
  return st == R(0) ? R(1) : clampr(v.y / st, R(-1), R(1));
}
template <typename R> inline R cos_phi(V3<R> v) {  // Frame.cpp:78-86
  R st = sin_theta(v);
  return st == R(0) ? R(1) : clampr(v.x / st, R(-1), R(1));
}
template <typename R> inline R tan_theta(V3<R> v) {  // Frame.cpp:48-56
  R tmp = R(1) - v.z * v.z;
  if (tmp <= R(0)) return R(0);
  return std::sqrt(tmp) / v.z;
}

// ---------------------------------------------------------------- fresnel etc.
// src/core/Common.cpp:494-527
template <typename R>
inline R fresnel_dielectric(R cos_i, R eta, R inv_eta, R &cos_t) {
  if (eta == R(1)) { cos_t = -cos_i; return R(0); }
  R scale = (cos_i > R(0)) ? inv_eta : eta;
  R cos_t_sqr = R(1) - (R(1) - cos_i * cos_i) * (scale * scale);
  if (cos_t_sqr <= R(0)) { cos_t = R(0); return R(1); }
  R ci = std::abs(cos_i);
  R ct = std::sqrt(cos_t_sqr);
  R rs = (ci - eta * ct) / (ci + eta * ct);
  R rp = (eta * ci - ct) / (eta * ci + ct);
  cos_t = (cos_i > R(0)) ? -ct : ct;
  return R(0.5f) * (rs * rs + rp * rp);
}
// src/core/Common.cpp:529-548
template <typename R>
inline V3<R> fresnel_conductor(R cos_i, V3<R> eta, V3<R> eta_k) {
  R ci2 = cos_i * cos_i;
  R si2 = R(1) - ci2;
  V3<R> eta2 = eta * eta, k2 = eta_k * eta_k;
  V3<R> t0 = eta2 - k2 - V3<R>(si2);
  V3<R> a2b2 = t0 * t0 + R(4) * eta2 * k2;
  a2b2 = V3<R>(std::sqrt(a2b2.x), std::sqrt(a2b2.y), std::sqrt(a2b2.z));
  V3<R> t1 = a2b2 + V3<R>(ci2);
  V3<R> a = R(0.5f) * (a2b2 + t0);
  a = V3<R>(std::sqrt(a.x), std::sqrt(a.y), std::sqrt(a.z));
  V3<R> t2 = (R(slip() == kSlipConductorFresnel ? 1 : 2) * cos_i) * a;
  V3<R> rs = (t1 - t2) / (t1 + t2);
  V3<R> t3 = ci2 * a2b2 + V3<R>(si2 * si2);
  V3<R> t4 = t2 * si2;
  V3<R> rp = rs * ((t3 - t4) / (t3 + t4));
  return R(0.5f) * (rp + rs);
}
// src/core/Common.cpp:550-589
template <typename R>
inline R approx_fresnel_diffuse_reflectance(R eta) {
  if (eta < R(1)) {
    return R(-1.4399f) * (eta * eta) + R(0.7099f) * eta + R(0.6681f) + R(0.0636f) / eta;
  }
  R i1 = R(1) / eta, i2 = i1 * i1, i3 = i2 * i1, i4 = i3 * i1, i5 = i4 * i1;
  return R(0.919317f) - R(3.4793f) * i1 + R(6.75335f) * i2 - R(7.80989f) * i3 + R(4.98554f) * i4 -
         R(1.36881f) * i5;
}
template <typename R> inline V3<R> reflect(V3<R> wi) { return {-wi.x, -wi.y, wi.z}; }  // Common.cpp:606-609
template <typename R> inline V3<R> refract(V3<R> wi, R cos_t, R eta, R inv_eta) {       // Common.cpp:611-615
  R scale = -(cos_t < R(0) ? inv_eta : eta);
  return {scale * wi.x, scale * wi.y, cos_t};
}
template <typename R> inline V3<R> reflect(V3<R> wi, V3<R> m) {  // Common.cpp:617-620
  return (R(2) * dot(wi, m)) * m - wi;
}
template <typename R> inline V3<R> refract(V3<R> wi, V3<R> m, R cos_t, R eta, R inv_eta) {  // Common.cpp:622-626
  eta = (cos_t < R(0) ? inv_eta : eta);
  return m * (dot(wi, m) * eta + cos_t) - wi * eta;
}
template <typename R> inline V3<R> spherical_direction(R theta, R phi) {  // Common.cpp:465-478
  return {std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)};
}

// ---------------------------------------------------------------- warps
// src/core/Sampling.cpp:29-53
template <typename R>
inline V3<R> square_to_uniform_disk(R sx, R sy) {
  R ox = R(2) * sx - R(1), oy = R(2) * sy - R(1);
  if (ox == R(0) && oy == R(0)) return V3<R>(R(0));
  R theta, r;
  if (std::abs(ox) > std::abs(oy)) {
    r = ox;
    theta = R(float(kPi)) * R(0.25f) * (oy / ox);
  } else {
    r = oy;
    theta = R(float(kPi)) * R(0.5f) - R(float(kPi)) * R(0.25f) * (ox / oy);
  }
  return {r * std::cos(theta), r * std::sin(theta), R(0)};
}
// src/core/Sampling.cpp:92-107
template <typename R>
inline V3<R> square_to_cosine_hemisphere(R sx, R sy) {
  V3<R> d = square_to_uniform_disk(sx, sy);
  R z = std::sqrt(R(1) - d.x * d.x - d.y * d.y);
  if (std::isnan(z)) z = R(0);
  return {d.x, d.y, z};
}
template <typename R>
inline R square_to_cosine_hemisphere_pdf(V3<R> v) {  // Sampling.cpp:109-112
  return (norm(v) - R(1) <= R(1e-6f) && v.z >= R(0)) ? v.z / R(float(kPi)) : R(0);
}
// Sampling.cpp:115-133: tan^2(theta) = -alpha^2 log(1 - x), phi = 2 pi y
template <typename R>
inline V3<R> square_to_beckmann(R sx, R sy, R alpha) {
  R ln = std::log(R(1) - sx);
  if (std::isnan(ln)) ln = R(1);
  R tan2 = R(-1) * alpha * alpha * ln;
  R c = R(1) / std::sqrt(R(1) + tan2);
  R st = std::sqrt(R(1) - c * c);
  R phi = R(2) * R(float(kPi)) * sy;
  return {st * std::cos(phi), st * std::sin(phi), c};
}
// Sampling.cpp:135-159
template <typename R>
inline R square_to_beckmann_pdf(V3<R> m, R alpha) {
  if (double(norm(m) - R(1)) <= 1e-6 && m.z >= R(0)) {
    R c = cos_theta(m);
    if (c == R(0)) return R(0);
    R tt = tan_theta(m);
    if (std::isinf(tt)) return R(0);
    R a2 = alpha * alpha;
    return std::exp(R(-1) * std::pow(tt, R(2)) / a2) / (R(float(kPi)) * a2 * std::pow(c, R(3)));
  }
  return R(0);
}

// ---------------------------------------------------------------- microfacet
// src/core/MicrofacetDistribution.cpp
enum { kBeckmann = 0, kGGX = 1 };
template <typename R>
struct Microfacet {
  int type;
  R au, av;
  Microfacet(int t, R a, R b) : type(t), au(clampr(a, R(5e-4f), R(1))), av(clampr(b, R(5e-4f), R(1))) {}  // :15-25
  bool iso() const { return au == av; }
  void scale_alpha(R s) { au *= s; av *= s; }  // :27-31
  R eval(V3<R> m) const {  // :33-67
    if (m.z <= R(0)) return R(0);
    R c2 = m.z * m.z;
    R be = ((m.x * m.x) / (au * au) + (m.y * m.y) / (av * av)) / c2;
    R res;
    if (type == kBeckmann) {
      res = std::exp(-be) / (R(float(kPi)) * au * av * c2 * c2);
    } else {
      R root = (R(1) + be) * c2;
      res = R(1) / (R(float(kPi)) * au * av * root * root);
    }
    if (res * m.z < R(1e-20f)) res = R(0);
    return res;
  }
  R pdf(V3<R> m) const { return eval(m) * m.z; }  // :236-240
  R project_roughness(V3<R> v) const {  // :242-250
    R inv_s2 = R(1) / sin_theta2(v);
    if (iso() || inv_s2 <= R(0)) return au;
    R c2 = v.x * v.x * inv_s2, s2 = v.y * v.y * inv_s2;
    return std::sqrt(c2 * au * au + s2 * av * av);
  }
  R smith_g1(V3<R> v, V3<R> m) const {  // :79-112
    if (dot(v, m) * v.z <= R(0)) return R(0);
    R tt = std::abs(tan_theta(v));
    if (tt == R(0)) return R(1);
    R alpha = project_roughness(v);
    if (type == kBeckmann) {
      R a = R(1) / (alpha * tt);
      if (a >= R(1.6f)) return R(1);
      R a2 = a * a;
      if (slip() == kSlipBeckmannG1) return (R(3.535f) * a) / (R(1) + R(2.276f) * a + R(2.577f) * a2);
      return (R(3.535f) * a + R(2.181f) * a2) / (R(1) + R(2.276f) * a + R(2.577f) * a2);
    }
    R root = alpha * tt;
    if (slip() == kSlipGgxG1) root = root * root;
    // Common.hpp:608-626 Hypot2(1, root)
    R h;
    if (R(1) > std::abs(root)) { R q = root / R(1); h = R(1) * std::sqrt(R(1) + q * q); }
    else if (root != R(0)) { R q = R(1) / root; h = std::abs(root) * std::sqrt(R(1) + q * q); }
    else h = R(0);
    return R(2) / (R(1) + h);
  }
  R g(V3<R> wi, V3<R> wo, V3<R> m) const { return smith_g1(wi, m) * smith_g1(wo, m); }
  V3<R> sample(R sx, R sy, R &pdf) const {  // :148-234 (SampleAll)
    R cos_m = 0, sin_phi, cos_phi, alpha_sqr;
    if (type == kBeckmann) {
      if (iso()) {
        R phi = (R(2) * R(float(kPi))) * sy;
        sin_phi = std::sin(phi); cos_phi = std::cos(phi);
        alpha_sqr = au * au;
      } else {
        R phi = std::atan(av / au * std::tan(R(float(kPi)) + R(2) * R(float(kPi)) * sy)) +
                R(float(kPi)) * std::floor(R(2) * sy + R(0.5f));
        sin_phi = std::sin(phi); cos_phi = std::cos(phi);
        R cs = cos_phi / au, ss = sin_phi / av;
        alpha_sqr = R(1) / (cs * cs + ss * ss);
      }
      R tan2 = alpha_sqr * -std::log(R(1) - sx);
      cos_m = R(1) / std::sqrt(R(1) + tan2);
      pdf = (R(1) - sx) / (R(float(kPi)) * au * av * cos_m * cos_m * cos_m);
    } else {
      if (iso()) {
        R phi = (R(2) * R(float(kPi))) * sy;
        sin_phi = std::sin(phi); cos_phi = std::cos(phi);
        alpha_sqr = au * au;
      } else {
        R phi = std::atan(av / au * std::tan(R(float(kPi)) + R(2) * R(float(kPi)) * sy)) +
                R(float(kPi)) * std::floor(R(2) * sy + R(0.5f));
        sin_phi = std::sin(phi); cos_phi = std::cos(phi);
        R cs = cos_phi / au, ss = sin_phi / av;
        alpha_sqr = R(1) / (cs * cs + ss * ss);
      }
      R tan2 = alpha_sqr * sx / (R(1) - sx);
      cos_m = R(1) / std::sqrt(R(1) + tan2);
      R tmp = R(1) + tan2 / alpha_sqr;
      pdf = R(float(kInvPi)) / (au * av * cos_m * cos_m * cos_m * tmp * tmp);
    }
    if (pdf < R(1e-20f)) pdf = R(0);
    R sin_m = std::sqrt(std::max(R(0), R(1) - cos_m * cos_m));
    return {sin_m * cos_phi, sin_m * sin_phi, cos_m};
  }
};

// ---------------------------------------------------------------- cubic interpolation
// src/core/Common.cpp:230-463
template <typename R>
inline R cubic1d(R x, const float *v, int n, R mn, R mx) {
  if (!(x >= mn && x <= mx)) return R(0);
  R t = ((x - mn) * R(n - 1)) / (mx - mn);
  int k = std::max(0, std::min(int(t), n - 2));
  R f0 = v[k], f1 = v[k + 1], d0, d1;
  d0 = (k > 0) ? R(0.5f) * (R(v[k + 1]) - R(v[k - 1])) : R(v[k + 1]) - R(v[k]);
  d1 = (k + 2 < n) ? R(0.5f) * (R(v[k + 2]) - R(v[k])) : R(v[k + 1]) - R(v[k]);
  t = t - R(k);
  R t2 = t * t, t3 = t2 * t;
  return (R(2) * t3 - R(3) * t2 + R(1)) * f0 + (R(-2) * t3 + R(3) * t2) * f1 + (t3 - R(2) * t2 + t) * d0 +
         (t3 - t2) * d1;
}
template <typename R>
inline bool cubic_weights(R p, int size, R mn, R mx, R w[4], int &knot) {
  if (!(p >= mn && p <= mx)) return false;
  R t = ((p - mn) * R(size - 1)) / (mx - mn);
  knot = std::min(int(t), size - 2);
  t = t - R(knot);
  R t2 = t * t, t3 = t2 * t;
  w[0] = R(0);
  w[1] = R(2) * t3 - R(3) * t2 + R(1);
  w[2] = R(-2) * t3 + R(3) * t2;
  w[3] = R(0);
  R d0 = t3 - R(2) * t2 + t, d1 = t3 - t2;
  if (knot > 0) { w[2] += R(0.5f) * d0; w[0] -= R(0.5f) * d0; }
  else { w[2] += d0; w[1] -= d0; }
  if (knot + 2 < size) { w[3] += R(0.5f) * d1; w[1] -= R(0.5f) * d1; }
  else { w[2] += d1; w[1] -= d1; }
  return true;
}
template <typename R>
inline R cubic2d(R px, R py, const float *v, int sx, int sy) {
  R wx[4], wy[4];
  int kx, ky;
  if (!cubic_weights(px, sx, R(0), R(1), wx, kx)) return R(0);
  if (!cubic_weights(py, sy, R(0), R(1), wy, ky)) return R(0);
  R res = 0;
  for (int y = -1; y <= 2; y++) {
    R wY = wy[y + 1];
    for (int x = -1; x <= 2; x++) {
      R wxy = wx[x + 1] * wY;
      if (wxy == R(0)) continue;
      res += R(v[(ky + y) * sx + kx + x]) * wxy;
    }
  }
  return res;
}
template <typename R>
inline R cubic3d(R px, R py, R pz, const float *v, int sx, int sy, int sz) {
  R wx[4], wy[4], wz[4];
  int kx, ky, kz;
  if (!cubic_weights(px, sx, R(0), R(1), wx, kx)) return R(0);
  if (!cubic_weights(py, sy, R(0), R(1), wy, ky)) return R(0);
  if (!cubic_weights(pz, sz, R(0), R(1), wz, kz)) return R(0);
  R res = 0;
  for (int z = -1; z <= 2; z++) {
    R wZ = wz[z + 1];
    for (int y = -1; y <= 2; y++) {
      R wyz = wy[y + 1] * wZ;
      for (int x = -1; x <= 2; x++) {
        R w = wx[x + 1] * wyz;
        if (w == R(0)) continue;
        res += R(v[((kz + z) * sy + (ky + y)) * sx + kx + x]) * w;
      }
    }
  }
  return res;
}

// ---------------------------------------------------------------- rough transmittance
// src/core/RoughTransmitance.cpp:9-340.  The reduced tables are kept as float,
// as in the reference (their entries are computed in R and stored as float).
struct RftFile {
  uint64_t eta_n = 0, alpha_n = 0, theta_n = 0, t_size = 0, d_size = 0;
  float eta_min = 0, eta_max = 0, alpha_min = 0, alpha_max = 0;
  std::vector<float> t, d;
  void load(const std::string &path) {  // :9-48
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open RFT data " + path);
    uint64_t h[5];
    float r[4];
    if (fread(h, 8, 5, f) != 5 || fread(r, 4, 4, f) != 4) { fclose(f); throw std::runtime_error("bad RFT header"); }
    eta_n = h[0]; alpha_n = h[1]; theta_n = h[2]; t_size = h[3]; d_size = h[4];
    eta_min = r[0]; eta_max = r[1]; alpha_min = r[2]; alpha_max = r[3];
    t.resize(t_size); d.resize(d_size);
    if (fread(t.data(), 4, t_size, f) != t_size || fread(d.data(), 4, d_size, f) != d_size) {
      fclose(f); throw std::runtime_error("short RFT data");
    }
    fclose(f);
  }
};
template <typename R>
struct Rft {
  int eta_n, alpha_n, theta_n;
  R eta_min, eta_max, alpha_min, alpha_max;
  bool eta_fixed = false, alpha_fixed = false;
  std::vector<float> t, d;
  explicit Rft(const RftFile &f)
      : eta_n(int(f.eta_n)), alpha_n(int(f.alpha_n)), theta_n(int(f.theta_n)), eta_min(f.eta_min),
        eta_max(f.eta_max), alpha_min(f.alpha_min), alpha_max(f.alpha_max), t(f.t), d(f.d) {}
  R warp(R v, R mn, R mx) const { return std::pow((v - mn) / (mx - mn), R(0.25f)); }
  void set_eta(R eta) {  // :170-232
    if (eta_fixed) return;
    const float *pt = t.data(), *pd = d.data();
    if ((eta < R(1)) != (slip() == kSlipRftDiffuse)) pd += size_t(eta_n) * alpha_n;
    if (eta < R(1)) { pt += size_t(eta_n) * alpha_n * theta_n; eta = R(1) / eta; }
    if (eta < eta_min) eta = eta_min;
    R we = warp(eta, eta_min, eta_max);
    std::vector<float> nt(size_t(alpha_n) * theta_n), nd(alpha_n);
    R da = R(1) / (R(alpha_n) - R(1)), dt = R(1) / (R(theta_n) - R(1));
    for (int i = 0; i < alpha_n; i++) {
      for (int j = 0; j < theta_n; j++)
        nt[size_t(i) * theta_n + j] = float(cubic3d(R(j) * dt, R(i) * da, we, pt, theta_n, alpha_n, eta_n));
      nd[i] = float(cubic2d(R(i) * da, we, pd, alpha_n, eta_n));
    }
    t.swap(nt); d.swap(nd);
    eta_fixed = true;
  }
  void set_alpha(R alpha) {  // :234-288
    if (!eta_fixed) throw std::runtime_error("SetAlpha(): needs a preceding call to SetEta()!");
    if (alpha_fixed) return;
    R wa = warp(alpha, alpha_min, alpha_max);
    std::vector<float> nt(theta_n), nd(1);
    R dt = R(1) / (R(theta_n) - R(1));
    for (int i = 0; i < theta_n; i++) nt[i] = float(cubic2d(R(i) * dt, wa, t.data(), theta_n, alpha_n));
    nd[0] = float(cubic1d(wa, d.data(), alpha_n, R(0), R(1)));
    t.swap(nt); d.swap(nd);
    alpha_fixed = true;
  }
  R eval(R cos_t, R alpha = 0, R eta = 0) const {  // :50-112
    R wc = std::pow(std::abs(cos_t), R(0.25f));
    R res;
    if (alpha_fixed && eta_fixed) {
      if (cos_t < R(0)) return R(0);
      res = cubic1d(wc, t.data(), theta_n, R(0), R(1));
    } else if (eta_fixed) {
      if (cos_t < R(0)) return R(0);
      res = cubic2d(wc, warp(alpha, alpha_min, alpha_max), t.data(), theta_n, alpha_n);
    } else {
      if (cos_t < R(0)) { cos_t = -cos_t; eta = R(1) / eta; }
      const float *p = t.data();
      if (eta < R(1)) { p += size_t(eta_n) * alpha_n * theta_n; eta = R(1) / eta; }
      if (eta < eta_min) eta = eta_min;
      res = cubic3d(wc, warp(alpha, alpha_min, alpha_max), warp(eta, eta_min, eta_max), p, theta_n, alpha_n,
                    eta_n);
    }
    return clampr(res, R(0), R(1));
  }
  R eval_diffuse(R alpha = 0, R eta = 0) const {  // :114-150
    R res;
    if (alpha_fixed && eta_fixed) res = d[0];
    else if (eta_fixed) res = cubic1d(warp(alpha, alpha_min, alpha_max), d.data(), alpha_n, R(0), R(1));
    else {
      const float *p = d.data();
      if ((eta < R(1)) != (slip() == kSlipRftDiffuse)) p += size_t(eta_n) * alpha_n;
      if (eta < R(1)) eta = R(1) / eta;
      if (eta < eta_min) eta = eta_min;
      res = cubic2d(warp(alpha, alpha_min, alpha_max), warp(eta, eta_min, eta_max), p, alpha_n, eta_n);
    }
    return clampr(res, R(0), R(1));
  }
  void check(R eta, R alpha_lo, R alpha_hi) const {  // :290-320
    if (eta < R(1)) eta = R(1) / eta;
    if (eta < eta_min || eta > eta_max) throw std::runtime_error("eta outside RFT range");
    if (alpha_lo < alpha_min || alpha_hi > alpha_max) throw std::runtime_error("alpha outside RFT range");
  }
};

// ---------------------------------------------------------------- BSDFs
// Flat material record shared with the HIP path (32 floats, see include/hikari_amd.h).
enum MatType {
  kDiffuse = 0, kConductor = 1, kDielectric = 2, kPlastic = 3, kRoughConductor = 4,
  kRoughDielectric = 5, kRoughPlastic = 6, kCoating = 7, kRoughCoating = 8,
  kMirror = 9, kMicrofacet = 10, kRoughDiffuse = 11, kTwoSided = 12
};
enum Measure { kUnknownMeasure = 0, kSolidAngle = 1, kDiscrete = 2 };

template <typename R>
struct BRec {  // include/core/BSDF.hpp:13-51 (Mode is always ERadiance on this path)
  V3<R> wi, wo;
  R eta = 1;
  int measure = kUnknownMeasure;
};

// Callback for the extra uniform RoughDielectric::Sample draws from the sampler
// (src/bsdf/RoughDielectricBSDF.cpp:91-101).
struct U1Source {
  virtual float next() = 0;
  virtual ~U1Source() {}
};

template <typename R>
struct Material {
  int type = kDiffuse;
  int dist = kBeckmann;
  bool nonlinear = false, aniso = false;
  const Material *nested = nullptr;
  const Material *nested_back = nullptr;  // twoSided: the BSDF seen from below
  bool fast = false;                      // roughDiffuse fastApprox
  V3<R> a, b;  // colour parameters (see include/hikari_amd.h for their meaning per type)
  R int_ior = 0, ext_ior = 0, alpha_u = 0, alpha_v = 0, thickness = 0;
  // derived (constructor + Activate() of each class)
  R eta = 1, inv_eta = 1, inv_eta2 = 1, spec_weight = 0, fdr_int = 0, fdr_ext = 0;
  V3<R> eta3, eta_k;
  std::vector<float> dummy;
  Rft<R> *ext = nullptr, *intd = nullptr;  // owned by the scene

  static constexpr int kStride = 32;
  // Fill from the flat 32-float record.  Parameter meaning mirrors the XML
  // properties read by each constructor in src/bsdf/*.cpp.
  void from_record(const float *r) {
    type = int(r[0]); dist = int(r[1]);
    int flags = int(r[3]);
    nonlinear = (flags & 1) != 0; aniso = (flags & 2) != 0; fast = (flags & 4) != 0;
    a = V3<R>(r[4], r[5], r[6]);
    b = V3<R>(r[8], r[9], r[10]);
    int_ior = r[20]; ext_ior = r[21]; alpha_u = r[22]; alpha_v = r[23]; thickness = r[24];
  }
  R clamp_alpha(R v) const { return clampr(v, R(5e-4f), R(1)); }
  // BSDF::IsDiffuse: true for DiffuseBSDF.cpp:71, PlasticBSDF.cpp:165,
  // RoughPlasticBSDF.cpp:198; the coatings forward to their nested BSDF
  // (CoatingBSDF.cpp:194, RoughCoatingBSDF.cpp:223); false otherwise
  // (the BSDF base class).
  // MicrofacetBSDF.cpp:167 true; TwoSidedBSDF.cpp:77-80 either side.
  bool is_diffuse() const {
    if (type == kDiffuse || type == kPlastic || type == kRoughPlastic || type == kMicrofacet) return true;
    if ((type == kCoating || type == kRoughCoating) && nested) return nested->is_diffuse();
    if (type == kTwoSided) return nested->is_diffuse() || nested_back->is_diffuse();
    return false;
  }
  // Constructor/Activate() derived quantities.
  void activate(const RftFile *beck, const RftFile *ggx, std::vector<Rft<R> *> &pool) {
    eta = int_ior / ext_ior;
    inv_eta = R(1) / eta;
    switch (type) {
      case kConductor: case kRoughConductor:  // ConductorBSDF.cpp:22-23, RoughConductorBSDF.cpp:47-48
        eta3 = V3<R>(int_ior / ext_ior);
        eta_k = b / ext_ior;
        if (type == kRoughConductor) {
          alpha_u = clamp_alpha(alpha_u);
          alpha_v = aniso ? clamp_alpha(alpha_v) : alpha_u;
        }
        break;
      case kRoughDielectric:  // RoughDielectricBSDF.cpp:36-47
        alpha_u = clamp_alpha(alpha_u);
        alpha_v = aniso ? clamp_alpha(alpha_v) : alpha_u;
        break;
      case kPlastic: {  // PlasticBSDF.cpp:25-29, 215-220
        inv_eta2 = inv_eta * inv_eta;
        fdr_int = approx_fresnel_diffuse_reflectance(inv_eta);
        fdr_ext = approx_fresnel_diffuse_reflectance(eta);
        R ks = luminance(a), kd = luminance(b);
        spec_weight = ks / (kd + ks);
        break;
      }
      case kRoughPlastic: {  // RoughPlasticBSDF.cpp:22-47, 261-287
        inv_eta2 = inv_eta * inv_eta;
        alpha_u = alpha_v = clamp_alpha(alpha_u);
        R ks = luminance(a), kd = luminance(b);
        spec_weight = ks / (kd + ks);
        const RftFile *src = dist == kGGX ? ggx : beck;
        ext = new Rft<R>(*src); pool.push_back(ext);
        ext->check(eta, alpha_u, alpha_u);
        intd = new Rft<R>(*src); pool.push_back(intd);
        ext->set_eta(eta);
        intd->set_eta(inv_eta);
        ext->set_alpha(alpha_u);
        break;
      }
      case kCoating: case kRoughCoating: {  // CoatingBSDF.cpp:207-213, RoughCoatingBSDF.cpp:293-306
        V3<R> tmp = vexp(a * (R(-2) * thickness));
        R avg = (tmp.x + tmp.y + tmp.z) / R(3);
        spec_weight = R(1) / (avg + R(1));
        if (type == kRoughCoating) {
          alpha_u = alpha_v = clamp_alpha(alpha_u);
          const RftFile *src = dist == kGGX ? ggx : beck;
          ext = new Rft<R>(*src); pool.push_back(ext);
          ext->check(eta, alpha_u, alpha_u);
          ext->set_eta(eta);
          ext->set_alpha(alpha_u);
        }
        break;
      }
      case kMicrofacet:  // MicrofacetBSDF.cpp:10-28: alpha clamped, eta = int/ext
        alpha_u = alpha_v = clamp_alpha(alpha_u);
        spec_weight = R(1) - std::max(a.x, std::max(a.y, a.z));  // Ks = 1 - max(kd)
        break;
      case kRoughDiffuse:  // RoughDiffuseBSDF.cpp:10-21
        alpha_u = alpha_v = clamp_alpha(alpha_u);
        break;
      default: break;
    }
  }

  // ---- Sample / Eval / Pdf dispatch
  V3<R> sample(BRec<R> &rec, R sx, R sy, U1Source *u1) const {
    switch (type) {
      case kDiffuse: {  // DiffuseBSDF.cpp:20-38
        if (cos_theta(rec.wi) <= R(0)) return V3<R>(R(0));
        rec.measure = kSolidAngle;
        rec.wo = square_to_cosine_hemisphere(sx, sy);
        rec.eta = R(1);
        return a;
      }
      case kConductor: {  // ConductorBSDF.cpp:32-49
        rec.measure = kDiscrete;
        R ci = cos_theta(rec.wi);
        if (ci <= R(0)) return V3<R>(R(0));
        V3<R> f = fresnel_conductor(ci, eta3, eta_k);
        rec.wo = reflect(rec.wi);
        rec.eta = R(1);
        return a * f;
      }
      case kDielectric: {  // DielectricBSDF.cpp:33-58
        rec.measure = kDiscrete;
        R ci = cos_theta(rec.wi), ct;
        R fr = fresnel_dielectric(ci, eta, inv_eta, ct);
        if (sx < fr) {
          rec.wo = reflect(rec.wi);
          rec.eta = R(1);
          return a;
        }
        rec.wo = refract(rec.wi, ct, eta, inv_eta);
        rec.eta = ct < R(0) ? eta : inv_eta;
        R factor = ct < R(0) ? inv_eta : eta;
        return b * (factor * factor);
      }
      case kPlastic: {  // PlasticBSDF.cpp:41-88
        R ci = cos_theta(rec.wi);
        if (ci <= R(0)) return V3<R>(R(0));
        R ct;
        R fi = fresnel_dielectric(ci, eta, inv_eta, ct);
        R sp = (fi * spec_weight) / (fi * spec_weight + (R(1) - fi) * (R(1) - spec_weight));
        rec.eta = R(1);
        if (sx < sp) {
          rec.measure = kDiscrete;
          rec.wo = reflect(rec.wi);
          return a * (slip() == kSlipPlasticFresnel ? R(1) - fi : fi) / sp;
        }
        rec.measure = kSolidAngle;
        rec.wo = square_to_cosine_hemisphere((sx - sp) / (R(1) - sp), sy);
        R fo = fresnel_dielectric(cos_theta(rec.wo), eta, inv_eta, ct);
        V3<R> diff = b;
        if (nonlinear) diff = diff / (V3<R>(R(1)) - diff * fdr_int);
        else diff = diff / (R(1) - fdr_int);
        return diff * ((slip() == kSlipPlasticEta2 ? R(1) : inv_eta2) * (R(1) - fi) * (R(1) - fo) / (R(1) - sp));
      }
      case kRoughConductor: {  // RoughConductorBSDF.cpp:54-90
        rec.measure = kSolidAngle;
        R ci = cos_theta(rec.wi);
        if (ci <= R(0)) return V3<R>(R(0));
        Microfacet<R> md(dist, alpha_u, alpha_v);
        R pdf;
        V3<R> m = md.sample(sx, sy, pdf);
        if (pdf == R(0)) return V3<R>(R(0));
        rec.wo = reflect(rec.wi, m);
        rec.eta = R(1);
        if (cos_theta(rec.wo) < R(0)) return V3<R>(R(0));
        R mdwi = dot(m, rec.wi);
        V3<R> f = fresnel_conductor(mdwi, eta3, eta_k);
        R gg = slip() == kSlipRoughConductorG ? R(1) : md.g(rec.wi, rec.wo, m);
        return a * f * md.eval(m) * gg * mdwi / (pdf * ci);
      }
      case kRoughDielectric: return sample_rough_dielectric(rec, sx, sy, u1);
      case kRoughPlastic: return sample_rough_plastic(rec, sx, sy);
      case kCoating: return sample_coating(rec, sx, sy, u1);
      case kRoughCoating: return sample_rough_coating(rec, sx, sy, u1);
      case kMirror: {  // MirrorBSDF.cpp:12-27
        if (cos_theta(rec.wi) <= R(0)) return V3<R>(R(0));
        rec.wo = reflect(rec.wi);
        rec.measure = kDiscrete;
        rec.eta = R(1);
        return V3<R>(R(1));
      }
      case kMicrofacet: return sample_microfacet(rec, sx, sy);
      case kRoughDiffuse: {  // RoughDiffuseBSDF.cpp:28-49
        R ci = cos_theta(rec.wi);
        rec.measure = kSolidAngle;
        if (ci <= R(0)) return V3<R>(R(0));
        rec.wo = square_to_cosine_hemisphere(sx, sy);
        rec.eta = R(1);
        R p = square_to_cosine_hemisphere_pdf(rec.wo);
        if (p == R(0)) return V3<R>(R(0));
        return eval_rough_diffuse(rec) / p;
      }
      case kTwoSided: {  // TwoSidedBSDF.cpp:21-43
        bool flip = false;
        if (cos_theta(rec.wi) < R(0)) {
          rec.wi.z = -rec.wi.z;
          flip = true;
        }
        V3<R> res = (flip ? nested_back : nested)->sample(rec, sx, sy, u1);
        if (flip) {
          rec.wi.z = -rec.wi.z;
          if (!is_zero_fuzzy(res)) rec.wo.z = -rec.wo.z;
        }
        return res;
      }
    }
    return V3<R>(R(0));
  }

  V3<R> eval(const BRec<R> &rec) const {
    switch (type) {
      case kDiffuse:  // DiffuseBSDF.cpp:40-52
        if (rec.measure != kSolidAngle || cos_theta(rec.wi) <= R(0) || cos_theta(rec.wo) <= R(0))
          return V3<R>(R(0));
        return a * R(float(kInvPi));
      case kConductor: {  // ConductorBSDF.cpp:51-62
        R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
        if (ci > R(0) && co > R(0) && rec.measure == kDiscrete &&
            std::abs(dot(reflect(rec.wi), rec.wo) - R(1)) <= R(1e-3f))
          return a * fresnel_conductor(ci, eta3, eta_k);
        return V3<R>(R(0));
      }
      case kDielectric: {  // DielectricBSDF.cpp:60-84
        R ci = cos_theta(rec.wi), co = cos_theta(rec.wo), ct;
        R fr = fresnel_dielectric(ci, eta, inv_eta, ct);
        if (ci * co >= R(0)) {
          if (std::abs(dot(reflect(rec.wi), rec.wo) - R(1)) <= R(1e-3f)) return a * fr;
        } else {
          if (std::abs(dot(refract(rec.wi, ct, eta, inv_eta), rec.wo) - R(1)) <= R(1e-3f)) {
            R factor = ct < R(0) ? inv_eta : eta;
            return b * (factor * factor) * (R(1) - fr);
          }
        }
        return V3<R>(R(0));
      }
      case kPlastic: {  // PlasticBSDF.cpp:90-128
        R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
        if (rec.measure == kUnknownMeasure || ci <= R(0) || co <= R(0)) return V3<R>(R(0));
        R ct;
        R fi = fresnel_dielectric(ci, eta, inv_eta, ct);
        if (rec.measure == kDiscrete) {
          if (std::abs(dot(reflect(rec.wi), rec.wo) - R(1)) <= R(1e-3f)) return a * fi;
          return V3<R>(R(0));
        }
        R fo = fresnel_dielectric(co, eta, inv_eta, ct);
        V3<R> diff = b;
        if (nonlinear) diff = diff / (V3<R>(R(1)) - diff * fdr_int);
        else diff = diff / (R(1) - fdr_int);
        return diff * (square_to_cosine_hemisphere_pdf(rec.wo) * inv_eta2 * (R(1) - fi) * (R(1) - fo));
      }
      case kRoughConductor: {  // RoughConductorBSDF.cpp:92-126
        R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
        if (rec.measure != kSolidAngle || ci <= R(0) || co <= R(0)) return V3<R>(R(0));
        V3<R> h = normalized(rec.wi + rec.wo);
        Microfacet<R> md(dist, alpha_u, alpha_v);
        R dd = md.eval(h);
        if (dd == R(0)) return V3<R>(R(0));
        V3<R> f = fresnel_conductor(dot(h, rec.wi), eta3, eta_k) * a;
        R gg = md.g(rec.wi, rec.wo, h);
        return dd * gg * f / (R(4) * ci);
      }
      case kRoughDielectric: return eval_rough_dielectric(rec);
      case kRoughPlastic: return eval_rough_plastic(rec);
      case kCoating: return eval_coating(rec);
      case kRoughCoating: return eval_rough_coating(rec);
      case kMirror:  // MirrorBSDF.cpp:29-41
        if (cos_theta(rec.wi) > R(0) && cos_theta(rec.wo) > R(0) && rec.measure == kDiscrete &&
            std::abs(dot(reflect(rec.wi), rec.wo) - R(1)) <= R(1e-3f))
          return V3<R>(R(1));
        return V3<R>(R(0));
      case kMicrofacet: {  // MicrofacetBSDF.cpp:96-135
        if (rec.measure != kSolidAngle || cos_theta(rec.wi) <= R(0) || cos_theta(rec.wo) <= R(0))
          return V3<R>(R(0));
        return microfacet_value(rec);
      }
      case kRoughDiffuse: return eval_rough_diffuse(rec);
      case kTwoSided: {  // TwoSidedBSDF.cpp:45-59
        BRec<R> c = rec;
        if (cos_theta(c.wi) > R(0)) return nested->eval(c);
        c.wi.z = -c.wi.z;
        c.wo.z = -c.wo.z;
        return nested_back->eval(c);
      }
    }
    return V3<R>(R(0));
  }

  R pdf(const BRec<R> &rec) const {
    switch (type) {
      case kDiffuse:  // DiffuseBSDF.cpp:54-69
        if (rec.measure != kSolidAngle || cos_theta(rec.wi) <= R(0) || cos_theta(rec.wo) <= R(0)) return R(0);
        return R(float(kInvPi)) * cos_theta(rec.wo);
      case kConductor: {  // ConductorBSDF.cpp:64-75
        R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
        if (ci > R(0) && co > R(0) && rec.measure == kDiscrete &&
            std::abs(dot(reflect(rec.wi), rec.wo) - R(1)) <= R(1e-3f))
          return R(1);
        return R(0);
      }
      case kDielectric: {  // DielectricBSDF.cpp:86-110
        R ci = cos_theta(rec.wi), co = cos_theta(rec.wo), ct;
        R fr = fresnel_dielectric(ci, eta, inv_eta, ct);
        if (ci * co >= R(0)) {
          if (std::abs(dot(reflect(rec.wi), rec.wo) - R(1)) <= R(1e-3f)) return fr;
        } else {
          if (std::abs(dot(refract(rec.wi, ct, eta, inv_eta), rec.wo) - R(1)) <= R(1e-3f)) return R(1) - fr;
        }
        return R(0);
      }
      case kPlastic: {  // PlasticBSDF.cpp:130-163
        R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
        if (rec.measure == kUnknownMeasure || ci <= R(0) || co <= R(0)) return R(0);
        R ct;
        R fi = fresnel_dielectric(ci, eta, inv_eta, ct);
        R sp = (fi * spec_weight) / (fi * spec_weight + (R(1) - fi) * (R(1) - spec_weight));
        if (rec.measure == kDiscrete) {
          if (std::abs(dot(reflect(rec.wi), rec.wo) - R(1)) <= R(1e-3f)) return R(1) - sp;
          return R(0);
        }
        return square_to_cosine_hemisphere_pdf(rec.wo) * (R(1) - sp);
      }
      case kRoughConductor: {  // RoughConductorBSDF.cpp:128-154
        R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
        if (rec.measure != kSolidAngle || ci <= R(0) || co <= R(0)) return R(0);
        V3<R> h = normalized(rec.wi + rec.wo);
        Microfacet<R> md(dist, alpha_u, alpha_v);
        return md.pdf(h) / (R(4) * std::abs(dot(h, rec.wo)));
      }
      case kRoughDielectric: return pdf_rough_dielectric(rec);
      case kRoughPlastic: return pdf_rough_plastic(rec);
      case kCoating: return pdf_coating(rec);
      case kRoughCoating: return pdf_rough_coating(rec);
      case kMirror:  // MirrorBSDF.cpp:43-55
        if (cos_theta(rec.wi) > R(0) && cos_theta(rec.wo) > R(0) && rec.measure == kDiscrete &&
            std::abs(dot(reflect(rec.wi), rec.wo) - R(1)) <= R(1e-3f))
          return R(1);
        return R(0);
      case kMicrofacet: {  // MicrofacetBSDF.cpp:137-165
        if (rec.measure != kSolidAngle || cos_theta(rec.wi) <= R(0) || cos_theta(rec.wo) <= R(0)) return R(0);
        R ks = spec_weight;
        V3<R> wh = normalized(rec.wi + rec.wo);
        R j = R(0.25f) / dot(wh, rec.wo);
        R spec = ks * square_to_beckmann_pdf(wh, alpha_u) * j;
        R diff = (R(1) - ks) * square_to_cosine_hemisphere_pdf(rec.wo);
        return spec + diff;
      }
      case kRoughDiffuse: {  // RoughDiffuseBSDF.cpp:161-172
        if (cos_theta(rec.wi) <= R(0) || cos_theta(rec.wo) <= R(0) || rec.measure != kSolidAngle) return R(0);
        return square_to_cosine_hemisphere_pdf(rec.wo);
      }
      case kTwoSided: {  // TwoSidedBSDF.cpp:61-75
        BRec<R> c = rec;
        if (cos_theta(c.wi) > R(0)) return nested->pdf(c);
        c.wi.z = -c.wi.z;
        c.wo.z = -c.wo.z;
        return nested_back->pdf(c);
      }
    }
    return R(0);
  }

  // ---- MicrofacetBSDF (src/bsdf/MicrofacetBSDF.cpp): Beckmann specular lobe
  // weighted by Ks = 1 - max(kd) over a Lambertian base.
  static R beckmann_d(V3<R> m, R alpha) {  // :240-246
    R expon = tan_theta(m) / alpha;
    R c2 = cos_theta(m) * cos_theta(m);
    return std::exp(-expon * expon) / (R(float(kPi)) * alpha * alpha * c2 * c2);
  }
  static R smith_beckmann_g1(V3<R> v, V3<R> m, R alpha) {  // :248-272
    R tt = tan_theta(v);
    if (tt == R(0)) return R(1);
    if (dot(m, v) * cos_theta(v) <= R(0)) return R(0);
    R bb = R(1) / (alpha * tt);
    if (bb >= R(1.6f)) return R(1);
    R b2 = bb * bb;
    return (R(3.535f) * bb + R(2.181f) * b2) / (R(1) + R(2.276f) * bb + R(2.577f) * b2);
  }
  // Ks F D G / (4 cos_i cos_o) (zeroed when not a valid colour) + kd / pi.
  V3<R> microfacet_value(const BRec<R> &rec) const {
    R alpha = alpha_u, ks = spec_weight, ct;
    V3<R> wh = normalized(rec.wi + rec.wo);
    R d = beckmann_d(wh, alpha);
    R f = fresnel_dielectric(dot(wh, rec.wi), eta, inv_eta, ct);
    R g = smith_beckmann_g1(rec.wi, wh, alpha) * smith_beckmann_g1(rec.wo, wh, alpha);
    R spec = ks * f * d * g / (R(4) * cos_theta(rec.wi) * cos_theta(rec.wo));
    if (spec < R(0) || !std::isfinite(spec)) spec = R(0);
    return V3<R>(spec) + a / R(float(kPi));
  }
  V3<R> sample_microfacet(BRec<R> &rec, R sx, R sy) const {  // :34-94
    if (cos_theta(rec.wi) <= R(0)) return V3<R>(R(0));
    rec.measure = kSolidAngle;
    R ks = spec_weight;
    if (sx < ks) {
      V3<R> wh = square_to_beckmann(sx / ks, sy, alpha_u);
      rec.wo = reflect(rec.wi, wh);
    } else {
      rec.wo = square_to_cosine_hemisphere((sx - ks) / (R(1) - ks), sy);
    }
    rec.eta = R(1);
    R p = pdf(rec);
    if (p == R(0)) return V3<R>(R(0));
    return microfacet_value(rec) * cos_theta(rec.wo) / p;
  }

  // ---- RoughDiffuseBSDF::Eval (src/bsdf/RoughDiffuseBSDF.cpp:51-159): the
  // Oren-Nayar model; the returned value already carries the cos(theta_o)
  // factor (the integrators multiply by |cos theta_o| once more).
  V3<R> eval_rough_diffuse(const BRec<R> &rec) const {
    R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
    if (ci <= R(0) || co <= R(0) || rec.measure != kSolidAngle) return V3<R>(R(0));
    R sigma = alpha_u * R(0.70710678118655f);
    R s2 = sigma * sigma;
    R sti = sin_theta(rec.wi), sto = sin_theta(rec.wo);
    R cpd = R(0);
    if (sti > R(1e-4f) && sto > R(1e-4f))
      cpd = cos_phi(rec.wi) * cos_phi(rec.wo) + sin_phi(rec.wi) * sin_phi(rec.wo);
    const R inv_pi = R(float(kInvPi));
    if (fast) {
      R A = R(1) - R(0.5f) * s2 / (s2 + R(0.33f));
      R B = R(0.45f) * s2 / (s2 + R(0.09f));
      R sa, tb;
      if (ci > co) { sa = sto; tb = sti / ci; }
      else { sa = sti; tb = sto / co; }
      return a * (inv_pi * co * (A + B * std::max(cpd, R(0)) * sa * tb));
    }
    R ti = std::acos(clampr(ci, R(-1), R(1))), to = std::acos(clampr(co, R(-1), R(1)));
    R al = std::max(ti, to), be = std::min(ti, to);
    R sa, sb, tb;
    if (ci > co) { sa = sto; sb = sti; tb = sti / ci; }
    else { sa = sti; sb = sto; tb = sto / co; }
    R t1 = s2 / (s2 + R(0.09f));
    R t2 = R(float(4.0 * kInvPi * kInvPi)) * al * be;
    R t3 = R(2) * be * inv_pi;
    R c1 = R(1) - R(0.5f) * s2 / (s2 + R(0.33f));
    R c2 = R(0.45f) * t1;
    R c3 = R(0.125f) * t1 * t2 * t2;
    R c4 = R(0.17f) * s2 / (s2 + R(0.13f));
    if (cpd > R(0)) c2 *= sa;
    else c2 *= sa - t3 * t3 * t3;
    R th = (sa + sb) / (std::sqrt(std::max(R(0), R(1) - sa * sa)) + std::sqrt(std::max(R(0), R(1) - sb * sb)));
    if (std::isinf(th) || std::isnan(th)) return V3<R>(R(0));
    V3<R> lr1 = a * (c1 + cpd * c2 * tb + (R(1) - std::abs(cpd)) * c3 * th);
    V3<R> lr2 = a * a * (c4 * (R(1) - cpd * t3 * t3));
    return (lr1 + lr2) * (inv_pi * co);
  }

  // src/bsdf/RoughDielectricBSDF.cpp:57-157
  V3<R> sample_rough_dielectric(BRec<R> &rec, R sx, R sy, U1Source *u1) const {
    rec.measure = kSolidAngle;
    R ci = cos_theta(rec.wi);
    if (ci == R(0)) return V3<R>(R(0));
    Microfacet<R> md(dist, alpha_u, alpha_v);
    Microfacet<R> sd(md);
    sd.scale_alpha(R(1.2f) - R(0.2f) * std::sqrt(std::abs(ci)));
    R pdf;
    V3<R> m = sd.sample(sx, sy, pdf);
    if (pdf == R(0)) return V3<R>(R(0));
    R ct;
    R f = fresnel_dielectric(dot(rec.wi, m), eta, inv_eta, ct);
    R s1 = u1 ? R(u1->next()) : R(Pcg32().next_float());
    bool refl = !(s1 > f);
    V3<R> w(R(1));
    if (refl) {
      rec.wo = reflect(rec.wi, m);
      rec.eta = R(1);
      R co = cos_theta(rec.wo);
      if (co * ci <= R(0)) return V3<R>(R(0));
      w = w * a;
    } else {
      if (ct == R(0)) return V3<R>(R(0));
      rec.wo = refract(rec.wi, m, ct, eta, inv_eta);
      rec.eta = ct < R(0) ? eta : inv_eta;
      R co = cos_theta(rec.wo);
      if (co * ci >= R(0)) return V3<R>(R(0));
      R factor = ct < R(0) ? inv_eta : eta;
      if (slip() == kSlipRoughDielectricEta) factor = R(1);
      w = w * (b * (factor * factor));
    }
    R s = std::abs(md.eval(m) * md.g(rec.wi, rec.wo, m) * dot(rec.wi, m) / (pdf * ci));
    return w * s;
  }
  // src/bsdf/RoughDielectricBSDF.cpp:159-221
  V3<R> eval_rough_dielectric(const BRec<R> &rec) const {
    R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
    if (rec.measure != kSolidAngle || ci == R(0)) return V3<R>(R(0));
    bool refl = ci * co > R(0);
    V3<R> h;
    if (refl) h = normalized(rec.wi + rec.wo);
    else {
      R e = ci > R(0) ? eta : inv_eta;
      h = normalized(rec.wi + rec.wo * e);
    }
    h = h * signum(cos_theta(h));
    Microfacet<R> md(dist, alpha_u, alpha_v);
    R dd = md.eval(h);
    if (dd == R(0)) return V3<R>(R(0));
    R ct;
    R f = fresnel_dielectric(dot(rec.wi, h), eta, inv_eta, ct);
    R gg = md.g(rec.wi, rec.wo, h);
    if (refl) return (f * dd * gg / (R(4) * std::abs(ci))) * a;
    R e = ci > R(0) ? eta : inv_eta;
    R sd = dot(rec.wi, h) + e * dot(rec.wo, h);
    R value = ((R(1) - f) * dd * gg * e * e * dot(rec.wi, h) * dot(rec.wo, h)) / (ci * sd * sd);
    R factor = ci < R(0) ? inv_eta : eta;
    return std::abs(value * factor * factor) * b;
  }
  // src/bsdf/RoughDielectricBSDF.cpp:223-281
  R pdf_rough_dielectric(const BRec<R> &rec) const {
    if (rec.measure != kSolidAngle) return R(0);
    R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
    if (ci == R(0) || co == R(0)) return R(0);
    bool refl = ci * co > R(0);
    V3<R> h;
    R j;
    if (refl) {
      h = normalized(rec.wi + rec.wo);
      j = R(1) / (R(4) * dot(rec.wo, h));
    } else {
      R e = ci > R(0) ? eta : inv_eta;
      h = normalized(rec.wi + rec.wo * e);
      R sd = dot(rec.wi, h) + e * dot(rec.wo, h);
      j = (e * e * dot(rec.wo, h)) / (sd * sd);
    }
    h = h * signum(cos_theta(h));
    Microfacet<R> md(dist, alpha_u, alpha_v);
    md.scale_alpha(R(1.2f) - R(0.2f) * std::sqrt(std::abs(ci)));
    R ct;
    R f = fresnel_dielectric(dot(rec.wi, h), eta, inv_eta, ct);
    R p = md.pdf(h) * (refl ? f : (R(1) - f));
    return std::abs(p * j);
  }

  // src/bsdf/RoughPlasticBSDF.cpp:66-123
  V3<R> sample_rough_plastic(BRec<R> &rec, R sx, R sy) const {
    R ci = cos_theta(rec.wi);
    if (ci <= R(0)) return V3<R>(R(0));
    Microfacet<R> md(dist, alpha_u, alpha_u);
    R sp = R(1) - ext->eval(ci, md.au);
    sp = (sp * spec_weight) / (sp * spec_weight + (R(1) - sp) * (R(1) - spec_weight));
    R dp = R(1) - sp;
    bool spec;
    if (sy < sp) { sy /= sp; spec = true; }
    else { sy = (sy - sp) / (R(1) - sp); spec = false; }
    if (spec) {
      R pdf;
      V3<R> m = md.sample(sx, sy, pdf);
      rec.wo = reflect(rec.wi, m);
      rec.measure = kDiscrete;
    } else {
      rec.wo = square_to_cosine_hemisphere(sx, sy);
      rec.measure = kSolidAngle;
    }
    rec.eta = R(1);
    V3<R> h = normalized(rec.wi + rec.wo);
    R j = R(1) / (R(4) * dot(rec.wo, h));
    R p = md.pdf(h) * j * sp + dp * square_to_cosine_hemisphere_pdf(rec.wo);
    if (p == R(0)) return V3<R>(R(0));
    return eval_rough_plastic(rec) / p;
  }
  // src/bsdf/RoughPlasticBSDF.cpp:125-167 (note the reference's G(Wi, Wi, H))
  V3<R> eval_rough_plastic(const BRec<R> &rec) const {
    R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
    if (rec.measure == kUnknownMeasure || ci <= R(0) || co <= R(0)) return V3<R>(R(0));
    Microfacet<R> md(dist, alpha_u, alpha_u);
    V3<R> res(R(0));
    R ct;
    V3<R> h = normalized(rec.wi + rec.wo);
    R dd = md.eval(h);
    R f = slip() == kSlipRoughPlasticEta ? fresnel_dielectric(dot(rec.wi, h), inv_eta, eta, ct)
                                         : fresnel_dielectric(dot(rec.wi, h), eta, inv_eta, ct);
    R gg = md.g(rec.wi, rec.wi, h);
    R value = f * dd * gg / (R(4) * ci);
    res = res + a * value;
    V3<R> diff = b;
    R t12 = ext->eval(ci, md.au);
    R t21 = ext->eval(co, md.au);
    R fdr = R(1) - intd->eval_diffuse(md.au);
    if (nonlinear) diff = diff / (V3<R>(R(1)) - diff * fdr);
    else diff = diff / (R(1) - fdr);
    res = res + diff * (R(float(kInvPi)) * co * t12 * t21 * inv_eta2);
    return res;
  }
  // src/bsdf/RoughPlasticBSDF.cpp:169-196
  R pdf_rough_plastic(const BRec<R> &rec) const {
    R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
    if (rec.measure == kUnknownMeasure || ci <= R(0) || co <= R(0)) return R(0);
    Microfacet<R> md(dist, alpha_u, alpha_u);
    V3<R> h = normalized(rec.wi + rec.wo);
    R sp = R(1) - ext->eval(ci, md.au);
    sp = (sp * spec_weight) / (sp * spec_weight + (R(1) - sp) * (R(1) - spec_weight));
    R dp = R(1) - sp;
    R j = R(1) / (R(4) * dot(rec.wo, h));
    R res = md.pdf(h) * j * sp;
    res += dp * square_to_cosine_hemisphere_pdf(rec.wo);
    return res;
  }

  // src/bsdf/CoatingBSDF.cpp:290-302
  V3<R> coat_refract_in(V3<R> wi, R &refl) const {
    R ct;
    refl = fresnel_dielectric(std::abs(cos_theta(wi)), eta, inv_eta, ct);
    return {inv_eta * wi.x, inv_eta * wi.y, -signum(cos_theta(wi)) * ct};
  }
  V3<R> coat_refract_out(V3<R> wi, R &refl) const {
    R ct;
    refl = slip() == kSlipCoatingExit ? fresnel_dielectric(std::abs(cos_theta(wi)), eta, inv_eta, ct)
                                      : fresnel_dielectric(std::abs(cos_theta(wi)), inv_eta, eta, ct);
    return {eta * wi.x, eta * wi.y, -signum(cos_theta(wi)) * ct};
  }
  // src/bsdf/CoatingBSDF.cpp:39-104
  V3<R> sample_coating(BRec<R> &rec, R sx, R sy, U1Source *u1) const {
    R r12;
    V3<R> wip = coat_refract_in(rec.wi, r12);
    R sp = (r12 * spec_weight) / (r12 * spec_weight + (R(1) - r12) * (R(1) - spec_weight));
    bool spec;
    if (sx < sp) { spec = true; sx /= sp; }
    else { spec = false; sx = (sx - sp) / (R(1) - sp); }
    if (spec) {
      rec.wo = reflect(rec.wi);
      rec.eta = R(1);
      rec.measure = kDiscrete;
      return b * (r12 / sp);
    }
    if (r12 == R(1)) return V3<R>(R(0));
    V3<R> wib = rec.wi;
    rec.wi = wip;
    V3<R> res = nested->sample(rec, sx, sy, u1);
    rec.wi = wib;
    if (is_zero_fuzzy(res)) return V3<R>(R(0));
    V3<R> wop = rec.wo;
    V3<R> sig = a * thickness;
    if (!is_zero_fuzzy(sig))
      res = res * vexp(-sig * (R(1) / std::abs(cos_theta(wip)) +
                               (slip() == kSlipCoatingSigma ? R(0) : R(1) / std::abs(cos_theta(wop)))));
    R r21;
    rec.wo = coat_refract_out(wop, r21);
    if (r21 == R(1)) return V3<R>(R(0));
    res = res / (R(1) - sp);
    res = res * ((R(1) - r12) * (R(1) - r21));
    return res;
  }
  // src/bsdf/CoatingBSDF.cpp:106-141
  V3<R> eval_coating(const BRec<R> &rec) const {
    if (rec.measure == kDiscrete) {
      if (std::abs(dot(reflect(rec.wi), rec.wo) - R(1)) <= R(1e-3f)) {
        R ct;
        return b * fresnel_dielectric(std::abs(cos_theta(rec.wi)), eta, inv_eta, ct);
      }
      return V3<R>(R(0));
    }
    R r12, r21;
    BRec<R> ri = rec;
    ri.wi = coat_refract_in(rec.wi, r12);
    ri.wo = coat_refract_in(rec.wo, r21);
    if (r12 == R(1) || r21 == R(1)) return V3<R>(R(0));
    V3<R> res = nested->eval(ri) * (R(1) - r12) * (R(1) - r21);
    V3<R> sig = a * thickness;
    if (!is_zero_fuzzy(sig))
      res = res * vexp(-sig * (R(1) / std::abs(cos_theta(ri.wi)) +
                               (slip() == kSlipCoatingSigma ? R(0) : R(1) / std::abs(cos_theta(ri.wo)))));
    if (rec.measure == kSolidAngle)
      res = res * (inv_eta * inv_eta * cos_theta(rec.wo) / cos_theta(ri.wo));
    return res;
  }
  // src/bsdf/CoatingBSDF.cpp:143-174
  R pdf_coating(const BRec<R> &rec) const {
    R r12;
    V3<R> wip = coat_refract_in(rec.wi, r12);
    R sp = (r12 * spec_weight) / (r12 * spec_weight + (R(1) - r12) * (R(1) - spec_weight));
    if (rec.measure == kDiscrete) {
      if (std::abs(dot(reflect(rec.wi), rec.wo) - R(1)) <= R(1e-3f)) return sp;
      return R(0);
    }
    R r21;
    BRec<R> ri = rec;
    ri.wi = wip;
    ri.wo = coat_refract_in(rec.wo, r21);
    if (r12 == R(1) || r21 == R(1)) return R(0);
    R p = nested->pdf(ri);
    if (rec.measure == kSolidAngle) p *= (inv_eta * inv_eta * cos_theta(rec.wo) / cos_theta(ri.wo));
    return p * (R(1) - sp);
  }

  // src/bsdf/RoughCoatingBSDF.cpp:342-376
  V3<R> rc_refract(V3<R> wi, R ie) const {
    R ci = cos_theta(wi);
    bool entering = ci > R(0);
    R s2 = ie * ie * sin_theta2(wi);
    if (s2 >= R(1)) return V3<R>(R(0));
    R ct = std::sqrt(R(1) - s2);
    return {ie * wi.x, ie * wi.y, entering ? ct : -ct};
  }
  // src/bsdf/RoughCoatingBSDF.cpp:62-127
  V3<R> sample_rough_coating(BRec<R> &rec, R sx, R sy, U1Source *u1) const {
    Microfacet<R> md(dist, alpha_u, alpha_u);
    R aci = std::abs(cos_theta(rec.wi));
    R r12 = R(1) - ext->eval(aci, md.au);
    R sp = (r12 * spec_weight) / (r12 * spec_weight + (R(1) - r12) * (R(1) - spec_weight));
    bool spec;
    if (sx < sp) { spec = true; sx /= sp; }
    else { spec = false; sx = (sx - sp) / (R(1) - sp); }
    if (spec) {
      R pdf;
      V3<R> m = md.sample(sx, sy, pdf);
      rec.wo = reflect(rec.wi, m);
      rec.measure = kDiscrete;
      rec.eta = R(1);
      if (cos_theta(rec.wo) * cos_theta(rec.wo) <= R(0)) return V3<R>(R(0));
    } else {
      V3<R> wib = rec.wi;
      rec.wi = rc_refract(rec.wi, inv_eta);
      V3<R> res = nested->sample(rec, sx, sy, u1);
      rec.wi = wib;
      if (is_zero_fuzzy(res)) return V3<R>(R(0));
      rec.wo = rc_refract(rec.wo, eta);
      if (is_zero_fuzzy(rec.wo)) return V3<R>(R(0));
    }
    R p = pdf_rough_coating(rec);
    if (p == R(0)) return V3<R>(R(0));
    return eval_rough_coating(rec) / p;
  }
  // src/bsdf/RoughCoatingBSDF.cpp:129-184
  V3<R> eval_rough_coating(const BRec<R> &rec) const {
    R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
    if (ci * co <= R(0)) return V3<R>(R(0));
    V3<R> res(R(0));
    Microfacet<R> md(dist, alpha_u, alpha_u);
    V3<R> h = normalized(rec.wi + rec.wo) * signum(cos_theta(rec.wo));
    R ct;
    R dd = md.eval(h);
    R f = fresnel_dielectric(std::abs(dot(rec.wi, h)), eta, inv_eta, ct);
    R gg = md.g(rec.wi, rec.wo, h);
    V3<R> spec = b * f * dd * gg / (R(4) * std::abs(ci));
    res = res + spec;
    BRec<R> ri = rec;
    ri.wi = rc_refract(rec.wi, inv_eta);
    ri.wo = rc_refract(rec.wo, inv_eta);
    V3<R> nv = nested->eval(ri) * ext->eval(std::abs(ci), md.au) *
               ext->eval(std::abs(slip() == kSlipRoughCoatingT ? ci : co), md.au);
    V3<R> sig = a * thickness;
    if (!is_zero_fuzzy(sig))
      nv = nv * vexp(-sig * (R(1) / std::abs(cos_theta(ri.wi)) + R(1) / std::abs(cos_theta(ri.wo))));
    if (rec.measure == kSolidAngle) nv = nv * (inv_eta * inv_eta * cos_theta(rec.wo) / cos_theta(ri.wo));
    res = res + nv;
    return res;
  }
  // src/bsdf/RoughCoatingBSDF.cpp:186-221
  R pdf_rough_coating(const BRec<R> &rec) const {
    V3<R> h = normalized(rec.wi + rec.wo) * signum(cos_theta(rec.wo));
    Microfacet<R> md(dist, alpha_u, alpha_u);
    R ci = cos_theta(rec.wi), co = cos_theta(rec.wo);
    R aci = std::abs(ci);
    R r12 = R(1) - ext->eval(aci, md.au);
    R sp = (r12 * spec_weight) / (r12 * spec_weight + (R(1) - r12) * (R(1) - spec_weight));
    R np = R(1) - sp;
    R res = R(0);
    if (ci * co > R(0)) {
      R j = R(1) / (R(4) * std::abs(dot(rec.wo, h)));
      res = md.pdf(h) * j * sp;
    }
    BRec<R> ri = rec;
    ri.wi = rc_refract(rec.wi, inv_eta);
    ri.wo = rc_refract(rec.wo, inv_eta);
    R p = nested->pdf(ri);
    if (ri.measure == kSolidAngle) p *= inv_eta * inv_eta * co / cos_theta(ri.wo);
    res += p * np;
    return res;
  }
};

// Links nested BSDFs by record index: [2] the nested BSDF of a coating, or
// the upper-side BSDF of a twoSided; [12] a twoSided's lower-side BSDF (-1:
// the same as the upper, TwoSidedBSDF.cpp:106-112).  A twoSided may not nest
// a transmissive BSDF (:113-122) or another twoSided.
template <typename R>
inline void link_materials(std::vector<Material<R>> &m, const float *recs) {
  int n = int(m.size());
  auto at = [&](int i, int k) -> const Material<R> * {
    if (k < 0 || k >= n) throw std::runtime_error("nested BSDF index out of range");
    return &m[k];
  };
  for (int i = 0; i < n; i++) {
    int k = int(recs[i * 32 + 2]), kb = int(recs[i * 32 + 12]);
    if (m[i].type == kCoating || m[i].type == kRoughCoating) {
      m[i].nested = at(i, k);
    } else if (m[i].type == kTwoSided) {
      m[i].nested = at(i, k);
      m[i].nested_back = kb < 0 ? m[i].nested : at(i, kb);
      for (int c : {k, kb < 0 ? k : kb}) {
        int t = m[c].type;
        if (t == kTwoSided) throw std::runtime_error("twoSided cannot nest a twoSided BSDF");
        if (t == kCoating || t == kRoughCoating) t = m[at(c, int(recs[c * 32 + 2])) - m.data()].type;
        if (t == kDielectric || t == kRoughDielectric)
          throw std::runtime_error("Only BSDF without transmission can be nested!");
      }
    }
  }
}

}  // namespace hko

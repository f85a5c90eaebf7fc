// hk_math.h -- FP32 math of the path-tracing core: vectors, pcg32 streams,
// Moller-Trumbore, frames, Fresnel, microfacet distributions, cubic
// interpolation of rough-transmittance tables and the nine BSDFs of the
// pathMIS hot path.  Compiled for gfx950 (device) and for the host where the
// runtime needs the same routine (table reduction, host-side checks).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "hk_shared.h"

#define HK_HD __host__ __device__ __forceinline__

namespace hk {

constexpr float kPiF = 3.14159265358979323846f;
constexpr float kInvPiF = 0.31830988618379067154f;

struct vec3 {
  float x, y, z;
};
HK_HD vec3 v3(float a, float b, float c) { return vec3{a, b, c}; }
HK_HD vec3 v3(float a) { return vec3{a, a, a}; }
HK_HD vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
HK_HD vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
HK_HD vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
HK_HD vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
HK_HD vec3 operator/(vec3 a, vec3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
HK_HD vec3 operator*(float s, vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
HK_HD vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
HK_HD vec3 operator/(vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
// Reductions in the reference's order (Eigen: x0 + (x1 + x2)).
HK_HD float dot(vec3 a, vec3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
HK_HD vec3 cross(vec3 a, vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
HK_HD float sqnorm(vec3 a) { return a.x * a.x + (a.y * a.y + a.z * a.z); }
HK_HD float norm(vec3 a) { return sqrtf(sqnorm(a)); }
HK_HD vec3 normalized(vec3 a) {
  float z = sqnorm(a);
  if (z > 0.0f) {
    float s = sqrtf(z);
    return {a.x / s, a.y / s, a.z / s};
  }
  return a;
}
HK_HD bool is_zero_fuzzy(vec3 c) {  // Eigen isZero(): every |c| <= 1e-5
  return fabsf(c.x) <= 1e-5f && fabsf(c.y) <= 1e-5f && fabsf(c.z) <= 1e-5f;
}
HK_HD vec3 vexp(vec3 a) { return {expf(a.x), expf(a.y), expf(a.z)}; }
HK_HD float luminance(vec3 c) { return c.x * 0.212671f + c.y * 0.715160f + c.z * 0.072169f; }
HK_HD bool color_valid(vec3 c) {
  return c.x >= 0.0f && c.y >= 0.0f && c.z >= 0.0f && isfinite(c.x) && isfinite(c.y) && isfinite(c.z);
}
HK_HD float signumf(float v) { return copysignf(1.0f, v); }
HK_HD float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
HK_HD vec3 ld3(const float *p) { return {p[0], p[1], p[2]}; }

// ------------------------------------------------------------------ pcg32
// PCG32 XSH-RR (O'Neill 2014) as vendored by the reference (ext/pcg32/pcg32.h).
// One stream per path: initseq = pixel index, initstate = splitmix64 of
// (request seed, sample index).
struct Rng {
  uint64_t state, inc;
  HK_HD uint32_t next_uint() {
    uint64_t old = state;
    state = old * 0x5851f42d4c957f2dULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31));
  }
  HK_HD float next() {
    union { uint32_t u; float f; } x;
    x.u = (next_uint() >> 9) | 0x3f800000u;
    return x.f - 1.0f;
  }
};
HK_HD uint64_t splitmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
HK_HD uint64_t rng_inc(uint32_t pixel) { return ((uint64_t)pixel << 1u) | 1u; }
HK_HD Rng rng_for_path(uint32_t seed, uint32_t pixel, uint32_t sample) {
  Rng r;
  r.inc = rng_inc(pixel);
  r.state = 0u;
  r.next_uint();
  r.state += splitmix64(((uint64_t)seed << 32) | sample);
  r.next_uint();
  return r;
}

// ------------------------------------------------------------------ ray / triangle
// Moller-Trumbore with the reference's operation order (src/core/Mesh.cpp:
// 340-384); contraction off so every product/sum rounds like the FP32
// reference.  e1, e2 come precomputed (same FP32 values).
HK_HD bool mt_test(vec3 o, vec3 d, vec3 p0, vec3 e1, vec3 e2, float tmin, float tmax, float &u, float &v,
                   float &t) {
  // Every operation is written out here so that `contract(off)` covers all of
  // them (a pragma does not reach into inlined helpers such as dot/cross).
#pragma clang fp contract(off)
  float px = d.y * e2.z - d.z * e2.y, py = d.z * e2.x - d.x * e2.z, pz = d.x * e2.y - d.y * e2.x;
  float det = e1.x * px + (e1.y * py + e1.z * pz);
  if (det > -1e-8f && det < 1e-8f) return false;
  float inv = 1.0f / det;
  float tx = o.x - p0.x, ty = o.y - p0.y, tz = o.z - p0.z;
  u = (tx * px + (ty * py + tz * pz)) * inv;
  if (u < 0.0f || u > 1.0f) return false;
  float qx = ty * e1.z - tz * e1.y, qy = tz * e1.x - tx * e1.z, qz = tx * e1.y - ty * e1.x;
  v = (d.x * qx + (d.y * qy + d.z * qz)) * inv;
  if (v < 0.0f || u + v > 1.0f) return false;
  t = (e2.x * qx + (e2.y * qy + e2.z * qz)) * inv;
  return t >= tmin && t <= tmax;
}

// ------------------------------------------------------------------ frames
struct Frame {
  vec3 s, t, n;
};
HK_HD Frame frame_from_n(vec3 a) {  // CoordinateSystem, src/core/Common.cpp:591-604
  vec3 c;
  if (fabsf(a.x) > fabsf(a.y)) {
    float inv = 1.0f / sqrtf(a.x * a.x + a.z * a.z);
    c = v3(a.z * inv, 0.0f, -a.x * inv);
  } else {
    float inv = 1.0f / sqrtf(a.y * a.y + a.z * a.z);
    c = v3(0.0f, a.z * inv, -a.y * inv);
  }
  return Frame{cross(c, a), c, a};
}
HK_HD Frame frame_from_n_dpdu(vec3 n, vec3 dpdu) {  // src/core/Frame.cpp:9-14
  vec3 s = normalized(dpdu - n * dot(n, dpdu));
  return Frame{s, cross(n, s), n};
}
HK_HD vec3 to_local(const Frame &f, vec3 v) { return {dot(v, f.s), dot(v, f.t), dot(v, f.n)}; }
HK_HD vec3 to_world(const Frame &f, vec3 v) { return (f.s * v.x + f.t * v.y) + f.n * v.z; }
HK_HD float sin_theta2(vec3 v) { return 1.0f - v.z * v.z; }
HK_HD float sin_theta(vec3 v) {  // Frame.cpp:38-46
  float t = sin_theta2(v);
  return t <= 0.0f ? 0.0f : sqrtf(t);
}
HK_HD float sin_phi(vec3 v) {  // Frame.cpp:68-76
  float st = sin_theta(v);
  return st == 0.0f ? 1.0f : clampf(v.y / st, -1.0f, 1.0f);
}
HK_HD float cos_phi(vec3 v) {  // Frame.cpp:78-86
  float st = sin_theta(v);
  return st == 0.0f ? 1.0f : clampf(v.x / st, -1.0f, 1.0f);
}
HK_HD float tan_theta(vec3 v) {
  float tmp = 1.0f - v.z * v.z;
  if (tmp <= 0.0f) return 0.0f;
  return sqrtf(tmp) / v.z;
}

// ------------------------------------------------------------------ fresnel
HK_HD float fresnel_dielectric(float cos_i, float eta, float inv_eta, float &cos_t) {  // Common.cpp:494-527
  if (eta == 1.0f) {
    cos_t = -cos_i;
    return 0.0f;
  }
  float scale = (cos_i > 0.0f) ? inv_eta : eta;
  float cts = 1.0f - (1.0f - cos_i * cos_i) * (scale * scale);
  if (cts <= 0.0f) {
    cos_t = 0.0f;
    return 1.0f;
  }
  float ci = fabsf(cos_i), ct = sqrtf(cts);
  float rs = (ci - eta * ct) / (ci + eta * ct);
  float rp = (eta * ci - ct) / (eta * ci + ct);
  cos_t = (cos_i > 0.0f) ? -ct : ct;
  return 0.5f * (rs * rs + rp * rp);
}
HK_HD vec3 fresnel_conductor(float cos_i, vec3 eta, vec3 k) {  // Common.cpp:529-548
  float ci2 = cos_i * cos_i, si2 = 1.0f - ci2;
  vec3 eta2 = eta * eta, k2 = k * k;
  vec3 t0 = eta2 - k2 - v3(si2);
  vec3 a2b2 = t0 * t0 + 4.0f * eta2 * k2;
  a2b2 = v3(sqrtf(a2b2.x), sqrtf(a2b2.y), sqrtf(a2b2.z));
  vec3 t1 = a2b2 + v3(ci2);
  vec3 a = 0.5f * (a2b2 + t0);
  a = v3(sqrtf(a.x), sqrtf(a.y), sqrtf(a.z));
  vec3 t2 = (2.0f * cos_i) * a;
  vec3 rs = (t1 - t2) / (t1 + t2);
  vec3 t3 = ci2 * a2b2 + v3(si2 * si2);
  vec3 t4 = t2 * si2;
  vec3 rp = rs * ((t3 - t4) / (t3 + t4));
  return 0.5f * (rp + rs);
}
HK_HD float approx_fresnel_diffuse_reflectance(float eta) {  // Common.cpp:550-589
  if (eta < 1.0f) return -1.4399f * (eta * eta) + 0.7099f * eta + 0.6681f + 0.0636f / eta;
  float i1 = 1.0f / eta, i2 = i1 * i1, i3 = i2 * i1, i4 = i3 * i1, i5 = i4 * i1;
  return 0.919317f - 3.4793f * i1 + 6.75335f * i2 - 7.80989f * i3 + 4.98554f * i4 - 1.36881f * i5;
}
HK_HD vec3 reflect_z(vec3 wi) { return {-wi.x, -wi.y, wi.z}; }
HK_HD vec3 refract_z(vec3 wi, float cos_t, float eta, float inv_eta) {
  float scale = -(cos_t < 0.0f ? inv_eta : eta);
  return {scale * wi.x, scale * wi.y, cos_t};
}
HK_HD vec3 reflect_m(vec3 wi, vec3 m) { return (2.0f * dot(wi, m)) * m - wi; }
HK_HD vec3 refract_m(vec3 wi, vec3 m, float cos_t, float eta, float inv_eta) {
  eta = (cos_t < 0.0f ? inv_eta : eta);
  return m * (dot(wi, m) * eta + cos_t) - wi * eta;
}

// ------------------------------------------------------------------ warps
HK_HD vec3 cosine_hemisphere(float sx, float sy) {  // Sampling.cpp:29-53, 92-107
  float ox = 2.0f * sx - 1.0f, oy = 2.0f * sy - 1.0f;
  float dx = 0.0f, dy = 0.0f;
  if (!(ox == 0.0f && oy == 0.0f)) {
    float th, r;
    if (fabsf(ox) > fabsf(oy)) {
      r = ox;
      th = kPiF * 0.25f * (oy / ox);
    } else {
      r = oy;
      th = kPiF * 0.5f - kPiF * 0.25f * (ox / oy);
    }
    dx = r * cosf(th);
    dy = r * sinf(th);
    // 1 - dx^2 - dy^2 of the reference is 1 - r^2; (1 - |r|)(1 + |r|) keeps
    // its digits near the rim of the disk
    float ar = fabsf(r);
    return {dx, dy, sqrtf(fmaxf(0.0f, (1.0f - ar) * (1.0f + ar)))};
  }
  return {dx, dy, 1.0f};
}
HK_HD float cosine_hemisphere_pdf(vec3 v) {
  return (norm(v) - 1.0f <= 1e-6f && v.z >= 0.0f) ? v.z / kPiF : 0.0f;
}

// ------------------------------------------------------------------ microfacet
// src/core/MicrofacetDistribution.cpp (full-distribution sampling, Smith G1).
struct Micro {
  int type;
  float au, av;
};
HK_HD Micro micro(int type, float a, float b) { return Micro{type, clampf(a, 5e-4f, 1.0f), clampf(b, 5e-4f, 1.0f)}; }
HK_HD float micro_eval(const Micro &m, vec3 h) {
  if (h.z <= 0.0f) return 0.0f;
  float c2 = h.z * h.z;
  float be = ((h.x * h.x) / (m.au * m.au) + (h.y * h.y) / (m.av * m.av)) / c2;
  float res;
  if (m.type == kBeckmann) {
    res = expf(-be) / (kPiF * m.au * m.av * c2 * c2);
  } else {
    float root = (1.0f + be) * c2;
    res = 1.0f / (kPiF * m.au * m.av * root * root);
  }
  if (res * h.z < 1e-20f) res = 0.0f;
  return res;
}
HK_HD float micro_pdf(const Micro &m, vec3 h) { return micro_eval(m, h) * h.z; }
HK_HD float micro_g1(const Micro &m, vec3 v, vec3 h) {
  if (dot(v, h) * v.z <= 0.0f) return 0.0f;
  float tt = fabsf(tan_theta(v));
  if (tt == 0.0f) return 1.0f;
  float alpha = m.au;
  float inv_s2 = 1.0f / sin_theta2(v);
  if (!(m.au == m.av || inv_s2 <= 0.0f)) {
    float c2 = v.x * v.x * inv_s2, s2 = v.y * v.y * inv_s2;
    alpha = sqrtf(c2 * m.au * m.au + s2 * m.av * m.av);
  }
  if (m.type == kBeckmann) {
    float a = 1.0f / (alpha * tt);
    if (a >= 1.6f) return 1.0f;
    float a2 = a * a;
    return (3.535f * a + 2.181f * a2) / (1.0f + 2.276f * a + 2.577f * a2);
  }
  float root = alpha * tt;
  float hyp;  // Hypot2(1, root), include/core/Common.hpp:608-626
  if (1.0f > fabsf(root)) {
    float q = root / 1.0f;
    hyp = 1.0f * sqrtf(1.0f + q * q);
  } else if (root != 0.0f) {
    float q = 1.0f / root;
    hyp = fabsf(root) * sqrtf(1.0f + q * q);
  } else {
    hyp = 0.0f;
  }
  return 2.0f / (1.0f + hyp);
}
HK_HD float micro_g(const Micro &m, vec3 wi, vec3 wo, vec3 h) { return micro_g1(m, wi, h) * micro_g1(m, wo, h); }
HK_HD vec3 micro_sample(const Micro &m, float sx, float sy, float &pdf) {
  float sin_phi, cos_phi, alpha_sqr;
  if (m.au == m.av) {
    float phi = (2.0f * kPiF) * sy;
    sin_phi = sinf(phi);
    cos_phi = cosf(phi);
    alpha_sqr = m.au * m.au;
  } else {
    float phi = atanf(m.av / m.au * tanf(kPiF + 2.0f * kPiF * sy)) + kPiF * floorf(2.0f * sy + 0.5f);
    sin_phi = sinf(phi);
    cos_phi = cosf(phi);
    float cs = cos_phi / m.au, ss = sin_phi / m.av;
    alpha_sqr = 1.0f / (cs * cs + ss * ss);
  }
  float cos_m, tan2;
  if (m.type == kBeckmann) {
    tan2 = alpha_sqr * -logf(1.0f - sx);
    cos_m = 1.0f / sqrtf(1.0f + tan2);
    pdf = (1.0f - sx) / (kPiF * m.au * m.av * cos_m * cos_m * cos_m);
  } else {
    tan2 = alpha_sqr * sx / (1.0f - sx);
    cos_m = 1.0f / sqrtf(1.0f + tan2);
    float tmp = 1.0f + tan2 / alpha_sqr;
    pdf = kInvPiF / (m.au * m.av * cos_m * cos_m * cos_m * tmp * tmp);
  }
  if (pdf < 1e-20f) pdf = 0.0f;
  // sin = sqrt(1 - cos^2) of the reference, evaluated as tan * cos: the same
  // value without the cancellation that costs FP32 digits near the normal
  // (a few ulp from the exact sine of the FP32 angle, where the reference's
  // FP32 expression loses up to half its digits at cos -> 1).  At sx = 1
  // tan^2 is infinite
  // and the reference's expression gives sin = 1, cos = 0; so does this one.
  float sin_m = isinf(tan2) ? 1.0f : sqrtf(tan2) * cos_m;
  return {sin_m * cos_phi, sin_m * sin_phi, cos_m};
}

// ------------------------------------------------------------------ cubic interpolation
// Catmull-Rom splines on uniform knots (src/core/Common.cpp:230-463).
HK_HD float cubic1d(float x, const float *v, int n, float mn, float mx) {
  if (!(x >= mn && x <= mx)) return 0.0f;
  float t = ((x - mn) * (float)(n - 1)) / (mx - mn);
  int k = t < 0.0f ? 0 : (int)t;
  k = k > n - 2 ? n - 2 : k;
  float f0 = v[k], f1 = v[k + 1];
  float d0 = (k > 0) ? 0.5f * (v[k + 1] - v[k - 1]) : v[k + 1] - v[k];
  float d1 = (k + 2 < n) ? 0.5f * (v[k + 2] - v[k]) : v[k + 1] - v[k];
  t = t - (float)k;
  float t2 = t * t, t3 = t2 * t;
  return (2.0f * t3 - 3.0f * t2 + 1.0f) * f0 + (-2.0f * t3 + 3.0f * t2) * f1 + (t3 - 2.0f * t2 + t) * d0 +
         (t3 - t2) * d1;
}
HK_HD bool cubic_weights(float p, int size, float w[4], int &knot) {
  if (!(p >= 0.0f && p <= 1.0f)) return false;
  float t = (p * (float)(size - 1)) / 1.0f;
  knot = (int)t < size - 2 ? (int)t : size - 2;
  t = t - (float)knot;
  float t2 = t * t, t3 = t2 * t;
  w[0] = 0.0f;
  w[1] = 2.0f * t3 - 3.0f * t2 + 1.0f;
  w[2] = -2.0f * t3 + 3.0f * t2;
  w[3] = 0.0f;
  float d0 = t3 - 2.0f * t2 + t, d1 = t3 - t2;
  if (knot > 0) { w[2] += 0.5f * d0; w[0] -= 0.5f * d0; }
  else { w[2] += d0; w[1] -= d0; }
  if (knot + 2 < size) { w[3] += 0.5f * d1; w[1] -= 0.5f * d1; }
  else { w[2] += d1; w[1] -= d1; }
  return true;
}
inline float cubic2d_host(float px, float py, const float *v, int sx, int sy) {
  float wx[4], wy[4];
  int kx, ky;
  if (!cubic_weights(px, sx, wx, kx) || !cubic_weights(py, sy, wy, ky)) return 0.0f;
  float res = 0.0f;
  for (int y = -1; y <= 2; y++) {
    float wY = wy[y + 1];
    for (int x = -1; x <= 2; x++) {
      float wxy = wx[x + 1] * wY;
      if (wxy == 0.0f) continue;
      res += v[(ky + y) * sx + kx + x] * wxy;
    }
  }
  return res;
}
inline float cubic3d_host(float px, float py, float pz, const float *v, int sx, int sy, int sz) {
  float wx[4], wy[4], wz[4];
  int kx, ky, kz;
  if (!cubic_weights(px, sx, wx, kx) || !cubic_weights(py, sy, wy, ky) || !cubic_weights(pz, sz, wz, kz))
    return 0.0f;
  float res = 0.0f;
  for (int z = -1; z <= 2; z++) {
    float wZ = wz[z + 1];
    for (int y = -1; y <= 2; y++) {
      float wyz = wy[y + 1] * wZ;
      for (int x = -1; x <= 2; x++) {
        float w = wx[x + 1] * wyz;
        if (w == 0.0f) continue;
        res += v[((kz + z) * sy + (ky + y)) * sx + kx + x] * w;
      }
    }
  }
  return res;
}
// Reduced (eta and alpha fixed) rough transmittance: RoughTransmittance::Eval
// with both dimensions fixed (src/core/RoughTransmitance.cpp:53-65).
HK_HD float rft_eval1(const float *table, float cos_t) {
  if (cos_t < 0.0f) return 0.0f;
  float wc = powf(fabsf(cos_t), 0.25f);
  return clampf(cubic1d(wc, table, kRftThetaN, 0.0f, 1.0f), 0.0f, 1.0f);
}

}  // namespace hk

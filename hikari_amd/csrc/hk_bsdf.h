// hk_bsdf.h -- Sample / Eval / Pdf of the BSDFs (src/bsdf/{Diffuse,Conductor,
// Dielectric,Plastic,RoughConductor,RoughDielectric,RoughPlastic,Coating,
// RoughCoating,Mirror,Microfacet,RoughDiffuse,TwoSided}BSDF.cpp), FP32 on the
// GPU.  Directions are in the local shading frame (z = shading normal).  A
// coating nests a base (non-coating, non-twoSided) BSDF; a twoSided nests
// base BSDFs or coatings without transmission (checked on upload).
#pragma once
#include "hk_math.h"

namespace hk {

struct BRec {
  vec3 wi, wo;
  float eta;
  int measure;
};

// Source of the extra uniform draws Sample() takes from the sampler
// (RoughDielectricBSDF.cpp:91-101): the path's pcg32 stream in the renderer,
// a caller array in the batch test entry.
struct U1 {
  Rng *rng;
  const float *arr;
  int k, n;
  HK_HD float next() {
    if (rng) return rng->next();
    return k < n ? arr[k++] : 0.5f;
  }
};

HK_HD vec3 mat_a(const DevMaterial &m) { return v3(m.a[0], m.a[1], m.a[2]); }
HK_HD vec3 mat_b(const DevMaterial &m) { return v3(m.b[0], m.b[1], m.b[2]); }
HK_HD bool near_dir(vec3 a, vec3 b) { return fabsf(dot(a, b) - 1.0f) <= 1e-3f; }  // DeltaEpsilon
HK_HD float spec_pick(float r, float w) { return (r * w) / (r * w + (1.0f - r) * (1.0f - w)); }

// ---------------------------------------------------------------- base BSDFs
HK_HD vec3 rough_dielectric_sample(const DevMaterial &m, BRec &rec, float sx, float sy, U1 &u1) {
  rec.measure = kSolidAngle;
  float ci = rec.wi.z;
  if (ci == 0.0f) return v3(0.0f);
  Micro md = micro(m.dist, m.alpha_u, m.alpha_v);
  Micro sd = md;
  float sc = 1.2f - 0.2f * sqrtf(fabsf(ci));
  sd.au *= sc;
  sd.av *= sc;
  float pdf;
  vec3 h = micro_sample(sd, sx, sy, pdf);
  if (pdf == 0.0f) return v3(0.0f);
  float ct;
  float f = fresnel_dielectric(dot(rec.wi, h), m.eta, m.inv_eta, ct);
  float s1 = u1.next();
  vec3 w = v3(1.0f);
  if (!(s1 > f)) {
    rec.wo = reflect_m(rec.wi, h);
    rec.eta = 1.0f;
    if (rec.wo.z * ci <= 0.0f) return v3(0.0f);
    w = w * mat_a(m);
  } else {
    if (ct == 0.0f) return v3(0.0f);
    rec.wo = refract_m(rec.wi, h, ct, m.eta, m.inv_eta);
    rec.eta = ct < 0.0f ? m.eta : m.inv_eta;
    if (rec.wo.z * ci >= 0.0f) return v3(0.0f);
    float factor = ct < 0.0f ? m.inv_eta : m.eta;
    w = w * (mat_b(m) * (factor * factor));
  }
  float s = fabsf(micro_eval(md, h) * micro_g(md, rec.wi, rec.wo, h) * dot(rec.wi, h) / (pdf * ci));
  return w * s;
}
HK_HD vec3 rough_dielectric_eval(const DevMaterial &m, const BRec &rec) {
  float ci = rec.wi.z, co = rec.wo.z;
  if (rec.measure != kSolidAngle || ci == 0.0f) return v3(0.0f);
  bool refl = ci * co > 0.0f;
  float e = ci > 0.0f ? m.eta : m.inv_eta;
  vec3 h = refl ? normalized(rec.wi + rec.wo) : normalized(rec.wi + rec.wo * e);
  h = h * signumf(h.z);
  Micro md = micro(m.dist, m.alpha_u, m.alpha_v);
  float dd = micro_eval(md, h);
  if (dd == 0.0f) return v3(0.0f);
  float ct;
  float f = fresnel_dielectric(dot(rec.wi, h), m.eta, m.inv_eta, ct);
  float g = micro_g(md, rec.wi, rec.wo, h);
  if (refl) return (f * dd * g / (4.0f * fabsf(ci))) * mat_a(m);
  float sd = dot(rec.wi, h) + e * dot(rec.wo, h);
  float value = ((1.0f - f) * dd * g * e * e * dot(rec.wi, h) * dot(rec.wo, h)) / (ci * sd * sd);
  float factor = ci < 0.0f ? m.inv_eta : m.eta;
  return fabsf(value * factor * factor) * mat_b(m);
}
HK_HD float rough_dielectric_pdf(const DevMaterial &m, const BRec &rec) {
  if (rec.measure != kSolidAngle) return 0.0f;
  float ci = rec.wi.z, co = rec.wo.z;
  if (ci == 0.0f || co == 0.0f) return 0.0f;
  bool refl = ci * co > 0.0f;
  vec3 h;
  float j;
  if (refl) {
    h = normalized(rec.wi + rec.wo);
    j = 1.0f / (4.0f * dot(rec.wo, h));
  } else {
    float e = ci > 0.0f ? m.eta : m.inv_eta;
    h = normalized(rec.wi + rec.wo * e);
    float sd = dot(rec.wi, h) + e * dot(rec.wo, h);
    j = (e * e * dot(rec.wo, h)) / (sd * sd);
  }
  h = h * signumf(h.z);
  Micro md = micro(m.dist, m.alpha_u, m.alpha_v);
  float sc = 1.2f - 0.2f * sqrtf(fabsf(ci));
  md.au *= sc;
  md.av *= sc;
  float ct;
  float f = fresnel_dielectric(dot(rec.wi, h), m.eta, m.inv_eta, ct);
  float p = micro_pdf(md, h) * (refl ? f : (1.0f - f));
  return fabsf(p * j);
}

HK_HD vec3 rough_plastic_eval(const DevMaterial &m, const float *rft, const BRec &rec) {
  float ci = rec.wi.z, co = rec.wo.z;
  if (rec.measure == kUnknownMeasure || ci <= 0.0f || co <= 0.0f) return v3(0.0f);
  Micro md = micro(m.dist, m.alpha_u, m.alpha_u);
  float ct;
  vec3 h = normalized(rec.wi + rec.wo);
  float dd = micro_eval(md, h);
  float f = fresnel_dielectric(dot(rec.wi, h), m.eta, m.inv_eta, ct);
  float g = micro_g(md, rec.wi, rec.wi, h);  // the reference's G(Wi, Wi, H), RoughPlasticBSDF.cpp:146
  float value = f * dd * g / (4.0f * ci);
  vec3 res = v3(0.0f) + mat_a(m) * value;
  vec3 diff = mat_b(m);
  const float *tab = rft + m.rft_off;
  float t12 = rft_eval1(tab, ci), t21 = rft_eval1(tab, co);
  float fdr = m.fdr_int;
  if (m.flags & 1) diff = diff / (v3(1.0f) - diff * fdr);
  else diff = diff / (1.0f - fdr);
  return res + diff * (kInvPiF * co * t12 * t21 * m.inv_eta2);
}
HK_HD float rough_plastic_pdf(const DevMaterial &m, const float *rft, const BRec &rec) {
  float ci = rec.wi.z, co = rec.wo.z;
  if (rec.measure == kUnknownMeasure || ci <= 0.0f || co <= 0.0f) return 0.0f;
  Micro md = micro(m.dist, m.alpha_u, m.alpha_u);
  vec3 h = normalized(rec.wi + rec.wo);
  float sp = spec_pick(1.0f - rft_eval1(rft + m.rft_off, ci), m.spec_weight);
  float j = 1.0f / (4.0f * dot(rec.wo, h));
  float res = micro_pdf(md, h) * j * sp;
  res += (1.0f - sp) * cosine_hemisphere_pdf(rec.wo);
  return res;
}
HK_HD vec3 rough_plastic_sample(const DevMaterial &m, const float *rft, BRec &rec, float sx, float sy) {
  float ci = rec.wi.z;
  if (ci <= 0.0f) return v3(0.0f);
  Micro md = micro(m.dist, m.alpha_u, m.alpha_u);
  float sp = spec_pick(1.0f - rft_eval1(rft + m.rft_off, ci), m.spec_weight);
  float dp = 1.0f - sp;
  bool spec;
  if (sy < sp) { sy /= sp; spec = true; }
  else { sy = (sy - sp) / (1.0f - sp); spec = false; }
  if (spec) {
    float pdf;
    vec3 h = micro_sample(md, sx, sy, pdf);
    rec.wo = reflect_m(rec.wi, h);
    rec.measure = kDiscrete;
  } else {
    rec.wo = cosine_hemisphere(sx, sy);
    rec.measure = kSolidAngle;
  }
  rec.eta = 1.0f;
  vec3 h = normalized(rec.wi + rec.wo);
  float j = 1.0f / (4.0f * dot(rec.wo, h));
  float p = micro_pdf(md, h) * j * sp + dp * cosine_hemisphere_pdf(rec.wo);
  if (p == 0.0f) return v3(0.0f);
  return rough_plastic_eval(m, rft, rec) / p;
}

// ---------------------------------------------------------------- microfacet BSDF
// src/bsdf/MicrofacetBSDF.cpp: a Beckmann lobe weighted by Ks = 1 - max(kd)
// (held in spec_weight) over a Lambertian base.  sin^2 of a unit direction
// is evaluated as x^2 + y^2 (the reference's 1 - z^2 loses FP32 digits near
// the normal, where alpha = 0.005 plates put the lobe).
HK_HD float mf_beckmann_d(vec3 h, float alpha) {  // :240-246
  float c2 = h.z * h.z;
  float e2 = (h.x * h.x + h.y * h.y) / c2 / (alpha * alpha);
  return expf(-e2) / (kPiF * alpha * alpha * c2 * c2);
}
HK_HD float mf_smith_g1(vec3 v, vec3 h, float alpha) {  // :248-272
  float tt = tan_theta(v);
  if (tt == 0.0f) return 1.0f;
  if (dot(h, v) * v.z <= 0.0f) return 0.0f;
  float b = 1.0f / (alpha * tt);
  if (b >= 1.6f) return 1.0f;
  float b2 = b * b;
  return (3.535f * b + 2.181f * b2) / (1.0f + 2.276f * b + 2.577f * b2);
}
HK_HD float mf_beckmann_pdf(vec3 h, float alpha) {  // Sampling.cpp:135-159
  if (!(norm(h) - 1.0f <= 1e-6f && h.z >= 0.0f)) return 0.0f;
  float c = h.z;
  if (c == 0.0f) return 0.0f;
  float t2 = (h.x * h.x + h.y * h.y) / (c * c);
  if (isinf(t2)) return 0.0f;
  float a2 = alpha * alpha;
  return expf(-t2 / a2) / (kPiF * a2 * (c * c * c));
}
HK_HD vec3 mf_beckmann_sample(float sx, float sy, float alpha) {  // Sampling.cpp:115-133
  float ln = logf(1.0f - sx);
  if (isnan(ln)) ln = 1.0f;
  float t2 = -alpha * alpha * ln;
  float c = 1.0f / sqrtf(1.0f + t2);
  // sqrt(1 - c^2) without the cancellation; 1 where t2 is infinite (sx = 1),
  // as the reference's expression gives
  float st = isinf(t2) ? 1.0f : sqrtf(t2) * c;
  float phi = 2.0f * kPiF * sy;
  return v3(st * cosf(phi), st * sinf(phi), c);
}
// h: the half vector normalize(wi + wo) of the reference.  Sample() passes the
// sampled microfacet normal instead, which is that vector exactly: at grazing
// incidence wi + wo cancels to a few ulps of its terms and the FP32 sum would
// cost the sharp lobes (alpha = 0.005) percents.
HK_HD vec3 microfacet_value(const DevMaterial &m, const BRec &rec, vec3 h) {  // :96-135 (after its checks)
  float a = m.alpha_u, ct;
  float d = mf_beckmann_d(h, a);
  float f = fresnel_dielectric(dot(h, rec.wi), m.eta, m.inv_eta, ct);
  float g = mf_smith_g1(rec.wi, h, a) * mf_smith_g1(rec.wo, h, a);
  float spec = m.spec_weight * f * d * g / (4.0f * rec.wi.z * rec.wo.z);
  if (!(spec >= 0.0f) || isinf(spec)) spec = 0.0f;
  return v3(spec) + mat_a(m) * kInvPiF;
}
HK_HD float microfacet_pdf(const DevMaterial &m, const BRec &rec, vec3 h) {  // :137-165
  if (rec.measure != kSolidAngle || rec.wi.z <= 0.0f || rec.wo.z <= 0.0f) return 0.0f;
  float j = 0.25f / dot(h, rec.wo);
  float ks = m.spec_weight;
  return ks * mf_beckmann_pdf(h, m.alpha_u) * j + (1.0f - ks) * cosine_hemisphere_pdf(rec.wo);
}
HK_HD vec3 microfacet_sample(const DevMaterial &m, BRec &rec, float sx, float sy) {  // :34-94
  if (rec.wi.z <= 0.0f) return v3(0.0f);
  rec.measure = kSolidAngle;
  float ks = m.spec_weight;
  vec3 h;
  if (sx < ks) {
    h = mf_beckmann_sample(sx / ks, sy, m.alpha_u);
    rec.wo = reflect_m(rec.wi, h);
    if (dot(rec.wi, h) < 0.0f) h = -h;  // normalize(wi + wo) = sign(wi.h) h
  } else {
    rec.wo = cosine_hemisphere((sx - ks) / (1.0f - ks), sy);
    h = normalized(rec.wi + rec.wo);
  }
  rec.eta = 1.0f;
  float p = microfacet_pdf(m, rec, h);
  if (p == 0.0f) return v3(0.0f);
  return microfacet_value(m, rec, h) * rec.wo.z / p;
}

// ---------------------------------------------------------------- rough diffuse
// RoughDiffuseBSDF::Eval (src/bsdf/RoughDiffuseBSDF.cpp:51-159): Oren-Nayar,
// full or fast form; the value carries the cos(theta_o) factor.
HK_HD vec3 rough_diffuse_eval(const DevMaterial &m, const BRec &rec) {
  float ci = rec.wi.z, co = rec.wo.z;
  if (ci <= 0.0f || co <= 0.0f || rec.measure != kSolidAngle) return v3(0.0f);
  float sigma = m.alpha_u * 0.70710678118655f;
  float s2 = sigma * sigma;
  float sti = sin_theta(rec.wi), sto = sin_theta(rec.wo);
  float cpd = 0.0f;
  if (sti > 1e-4f && sto > 1e-4f) cpd = cos_phi(rec.wi) * cos_phi(rec.wo) + sin_phi(rec.wi) * sin_phi(rec.wo);
  if (m.flags & kMatFlagFast) {
    float A = 1.0f - 0.5f * s2 / (s2 + 0.33f);
    float B = 0.45f * s2 / (s2 + 0.09f);
    float sa = ci > co ? sto : sti;
    float tb = ci > co ? sti / ci : sto / co;
    return mat_a(m) * (kInvPiF * co * (A + B * fmaxf(cpd, 0.0f) * sa * tb));
  }
  float ti = acosf(clampf(ci, -1.0f, 1.0f)), to = acosf(clampf(co, -1.0f, 1.0f));
  float al = fmaxf(ti, to), be = fminf(ti, to);
  float sa, sb, tb;
  if (ci > co) { sa = sto; sb = sti; tb = sti / ci; }
  else { sa = sti; sb = sto; tb = sto / co; }
  float t1 = s2 / (s2 + 0.09f);
  float t2 = (4.0f * kInvPiF * kInvPiF) * al * be;
  float t3 = 2.0f * be * kInvPiF;
  float c1 = 1.0f - 0.5f * s2 / (s2 + 0.33f);
  float c2 = 0.45f * t1;
  float c3 = 0.125f * t1 * t2 * t2;
  float c4 = 0.17f * s2 / (s2 + 0.13f);
  c2 *= cpd > 0.0f ? sa : sa - t3 * t3 * t3;
  float th = (sa + sb) / (sqrtf(fmaxf(0.0f, 1.0f - sa * sa)) + sqrtf(fmaxf(0.0f, 1.0f - sb * sb)));
  if (isinf(th) || isnan(th)) return v3(0.0f);
  vec3 rho = mat_a(m);
  vec3 lr1 = rho * (c1 + cpd * c2 * tb + (1.0f - fabsf(cpd)) * c3 * th);
  vec3 lr2 = rho * rho * (c4 * (1.0f - cpd * t3 * t3));
  return (lr1 + lr2) * (kInvPiF * co);
}

HK_HD vec3 base_sample(const DevMaterial &m, const float *rft, BRec &rec, float sx, float sy, U1 &u1) {
  switch (m.type) {
    case kDiffuse: {
      if (rec.wi.z <= 0.0f) return v3(0.0f);
      rec.measure = kSolidAngle;
      rec.wo = cosine_hemisphere(sx, sy);
      rec.eta = 1.0f;
      return mat_a(m);
    }
    case kConductor: {
      rec.measure = kDiscrete;
      float ci = rec.wi.z;
      if (ci <= 0.0f) return v3(0.0f);
      vec3 f = fresnel_conductor(ci, ld3(m.eta3), ld3(m.eta_k));
      rec.wo = reflect_z(rec.wi);
      rec.eta = 1.0f;
      return mat_a(m) * f;
    }
    case kDielectric: {
      rec.measure = kDiscrete;
      float ct;
      float fr = fresnel_dielectric(rec.wi.z, m.eta, m.inv_eta, ct);
      if (sx < fr) {
        rec.wo = reflect_z(rec.wi);
        rec.eta = 1.0f;
        return mat_a(m);
      }
      rec.wo = refract_z(rec.wi, ct, m.eta, m.inv_eta);
      rec.eta = ct < 0.0f ? m.eta : m.inv_eta;
      float factor = ct < 0.0f ? m.inv_eta : m.eta;
      return mat_b(m) * (factor * factor);
    }
    case kPlastic: {
      float ci = rec.wi.z;
      if (ci <= 0.0f) return v3(0.0f);
      float ct;
      float fi = fresnel_dielectric(ci, m.eta, m.inv_eta, ct);
      float sp = spec_pick(fi, m.spec_weight);
      rec.eta = 1.0f;
      if (sx < sp) {
        rec.measure = kDiscrete;
        rec.wo = reflect_z(rec.wi);
        return mat_a(m) * fi / sp;
      }
      rec.measure = kSolidAngle;
      rec.wo = cosine_hemisphere((sx - sp) / (1.0f - sp), sy);
      float fo = fresnel_dielectric(rec.wo.z, m.eta, m.inv_eta, ct);
      vec3 diff = mat_b(m);
      if (m.flags & 1) diff = diff / (v3(1.0f) - diff * m.fdr_int);
      else diff = diff / (1.0f - m.fdr_int);
      return diff * (m.inv_eta2 * (1.0f - fi) * (1.0f - fo) / (1.0f - sp));
    }
    case kRoughConductor: {
      rec.measure = kSolidAngle;
      float ci = rec.wi.z;
      if (ci <= 0.0f) return v3(0.0f);
      Micro md = micro(m.dist, m.alpha_u, m.alpha_v);
      float pdf;
      vec3 h = micro_sample(md, sx, sy, pdf);
      if (pdf == 0.0f) return v3(0.0f);
      rec.wo = reflect_m(rec.wi, h);
      rec.eta = 1.0f;
      if (rec.wo.z < 0.0f) return v3(0.0f);
      float hdwi = dot(h, rec.wi);
      vec3 f = fresnel_conductor(hdwi, ld3(m.eta3), ld3(m.eta_k));
      float g = micro_g(md, rec.wi, rec.wo, h);
      return mat_a(m) * f * micro_eval(md, h) * g * hdwi / (pdf * ci);
    }
    case kRoughDielectric: return rough_dielectric_sample(m, rec, sx, sy, u1);
    case kRoughPlastic: return rough_plastic_sample(m, rft, rec, sx, sy);
    case kMirror: {  // MirrorBSDF.cpp:12-27
      if (rec.wi.z <= 0.0f) return v3(0.0f);
      rec.wo = reflect_z(rec.wi);
      rec.measure = kDiscrete;
      rec.eta = 1.0f;
      return v3(1.0f);
    }
    case kMicrofacet: return microfacet_sample(m, rec, sx, sy);
    case kRoughDiffuse: {  // RoughDiffuseBSDF.cpp:28-49
      rec.measure = kSolidAngle;
      if (rec.wi.z <= 0.0f) return v3(0.0f);
      rec.wo = cosine_hemisphere(sx, sy);
      rec.eta = 1.0f;
      float p = cosine_hemisphere_pdf(rec.wo);
      if (p == 0.0f) return v3(0.0f);
      return rough_diffuse_eval(m, rec) / p;
    }
    default: return v3(0.0f);
  }
}

HK_HD vec3 base_eval(const DevMaterial &m, const float *rft, const BRec &rec) {
  switch (m.type) {
    case kDiffuse:
      if (rec.measure != kSolidAngle || rec.wi.z <= 0.0f || rec.wo.z <= 0.0f) return v3(0.0f);
      return mat_a(m) * kInvPiF;
    case kConductor: {
      float ci = rec.wi.z, co = rec.wo.z;
      if (ci > 0.0f && co > 0.0f && rec.measure == kDiscrete && near_dir(reflect_z(rec.wi), rec.wo))
        return mat_a(m) * fresnel_conductor(ci, ld3(m.eta3), ld3(m.eta_k));
      return v3(0.0f);
    }
    case kDielectric: {  // no measure check in the reference (DielectricBSDF.cpp:60-84)
      float ci = rec.wi.z, co = rec.wo.z, ct;
      float fr = fresnel_dielectric(ci, m.eta, m.inv_eta, ct);
      if (ci * co >= 0.0f) {
        if (near_dir(reflect_z(rec.wi), rec.wo)) return mat_a(m) * fr;
      } else if (near_dir(refract_z(rec.wi, ct, m.eta, m.inv_eta), rec.wo)) {
        float factor = ct < 0.0f ? m.inv_eta : m.eta;
        return mat_b(m) * (factor * factor) * (1.0f - fr);
      }
      return v3(0.0f);
    }
    case kPlastic: {
      float ci = rec.wi.z, co = rec.wo.z;
      if (rec.measure == kUnknownMeasure || ci <= 0.0f || co <= 0.0f) return v3(0.0f);
      float ct;
      float fi = fresnel_dielectric(ci, m.eta, m.inv_eta, ct);
      if (rec.measure == kDiscrete) {
        if (near_dir(reflect_z(rec.wi), rec.wo)) return mat_a(m) * fi;
        return v3(0.0f);
      }
      float fo = fresnel_dielectric(co, m.eta, m.inv_eta, ct);
      vec3 diff = mat_b(m);
      if (m.flags & 1) diff = diff / (v3(1.0f) - diff * m.fdr_int);
      else diff = diff / (1.0f - m.fdr_int);
      return diff * (cosine_hemisphere_pdf(rec.wo) * m.inv_eta2 * (1.0f - fi) * (1.0f - fo));
    }
    case kRoughConductor: {
      float ci = rec.wi.z, co = rec.wo.z;
      if (rec.measure != kSolidAngle || ci <= 0.0f || co <= 0.0f) return v3(0.0f);
      vec3 h = normalized(rec.wi + rec.wo);
      Micro md = micro(m.dist, m.alpha_u, m.alpha_v);
      float dd = micro_eval(md, h);
      if (dd == 0.0f) return v3(0.0f);
      vec3 f = fresnel_conductor(dot(h, rec.wi), ld3(m.eta3), ld3(m.eta_k)) * mat_a(m);
      float g = micro_g(md, rec.wi, rec.wo, h);
      return dd * g * f / (4.0f * ci);
    }
    case kRoughDielectric: return rough_dielectric_eval(m, rec);
    case kRoughPlastic: return rough_plastic_eval(m, rft, rec);
    case kMirror:  // MirrorBSDF.cpp:29-41
      if (rec.wi.z > 0.0f && rec.wo.z > 0.0f && rec.measure == kDiscrete && near_dir(reflect_z(rec.wi), rec.wo))
        return v3(1.0f);
      return v3(0.0f);
    case kMicrofacet:
      if (rec.measure != kSolidAngle || rec.wi.z <= 0.0f || rec.wo.z <= 0.0f) return v3(0.0f);
      return microfacet_value(m, rec, normalized(rec.wi + rec.wo));
    case kRoughDiffuse: return rough_diffuse_eval(m, rec);
    default: return v3(0.0f);
  }
}

HK_HD float base_pdf(const DevMaterial &m, const float *rft, const BRec &rec) {
  switch (m.type) {
    case kDiffuse:
      if (rec.measure != kSolidAngle || rec.wi.z <= 0.0f || rec.wo.z <= 0.0f) return 0.0f;
      return kInvPiF * rec.wo.z;
    case kConductor: {
      float ci = rec.wi.z, co = rec.wo.z;
      if (ci > 0.0f && co > 0.0f && rec.measure == kDiscrete && near_dir(reflect_z(rec.wi), rec.wo)) return 1.0f;
      return 0.0f;
    }
    case kDielectric: {
      float ci = rec.wi.z, co = rec.wo.z, ct;
      float fr = fresnel_dielectric(ci, m.eta, m.inv_eta, ct);
      if (ci * co >= 0.0f) {
        if (near_dir(reflect_z(rec.wi), rec.wo)) return fr;
      } else if (near_dir(refract_z(rec.wi, ct, m.eta, m.inv_eta), rec.wo)) {
        return 1.0f - fr;
      }
      return 0.0f;
    }
    case kPlastic: {
      float ci = rec.wi.z, co = rec.wo.z;
      if (rec.measure == kUnknownMeasure || ci <= 0.0f || co <= 0.0f) return 0.0f;
      float ct;
      float fi = fresnel_dielectric(ci, m.eta, m.inv_eta, ct);
      float sp = spec_pick(fi, m.spec_weight);
      if (rec.measure == kDiscrete) return near_dir(reflect_z(rec.wi), rec.wo) ? 1.0f - sp : 0.0f;
      return cosine_hemisphere_pdf(rec.wo) * (1.0f - sp);
    }
    case kRoughConductor: {
      float ci = rec.wi.z, co = rec.wo.z;
      if (rec.measure != kSolidAngle || ci <= 0.0f || co <= 0.0f) return 0.0f;
      vec3 h = normalized(rec.wi + rec.wo);
      Micro md = micro(m.dist, m.alpha_u, m.alpha_v);
      return micro_pdf(md, h) / (4.0f * fabsf(dot(h, rec.wo)));
    }
    case kRoughDielectric: return rough_dielectric_pdf(m, rec);
    case kRoughPlastic: return rough_plastic_pdf(m, rft, rec);
    case kMirror:  // MirrorBSDF.cpp:43-55
      if (rec.wi.z > 0.0f && rec.wo.z > 0.0f && rec.measure == kDiscrete && near_dir(reflect_z(rec.wi), rec.wo))
        return 1.0f;
      return 0.0f;
    case kMicrofacet: return microfacet_pdf(m, rec, normalized(rec.wi + rec.wo));
    case kRoughDiffuse:  // RoughDiffuseBSDF.cpp:161-172
      if (rec.wi.z <= 0.0f || rec.wo.z <= 0.0f || rec.measure != kSolidAngle) return 0.0f;
      return cosine_hemisphere_pdf(rec.wo);
    default: return 0.0f;
  }
}

// ---------------------------------------------------------------- coatings
// src/bsdf/CoatingBSDF.cpp:290-302
HK_HD vec3 coat_in(const DevMaterial &m, vec3 wi, float &refl) {
  float ct;
  refl = fresnel_dielectric(fabsf(wi.z), m.eta, m.inv_eta, ct);
  return {m.inv_eta * wi.x, m.inv_eta * wi.y, -signumf(wi.z) * ct};
}
HK_HD vec3 coat_out(const DevMaterial &m, vec3 wi, float &refl) {
  float ct;
  refl = fresnel_dielectric(fabsf(wi.z), m.inv_eta, m.eta, ct);
  return {m.eta * wi.x, m.eta * wi.y, -signumf(wi.z) * ct};
}
HK_HD vec3 coat_sigma(const DevMaterial &m) { return mat_a(m) * m.thickness; }
// src/bsdf/RoughCoatingBSDF.cpp:342-376
HK_HD vec3 rc_refract(vec3 wi, float ie) {
  bool entering = wi.z > 0.0f;
  float s2 = ie * ie * sin_theta2(wi);
  if (s2 >= 1.0f) return v3(0.0f);
  float ct = sqrtf(1.0f - s2);
  return {ie * wi.x, ie * wi.y, entering ? ct : -ct};
}

HK_HD vec3 coating_sample(const DevMaterial &m, const DevMaterial &nm, const float *rft, BRec &rec, float sx, float sy,
                          U1 &u1) {  // CoatingBSDF.cpp:39-104
  float r12;
  vec3 wip = coat_in(m, rec.wi, r12);
  float sp = spec_pick(r12, m.spec_weight);
  bool spec;
  if (sx < sp) { spec = true; sx /= sp; }
  else { spec = false; sx = (sx - sp) / (1.0f - sp); }
  if (spec) {
    rec.wo = reflect_z(rec.wi);
    rec.eta = 1.0f;
    rec.measure = kDiscrete;
    return mat_b(m) * (r12 / sp);
  }
  if (r12 == 1.0f) return v3(0.0f);
  vec3 wib = rec.wi;
  rec.wi = wip;
  vec3 res = base_sample(nm, rft, rec, sx, sy, u1);
  rec.wi = wib;
  if (is_zero_fuzzy(res)) return v3(0.0f);
  vec3 wop = rec.wo;
  vec3 sig = coat_sigma(m);
  if (!is_zero_fuzzy(sig)) res = res * vexp(-sig * (1.0f / fabsf(wip.z) + 1.0f / fabsf(wop.z)));
  float r21;
  rec.wo = coat_out(m, wop, r21);
  if (r21 == 1.0f) return v3(0.0f);
  res = res / (1.0f - sp);
  res = res * ((1.0f - r12) * (1.0f - r21));
  return res;
}
HK_HD vec3 coating_eval(const DevMaterial &m, const DevMaterial &nm, const float *rft, const BRec &rec) {
  if (rec.measure == kDiscrete) {  // CoatingBSDF.cpp:106-141
    if (near_dir(reflect_z(rec.wi), rec.wo)) {
      float ct;
      return mat_b(m) * fresnel_dielectric(fabsf(rec.wi.z), m.eta, m.inv_eta, ct);
    }
    return v3(0.0f);
  }
  float r12, r21;
  BRec ri = rec;
  ri.wi = coat_in(m, rec.wi, r12);
  ri.wo = coat_in(m, rec.wo, r21);
  if (r12 == 1.0f || r21 == 1.0f) return v3(0.0f);
  vec3 res = base_eval(nm, rft, ri) * (1.0f - r12) * (1.0f - r21);
  vec3 sig = coat_sigma(m);
  if (!is_zero_fuzzy(sig)) res = res * vexp(-sig * (1.0f / fabsf(ri.wi.z) + 1.0f / fabsf(ri.wo.z)));
  if (rec.measure == kSolidAngle) res = res * (m.inv_eta * m.inv_eta * rec.wo.z / ri.wo.z);
  return res;
}
HK_HD float coating_pdf(const DevMaterial &m, const DevMaterial &nm, const float *rft, const BRec &rec) {
  float r12;  // CoatingBSDF.cpp:143-174
  vec3 wip = coat_in(m, rec.wi, r12);
  float sp = spec_pick(r12, m.spec_weight);
  if (rec.measure == kDiscrete) return near_dir(reflect_z(rec.wi), rec.wo) ? sp : 0.0f;
  float r21;
  BRec ri = rec;
  ri.wi = wip;
  ri.wo = coat_in(m, rec.wo, r21);
  if (r12 == 1.0f || r21 == 1.0f) return 0.0f;
  float p = base_pdf(nm, rft, ri);
  if (rec.measure == kSolidAngle) p *= (m.inv_eta * m.inv_eta * rec.wo.z / ri.wo.z);
  return p * (1.0f - sp);
}

HK_HD vec3 rough_coating_eval(const DevMaterial &m, const DevMaterial &nm, const float *rft, const BRec &rec) {
  float ci = rec.wi.z, co = rec.wo.z;  // RoughCoatingBSDF.cpp:129-184
  if (ci * co <= 0.0f) return v3(0.0f);
  Micro md = micro(m.dist, m.alpha_u, m.alpha_u);
  vec3 h = normalized(rec.wi + rec.wo) * signumf(rec.wo.z);
  float ct;
  float dd = micro_eval(md, h);
  float f = fresnel_dielectric(fabsf(dot(rec.wi, h)), m.eta, m.inv_eta, ct);
  float g = micro_g(md, rec.wi, rec.wo, h);
  vec3 res = v3(0.0f) + mat_b(m) * f * dd * g / (4.0f * fabsf(ci));
  BRec ri = rec;
  ri.wi = rc_refract(rec.wi, m.inv_eta);
  ri.wo = rc_refract(rec.wo, m.inv_eta);
  const float *tab = rft + m.rft_off;
  vec3 nv = base_eval(nm, rft, ri) * rft_eval1(tab, fabsf(ci)) * rft_eval1(tab, fabsf(co));
  vec3 sig = coat_sigma(m);
  if (!is_zero_fuzzy(sig)) nv = nv * vexp(-sig * (1.0f / fabsf(ri.wi.z) + 1.0f / fabsf(ri.wo.z)));
  if (rec.measure == kSolidAngle) nv = nv * (m.inv_eta * m.inv_eta * rec.wo.z / ri.wo.z);
  return res + nv;
}
HK_HD float rough_coating_pdf(const DevMaterial &m, const DevMaterial &nm, const float *rft, const BRec &rec) {
  vec3 h = normalized(rec.wi + rec.wo) * signumf(rec.wo.z);  // RoughCoatingBSDF.cpp:186-221
  Micro md = micro(m.dist, m.alpha_u, m.alpha_u);
  float ci = rec.wi.z, co = rec.wo.z;
  float sp = spec_pick(1.0f - rft_eval1(rft + m.rft_off, fabsf(ci)), m.spec_weight);
  float res = 0.0f;
  if (ci * co > 0.0f) {
    float j = 1.0f / (4.0f * fabsf(dot(rec.wo, h)));
    res = micro_pdf(md, h) * j * sp;
  }
  BRec ri = rec;
  ri.wi = rc_refract(rec.wi, m.inv_eta);
  ri.wo = rc_refract(rec.wo, m.inv_eta);
  float p = base_pdf(nm, rft, ri);
  if (ri.measure == kSolidAngle) p *= m.inv_eta * m.inv_eta * co / ri.wo.z;
  return res + p * (1.0f - sp);
}
HK_HD vec3 rough_coating_sample(const DevMaterial &m, const DevMaterial &nm, const float *rft, BRec &rec, float sx,
                                float sy, U1 &u1) {  // RoughCoatingBSDF.cpp:62-127
  Micro md = micro(m.dist, m.alpha_u, m.alpha_u);
  float sp = spec_pick(1.0f - rft_eval1(rft + m.rft_off, fabsf(rec.wi.z)), m.spec_weight);
  if (sx < sp) {
    sx /= sp;
    float pdf;
    vec3 h = micro_sample(md, sx, sy, pdf);
    rec.wo = reflect_m(rec.wi, h);
    rec.measure = kDiscrete;
    rec.eta = 1.0f;
    if (rec.wo.z * rec.wo.z <= 0.0f) return v3(0.0f);
  } else {
    sx = (sx - sp) / (1.0f - sp);
    vec3 wib = rec.wi;
    rec.wi = rc_refract(rec.wi, m.inv_eta);
    vec3 res = base_sample(nm, rft, rec, sx, sy, u1);
    rec.wi = wib;
    if (is_zero_fuzzy(res)) return v3(0.0f);
    rec.wo = rc_refract(rec.wo, m.eta);
    if (is_zero_fuzzy(rec.wo)) return v3(0.0f);
  }
  float p = rough_coating_pdf(m, nm, rft, rec);
  if (p == 0.0f) return v3(0.0f);
  return rough_coating_eval(m, nm, rft, rec) / p;
}

// ---------------------------------------------------------------- dispatch
// Three levels, no recursion: a twoSided picks its side's material, which may
// be a coating over a base BSDF (both nestings are checked on upload).
HK_HD vec3 coated_sample(const DevMaterial *mats, const float *rft, int idx, BRec &rec, float sx, float sy, U1 &u1) {
  const DevMaterial &m = mats[idx];
  if (m.type == kCoating) return coating_sample(m, mats[m.nested], rft, rec, sx, sy, u1);
  if (m.type == kRoughCoating) return rough_coating_sample(m, mats[m.nested], rft, rec, sx, sy, u1);
  return base_sample(m, rft, rec, sx, sy, u1);
}
HK_HD vec3 coated_eval(const DevMaterial *mats, const float *rft, int idx, const BRec &rec) {
  const DevMaterial &m = mats[idx];
  if (m.type == kCoating) return coating_eval(m, mats[m.nested], rft, rec);
  if (m.type == kRoughCoating) return rough_coating_eval(m, mats[m.nested], rft, rec);
  return base_eval(m, rft, rec);
}
HK_HD float coated_pdf(const DevMaterial *mats, const float *rft, int idx, const BRec &rec) {
  const DevMaterial &m = mats[idx];
  if (m.type == kCoating) return coating_pdf(m, mats[m.nested], rft, rec);
  if (m.type == kRoughCoating) return rough_coating_pdf(m, mats[m.nested], rft, rec);
  return base_pdf(m, rft, rec);
}
// A twoSided is resolved to one side's material before the single call site
// of each entry point (one inlined copy of the BSDF code per caller).
// TwoSidedBSDF.cpp:21-75: a lower-side query is mirrored to the upper side;
// Sample() mirrors the sampled direction back unless the weight is (fuzzy)
// zero; Eval/Pdf take the upper side only for cos(wi) > 0.
HK_HD int two_sided_side(const DevMaterial *mats, int idx, bool lower) {
  const DevMaterial &m = mats[idx];
  if (m.type != kTwoSided) return idx;
  return lower ? m.nested_back : m.nested;
}
HK_HD vec3 bsdf_sample(const DevMaterial *mats, const float *rft, int idx, BRec &rec, float sx, float sy, U1 &u1) {
  bool flip = mats[idx].type == kTwoSided && rec.wi.z < 0.0f;
  int eff = two_sided_side(mats, idx, flip);
  if (flip) rec.wi.z = -rec.wi.z;
  vec3 res = coated_sample(mats, rft, eff, rec, sx, sy, u1);
  if (flip) {
    rec.wi.z = -rec.wi.z;
    if (!is_zero_fuzzy(res)) rec.wo.z = -rec.wo.z;
  }
  return res;
}
HK_HD vec3 bsdf_eval(const DevMaterial *mats, const float *rft, int idx, const BRec &rec) {
  bool flip = mats[idx].type == kTwoSided && !(rec.wi.z > 0.0f);
  BRec c = rec;
  if (flip) {
    c.wi.z = -c.wi.z;
    c.wo.z = -c.wo.z;
  }
  return coated_eval(mats, rft, two_sided_side(mats, idx, flip), c);
}
HK_HD float bsdf_pdf(const DevMaterial *mats, const float *rft, int idx, const BRec &rec) {
  bool flip = mats[idx].type == kTwoSided && !(rec.wi.z > 0.0f);
  BRec c = rec;
  if (flip) {
    c.wi.z = -c.wi.z;
    c.wo.z = -c.wo.z;
  }
  return coated_pdf(mats, rft, two_sided_side(mats, idx, flip), c);
}

}  // namespace hk

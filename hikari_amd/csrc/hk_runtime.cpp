// hk_runtime.cpp -- host runtime behind the C ABI (include/hikari_amd.h):
// contexts, device buffers, scene upload (the Activate() of every reference
// object, done once on the host), and the wavefront pass scheduler.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hikari_amd.h"
#include "hk_build.h"
#include "hk_kernels.h"
#include "hk_math.h"
#include "hk_runtime.h"

using namespace hk;

namespace hk {

thread_local std::string g_create_error;

HkError::HkError(int c, const char *fmt, ...) : code(c) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  msg = buf;
}

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw HkError(HK_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

void DBuf::alloc(size_t n) {
  if (n <= bytes && p) return;
  release();
  if (n == 0) return;
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    p = nullptr;
    bytes = 0;
    throw HkError(HK_ERR_OOM, "hipMalloc(%zu bytes): %s", n, hipGetErrorString(e));
  }
  bytes = n;
}
void DBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}
void DBuf::upload(const void *src, size_t n, hipStream_t st) {
  alloc(n);
  if (n) hip_check(hipMemcpyAsync(p, src, n, hipMemcpyHostToDevice, st), "hipMemcpyAsync H2D");
}

// ------------------------------------------------------------------ rough transmittance (host)
// src/core/RoughTransmitance.cpp:9-288: file load, SetEta (3-D -> 2-D) and
// SetAlpha (2-D -> 1-D) reductions, EvalDiffuse with eta fixed.
struct RftTable {
  uint64_t eta_n = 0, alpha_n = 0, theta_n = 0;
  float eta_min = 0, eta_max = 0, alpha_min = 0, alpha_max = 0;
  std::vector<float> t, d;
  bool eta_fixed = false, alpha_fixed = false;

  void load(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) throw HkError(HK_ERR_IO, "cannot open rough transmittance data \"%s\"", path.c_str());
    uint64_t h[5];
    float r[4];
    bool ok = fread(h, 8, 5, f) == 5 && fread(r, 4, 4, f) == 4;
    if (ok) {
      eta_n = h[0]; alpha_n = h[1]; theta_n = h[2];
      eta_min = r[0]; eta_max = r[1]; alpha_min = r[2]; alpha_max = r[3];
      ok = h[3] == 2 * eta_n * alpha_n * theta_n && h[4] == 2 * eta_n * alpha_n && theta_n == (uint64_t)kRftThetaN;
      if (ok) {
        t.resize(h[3]);
        d.resize(h[4]);
        ok = fread(t.data(), 4, h[3], f) == h[3] && fread(d.data(), 4, h[4], f) == h[4];
      }
    }
    fclose(f);
    if (!ok) throw HkError(HK_ERR_IO, "malformed rough transmittance data \"%s\"", path.c_str());
  }
  float warp(float v, float mn, float mx) const { return std::pow((v - mn) / (mx - mn), 0.25f); }
  void check(float eta, float alpha) const {
    if (eta < 1.0f) eta = 1.0f / eta;
    if (eta < eta_min || eta > eta_max)
      throw HkError(HK_ERR_INVALID, "relative IOR %f is outside the supported range [%f, %f]", eta, eta_min, eta_max);
    if (alpha < alpha_min || alpha > alpha_max)
      throw HkError(HK_ERR_INVALID, "roughness %f is outside the supported range [%f, %f]", alpha, alpha_min,
                    alpha_max);
  }
  void set_eta(float eta) {
    if (eta_fixed) return;
    const float *pt = t.data(), *pd = d.data();
    if (eta < 1.0f) {
      pt += eta_n * alpha_n * theta_n;
      pd += eta_n * alpha_n;
      eta = 1.0f / eta;
    }
    if (eta < eta_min) eta = eta_min;
    float we = warp(eta, eta_min, eta_max);
    std::vector<float> nt(alpha_n * theta_n), nd(alpha_n);
    float da = 1.0f / (alpha_n - 1.0f), dt = 1.0f / (theta_n - 1.0f);
    for (size_t i = 0; i < alpha_n; i++) {
      for (size_t j = 0; j < theta_n; j++)
        nt[i * theta_n + j] = cubic3d_host(j * dt, i * da, we, pt, int(theta_n), int(alpha_n), int(eta_n));
      nd[i] = cubic2d_host(i * da, we, pd, int(alpha_n), int(eta_n));
    }
    t.swap(nt);
    d.swap(nd);
    eta_fixed = true;
  }
  void set_alpha(float alpha) {
    if (alpha_fixed) return;
    float wa = warp(alpha, alpha_min, alpha_max);
    std::vector<float> nt(theta_n), nd(1);
    float dt = 1.0f / (theta_n - 1.0f);
    for (size_t i = 0; i < theta_n; i++) nt[i] = cubic2d_host(i * dt, wa, t.data(), int(theta_n), int(alpha_n));
    nd[0] = cubic1d(wa, d.data(), int(alpha_n), 0.0f, 1.0f);
    t.swap(nt);
    d.swap(nd);
    alpha_fixed = true;
  }
  float eval_diffuse(float alpha) const {  // eta fixed, alpha free
    float r = cubic1d(warp(alpha, alpha_min, alpha_max), d.data(), int(alpha_n), 0.0f, 1.0f);
    return clampf(r, 0.0f, 1.0f);
  }
};

// Directory of the rough-transmittance tables: data/ next to the library.
std::string default_data_dir() {
  Dl_info info;
  if (dladdr((void *)&hk_version, &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    size_t k = p.rfind('/');
    return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/data";
  }
  return "data";
}

// ------------------------------------------------------------------ materials
// Constructor + Activate() of each BSDF class (src/bsdf/*.cpp).
void build_materials(const float *recs, int n, const char *rft_dir, std::vector<DevMaterial> &out,
                     std::vector<float> &rft_pool) {
  out.assign(n, DevMaterial());
  rft_pool.clear();
  std::string dir = rft_dir ? std::string(rft_dir) : default_data_dir();
  std::unique_ptr<RftTable> beck, ggx;
  auto table = [&](int dist) -> const RftTable & {
    std::unique_ptr<RftTable> &p = dist == kGGX ? ggx : beck;
    if (!p) {
      p.reset(new RftTable());
      p->load(dir + (dist == kGGX ? "/GGXRFTData.bin" : "/BeckmannRFTData.bin"));
    }
    return *p;
  };
  auto lum = [](const float *c) { return c[0] * 0.212671f + c[1] * 0.715160f + c[2] * 0.072169f; };
  for (int i = 0; i < n; i++) {
    const float *r = recs + size_t(i) * HK_MATERIAL_STRIDE;
    DevMaterial m;
    std::memset(&m, 0, sizeof m);
    m.type = int(r[0]);
    m.dist = int(r[1]);
    m.nested = int(r[2]);
    m.flags = int(r[3]) & 3;
    if (int(r[3]) & 4) m.flags |= kMatFlagFast;
    m.rft_off = -1;
    m.nested_back = -1;
    if (m.type < 0 || m.type >= kMatTypeCount) throw HkError(HK_ERR_INVALID, "material %d: unknown type %d", i, m.type);
    if (m.dist != kBeckmann && m.dist != kGGX) throw HkError(HK_ERR_INVALID, "material %d: unknown distribution", i);
    for (int c = 0; c < 3; c++) { m.a[c] = r[4 + c]; m.b[c] = r[8 + c]; }
    float int_ior = r[20], ext_ior = r[21];
    m.eta = int_ior / ext_ior;
    m.inv_eta = 1.0f / m.eta;
    m.thickness = r[24];
    auto clamp_alpha = [](float a) { return clampf(a, float(5e-4), float(1.0)); };
    switch (m.type) {
      case kConductor:
      case kRoughConductor:
        for (int c = 0; c < 3; c++) { m.eta3[c] = int_ior / ext_ior; m.eta_k[c] = m.b[c] / ext_ior; }
        if (m.type == kRoughConductor) {
          m.alpha_u = clamp_alpha(r[22]);
          m.alpha_v = (m.flags & 2) ? clamp_alpha(r[23]) : m.alpha_u;
        }
        break;
      case kRoughDielectric:
        m.alpha_u = clamp_alpha(r[22]);
        m.alpha_v = (m.flags & 2) ? clamp_alpha(r[23]) : m.alpha_u;
        break;
      case kPlastic: {
        m.inv_eta2 = m.inv_eta * m.inv_eta;
        m.fdr_int = approx_fresnel_diffuse_reflectance(m.inv_eta);
        float ks = lum(m.a), kd = lum(m.b);
        m.spec_weight = ks / (kd + ks);
        break;
      }
      case kRoughPlastic: {
        m.inv_eta2 = m.inv_eta * m.inv_eta;
        m.alpha_u = m.alpha_v = clamp_alpha(r[22]);
        float ks = lum(m.a), kd = lum(m.b);
        m.spec_weight = ks / (kd + ks);
        RftTable ext = table(m.dist);
        ext.check(m.eta, m.alpha_u);
        RftTable in = ext;
        ext.set_eta(m.eta);
        in.set_eta(m.inv_eta);
        ext.set_alpha(m.alpha_u);
        m.fdr_int = 1.0f - in.eval_diffuse(m.alpha_u);
        m.rft_off = int32_t(rft_pool.size());
        rft_pool.insert(rft_pool.end(), ext.t.begin(), ext.t.end());
        break;
      }
      case kCoating:
      case kRoughCoating: {
        float tmp[3];
        for (int c = 0; c < 3; c++) tmp[c] = std::exp(m.a[c] * (-2.0f * m.thickness));
        float avg = (tmp[0] + tmp[1] + tmp[2]) / 3.0f;
        m.spec_weight = 1.0f / (avg + 1.0f);
        if (m.nested < 0 || m.nested >= n)
          throw HkError(HK_ERR_INVALID, "material %d: a coating needs a nested BSDF", i);
        int nt = int(recs[size_t(m.nested) * HK_MATERIAL_STRIDE]);
        if (nt == kCoating || nt == kRoughCoating || nt == kTwoSided)
          throw HkError(HK_ERR_UNSUPPORTED, "material %d: a coating over a coating or twoSided is not supported", i);
        if (m.type == kRoughCoating) {
          m.alpha_u = m.alpha_v = clamp_alpha(r[22]);
          RftTable ext = table(m.dist);
          ext.check(m.eta, m.alpha_u);
          ext.set_eta(m.eta);
          ext.set_alpha(m.alpha_u);
          m.rft_off = int32_t(rft_pool.size());
          rft_pool.insert(rft_pool.end(), ext.t.begin(), ext.t.end());
        }
        break;
      }
      case kMicrofacet: {  // MicrofacetBSDF.cpp:10-28; Ks = 1 - max(kd) (:56)
        m.alpha_u = m.alpha_v = clamp_alpha(r[22]);
        m.spec_weight = 1.0f - std::max(m.a[0], std::max(m.a[1], m.a[2]));
        break;
      }
      case kRoughDiffuse:  // RoughDiffuseBSDF.cpp:10-21
        m.alpha_u = m.alpha_v = clamp_alpha(r[22]);
        break;
      case kTwoSided: {  // TwoSidedBSDF.cpp:93-122
        int back = int(r[12]);
        if (m.nested < 0 || m.nested >= n)
          throw HkError(HK_ERR_INVALID, "material %d: TwoSidedBSDF needs at least one nested BSDF!", i);
        if (back >= n) throw HkError(HK_ERR_INVALID, "material %d: lower-side material out of range", i);
        m.nested_back = back < 0 ? m.nested : back;
        for (int side : {m.nested, m.nested_back}) {
          int t = int(recs[size_t(side) * HK_MATERIAL_STRIDE]);
          if (t == kTwoSided) throw HkError(HK_ERR_UNSUPPORTED, "material %d: twoSided inside twoSided", i);
          if (t == kCoating || t == kRoughCoating) {
            int inner = int(recs[size_t(side) * HK_MATERIAL_STRIDE + 2]);
            if (inner >= 0 && inner < n) t = int(recs[size_t(inner) * HK_MATERIAL_STRIDE]);
          }
          if (t == kDielectric || t == kRoughDielectric)
            throw HkError(HK_ERR_INVALID, "material %d: Only BSDF without transmission can be nested!", i);
        }
        break;
      }
      default:
        break;
    }
    out[i] = m;
  }
  // BSDF::IsDiffuse (DiffuseBSDF.cpp:71, PlasticBSDF.cpp:165,
  // RoughPlasticBSDF.cpp:198 true; coatings forward to the nested BSDF,
  // CoatingBSDF.cpp:194, RoughCoatingBSDF.cpp:223; false otherwise)
  // MicrofacetBSDF.cpp:167 true; a twoSided is diffuse if either side is
  // (TwoSidedBSDF.cpp:77-80)
  auto diffuse_type = [](int t) { return t == kDiffuse || t == kPlastic || t == kRoughPlastic || t == kMicrofacet; };
  auto coated_diffuse = [&](int k) {
    const DevMaterial &m = out[k];
    return diffuse_type(m.type) ||
           ((m.type == kCoating || m.type == kRoughCoating) && diffuse_type(out[m.nested].type));
  };
  for (int i = 0; i < n; i++) {
    DevMaterial &m = out[i];
    bool d = m.type == kTwoSided ? coated_diffuse(m.nested) || coated_diffuse(m.nested_back) : coated_diffuse(i);
    if (d) m.flags |= kMatFlagDiffuse;
  }
  if (rft_pool.empty()) rft_pool.push_back(0.0f);
}

// PerspectiveCamera::Activate (src/camera/PerspectiveCamera.cpp:32-75):
// sample space [0,1]^2 -> camera space on the near plane.
void camera_setup(const float *c2w, float fov, int w, int h, float near_clip, float far_clip, DevRequest &r) {
  double aspect = double(w) / double(h);
  double recip = 1.0 / (double(far_clip) - near_clip);
  double cot = 1.0 / std::tan(double(fov) / 2.0 * M_PI / 180.0);
  Mat4 P = Mat4::identity();
  P.m[0][0] = cot;
  P.m[1][1] = cot;
  P.m[2][2] = far_clip * recip;
  P.m[2][3] = -double(near_clip) * far_clip * recip;
  P.m[3][2] = 1;
  P.m[3][3] = 0;
  Mat4 S = Mat4::identity();
  S.m[0][0] = -0.5;
  S.m[1][1] = -0.5 * aspect;
  Mat4 T = Mat4::identity();
  T.m[0][3] = -1;
  T.m[1][3] = -1.0 / aspect;
  (S * T * P).inverse().to_float(r.s2c);
  std::memcpy(r.c2w, c2w, sizeof(float) * 16);
  r.width = w;
  r.height = h;
  r.inv_w = 1.0f / float(w);
  r.inv_h = 1.0f / float(h);
  r.near_clip = near_clip;
  r.far_clip = far_clip;
}

// Filter tabulation of ImageBlock (src/core/Block.cpp:9-33) for the
// reconstruction filters of src/filter/*.cpp.
void filter_setup(const hk_scene_desc &d, DevRequest &r) {
  int ft = d.filter;
  float radius = ft == 0 ? 0.5f : (ft == 2 ? 1.0f : d.filter_radius);
  if (!(radius > 0.0f)) throw HkError(HK_ERR_INVALID, "filter radius must be positive");
  for (int i = 0; i < 32; i++) {
    float x = (radius * i) / 32;
    float w;
    if (ft == 0) {
      w = 1.0f;
    } else if (ft == 2) {
      w = std::max(0.0f, 1.0f - std::fabs(x));
    } else if (ft == 3) {
      float X = std::fabs(2.0f * x / radius), X2 = X * X, X3 = X2 * X, B = d.filter_b, C = d.filter_c;
      if (X < 1.0f)
        w = 1.0f / 6.0f * ((12.0f - 9.0f * B - 6.0f * C) * X3 + (-18.0f + 12.0f * B + 6.0f * C) * X2 + (6.0f - 2.0f * B));
      else if (X < 2.0f)
        w = 1.0f / 6.0f * ((-B - 6 * C) * X3 + (6.0f * B + 30.0f * C) * X2 + (-12.0f * B - 48.0f * C) * X +
                           (8.0f * B + 24.0f * C));
      else
        w = 0.0f;
    } else if (ft == 1) {
      float al = -1.0f / (2.0f * d.filter_stddev * d.filter_stddev);
      w = std::max(0.0f, std::exp(al * x * x) - std::exp(al * radius * radius));
    } else {
      throw HkError(HK_ERR_INVALID, "unknown filter type %d", ft);
    }
    r.filter[i] = w;
  }
  r.filter[32] = 0.0f;
  r.filter_radius = radius;
  r.lookup = 32 / radius;
}

}  // namespace hk

// ====================================================================== C ABI
#define HK_API_BEGIN try {
#define HK_API_END(ctxp)                                   \
  }                                                        \
  catch (const HkError &e) {                               \
    if (ctxp) (ctxp)->err = e.msg;                         \
    else g_create_error = e.msg;                           \
    return e.code;                                         \
  }                                                        \
  catch (const std::exception &e) {                        \
    if (ctxp) (ctxp)->err = e.what();                      \
    else g_create_error = e.what();                        \
    return HK_ERR_INVALID;                                 \
  }

extern "C" {

const char *hk_version(void) { return "hikari_amd 0.1 (gfx950)"; }

const char *hk_last_error(const hk_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int hk_ctx_create(int device, hk_ctx **out) {
  hk_ctx *c = nullptr;
  HK_API_BEGIN
  if (!out) throw HkError(HK_ERR_INVALID, "out is null");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    throw HkError(HK_ERR_HIP, "no HIP device available (%s)", e == hipSuccess ? "0 devices" : hipGetErrorString(e));
  if (device < 0 || device >= n) throw HkError(HK_ERR_INVALID, "device %d out of range (%d devices)", device, n);
  hip_check(hipSetDevice(device), "hipSetDevice");
  // the device address of this device's copy of the no-alloc scratch
  hip_check(noalloc_scratch_setup(device), "no-alloc scratch address (hipGetSymbolAddress)");
  std::unique_ptr<hk_ctx> ctx(new hk_ctx());
  ctx->device = device;
  hip_check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  ctx->n_cu = prop.multiProcessorCount;
  ctx->arch = prop.gcnArchName;
  hip_check(hipEventCreate(&ctx->ev0), "hipEventCreate");
  hip_check(hipEventCreate(&ctx->ev1), "hipEventCreate");
  hip_check(hipEventCreate(&ctx->ev2), "hipEventCreate");
  hip_check(hipEventCreate(&ctx->ev3), "hipEventCreate");
  hip_check(hipHostMalloc((void **)&ctx->host_counters, 64, hipHostMallocDefault), "hipHostMalloc");
  *out = ctx.release();
  return HK_OK;
  HK_API_END(c)
}

void hk_ctx_destroy(hk_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->ws.release();
  if (ctx->host_counters) (void)hipHostFree(ctx->host_counters);
  (void)hipEventDestroy(ctx->ev0);
  (void)hipEventDestroy(ctx->ev1);
  (void)hipEventDestroy(ctx->ev2);
  (void)hipEventDestroy(ctx->ev3);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int hk_ctx_sync(hk_ctx *ctx) {
  HK_API_BEGIN
  if (!ctx) throw HkError(HK_ERR_INVALID, "ctx is null");
  hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return HK_OK;
  HK_API_END(ctx)
}

int hk_set_option(hk_ctx *ctx, const char *name, int64_t value) {
  HK_API_BEGIN
  if (!ctx || !name) throw HkError(HK_ERR_INVALID, "null argument");
  std::string n(name);
  if (n == "max_paths") {
    if (value < 0 || value > (int64_t(1) << 28)) throw HkError(HK_ERR_INVALID, "max_paths out of range");
    ctx->max_paths = value == 0 ? kDefaultMaxPaths : uint32_t(value);
    ctx->max_paths_set = value != 0;
  } else if (n == "max_samples") {
    if (value < 0 || value > (int64_t(1) << 31)) throw HkError(HK_ERR_INVALID, "max_samples out of range");
    ctx->max_samples = value == 0 ? kDefaultMaxSamples : uint32_t(value);
  } else if (n == "profile") {
    ctx->profile = value != 0;
  } else if (n == "grid") {
    if (value < 0 || value > (1 << 20)) throw HkError(HK_ERR_INVALID, "grid out of range");
    ctx->grid = uint32_t(value);
  } else if (n == "refill" || n == "refill_shadow") {
    if (value < 0 || value > 64) throw HkError(HK_ERR_INVALID, "%s: 1..64 (0: default)", name);
    uint32_t v = value == 0 ? (n == "refill" ? kDefaultRefill : kDefaultRefillShadow) : uint32_t(value);
    (n == "refill" ? ctx->refill_extend : ctx->refill_shadow) = v;
  } else if (n == "fetch_batch") {
    if (value < 0 || value > 64) throw HkError(HK_ERR_INVALID, "fetch_batch: 1..64 (0: default)");
    ctx->fetch_batch = value == 0 ? kDefaultFetchBatch : uint32_t(value);
  } else if (n == "schedule") {
    if (value < 0 || value > 1) throw HkError(HK_ERR_INVALID, "schedule: 0 (FIFO) or 1 (shortest first)");
    ctx->schedule = int(value);
  } else if (n == "express") {
    if (value < 0 || value > 1) throw HkError(HK_ERR_INVALID, "express: 0 (every request on the bulk lane) or 1");
    ctx->express = value != 0;
  } else if (n == "graph") {
    ctx->graph = value != 0;
  } else if (n == "tile") {
    if (!(value == 0 || value == 1 || value == 2 || value == 4 || value == 8 || value == 16 || value == 32))
      throw HkError(HK_ERR_INVALID, "tile: 1, 2, 4, 8, 16 or 32");
    ctx->tile = value == 0 ? 8u : uint32_t(value);
  } else if (n == "resolve_tiles") {
    ctx->resolve_tiles = value != 0;
  } else if (n == "bvh_builder") {
    if (value < 0 || value > 1) throw HkError(HK_ERR_INVALID, "bvh_builder: 0 (host SAH) or 1 (device LBVH)");
    ctx->bvh_builder = int(value);
  } else if (n == "bvh_cell_bits") {
    if (value < 0 || value > 30) throw HkError(HK_ERR_INVALID, "bvh_cell_bits: 0..30");
    ctx->bvh_cell_bits = uint32_t(value);
  } else {
    throw HkError(HK_ERR_INVALID, "unknown option \"%s\"", name);
  }
  return HK_OK;
  HK_API_END(ctx)
}

// ------------------------------------------------------------------ accel
int hk_accel_build(hk_ctx *ctx, const float *v, uint32_t nv, const uint32_t *idx, uint32_t nt, hk_accel **out) {
  HK_API_BEGIN
  if (!ctx || !v || !idx || !out) throw HkError(HK_ERR_INVALID, "null argument");
  if (nt == 0) throw HkError(HK_ERR_INVALID, "no triangles");
  std::vector<float> tri9(size_t(nt) * 9);
  for (uint32_t i = 0; i < nt; i++)
    for (int k = 0; k < 3; k++) {
      uint32_t vi = idx[3 * size_t(i) + k];
      if (vi >= nv) throw HkError(HK_ERR_INVALID, "triangle %u references vertex %u >= %u", i, vi, nv);
      std::memcpy(&tri9[9 * size_t(i) + 3 * k], v + 3 * size_t(vi), 12);
    }
  std::unique_ptr<hk_accel> a(new hk_accel());
  a->ctx = ctx;
  auto t0 = std::chrono::steady_clock::now();
  upload_bvh(ctx, tri9.data(), nt, a->bvh);
  a->fetch.alloc(256);
  a->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  hip_check(hipStreamSynchronize(ctx->stream), "accel upload");
  *out = a.release();
  return HK_OK;
  HK_API_END(ctx)
}

void hk_accel_destroy(hk_accel *a) { delete a; }

int hk_accel_get_info(const hk_accel *a, hk_accel_info *info) {
  if (!a || !info) return HK_ERR_INVALID;
  info->n_triangles = a->bvh.n_tris;
  info->n_nodes = a->bvh.n_nodes;
  info->n_leaves = a->bvh.n_leaves;
  info->max_depth = a->bvh.max_depth;
  info->node_bytes = uint64_t(a->bvh.n_nodes) * sizeof(TraceNode);
  info->triangle_bytes = uint64_t(a->bvh.n_tris) * sizeof(TraceTri);
  info->build_ms = a->build_ms;
  return HK_OK;
}

int hk_accel_intersect_device(hk_accel *a, const hk_ray *d_rays, uint32_t n, hk_hit *d_hits, int any_hit) {
  hk_ctx *ctx = a ? a->ctx : nullptr;
  HK_API_BEGIN
  if (!a || (n && (!d_rays || !d_hits))) throw HkError(HK_ERR_INVALID, "null argument");
  if (n > (1u << 31)) throw HkError(HK_ERR_INVALID, "at most 2^31 rays per call");
  hip_check(launch_intersect_persistent((const TraceNode *)a->bvh.nodes.p, (const TraceTri *)a->bvh.tris.p, 0, d_rays,
                                        d_hits, n, any_hit, a->fetch.as<uint32_t>(),
                                        any_hit ? ctx->refill_shadow : ctx->refill_extend, ctx->fetch_batch,
                                        ctx->stream),
            "k_intersect launch");
  return HK_OK;
  HK_API_END(ctx)
}

int hk_accel_intersect_device_noalloc(hk_accel *a, const hk_ray *d_rays, uint32_t n, hk_hit *d_hits, int any_hit) {
  hk_ctx *ctx = a ? a->ctx : nullptr;
  HK_API_BEGIN
  if (!a || (n && (!d_rays || !d_hits))) throw HkError(HK_ERR_INVALID, "null argument");
  if (n > HK_NOALLOC_MAX_RAYS)
    throw HkError(HK_ERR_INVALID, "hk_accel_intersect_device_noalloc: %u rays exceed the limit of %u per call", n,
                  HK_NOALLOC_MAX_RAYS);
  uint32_t *counter = noalloc_scratch(ctx->device);
  if (!counter) throw HkError(HK_ERR_HIP, "no-alloc scratch of device %d is not set up", ctx->device);
  hip_check(launch_intersect_persistent((const TraceNode *)a->bvh.nodes.p, (const TraceTri *)a->bvh.tris.p, 0, d_rays,
                                        d_hits, n, any_hit, counter,
                                        any_hit ? ctx->refill_shadow : ctx->refill_extend, ctx->fetch_batch,
                                        ctx->stream),
            "k_intersect launch");
  return HK_OK;
  HK_API_END(ctx)
}

int hk_accel_intersect(hk_accel *a, const hk_ray *rays, uint32_t n, hk_hit *hits, int any_hit) {
  hk_ctx *ctx = a ? a->ctx : nullptr;
  HK_API_BEGIN
  if (!a || (n && (!rays || !hits))) throw HkError(HK_ERR_INVALID, "null argument");
  if (n > (1u << 31)) throw HkError(HK_ERR_INVALID, "at most 2^31 rays per call");
  if (n == 0) return HK_OK;
  DBuf dr, dh;
  dr.upload(rays, sizeof(hk_ray) * size_t(n), ctx->stream);
  dh.alloc(sizeof(hk_hit) * size_t(n));
  hip_check(launch_intersect_persistent((const TraceNode *)a->bvh.nodes.p, (const TraceTri *)a->bvh.tris.p, 0, dr.p,
                                        dh.p, n, any_hit, a->fetch.as<uint32_t>(),
                                        any_hit ? ctx->refill_shadow : ctx->refill_extend, ctx->fetch_batch,
                                        ctx->stream),
            "k_intersect launch");
  hip_check(hipMemcpyAsync(hits, dh.p, sizeof(hk_hit) * size_t(n), hipMemcpyDeviceToHost, ctx->stream), "D2H");
  hip_check(hipStreamSynchronize(ctx->stream), "hk_accel_intersect");
  return HK_OK;
  HK_API_END(ctx)
}

// ------------------------------------------------------------------ BSDF batch
int hk_bsdf_batch(hk_ctx *ctx, const float *materials, uint32_t n_materials, uint32_t index, const char *rft_dir,
                  uint32_t n, const float *wi, const float *wo, const int32_t *measure, const float *u,
                  float *out_eval, float *out_pdf, float *out_f, float *out_wo, int32_t *out_measure,
                  float *out_eta, float *out_spdf) {
  HK_API_BEGIN
  if (!ctx || !materials || index >= n_materials) throw HkError(HK_ERR_INVALID, "bad material table / index");
  if (n == 0) return HK_OK;
  std::vector<DevMaterial> mats;
  std::vector<float> pool;
  build_materials(materials, int(n_materials), rft_dir, mats, pool);
  DBuf dm, dp, dwi, dwo, dme, du, oe, op, of, ow, om, oeta, osp;
  hipStream_t st = ctx->stream;
  dm.upload(mats.data(), mats.size() * sizeof(DevMaterial), st);
  dp.upload(pool.data(), pool.size() * 4, st);
  dwi.upload(wi, size_t(n) * 12, st);
  dwo.upload(wo, size_t(n) * 12, st);
  dme.upload(measure, size_t(n) * 4, st);
  du.upload(u, size_t(n) * 16, st);
  oe.alloc(size_t(n) * 12); op.alloc(size_t(n) * 4); of.alloc(size_t(n) * 12); ow.alloc(size_t(n) * 12);
  om.alloc(size_t(n) * 4); oeta.alloc(size_t(n) * 4); osp.alloc(size_t(n) * 4);
  hip_check(launch_bsdf_batch((const DevMaterial *)dm.p, (const float *)dp.p, int(index), n, (const float *)dwi.p,
                              (const float *)dwo.p, (const int32_t *)dme.p, (const float *)du.p, (float *)oe.p,
                              (float *)op.p, (float *)of.p, (float *)ow.p, (int32_t *)om.p, (float *)oeta.p,
                              (float *)osp.p, st),
            "k_bsdf_batch launch");
  struct { void *dst; DBuf *src; size_t n; } cp[] = {
      {out_eval, &oe, size_t(n) * 12}, {out_pdf, &op, size_t(n) * 4}, {out_f, &of, size_t(n) * 12},
      {out_wo, &ow, size_t(n) * 12},   {out_measure, &om, size_t(n) * 4}, {out_eta, &oeta, size_t(n) * 4},
      {out_spdf, &osp, size_t(n) * 4}};
  for (auto &c : cp)
    if (c.dst) hip_check(hipMemcpyAsync(c.dst, c.src->p, c.n, hipMemcpyDeviceToHost, st), "D2H");
  hip_check(hipStreamSynchronize(st), "hk_bsdf_batch");
  return HK_OK;
  HK_API_END(ctx)
}

}  // extern "C"

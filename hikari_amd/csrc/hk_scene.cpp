// hk_scene.cpp -- scene upload (Scene::Activate and the Activate() of meshes,
// emitters and camera, src/core/Scene.cpp:128-165) and the wavefront pass
// scheduler behind hk_render / hk_render_batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "../../include/hikari_amd.h"
#include "hk_build.h"
#include "hk_engine.h"
#include "hk_kernels.h"
#include "hk_math.h"
#include "hk_runtime.h"

using namespace hk;

namespace hk {

extern thread_local std::string g_create_error;

void upload_bvh(hk_ctx *ctx, const float *tri9, uint32_t nt, DevBvh &out) {
  if (ctx->bvh_builder == 1 && nt >= 2) {
    build_bvh4_device(ctx->stream, tri9, nt, ctx->bvh_cell_bits, out);
    if (3 * out.max_depth + 1 > 96)
      throw HkError(HK_ERR_UNSUPPORTED, "device BVH depth %u exceeds the traversal stack (use the host builder)",
                    out.max_depth);
    return;
  }
  Bvh4Result r;
  try {
    build_bvh4(tri9, nt, r);
  } catch (const std::runtime_error &e) {
    throw HkError(HK_ERR_INVALID, "BVH build: %s", e.what());
  }
  // the traversal stack holds at most 3 entries per level below the root
  if (3 * r.max_depth + 1 > 96)
    throw HkError(HK_ERR_UNSUPPORTED, "BVH depth %u exceeds the traversal stack", r.max_depth);
  out.nodes.upload(r.nodes.data(), r.nodes.size() * sizeof(Bvh4Node), ctx->stream);
  out.tris.upload(r.tris.data(), r.tris.size() * sizeof(TraceTri), ctx->stream);
  out.n_nodes = uint32_t(r.nodes.size());
  out.n_tris = nt;
  out.n_leaves = r.n_leaves;
  out.max_depth = r.max_depth;
  hip_check(hipStreamSynchronize(ctx->stream), "BVH upload");
}

static float bbox_radius(const float *lo, const float *hi) {
  float e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  return std::sqrt(e[0] * e[0] + (e[1] * e[1] + e[2] * e[2])) * 0.5f;  // BoundingBox::GetRadius
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

int hk_scene_create(hk_ctx *ctx, const hk_scene_desc *d, hk_scene **out) {
  HK_API_BEGIN
  if (!ctx || !d || !out) throw HkError(HK_ERR_INVALID, "null argument");
  if (d->n_meshes <= 0 || !d->meshes) throw HkError(HK_ERR_INVALID, "scene has no meshes");
  if (d->n_materials <= 0 || !d->materials) throw HkError(HK_ERR_INVALID, "scene has no materials");
  hipStream_t st = ctx->stream;
  std::unique_ptr<hk_scene> s(new hk_scene());
  s->ctx = ctx;
  s->desc = *d;
  s->desc.meshes = nullptr;
  s->desc.materials = nullptr;
  s->desc.env = nullptr;
  s->desc.rft_dir = nullptr;

  // materials (constructors + Activate of src/bsdf/*.cpp)
  std::vector<DevMaterial> mats;
  std::vector<float> pool;
  build_materials(d->materials, d->n_materials, d->rft_dir, mats, pool);

  // meshes: global vertex / face arrays, area CDFs (src/core/Mesh.cpp:233-261)
  std::vector<float> V, N, UV, cdf, tri9;
  std::vector<uint32_t> F, prim_mesh;
  std::vector<DevMesh> meshes(d->n_meshes);
  std::vector<DevEmitter> emitters;
  bool any_n = false, any_uv = false;
  for (int mi = 0; mi < d->n_meshes; mi++) {
    const hk_mesh_desc &m = d->meshes[mi];
    any_n |= m.n != nullptr;
    any_uv |= m.uv != nullptr;
  }
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t prim = 0;
  for (int mi = 0; mi < d->n_meshes; mi++) {
    const hk_mesh_desc &m = d->meshes[mi];
    if (!m.v || !m.f || m.nt == 0 || m.nv == 0) throw HkError(HK_ERR_INVALID, "mesh %d is empty", mi);
    if (m.material < 0 || m.material >= d->n_materials)
      throw HkError(HK_ERR_INVALID, "mesh %d: material %d out of range", mi, m.material);
    uint32_t vbase = uint32_t(V.size() / 3);
    V.insert(V.end(), m.v, m.v + 3 * size_t(m.nv));
    if (any_n) {
      if (m.n) N.insert(N.end(), m.n, m.n + 3 * size_t(m.nv));
      else N.resize(N.size() + 3 * size_t(m.nv), 0.0f);
    }
    if (any_uv) {
      if (m.uv) UV.insert(UV.end(), m.uv, m.uv + 2 * size_t(m.nv));
      else UV.resize(UV.size() + 2 * size_t(m.nv), 0.0f);
    }
    for (uint32_t i = 0; i < m.nv; i++)
      for (int a = 0; a < 3; a++) {
        lo[a] = std::min(lo[a], m.v[3 * size_t(i) + a]);
        hi[a] = std::max(hi[a], m.v[3 * size_t(i) + a]);
      }
    DevMesh &dm = meshes[mi];
    std::memset(&dm, 0, sizeof dm);
    dm.first_prim = prim;
    dm.nt = m.nt;
    dm.material = uint32_t(m.material);
    dm.has_n = m.n != nullptr;
    dm.has_uv = m.uv != nullptr;
    dm.emitter = -1;
    std::vector<float> areas(m.nt);
    float total = 0.0f;
    for (uint32_t t = 0; t < m.nt; t++) {
      const uint32_t *f = m.f + 3 * size_t(t);
      for (int k = 0; k < 3; k++)
        if (f[k] >= m.nv) throw HkError(HK_ERR_INVALID, "mesh %d: face %u references vertex %u >= %u", mi, t, f[k], m.nv);
      vec3 p0 = ld3(m.v + 3 * size_t(f[0])), p1 = ld3(m.v + 3 * size_t(f[1])), p2 = ld3(m.v + 3 * size_t(f[2]));
      areas[t] = 0.5f * norm(cross(p1 - p0, p2 - p0));  // Mesh::SurfaceArea(i)
      total += areas[t];
      F.push_back(vbase + f[0]);
      F.push_back(vbase + f[1]);
      F.push_back(vbase + f[2]);
      prim_mesh.push_back(uint32_t(mi));
      const float *pp[3] = {m.v + 3 * size_t(f[0]), m.v + 3 * size_t(f[1]), m.v + 3 * size_t(f[2])};
      for (int k = 0; k < 3; k++) tri9.insert(tri9.end(), pp[k], pp[k] + 3);
    }
    dm.inv_area = 1.0f / total;
    dm.cdf_off = uint32_t(cdf.size());
    cdf.resize(cdf.size() + m.nt + 1);
    float integral;
    build_cdf(areas.data(), int(m.nt), cdf.data() + dm.cdf_off, integral);
    if (m.is_emitter) {
      dm.emitter = int32_t(emitters.size());
      for (int c = 0; c < 3; c++) dm.radiance[c] = m.radiance[c];
      emitters.push_back(DevEmitter{0, mi});
    }
    prim += m.nt;
  }
  s->n_tris = prim;

  DevScene &ds = s->dev;
  std::memset(&ds, 0, sizeof ds);
  // environment map (src/emitter/EnvironmentLight.cpp:8-33)
  if (d->env) {
    int w = d->env_w, h = d->env_h;
    if (w <= 0 || h <= 0) throw HkError(HK_ERR_INVALID, "bad environment map size");
    std::vector<float> row_func(size_t(w) * h), row_cdf(size_t(w + 1) * h), row_int(h), marg_cdf(h + 1);
    std::vector<uint32_t> row_guide(size_t(w + 1) * h), marg_guide(h + 1);
    for (int y = 0; y < h; y++) {
      float theta = float(y + 0.5f) / float(h) * float(M_PI);
      float sth = std::sin(theta);
      for (int x = 0; x < w; x++) {
        const float *c = d->env + 3 * (size_t(y) * w + x);
        row_func[size_t(y) * w + x] = (c[0] * 0.212671f + c[1] * 0.715160f + c[2] * 0.072169f) * sth;
      }
      build_cdf(&row_func[size_t(y) * w], w, &row_cdf[size_t(y) * (w + 1)], row_int[y]);
      build_cdf_guide(&row_cdf[size_t(y) * (w + 1)], w, &row_guide[size_t(y) * (w + 1)]);
    }
    float marg_int;
    build_cdf(row_int.data(), h, marg_cdf.data(), marg_int);
    build_cdf_guide(marg_cdf.data(), h, marg_guide.data());
    s->env_tex.upload(d->env, size_t(w) * h * 12, st);
    s->env_row_cdf.upload(row_cdf.data(), row_cdf.size() * 4, st);
    s->env_row_func.upload(row_func.data(), row_func.size() * 4, st);
    s->env_row_int.upload(row_int.data(), row_int.size() * 4, st);
    s->env_marg_cdf.upload(marg_cdf.data(), marg_cdf.size() * 4, st);
    s->env_marg_func.upload(row_int.data(), row_int.size() * 4, st);
    s->env_row_guide.upload(row_guide.data(), row_guide.size() * 4, st);
    s->env_marg_guide.upload(marg_guide.data(), marg_guide.size() * 4, st);
    ds.env_tex = s->env_tex.as<float>();
    ds.env_row_cdf = s->env_row_cdf.as<float>();
    ds.env_row_func = s->env_row_func.as<float>();
    ds.env_row_int = s->env_row_int.as<float>();
    ds.env_marg_cdf = s->env_marg_cdf.as<float>();
    ds.env_marg_func = s->env_marg_func.as<float>();
    ds.env_row_guide = s->env_row_guide.as<uint32_t>();
    ds.env_marg_guide = s->env_marg_guide.as<uint32_t>();
    ds.env_marg_int = marg_int;
    ds.env_w = w;
    ds.env_h = h;
    ds.has_env = 1;
    ds.env_scale = d->env_scale;
    Mat4 tw = Mat4::from(d->env_to_world), tl = tw.inverse();
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        ds.env_to_world[3 * i + j] = float(tw.m[i][j]);
        ds.env_to_local[3 * i + j] = float(tl.m[i][j]);
      }
    emitters.push_back(DevEmitter{1, -1});
  }
  if (int(emitters.size()) > kMaxEmitters)
    throw HkError(HK_ERR_UNSUPPORTED, "%zu emitters (at most %d)", emitters.size(), kMaxEmitters);

  upload_bvh(ctx, tri9.data(), prim, s->bvh);
  s->meshes.upload(meshes.data(), meshes.size() * sizeof(DevMesh), st);
  {
    // shading records gathered on the device; the indexed arrays are only
    // the gather's input
    DBuf dV, dN, dUV, dF, dPM;
    dV.upload(V.data(), V.size() * 4, st);
    if (any_n) dN.upload(N.data(), N.size() * 4, st);
    if (any_uv) dUV.upload(UV.data(), UV.size() * 4, st);
    dF.upload(F.data(), F.size() * 4, st);
    dPM.upload(prim_mesh.data(), prim_mesh.size() * 4, st);
    s->shade.alloc(size_t(prim) * sizeof(ShadeTri));
    hip_check(launch_shade_tris(dV.as<float>(), dN.as<float>(), dUV.as<float>(), dF.as<uint32_t>(),
                                dPM.as<uint32_t>(), s->meshes.as<DevMesh>(), s->shade.as<ShadeTri>(), prim, st),
              "shading records");
    hip_check(hipStreamSynchronize(st), "shading records");
  }
  s->mats.upload(mats.data(), mats.size() * sizeof(DevMaterial), st);
  s->rft.upload(pool.data(), pool.size() * 4, st);
  if (!emitters.empty()) s->emitters.upload(emitters.data(), emitters.size() * sizeof(DevEmitter), st);
  s->cdf.upload(cdf.data(), cdf.size() * 4, st);
  {
    std::vector<uint32_t> guide(cdf.size());
    for (const DevMesh &dm : meshes) build_cdf_guide(cdf.data() + dm.cdf_off, int(dm.nt), guide.data() + dm.cdf_off);
    s->cdf_guide.upload(guide.data(), guide.size() * 4, st);
    hip_check(hipStreamSynchronize(st), "CDF guide upload");  // the host vector is freed at the end of the block
  }

  ds.nodes = s->bvh.nodes.as<TraceNode>();
  ds.tris = s->bvh.tris.as<TraceTri>();
  ds.shade = s->shade.as<ShadeTri>();
  ds.meshes = s->meshes.as<DevMesh>();
  ds.mats = s->mats.as<DevMaterial>();
  ds.rft = s->rft.as<float>();
  ds.emitters = s->emitters.as<DevEmitter>();
  ds.cdf = s->cdf.as<float>();
  ds.cdf_guide = s->cdf_guide.as<uint32_t>();
  ds.n_emitters = int32_t(emitters.size());
  ds.n_meshes = d->n_meshes;
  ds.bbox_radius = bbox_radius(lo, hi);
  // The depth property is a uint32 in the reference (PathMISIntegrator.cpp:13,
  // WhittedIntegrator.cpp:13): negative means unlimited.  The depth lives in
  // 18 bits of the path state, capping paths at 262143 vertices; RR ends a
  // path long before (0.95^262143 ~ 1e-5840).
  ds.max_depth = d->max_depth < 0 ? int32_t(kMaxPathDepth) : std::min(d->max_depth, int32_t(kMaxPathDepth));
  if (d->integrator < 0 || d->integrator > 3) throw HkError(HK_ERR_INVALID, "unknown integrator %d", d->integrator);
  ds.integrator = d->integrator;
  ds.rr_start = std::max(0, d->rr_start_depth);
  for (int c = 0; c < 3; c++) ds.background[c] = d->background[c];
  ds.force_background = d->force_background;
  ds.root_ref = 0;
  ds.n_nodes = int32_t(s->bvh.n_nodes);
  ds.n_mats = int32_t(mats.size());
  hip_check(hipStreamSynchronize(st), "scene upload");
  DBuf *all[] = {&s->bvh.nodes, &s->bvh.tris, &s->shade, &s->meshes, &s->mats,
                 &s->rft, &s->emitters, &s->cdf, &s->cdf_guide, &s->env_tex, &s->env_row_cdf, &s->env_row_func,
                 &s->env_row_int, &s->env_marg_cdf, &s->env_marg_func, &s->env_row_guide, &s->env_marg_guide};
  for (DBuf *b : all) s->bytes += b->bytes;
  *out = s.release();
  return HK_OK;
  HK_API_END(ctx)
}

void hk_scene_destroy(hk_scene *s) { delete s; }

uint64_t hk_scene_device_bytes(const hk_scene *s) { return s ? s->bytes : 0; }

int hk_render(hk_scene *s, const hk_request *req, float *out_rgb, hk_render_stats *stats) {
  hk_ctx *ctx = s ? s->ctx : nullptr;
  HK_API_BEGIN
  if (!s || !req) throw HkError(HK_ERR_INVALID, "null argument");
  return hk_render_batch(s, req, 1, &out_rgb, 0, stats);
  HK_API_END(ctx)
}

int hk_render_batch(hk_scene *s, const hk_request *reqs, uint32_t n, float **out_rgb, uint32_t max_paths,
                    hk_render_stats *stats) {
  hk_ctx *ctx = s ? s->ctx : nullptr;
  HK_API_BEGIN
  if (!s || (n && !reqs)) throw HkError(HK_ERR_INVALID, "null argument");
  if (stats) std::memset(stats, 0, sizeof *stats);
  uint32_t pool = max_paths ? max_paths : ctx->max_paths;
  // validate every request before any work starts
  std::vector<std::unique_ptr<RequestRec>> rs;
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; i++) {
    rs.push_back(prepare_request(s, reqs[i], out_rgb ? out_rgb[i] : nullptr, ctx->max_samples));
    total += uint64_t(rs.back()->npix) * rs.back()->spp;
  }
  uint32_t buf = uint32_t(std::min<uint64_t>(ctx->max_samples, std::max<uint64_t>(total, 1)));
  Engine eng(s, ctx->ws, ctx->stream, uint32_t(std::min<uint64_t>(pool, std::max<uint64_t>(total, 1))), buf,
             ctx->max_samples, ctx->profile, ctx->ev0, ctx->ev1);
  for (auto &r : rs) eng.submit(r.get());
  eng.run();
  eng.take_stats(stats);
  if (stats)
    for (auto &r : rs) stats->invalid_samples += r->invalid_samples;
  return HK_OK;
  HK_API_END(ctx)
}

// Coordinate search: wavefront size first (the largest effect), then the
// fetch batch and refill thresholds of the persistent kernels, the camera-ray
// tile edge and the queue-kernel grid cap, each with the best setting found
// so far for the others.  Every calibration render goes
// through hk_render_batch with the candidate in the context options.
int hk_autotune(hk_scene *s, const hk_request *req, uint64_t calib_samples, hk_tune_result *out) {
  hk_ctx *ctx = s ? s->ctx : nullptr;
  HK_API_BEGIN
  if (!s || !req || !out) throw HkError(HK_ERR_INVALID, "null argument");
  auto t0 = std::chrono::steady_clock::now();
  std::memset(out, 0, sizeof *out);
  const hk_scene_desc &d = s->desc;
  uint64_t npix = req->use_scene_camera ? uint64_t(d.width) * d.height : uint64_t(req->width) * req->height;
  if (npix == 0) throw HkError(HK_ERR_INVALID, "empty film");
  if (calib_samples == 0) calib_samples = uint64_t(64) << 20;
  hk_request cal = *req;
  cal.spp = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(req->spp ? req->spp : 1, calib_samples / npix)));
  cal.spp_offset = 0;
  std::vector<float> film(size_t(npix) * 3);
  float *films[1] = {film.data()};
  struct Setting {
    uint32_t paths, grid, tile, re, rs, fb;
  };
  Setting best{ctx->max_paths, ctx->grid, ctx->tile, ctx->refill_extend, ctx->refill_shadow, ctx->fetch_batch};
  bool saved_profile = ctx->profile;
  ctx->profile = false;
  auto measure = [&](const Setting &t) {
    ctx->max_paths = t.paths;
    ctx->grid = t.grid;
    ctx->tile = t.tile;
    ctx->refill_extend = t.re;
    ctx->refill_shadow = t.rs;
    ctx->fetch_batch = t.fb;
    hk_render_stats st;
    int rc = hk_render_batch(s, &cal, 1, films, 0, &st);
    if (rc != HK_OK) throw HkError(rc, "calibration render: %s", ctx->err.c_str());
    double rate = st.gpu_ms > 0 ? double(st.paths) / (st.gpu_ms * 1e-3) : 0.0;
    if (out->n_tried < HK_TUNE_MAX) {
      uint32_t k = out->n_tried++;
      out->tried_max_paths[k] = t.paths;
      out->tried_grid[k] = t.grid;
      out->tried_tile[k] = t.tile;
      out->tried_refill[k] = t.re;
      out->tried_refill_shadow[k] = t.rs;
      out->tried_fetch_batch[k] = t.fb;
      out->samples_per_s[k] = rate;
    }
    return rate;
  };
  try {
    measure(best);  // warm-up: workspace, code and caches
    out->n_tried = 0;
    double best_rate = 0.0;
    auto consider = [&](const Setting &t) {
      double r = measure(t);
      if (r > best_rate) {
        best_rate = r;
        best = t;
      }
    };
    uint64_t total = npix * cal.spp;
    for (uint32_t p : {1u << 22, 1u << 23, 1u << 24, 1u << 25, 1u << 26, 1u << 27, 1u << 28}) {
      if (p > total && p != (1u << 22)) break;  // a wavefront larger than the work is never filled
      Setting t = best;
      t.paths = p;
      consider(t);
    }
    for (uint32_t fb : {1u, 2u, 4u, 8u}) {
      if (fb == best.fb) continue;
      Setting t = best;
      t.fb = fb;
      consider(t);
    }
    for (uint32_t re : {24u, 32u, 40u, 48u}) {
      if (re == best.re) continue;
      Setting t = best;
      t.re = re;
      consider(t);
    }
    for (uint32_t rs : {16u, 24u, 32u, 48u}) {
      if (rs == best.rs) continue;
      Setting t = best;
      t.rs = rs;
      consider(t);
    }
    for (uint32_t tile : {1u, 4u, 8u, 16u}) {
      if (tile == best.tile) continue;
      Setting t = best;
      t.tile = tile;
      consider(t);
    }
    for (uint32_t g : {0u, uint32_t(ctx->n_cu) * 8u, uint32_t(ctx->n_cu) * 32u}) {
      if (g == best.grid) continue;
      Setting t = best;
      t.grid = g;
      consider(t);
    }
  } catch (...) {
    ctx->profile = saved_profile;
    throw;
  }
  ctx->profile = saved_profile;
  ctx->max_paths = best.paths;
  ctx->grid = best.grid;
  ctx->tile = best.tile;
  ctx->refill_extend = best.re;
  ctx->refill_shadow = best.rs;
  ctx->fetch_batch = best.fb;
  ctx->ws.release();  // the next render sizes the workspace for the chosen wavefront
  out->max_paths = best.paths;
  out->grid = best.grid;
  out->tile = best.tile;
  out->refill = best.re;
  out->refill_shadow = best.rs;
  out->fetch_batch = best.fb;
  out->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return HK_OK;
  HK_API_END(ctx)
}

int hk_render_pixel_samples(hk_scene *s, uint32_t seed, uint32_t x, uint32_t y, uint32_t s0, uint32_t n, float *out) {
  hk_ctx *ctx = s ? s->ctx : nullptr;
  HK_API_BEGIN
  if (!s || !out) throw HkError(HK_ERR_INVALID, "null argument");
  const hk_scene_desc &d = s->desc;
  if (x >= uint32_t(d.width) || y >= uint32_t(d.height)) throw HkError(HK_ERR_INVALID, "pixel out of range");
  uint32_t pix = y * uint32_t(d.width) + x;
  // a one-pixel request: the scene's film with the pixel range narrowed to
  // `pix` (pix0 shifts the pixel index of every sample)
  hk_request rq;
  std::memset(&rq, 0, sizeof rq);
  rq.use_scene_camera = 1;
  rq.seed = seed;
  rq.spp = n;
  rq.spp_offset = s0;
  uint32_t cap = std::max(ctx->max_samples, uint32_t(d.width) * uint32_t(d.height));
  std::unique_ptr<RequestRec> r = prepare_request(s, rq, nullptr, cap);
  r->dev.pix0 = pix;
  r->dev.tile = 0u;
  r->npix = 1;
  r->chunk = std::max(1u, std::min(n, ctx->max_samples));
  r->n_chunks = (n + r->chunk - 1) / r->chunk;
  r->chunk_item.assign(r->n_chunks, -1);
  r->keep_samples = true;
  Engine eng(s, ctx->ws, ctx->stream, std::min(ctx->max_paths, std::max(1u, r->chunk)), std::max(1u, r->chunk),
             r->chunk, false, ctx->ev0, ctx->ev1);
  eng.submit(r.get());
  eng.run();
  for (uint32_t i = 0; i < n; i++) {
    Rng rg = rng_for_path(seed, pix, s0 + i);
    float jx = rg.next(), jy = rg.next();
    float *o = out + 5 * size_t(i);
    o[0] = r->samples[4 * size_t(i)];
    o[1] = r->samples[4 * size_t(i) + 1];
    o[2] = r->samples[4 * size_t(i) + 2];
    o[3] = float(x) + jx;
    o[4] = float(y) + jy;
  }
  return HK_OK;
  HK_API_END(ctx)
}

}  // extern "C"

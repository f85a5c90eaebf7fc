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
#include "hk_kernels.h"
#include "hk_math.h"
#include "hk_runtime.h"

using namespace hk;

namespace hk {

extern thread_local std::string g_create_error;

void upload_bvh(hk_ctx *ctx, const float *tri9, uint32_t nt, DevBvh &out) {
  Bvh4Result r;
  try {
    build_bvh4(tri9, nt, r);
  } catch (const std::runtime_error &e) {
    throw HkError(HK_ERR_INVALID, "BVH build: %s", e.what());
  }
  // the traversal stack holds at most 3 entries per level below the root
  if (3 * r.max_depth + 1 > 96)
    throw HkError(HK_ERR_UNSUPPORTED, "BVH depth %u exceeds the traversal stack", r.max_depth);
#ifdef HK_NODE_FLOAT
  out.nodes.upload(r.nodes.data(), r.nodes.size() * sizeof(Bvh4Node), ctx->stream);
#else
  out.nodes.upload(r.qnodes.data(), r.qnodes.size() * sizeof(Bvh4QNode), ctx->stream);
#endif
  out.tris.upload(r.tris.data(), r.tris.size() * sizeof(TraceTri), ctx->stream);
  out.n_nodes = uint32_t(r.nodes.size());
  out.n_tris = nt;
  out.n_leaves = r.n_leaves;
  out.max_depth = r.max_depth;
  hip_check(hipStreamSynchronize(ctx->stream), "BVH upload");
}

void Workspace::ensure(uint32_t paths, uint32_t emitters, uint32_t samples) {
  emitters = std::max(1u, emitters);
  if (samples > cap_samples) {
    sample_L.alloc(size_t(samples) * 16);
    cap_samples = samples;
  }
  if (paths <= cap_paths && emitters <= cap_emitters) return;
  paths = std::max(paths, cap_paths);
  emitters = std::max(emitters, cap_emitters);
  size_t P = paths;
  for (int i = 0; i < 2; i++) {
    q_o[i].alloc(P * 16); q_d[i].alloc(P * 16); q_beta[i].alloc(P * 16); q_L[i].alloc(P * 16);
    q_aux[i].alloc(P * 16); q_meta[i].alloc(P * 4);
  }
  hit.alloc(P * 16);
  s_ent.alloc(P * 16);
  sh_o.alloc(P * emitters * 16); sh_d.alloc(P * emitters * 16); sh_c.alloc(P * emitters * 16);
  counters.alloc(64);
  stats.alloc(64);
  cap_paths = paths;
  cap_emitters = emitters;
}
void Workspace::release() {
  DBuf *all[] = {&sample_L, &q_o[0], &q_o[1], &q_d[0], &q_d[1], &q_beta[0], &q_beta[1], &q_L[0], &q_L[1],
                 &q_aux[0], &q_aux[1], &q_meta[0], &q_meta[1], &hit, &s_ent, &sh_o, &sh_d, &sh_c, &counters,
                 &stats, &reqs, &items};
  for (DBuf *b : all) b->release();
  cap_paths = cap_emitters = cap_samples = 0;
}

static float bbox_radius(const float *lo, const float *hi) {
  float e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  return std::sqrt(e[0] * e[0] + (e[1] * e[1] + e[2] * e[2])) * 0.5f;  // BoundingBox::GetRadius
}

static RenderArgs make_args(hk_scene *s) {
  Workspace &ws = s->ctx->ws;
  RenderArgs A;
  std::memset(&A, 0, sizeof A);
  A.sc = s->dev;
  A.reqs = ws.reqs.as<DevRequest>();
  A.items = ws.items.as<WorkItem>();
  A.sample_L = ws.sample_L.as<float4>();
  for (int i = 0; i < 2; i++) {
    A.q_o[i] = ws.q_o[i].as<float4>();
    A.q_d[i] = ws.q_d[i].as<float4>();
    A.q_beta[i] = ws.q_beta[i].as<float4>();
    A.q_L[i] = ws.q_L[i].as<float4>();
    A.q_aux[i] = ws.q_aux[i].as<uint4>();
    A.q_meta[i] = ws.q_meta[i].as<uint32_t>();
  }
  A.hit = ws.hit.as<float4>();
  A.s_ent = ws.s_ent.as<uint4>();
  A.sh_o = ws.sh_o.as<float4>();
  A.sh_d = ws.sh_d.as<float4>();
  A.sh_c = ws.sh_c.as<float4>();
  A.counters = ws.counters.as<uint32_t>();
  A.stats = ws.stats.as<unsigned long long>();
  return A;
}

struct ReqState {
  DevRequest dev;
  uint32_t npix, spp, spp_offset, chunk, done;
  DBuf accum;
};

static uint32_t kernel_grid(hk_ctx *ctx, uint32_t count) {
  uint32_t g = (count + 255) / 256;
  if (ctx->grid) g = std::min(g, ctx->grid);
  return std::max(1u, g);
}

// Renders the given requests.  Their samples are cut into fixed chunks
// (`chunk` samples per pixel, set by the sample capacity alone) and packed
// into segments of at most cap_samples samples; each segment runs the
// bounce loop with up to `pool` paths in flight (finished paths are replaced
// by new samples at every bounce) until all its samples have finished, then
// resolves its chunks into the films in order.  Neither the
// chunks nor a sample's path depend on the other requests, so every request
// accumulates the same values in the same order alone or in a batch.
static void run_passes(hk_scene *s, std::vector<ReqState> &rs, hk_render_stats *stats, uint32_t pool_paths,
                       uint32_t cap_samples, bool no_resolve = false) {
  hk_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  Workspace &ws = ctx->ws;
  std::vector<DevRequest> dreq(rs.size());
  for (size_t i = 0; i < rs.size(); i++) dreq[i] = rs[i].dev;
  uint64_t total = 0;
  for (const ReqState &q : rs) total += uint64_t(q.npix) * (q.spp - q.done);
  uint32_t seg_cap = uint32_t(std::min<uint64_t>(cap_samples, std::max<uint64_t>(total, 1)));
  uint32_t pool = std::max(1u, std::min(pool_paths, seg_cap));
  ws.ensure(pool, uint32_t(s->dev.n_emitters), seg_cap);
  ws.reqs.upload(dreq.data(), dreq.size() * sizeof(DevRequest), st);
  size_t max_items = 1;
  for (const ReqState &q : rs) max_items += (q.spp - q.done + q.chunk - 1) / q.chunk;
  ws.items.alloc(sizeof(WorkItem) * max_items);
  hip_check(hipMemsetAsync(ws.stats.p, 0, 64, st), "memset stats");
  RenderArgs A = make_args(s);
  A.n_paths = pool;
  double trace_ms = 0;
  uint64_t launches = 0, paths_total = 0, trace_launches = 0;
  hip_check(hipEventRecord(ctx->ev2, st), "event");
  std::vector<WorkItem> items;
  for (;;) {
    items.clear();
    uint32_t used = 0;
    for (uint32_t r = 0; r < rs.size(); r++) {
      ReqState &q = rs[r];
      while (q.done < q.spp) {
        uint32_t k = std::min(q.chunk, q.spp - q.done);
        uint64_t need = uint64_t(q.npix) * k;
        if (used + need > seg_cap) break;
        WorkItem it;
        std::memset(&it, 0, sizeof it);
        it.req = r;
        it.sample_begin = used;
        it.npix = q.npix;
        it.s0 = q.spp_offset + q.done;
        it.k = k;
        items.push_back(it);
        used += uint32_t(need);
        q.done += k;
      }
      if (used >= seg_cap) break;
    }
    if (items.empty()) break;
    hip_check(hipMemcpyAsync(ws.items.p, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, st),
              "items upload");
    A.n_items = uint32_t(items.size());
    A.n_samples = used;
    paths_total += used;
    uint32_t count = 0, next = 0;
    int qi = 0;
    for (;;) {
      uint32_t n_new = std::min(pool - count, used - next);
      if (n_new) {
        hip_check(launch_regen(A, qi, count, n_new, next, st), "k_regen");
        next += n_new;
        count += n_new;
      }
      if (count == 0) break;
      uint32_t g = kernel_grid(ctx, count);
      if (ctx->profile) hip_check(hipEventRecord(ctx->ev0, st), "event");
      hip_check(launch_extend(A, qi, g, st), "k_extend");
      if (ctx->profile) hip_check(hipEventRecord(ctx->ev1, st), "event");
      hip_check(launch_shade(A, qi, g, st), "k_shade");
      hip_check(hipMemcpyAsync(ctx->host_counters, A.counters, 16, hipMemcpyDeviceToHost, st), "counters D2H");
      hip_check(hipStreamSynchronize(st), "bounce sync");
      if (ctx->profile) {
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event time");
        trace_ms += ms;
        trace_launches++;
      }
      uint32_t nshadow = ctx->host_counters[2];
      if (nshadow) hip_check(launch_shadow(A, qi ^ 1, kernel_grid(ctx, nshadow), st), "k_shadow");
      launches++;
      qi ^= 1;
      count = ctx->host_counters[qi];
    }
    if (no_resolve) {  // caller reads the per-sample results of this segment
      hip_check(hipStreamSynchronize(st), "segment sync");
      break;
    }
    for (uint32_t i = 0; i < items.size(); i++)
      hip_check(launch_resolve(A, i, items[i].npix, st), "k_resolve");
  }
  hip_check(hipEventRecord(ctx->ev3, st), "event");
  unsigned long long hs[6] = {0, 0, 0, 0, 0, 0};
  hip_check(hipMemcpyAsync(hs, ws.stats.p, sizeof hs, hipMemcpyDeviceToHost, st), "stats D2H");
  hip_check(hipStreamSynchronize(st), "render sync");
  if (stats) {
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ctx->ev2, ctx->ev3), "event time");
    stats->paths += paths_total;
    stats->extension_rays += hs[0];
    stats->shadow_rays += hs[1];
    stats->bounces_launched += launches;
    stats->gpu_ms += ms;
    stats->trace_ms += trace_ms;
    stats->trace_launches += trace_launches;
    stats->node_visits += hs[2] + hs[4];
    stats->triangle_tests += hs[3] + hs[5];
  }
}

static void request_setup(hk_scene *s, const hk_request *req, ReqState &q, uint32_t cap_samples) {
  const hk_scene_desc &d = s->desc;
  const float *c2w = req->use_scene_camera ? d.cam_to_world : req->cam_to_world;
  float fov = req->use_scene_camera ? d.fov : req->fov;
  int w = req->use_scene_camera ? d.width : req->width;
  int h = req->use_scene_camera ? d.height : req->height;
  if (w <= 0 || h <= 0 || uint64_t(w) * uint64_t(h) > (1u << 26))
    throw HkError(HK_ERR_INVALID, "bad film size %dx%d", w, h);
  if (!(fov > 0.0f && fov < 180.0f)) throw HkError(HK_ERR_INVALID, "bad field of view %f", fov);
  std::memset(&q.dev, 0, sizeof q.dev);
  camera_setup(c2w, fov, w, h, d.near_clip, d.far_clip, q.dev);
  filter_setup(d, q.dev);
  q.dev.seed = req->seed;
  q.npix = uint32_t(w) * uint32_t(h);
  if (q.npix > cap_samples)
    throw HkError(HK_ERR_INVALID, "film of %u pixels exceeds the sample capacity (%u)", q.npix, cap_samples);
  q.spp = req->spp;
  q.spp_offset = req->spp_offset;
  q.chunk = std::max(1u, std::min(q.spp == 0 ? 1u : q.spp, cap_samples / q.npix));
  q.done = 0;
  q.accum.alloc(size_t(q.npix) * 16);
  hip_check(hipMemsetAsync(q.accum.p, 0, size_t(q.npix) * 16, s->ctx->stream), "accum clear");
  q.dev.accum = q.accum.as<float>();
}

static void finalize_to_host(hk_scene *s, ReqState &q, float *out_rgb) {
  if (!out_rgb) return;
  DBuf rgb;
  rgb.alloc(size_t(q.npix) * 12);
  hip_check(launch_finalize(q.accum.p, rgb.as<float>(), q.npix, s->ctx->stream), "k_finalize");
  hip_check(hipMemcpyAsync(out_rgb, rgb.p, size_t(q.npix) * 12, hipMemcpyDeviceToHost, s->ctx->stream), "D2H");
  hip_check(hipStreamSynchronize(s->ctx->stream), "finalize");
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
    for (int y = 0; y < h; y++) {
      float theta = float(y + 0.5f) / float(h) * float(M_PI);
      float sth = std::sin(theta);
      for (int x = 0; x < w; x++) {
        const float *c = d->env + 3 * (size_t(y) * w + x);
        row_func[size_t(y) * w + x] = (c[0] * 0.212671f + c[1] * 0.715160f + c[2] * 0.072169f) * sth;
      }
      build_cdf(&row_func[size_t(y) * w], w, &row_cdf[size_t(y) * (w + 1)], row_int[y]);
    }
    float marg_int;
    build_cdf(row_int.data(), h, marg_cdf.data(), marg_int);
    s->env_tex.upload(d->env, size_t(w) * h * 12, st);
    s->env_row_cdf.upload(row_cdf.data(), row_cdf.size() * 4, st);
    s->env_row_func.upload(row_func.data(), row_func.size() * 4, st);
    s->env_row_int.upload(row_int.data(), row_int.size() * 4, st);
    s->env_marg_cdf.upload(marg_cdf.data(), marg_cdf.size() * 4, st);
    s->env_marg_func.upload(row_int.data(), row_int.size() * 4, st);
    ds.env_tex = s->env_tex.as<float>();
    ds.env_row_cdf = s->env_row_cdf.as<float>();
    ds.env_row_func = s->env_row_func.as<float>();
    ds.env_row_int = s->env_row_int.as<float>();
    ds.env_marg_cdf = s->env_marg_cdf.as<float>();
    ds.env_marg_func = s->env_marg_func.as<float>();
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
  s->V.upload(V.data(), V.size() * 4, st);
  if (any_n) s->N.upload(N.data(), N.size() * 4, st);
  if (any_uv) s->UV.upload(UV.data(), UV.size() * 4, st);
  s->F.upload(F.data(), F.size() * 4, st);
  s->prim_mesh.upload(prim_mesh.data(), prim_mesh.size() * 4, st);
  s->meshes.upload(meshes.data(), meshes.size() * sizeof(DevMesh), st);
  s->mats.upload(mats.data(), mats.size() * sizeof(DevMaterial), st);
  s->rft.upload(pool.data(), pool.size() * 4, st);
  if (!emitters.empty()) s->emitters.upload(emitters.data(), emitters.size() * sizeof(DevEmitter), st);
  s->cdf.upload(cdf.data(), cdf.size() * 4, st);

  ds.nodes = s->bvh.nodes.as<TraceNode>();
  ds.tris = s->bvh.tris.as<TraceTri>();
  ds.V = s->V.as<float>();
  ds.N = s->N.as<float>();
  ds.UV = s->UV.as<float>();
  ds.F = s->F.as<uint32_t>();
  ds.prim_mesh = s->prim_mesh.as<uint32_t>();
  ds.meshes = s->meshes.as<DevMesh>();
  ds.mats = s->mats.as<DevMaterial>();
  ds.rft = s->rft.as<float>();
  ds.emitters = s->emitters.as<DevEmitter>();
  ds.cdf = s->cdf.as<float>();
  ds.n_emitters = int32_t(emitters.size());
  ds.n_meshes = d->n_meshes;
  ds.bbox_radius = bbox_radius(lo, hi);
  ds.max_depth = d->max_depth;
  for (int c = 0; c < 3; c++) ds.background[c] = d->background[c];
  ds.force_background = d->force_background;
  ds.root_ref = 0;
  hip_check(hipStreamSynchronize(st), "scene upload");
  DBuf *all[] = {&s->bvh.nodes, &s->bvh.tris, &s->V, &s->N, &s->UV, &s->F, &s->prim_mesh, &s->meshes, &s->mats,
                 &s->rft, &s->emitters, &s->cdf, &s->env_tex, &s->env_row_cdf, &s->env_row_func, &s->env_row_int,
                 &s->env_marg_cdf, &s->env_marg_func};
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
  if (stats) std::memset(stats, 0, sizeof *stats);
  std::vector<ReqState> rs(1);
  request_setup(s, req, rs[0], ctx->max_samples);
  run_passes(s, rs, stats, ctx->max_paths, ctx->max_samples);
  finalize_to_host(s, rs[0], out_rgb);
  return HK_OK;
  HK_API_END(ctx)
}

int hk_render_batch(hk_scene *s, const hk_request *reqs, uint32_t n, float **out_rgb, uint32_t max_paths,
                    hk_render_stats *stats) {
  hk_ctx *ctx = s ? s->ctx : nullptr;
  HK_API_BEGIN
  if (!s || (n && !reqs)) throw HkError(HK_ERR_INVALID, "null argument");
  if (stats) std::memset(stats, 0, sizeof *stats);
  uint32_t pool = max_paths ? max_paths : ctx->max_paths;
  std::vector<ReqState> rs(n);
  for (uint32_t i = 0; i < n; i++) request_setup(s, &reqs[i], rs[i], ctx->max_samples);
  run_passes(s, rs, stats, pool, ctx->max_samples);
  for (uint32_t i = 0; i < n; i++) finalize_to_host(s, rs[i], out_rgb ? out_rgb[i] : nullptr);
  return HK_OK;
  HK_API_END(ctx)
}

int hk_render_pixel_samples(hk_scene *s, uint32_t seed, uint32_t x, uint32_t y, uint32_t s0, uint32_t n, float *out) {
  hk_ctx *ctx = s ? s->ctx : nullptr;
  HK_API_BEGIN
  if (!s || !out) throw HkError(HK_ERR_INVALID, "null argument");
  const hk_scene_desc &d = s->desc;
  if (x >= uint32_t(d.width) || y >= uint32_t(d.height)) throw HkError(HK_ERR_INVALID, "pixel out of range");
  uint32_t cap = ctx->max_samples;
  uint32_t pix = y * uint32_t(d.width) + x;
  for (uint32_t done = 0; done < n;) {
    uint32_t k = std::min(cap, n - done);
    // a one-pixel "request": pixel index offset handled through the film of
    // width 1 and a pixel-index shift in the RNG key
    std::vector<ReqState> rs(1);
    hk_request rq;
    std::memset(&rq, 0, sizeof rq);
    rq.use_scene_camera = 1;
    rq.seed = seed;
    rq.spp = k;
    rq.spp_offset = s0 + done;
    request_setup(s, &rq, rs[0], std::max(cap, uint32_t(d.width) * uint32_t(d.height)));
    rs[0].dev.pix0 = pix;
    rs[0].npix = 1;
    rs[0].chunk = k;
    run_passes(s, rs, nullptr, ctx->max_paths, std::max(k, 1u), true);
    std::vector<float> L(size_t(k) * 4);
    hip_check(hipMemcpy(L.data(), ctx->ws.sample_L.p, size_t(k) * 16, hipMemcpyDeviceToHost), "D2H");
    for (uint32_t i = 0; i < k; i++) {
      Rng r = rng_for_path(seed, pix, s0 + done + i);
      float jx = r.next(), jy = r.next();
      float *o = out + 5 * size_t(done + i);
      o[0] = L[4 * i];
      o[1] = L[4 * i + 1];
      o[2] = L[4 * i + 2];
      o[3] = float(x) + jx;
      o[4] = float(y) + jy;
    }
    done += k;
  }
  return HK_OK;
  HK_API_END(ctx)
}

}  // extern "C"

// hk_oracle_scene.hpp -- TEST INFRASTRUCTURE ONLY (see hk_oracle.hpp header).
//
// Geometry, emitters, camera, film and the pathMIS estimator of the CPU
// restatement.  Ray/triangle arithmetic is always FP32 (the hit-ID contract);
// shading runs in R.
#pragma once

#include <thread>

#include "hk_oracle.hpp"

namespace hko {

// ---------------------------------------------------------------- rays / triangles
struct RayF {
  V3<float> o, d;
  float tmin, tmax;
};

// src/core/Mesh.cpp:340-384 (Moller-Trumbore, Eigen evaluation order).
inline bool mt_intersect(const V3<float> &p0, const V3<float> &p1, const V3<float> &p2, const RayF &r, float &u,
                         float &v, float &t) {
  V3<float> e1 = p1 - p0, e2 = p2 - p0;
  V3<float> pvec = cross(r.d, e2);
  float det = dot(e1, pvec);
  if (det > -1e-8f && det < 1e-8f) return false;
  float inv = 1.0f / det;
  V3<float> tvec = r.o - p0;
  u = dot(tvec, pvec) * inv;
  if (u < 0.0 || u > 1.0) return false;
  V3<float> qvec = cross(tvec, e1);
  v = dot(r.d, qvec) * inv;
  if (v < 0.0 || u + v > 1.0) return false;
  t = dot(e2, qvec) * inv;
  return t >= r.tmin && t <= r.tmax;
}

struct Hit {
  float t = std::numeric_limits<float>::infinity(), u = 0, v = 0;
  int64_t prim = -1;
};

// Triangle soup over all meshes in scene order (src/core/Acceleration.cpp:13-28:
// the accel's shape list is mesh-major, face-minor).
struct TriSoup {
  std::vector<V3<float>> p0, p1, p2;
  size_t size() const { return p0.size(); }
};

// src/core/Acceleration.cpp:45-80: closest hit keeps the LAST tested triangle
// among equal t (t <= MaxT is accepted), i.e. max prim id among ties.
inline bool brute_intersect(const TriSoup &s, const RayF &ray, Hit &h, bool any) {
  RayF r = ray;
  bool found = false;
  for (size_t i = 0; i < s.size(); i++) {
    float u, v, t;
    if (mt_intersect(s.p0[i], s.p1[i], s.p2[i], r, u, v, t)) {
      if (any) { h.prim = int64_t(i); h.t = t; h.u = u; h.v = v; return true; }
      r.tmax = t;
      h.t = t; h.u = u; h.v = v; h.prim = int64_t(i);
      found = true;
    }
  }
  return found;
}

// Binary BVH restating src/acceleration/BVHAcceleration.cpp:46-248 (centroid
// major axis, 12-bucket SAH, leaf size 10).  Traversal differs from the
// reference on purpose: the caller's Intersection.T is never read (the stale-T
// pruning of BVHAcceleration.cpp:276 is neutralised) and ties resolve to the
// max prim id, so hits equal brute_intersect bit for bit.  Box tests are done
// in double with the boxes padded, so that no triangle the FP32 M-T test
// accepts is culled.
struct Bvh {
  struct Node {
    double lo[3], hi[3];
    int32_t left = -1, right = -1;  // children (internal) or -1
    uint32_t start = 0, count = 0;  // leaf range into order[]
  };
  std::vector<Node> nodes;
  std::vector<uint32_t> order;
  const TriSoup *soup = nullptr;

  void build(const TriSoup &s, int leaf_size = 10) {
    soup = &s;
    size_t n = s.size();
    order.resize(n);
    for (size_t i = 0; i < n; i++) order[i] = uint32_t(i);
    std::vector<double> bl(n * 3), bh(n * 3), ce(n * 3);
    for (size_t i = 0; i < n; i++)
      for (int a = 0; a < 3; a++) {
        double x0 = s.p0[i][a], x1 = s.p1[i][a], x2 = s.p2[i][a];
        double lo = std::min(x0, std::min(x1, x2)), hi = std::max(x0, std::max(x1, x2));
        double pad = 1e-5 * std::max(std::abs(lo), std::abs(hi)) + 1e-30;
        bl[i * 3 + a] = lo - pad; bh[i * 3 + a] = hi + pad;
        ce[i * 3 + a] = (x0 + x1 + x2) / 3.0;
      }
    nodes.clear();
    nodes.reserve(2 * n / leaf_size + 8);
    if (n == 0) return;
    struct Item { uint32_t s, e; int32_t node; };
    std::vector<Item> st;
    nodes.push_back(Node());
    st.push_back({0, uint32_t(n), 0});
    while (!st.empty()) {
      Item it = st.back(); st.pop_back();
      Node nd;
      double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
      for (int a = 0; a < 3; a++) { nd.lo[a] = 1e300; nd.hi[a] = -1e300; }
      for (uint32_t i = it.s; i < it.e; i++) {
        uint32_t p = order[i];
        for (int a = 0; a < 3; a++) {
          nd.lo[a] = std::min(nd.lo[a], bl[p * 3 + a]); nd.hi[a] = std::max(nd.hi[a], bh[p * 3 + a]);
          clo[a] = std::min(clo[a], ce[p * 3 + a]); chi[a] = std::max(chi[a], ce[p * 3 + a]);
        }
      }
      uint32_t cnt = it.e - it.s;
      int dim = 0;
      for (int a = 1; a < 3; a++) if (chi[a] - clo[a] > chi[dim] - clo[dim]) dim = a;
      if (cnt <= uint32_t(leaf_size) || chi[dim] - clo[dim] <= 0) {
        nd.start = it.s; nd.count = cnt;
        nodes[it.node] = nd;
        continue;
      }
      const int B = 12;
      double inv = 1.0 / (chi[dim] - clo[dim]);
      auto bucket = [&](uint32_t p) {
        int b = int((B - 1) * (ce[p * 3 + dim] - clo[dim]) * inv);
        return std::min(std::max(b, 0), B - 1);
      };
      double blo[B][3], bhi[B][3];
      uint32_t bc[B] = {0};
      for (int b = 0; b < B; b++) for (int a = 0; a < 3; a++) { blo[b][a] = 1e300; bhi[b][a] = -1e300; }
      for (uint32_t i = it.s; i < it.e; i++) {
        uint32_t p = order[i];
        int b = bucket(p);
        bc[b]++;
        for (int a = 0; a < 3; a++) { blo[b][a] = std::min(blo[b][a], bl[p * 3 + a]); bhi[b][a] = std::max(bhi[b][a], bh[p * 3 + a]); }
      }
      auto area = [](const double lo[3], const double hi[3]) {
        if (lo[0] > hi[0]) return 0.0;
        double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
      };
      double best = 1e300;
      int bs = 0;
      for (int sidx = 0; sidx < B - 1; sidx++) {
        double llo[3] = {1e300, 1e300, 1e300}, lhi[3] = {-1e300, -1e300, -1e300};
        double rlo[3] = {1e300, 1e300, 1e300}, rhi[3] = {-1e300, -1e300, -1e300};
        uint32_t nl = 0, nr = 0;
        for (int b = 0; b <= sidx; b++) { nl += bc[b]; for (int a = 0; a < 3; a++) { llo[a] = std::min(llo[a], blo[b][a]); lhi[a] = std::max(lhi[a], bhi[b][a]); } }
        for (int b = sidx + 1; b < B; b++) { nr += bc[b]; for (int a = 0; a < 3; a++) { rlo[a] = std::min(rlo[a], blo[b][a]); rhi[a] = std::max(rhi[a], bhi[b][a]); } }
        double c = 0.125 + (area(llo, lhi) * nl + area(rlo, rhi) * nr) / std::max(area(nd.lo, nd.hi), 1e-300);
        if (c < best) { best = c; bs = sidx; }
      }
      uint32_t mid = uint32_t(std::partition(order.begin() + it.s, order.begin() + it.e,
                                             [&](uint32_t p) { return bucket(p) <= bs; }) - order.begin());
      if (mid == it.s || mid == it.e) {
        mid = it.s + cnt / 2;
        std::nth_element(order.begin() + it.s, order.begin() + mid, order.begin() + it.e,
                         [&](uint32_t x, uint32_t y) { return ce[x * 3 + dim] < ce[y * 3 + dim]; });
      }
      int32_t l = int32_t(nodes.size()); nodes.push_back(Node());
      int32_t r = int32_t(nodes.size()); nodes.push_back(Node());
      nd.left = l; nd.right = r;
      nodes[it.node] = nd;
      st.push_back({mid, it.e, r});
      st.push_back({it.s, mid, l});
    }
  }

  static bool box_hit(const Node &n, const double o[3], const double inv[3], double tmin, double tmax,
                      double &tnear) {
    double t0 = tmin, t1 = tmax;
    for (int a = 0; a < 3; a++) {
      double lo = (n.lo[a] - o[a]) * inv[a], hi = (n.hi[a] - o[a]) * inv[a];
      if (std::isnan(lo) || std::isnan(hi)) continue;  // origin on the slab plane with d == 0
      if (lo > hi) std::swap(lo, hi);
      t0 = std::max(t0, lo); t1 = std::min(t1, hi);
      if (t0 > t1) return false;
    }
    tnear = t0;
    return true;
  }

  bool intersect(const RayF &ray, Hit &h, bool any) const {
    if (nodes.empty()) return false;
    double o[3] = {ray.o.x, ray.o.y, ray.o.z}, inv[3];
    for (int a = 0; a < 3; a++) {
      double d = ray.d[a];
      inv[a] = d == 0.0 ? std::copysign(1e300, 1.0) : 1.0 / d;
    }
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    bool found = false;
    RayF r = ray;
    while (sp > 0) {
      const Node &n = nodes[stack[--sp]];
      double tn;
      double tmx = found ? double(h.t) * (1 + 1e-6) + 1e-30 : double(r.tmax) * (1 + 1e-6);
      if (!box_hit(n, o, inv, double(r.tmin) * (1 - 1e-6) - 1e-30, tmx, tn)) continue;
      if (n.left < 0) {
        for (uint32_t i = n.start; i < n.start + n.count; i++) {
          uint32_t p = order[i];
          float u, v, t;
          RayF rr = r;
          if (found) rr.tmax = h.t;
          if (mt_intersect(soup->p0[p], soup->p1[p], soup->p2[p], rr, u, v, t)) {
            if (any) { h.t = t; h.u = u; h.v = v; h.prim = p; return true; }
            if (!found || t < h.t || (t == h.t && int64_t(p) > h.prim)) {
              h.t = t; h.u = u; h.v = v; h.prim = p;
              found = true;
            }
          }
        }
      } else {
        if (sp + 2 > 128) throw std::runtime_error("oracle BVH stack overflow");
        stack[sp++] = n.right;
        stack[sp++] = n.left;
      }
    }
    return found;
  }
};

// ---------------------------------------------------------------- discrete pdfs
// src/core/DiscretePDF.cpp:5-91
template <typename R>
struct Pdf1D {
  std::vector<R> cdf, func;
  R integral = 0;
  void build(const R *f, int n) {
    func.assign(f, f + n);
    cdf.assign(n + 1, R(0));
    for (int i = 1; i < n + 1; i++) cdf[i] = cdf[i - 1] + func[i - 1] / R(n);
    integral = cdf[n];
    if (integral == R(0)) for (int i = 1; i < n + 1; i++) cdf[i] = R(i) / R(n);
    else for (int i = 1; i < n + 1; i++) cdf[i] /= integral;
  }
  int count() const { return int(func.size()); }
  int find(R s) const {
    auto it = std::lower_bound(cdf.begin(), cdf.end(), s);
    return int(std::min(size_t(std::max(std::ptrdiff_t(0), it - cdf.begin() - 1)), cdf.size() - 2));
  }
  R sample_continuous(R s, R *pdf, int *idx) const {  // :32-58
    int i = find(s);
    if (idx) *idx = i;
    R du = s - cdf[i];
    if (cdf[i + 1] - cdf[i] > R(0)) du /= (cdf[i + 1] - cdf[i]);
    if (pdf) *pdf = func[i] / integral;
    return (R(i) + du) / R(count());
  }
  int sample_discrete(R s) const { return find(s); }  // :60-78
};
// src/core/DiscretePDF.cpp:93-132
template <typename R>
struct Pdf2D {
  std::vector<Pdf1D<R>> rows;
  Pdf1D<R> marg;
  void build(const R *d, int w, int h) {
    rows.resize(h);
    std::vector<R> m(h);
    for (int i = 0; i < h; i++) { rows[i].build(d + size_t(i) * w, w); m[i] = rows[i].integral; }
    marg.build(m.data(), h);
  }
  void sample(R sx, R sy, R &u, R &v, R &pdf, int &ix, int &iy) const {
    R px, py;
    v = marg.sample_continuous(sy, &py, &iy);
    u = rows[iy].sample_continuous(sx, &px, &ix);
    pdf = px * py;
  }
  R pdf(R u, R v) const {
    int w = rows[0].count(), h = marg.count();
    int x = std::min(std::max(int(u * R(w)), 0), w - 1);
    int y = std::min(std::max(int(v * R(h)), 0), h - 1);
    return rows[y].func[x] / marg.integral;
  }
};

// ---------------------------------------------------------------- 4x4 helpers
struct M4 {
  double m[4][4];
  static M4 identity() { M4 r; for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) r.m[i][j] = i == j; return r; }
  M4 operator*(const M4 &b) const {
    M4 r;
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) {
      double s = 0;
      for (int k = 0; k < 4; k++) s += m[i][k] * b.m[k][j];
      r.m[i][j] = s;
    }
    return r;
  }
  M4 inverse() const {  // Gauss-Jordan with partial pivoting
    double a[4][8];
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) { a[i][j] = m[i][j]; a[i][j + 4] = i == j; }
    for (int c = 0; c < 4; c++) {
      int p = c;
      for (int r = c + 1; r < 4; r++) if (std::abs(a[r][c]) > std::abs(a[p][c])) p = r;
      for (int j = 0; j < 8; j++) std::swap(a[c][j], a[p][j]);
      double d = a[c][c];
      for (int j = 0; j < 8; j++) a[c][j] /= d;
      for (int r = 0; r < 4; r++) if (r != c) { double f = a[r][c]; for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j]; }
    }
    M4 r;
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) r.m[i][j] = a[i][j + 4];
    return r;
  }
  // src/core/Transform.cpp:46-50: Matrix4f * Vector4f accumulates column by
  // column, ((c0 x + c1 y) + c2 z) + c3 (pinned against oracle/_ref/eigen_xform)
  template <typename R> V3<R> point(V3<R> p) const {
    R v[4];
    for (int i = 0; i < 4; i++) v[i] = R(m[i][0]) * p.x + R(m[i][1]) * p.y + R(m[i][2]) * p.z + R(m[i][3]);
    return {v[0] / v[3], v[1] / v[3], v[2] / v[3]};
  }
  // Transform.cpp:36-39: Eigen sums each row of the 3x3 block as
  // a0 + (a1 + a2) (pinned against oracle/_ref/eigen_xform)
  template <typename R> V3<R> vec(V3<R> p) const {
    R v[3];
    for (int i = 0; i < 3; i++) v[i] = R(m[i][0]) * p.x + (R(m[i][1]) * p.y + R(m[i][2]) * p.z);
    return {v[0], v[1], v[2]};
  }
};

// ---------------------------------------------------------------- scene input
// Flat scene description shared with the HIP path (see include/hikari_amd.h).
struct MeshIn {
  const float *v = nullptr; uint32_t nv = 0;
  const float *n = nullptr;   // nullable, nv*3
  const float *uv = nullptr;  // nullable, nv*2
  const uint32_t *f = nullptr; uint32_t nt = 0;
  int material = 0;
  int is_emitter = 0;
  float radiance[3] = {0, 0, 0};
};
// The tabulated reconstruction filter of ImageBlock (src/core/Block.cpp:9-33:
// 32 entries on [0, r) plus a closing zero) for the box, Gaussian, tent and
// Mitchell-Netravali filters of src/filter/*.cpp.
struct FilmFilter {
  float table[33];
  float radius = 0.5f, lookup = 64;
  void setup(int ft, float r, float stddev, float b, float c) {
    radius = ft == 0 ? 0.5f : (ft == 2 ? 1.0f : r);
    for (int i = 0; i < 32; i++) {
      float x = (radius * i) / 32;
      float w;
      if (ft == 0) w = 1.0f;
      else if (ft == 2) w = std::max(0.0f, 1.0f - std::abs(x));
      else if (ft == 3) {
        float X = std::abs(2.0f * x / radius), X2 = X * X, X3 = X2 * X, B = b, C = c;
        if (X < 1.0f) w = 1.0f / 6.0f * ((12.0f - 9.0f * B - 6.0f * C) * X3 + (-18.0f + 12.0f * B + 6.0f * C) * X2 + (6.0f - 2.0f * B));
        else if (X < 2.0f) w = 1.0f / 6.0f * ((-B - 6 * C) * X3 + (6.0f * B + 30.0f * C) * X2 + (-12.0f * B - 48.0f * C) * X + (8.0f * B + 24.0f * C));
        else w = 0.0f;
      } else {
        float al = -1.0f / (2.0f * stddev * stddev);
        w = std::max(0.0f, std::exp(al * x * x) - std::exp(al * radius * radius));
      }
      table[i] = w;
    }
    table[32] = 0.0f;
    lookup = 32 / radius;
  }
};

struct SceneIn {
  std::vector<MeshIn> meshes;
  const float *materials = nullptr; int n_materials = 0;
  const float *env = nullptr; int env_w = 0, env_h = 0; float env_scale = 1; double env_to_world[16];
  double cam_to_world[16];
  float fov = 30, near_clip = 1e-4f, far_clip = 1e4f;
  int width = 1, height = 1;
  int filter = 1;  // 0 box, 1 gaussian, 2 tent, 3 mitchell
  float filter_radius = 2.0f, filter_stddev = 0.5f, filter_b = 1.0f / 3, filter_c = 1.0f / 3;
  int max_depth = 1000;
  float background[3] = {0, 0, 0};
  int force_background = 0;
  const char *rft_dir = nullptr;
  int integrator = 0;  // 0 pathMIS, 1 pathEMS, 2 pathMATS, 3 whitted
  int rr_start_depth = 0;
};

template <typename R>
struct Scene {
  struct MeshD {
    uint32_t first_prim = 0, nt = 0;
    std::vector<float> v, n, uv;
    std::vector<uint32_t> f;
    int material = 0, emitter = -1;
    V3<R> radiance;
    Pdf1D<R> area_pdf;
    R area = 0, inv_area = 0;
    V3<float> vtx(uint32_t i) const { return {v[3 * i], v[3 * i + 1], v[3 * i + 2]}; }
  };
  struct Emitter { int kind; int mesh; };  // kind 0 area, 1 env
  std::vector<MeshD> meshes;
  std::vector<uint32_t> prim_mesh;
  TriSoup soup;
  Bvh bvh;
  std::vector<Material<R>> mats;
  std::vector<Rft<R> *> rft_pool;
  RftFile beck, ggx;
  std::vector<Emitter> emitters;
  // environment
  bool has_env = false;
  int env_w = 0, env_h = 0;
  std::vector<float> env_tex;
  R env_scale = 1;
  M4 env_to_world, env_to_local;
  Pdf2D<R> env_pdf;
  // camera
  M4 sample_to_camera, cam_to_world;
  int width = 1, height = 1;
  R inv_w = 1, inv_h = 1, near_clip = 0, far_clip = 0;
  FilmFilter film;
  int max_depth = 1000;
  int integrator = 0;  // 0 pathMIS, 1 pathEMS, 2 pathMATS, 3 whitted
  // Russian roulette from this depth on; below it a vertex continues with
  // weight 1 (the draw is made all the same).  Not a reference parameter:
  // 0 is the reference's schedule (roulette at every vertex).
  int rr_start = 0;
  V3<R> background;
  bool force_background = false;
  double bbox_lo[3], bbox_hi[3];
  R bbox_radius = 0;

  ~Scene() { for (auto *p : rft_pool) delete p; }

  void load(const SceneIn &in) {
    // materials (src/bsdf/*: constructors + Activate)
    bool need_rft = false;
    for (int i = 0; i < in.n_materials; i++) {
      int t = int(in.materials[i * 32]);
      if (t == kRoughPlastic || t == kRoughCoating) need_rft = true;
    }
    if (need_rft) {
      std::string dir = in.rft_dir ? in.rft_dir : ".";
      beck.load(dir + "/BeckmannRFTData.bin");
      ggx.load(dir + "/GGXRFTData.bin");
    }
    mats.resize(in.n_materials);
    for (int i = 0; i < in.n_materials; i++) mats[i].from_record(in.materials + i * 32);
    link_materials(mats, in.materials);
    for (int i = 0; i < in.n_materials; i++) mats[i].activate(&beck, &ggx, rft_pool);
    // meshes (src/core/Mesh.cpp:233-261 Activate, Scene.cpp:128-165)
    for (int a = 0; a < 3; a++) { bbox_lo[a] = 1e300; bbox_hi[a] = -1e300; }
    float flo[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity()};
    float fhi[3] = {-flo[0], -flo[0], -flo[0]};
    uint32_t prim = 0;
    for (size_t mi = 0; mi < in.meshes.size(); mi++) {
      const MeshIn &m = in.meshes[mi];
      MeshD d;
      d.first_prim = prim; d.nt = m.nt;
      d.v.assign(m.v, m.v + 3 * m.nv);
      if (m.n) d.n.assign(m.n, m.n + 3 * m.nv);
      if (m.uv) d.uv.assign(m.uv, m.uv + 2 * m.nv);
      d.f.assign(m.f, m.f + 3 * m.nt);
      d.material = m.material;
      d.radiance = V3<R>(m.radiance[0], m.radiance[1], m.radiance[2]);
      for (uint32_t i = 0; i < m.nv; i++)
        for (int a = 0; a < 3; a++) { flo[a] = std::min(flo[a], m.v[3 * i + a]); fhi[a] = std::max(fhi[a], m.v[3 * i + a]); }
      std::vector<R> areas(m.nt);
      R total = 0;
      for (uint32_t i = 0; i < m.nt; i++) {  // Mesh.cpp:310-315 SurfaceArea
        V3<float> p0 = d.vtx(d.f[3 * i]), p1 = d.vtx(d.f[3 * i + 1]), p2 = d.vtx(d.f[3 * i + 2]);
        V3<R> e1 = V3<R>(p1.x, p1.y, p1.z) - V3<R>(p0.x, p0.y, p0.z);
        V3<R> e2 = V3<R>(p2.x, p2.y, p2.z) - V3<R>(p0.x, p0.y, p0.z);
        areas[i] = R(0.5f) * norm(cross(e1, e2));
        total += areas[i];
        soup.p0.push_back(p0); soup.p1.push_back(p1); soup.p2.push_back(p2);
        prim_mesh.push_back(uint32_t(mi));
      }
      d.area = total;
      d.inv_area = R(1) / total;
      if (m.nt) d.area_pdf.build(areas.data(), int(m.nt));
      if (m.is_emitter) {
        d.emitter = int(emitters.size());
        emitters.push_back({0, int(mi)});
      }
      prim += m.nt;
      meshes.push_back(std::move(d));
    }
    for (int a = 0; a < 3; a++) { bbox_lo[a] = flo[a]; bbox_hi[a] = fhi[a]; }
    {  // BoundingBox.hpp:452-455 GetRadius in float
      V3<float> ext(fhi[0] - flo[0], fhi[1] - flo[1], fhi[2] - flo[2]);
      bbox_radius = R(norm(ext) * 0.5f);
    }
    bvh.build(soup);
    // environment (src/emitter/EnvironmentLight.cpp:8-33); emitted after area lights
    if (in.env) {
      has_env = true;
      env_w = in.env_w; env_h = in.env_h;
      env_tex.assign(in.env, in.env + size_t(env_w) * env_h * 3);
      env_scale = in.env_scale;
      for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) env_to_world.m[i][j] = in.env_to_world[i * 4 + j];
      env_to_local = env_to_world.inverse();
      std::vector<R> lum(size_t(env_w) * env_h);
      for (int y = 0; y < env_h; y++) {
        R theta = R(float(y + 0.5f)) / R(env_h) * R(float(kPi));
        R st = std::sin(theta);
        for (int x = 0; x < env_w; x++) {
          const float *c = &env_tex[(size_t(y) * env_w + x) * 3];
          lum[size_t(y) * env_w + x] = luminance(V3<R>(c[0], c[1], c[2])) * st;
        }
      }
      env_pdf.build(lum.data(), env_w, env_h);
      emitters.push_back({1, -1});
    }
    // camera (src/camera/PerspectiveCamera.cpp:8-75)
    width = in.width; height = in.height;
    inv_w = R(1) / R(float(width)); inv_h = R(1) / R(float(height));
    near_clip = in.near_clip; far_clip = in.far_clip;
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) cam_to_world.m[i][j] = in.cam_to_world[i * 4 + j];
    {
      double aspect = double(width) / double(height);
      double recip = 1.0 / (double(in.far_clip) - in.near_clip);
      double cot = 1.0 / std::tan(double(in.fov) / 2.0 * kPi / 180.0);
      M4 P = M4::identity();
      P.m[0][0] = cot; P.m[1][1] = cot; P.m[2][2] = in.far_clip * recip; P.m[2][3] = -double(in.near_clip) * in.far_clip * recip;
      P.m[3][2] = 1; P.m[3][3] = 0;
      M4 S = M4::identity(); S.m[0][0] = -0.5; S.m[1][1] = -0.5 * aspect;
      M4 T = M4::identity(); T.m[0][3] = -1; T.m[1][3] = -1.0 / aspect;
      sample_to_camera = (S * T * P).inverse();
    }
    film.setup(in.filter, in.filter_radius, in.filter_stddev, in.filter_b, in.filter_c);
    // the depth property is read as uint32 (PathMISIntegrator.cpp:13,
    // WhittedIntegrator.cpp:13): a negative value means no limit; paths are
    // capped at 262143 vertices as on the device
    max_depth = in.max_depth < 0 ? 262143 : std::min(in.max_depth, 262143);
    integrator = in.integrator;
    rr_start = std::max(0, in.rr_start_depth);
    background = V3<R>(in.background[0], in.background[1], in.background[2]);
    force_background = in.force_background != 0;
  }

  // ---- camera ray (PerspectiveCamera.cpp:77-102)
  void camera_ray(R px, R py, V3<R> &o, V3<R> &d, R &tmin, R &tmax) const {
    V3<R> nearp = sample_to_camera.point(V3<R>(px * inv_w, py * inv_h, R(0)));
    V3<R> dir = normalized(nearp);
    R inv_z = R(1) / dir.z;
    o = cam_to_world.point(V3<R>(R(0)));
    d = cam_to_world.vec(dir);
    tmin = near_clip * inv_z;
    tmax = far_clip * inv_z;
  }

  // ---- intersection + PostIntersect (src/core/Mesh.cpp:67-172)
  struct Isect {
    V3<R> p;
    Frame<R> sh, geo;
    int mesh = -1;
    uint32_t prim = 0;
  };
  bool intersect(const RayF &r, Isect &it, Hit *raw = nullptr) const {
    Hit h;
    if (!bvh.intersect(r, h, false)) return false;
    if (raw) *raw = h;
    post_intersect(h, it);
    return true;
  }
  bool occluded(const RayF &r) const {
    Hit h;
    return bvh.intersect(r, h, true);
  }
  void post_intersect(const Hit &h, Isect &it) const {
    uint32_t prim = uint32_t(h.prim);
    const MeshD &m = meshes[prim_mesh[prim]];
    uint32_t f = prim - m.first_prim;
    it.mesh = int(prim_mesh[prim]);
    it.prim = prim;
    R u = h.u, v = h.v;
    R b0 = R(1) - (u + v), b1 = u, b2 = v;  // Mesh.cpp:82-83: (1 - UV.sum(), UV)
    uint32_t i0 = m.f[3 * f], i1 = m.f[3 * f + 1], i2 = m.f[3 * f + 2];
    auto P = [&](uint32_t i) { return V3<R>(m.v[3 * i], m.v[3 * i + 1], m.v[3 * i + 2]); };
    auto N = [&](uint32_t i) { return V3<R>(m.n[3 * i], m.n[3 * i + 1], m.n[3 * i + 2]); };
    V3<R> p0 = P(i0), p1 = P(i1), p2 = P(i2);
    it.p = (b0 * p0 + b1 * p1) + b2 * p2;
    it.geo = Frame<R>(normalized(cross(p1 - p0, p2 - p0)));
    bool has_n = !m.n.empty(), has_uv = !m.uv.empty();
    if (has_n && has_uv) {
      R uv0x = m.uv[2 * i0], uv0y = m.uv[2 * i0 + 1], uv1x = m.uv[2 * i1], uv1y = m.uv[2 * i1 + 1];
      R uv2x = m.uv[2 * i2], uv2y = m.uv[2 * i2 + 1];
      V3<R> dp1 = p1 - p0, dp2 = p2 - p0;
      R du1x = uv1x - uv0x, du1y = uv1y - uv0y, du2x = uv2x - uv0x, du2y = uv2y - uv0y;
      R len = norm(cross(dp1, dp2));
      if (len == R(0)) {
        it.sh = it.geo;
      } else {
        R det = du1x * du2y - du1y * du2x;
        V3<R> ns = (b0 * N(i0) + b1 * N(i1)) + b2 * N(i2);
        if (det == R(0)) {
          it.sh = Frame<R>(normalized(ns));
        } else {
          R idet = R(1) / det;
          V3<R> dpdu = (du2y * dp1 - du1y * dp2) * idet;
          R ln = norm(ns);
          R iln = R(1) / ln;
          ns = ns * iln;
          it.sh = Frame<R>(ns, dpdu);
        }
      }
    } else if (has_n) {
      it.sh = Frame<R>(normalized((b0 * N(i0) + b1 * N(i1)) + b2 * N(i2)));
    } else {
      it.sh = it.geo;
    }
  }

  // ---- emitters
  struct ERec {
    V3<R> ref, p, n, wi;
    R pdf = 0, dist = 0;
  };
  // src/emitter/AreaLight.cpp:14-74 + Mesh::SamplePosition (Mesh.cpp:273-303)
  V3<R> area_sample(const MeshD &m, ERec &rec, R sx, R sy, R s1) const {
    int tri = m.area_pdf.sample_discrete(s1);
    R sq = std::sqrt(R(1) - sx);
    R al = R(1) - sq, be = sy * sq, ga = R(1) - al - be;
    uint32_t i0 = m.f[3 * tri], i1 = m.f[3 * tri + 1], i2 = m.f[3 * tri + 2];
    auto P = [&](uint32_t i) { return V3<R>(m.v[3 * i], m.v[3 * i + 1], m.v[3 * i + 2]); };
    V3<R> p0 = P(i0), p1 = P(i1), p2 = P(i2);
    rec.p = (ga * p0 + al * p1) + be * p2;
    if (!m.n.empty()) {
      auto N = [&](uint32_t i) { return V3<R>(m.n[3 * i], m.n[3 * i + 1], m.n[3 * i + 2]); };
      rec.n = (ga * N(i0) + al * N(i1)) + be * N(i2);
    } else {
      rec.n = normalized(cross(p1 - p0, p2 - p0));
    }
    V3<R> wi = rec.p - rec.ref;
    rec.dist = norm(wi);
    rec.wi = normalized(wi);
    rec.pdf = area_pdf_eval(m, rec);
    if (rec.pdf == R(0) || std::isinf(rec.pdf)) return V3<R>(R(0));
    return area_eval(m, rec) / rec.pdf;
  }
  R area_pdf_eval(const MeshD &m, const ERec &rec) const {
    R gd = std::abs(dot(-rec.wi, rec.n));
    if (gd == R(0)) return R(0);
    R gn = rec.dist * rec.dist;
    return m.inv_area * (gn / gd);
  }
  V3<R> area_eval(const MeshD &m, const ERec &rec) const {
    if (dot(rec.n, rec.wi) < R(0)) return m.radiance;
    return V3<R>(R(0));
  }
  // src/emitter/EnvironmentLight.cpp:40-111
  V3<R> env_sample(ERec &rec, R sx, R sy) const {
    R map_pdf, u, v;
    int ix, iy;
    env_pdf.sample(sx, sy, u, v, map_pdf, ix, iy);
    R theta = v * R(float(kPi));
    R st = std::sin(theta);
    if (map_pdf == R(0)) return V3<R>(R(0));
    if (st != R(0)) rec.pdf = map_pdf / (R(2) * R(float(kPi * kPi)) * st);
    else return V3<R>(R(0));
    R phi = u * R(float(2.0 * kPi));
    rec.wi = env_to_world.vec(spherical_direction(theta, phi));
    rec.n = -rec.wi;
    rec.p = rec.ref + rec.wi * rec.dist * R(2);
    const float *c = &env_tex[(size_t(iy) * env_w + ix) * 3];
    V3<R> rad(c[0], c[1], c[2]);
    if (!color_valid(rad)) return V3<R>(R(0));
    return rad * env_scale / rec.pdf;
  }
  V3<R> env_eval(V3<R> wi) const {
    V3<R> l = env_to_local.vec(wi);
    R theta = std::acos(l.z), phi = std::atan2(l.y, l.x);
    if (phi < R(0)) phi += R(2) * R(float(kPi));
    if (std::isnan(theta) || std::isnan(phi)) return V3<R>(R(0));
    R X = clampr(phi / R(float(2.0 * kPi)) * R(env_w), R(0), R(float(env_w - 1.0f)));
    R Y = clampr(theta / R(float(kPi)) * R(env_h), R(0), R(float(env_h - 1.0f)));
    const float *c = &env_tex[(size_t(int(Y)) * env_w + int(X)) * 3];
    V3<R> rad(c[0], c[1], c[2]);
    if (!color_valid(rad)) return V3<R>(R(0));
    return rad * env_scale;
  }
  R env_pdf_eval(V3<R> wi) const {
    V3<R> l = env_to_local.vec(wi);
    R theta = std::acos(l.z), phi = std::atan2(l.y, l.x);
    if (phi < R(0)) phi += R(2) * R(float(kPi));
    R st = std::sin(theta);
    if (st == R(0)) return R(0);
    return env_pdf.pdf(phi / R(float(2.0 * kPi)), theta / R(float(kPi))) / (R(2) * R(float(kPi * kPi)) * st);
  }

  static RayF to_rayf(V3<R> o, V3<R> d, R tmin, R tmax) {
    RayF r;
    r.o = V3<float>(float(o.x), float(o.y), float(o.z));
    r.d = V3<float>(float(d.x), float(d.y), float(d.z));
    r.tmin = float(tmin); r.tmax = float(tmax);
    return r;
  }
};

// ---------------------------------------------------------------- estimator
struct PathStats {
  uint64_t ext = 0, shadow = 0, paths = 0;
};
struct PcgU1 : U1Source {
  Pcg32 *r;
  explicit PcgU1(Pcg32 *p) : r(p) {}
  float next() override { return r->next_float(); }
};

// src/integrator/PathMISIntegrator.cpp:16-157.  Random draw order (the
// reference's is compiler dependent, src/integrator/PathMISIntegrator.cpp:92):
// per emitter (x, y, s1); BSDF (x, y [, extra draws inside Sample]); RR.
// Russian roulette of the path estimators (PathMISIntegrator.cpp:143-151,
// PathEMSIntegrator.cpp, PathMATSIntegrator.cpp): continue with probability
// 0.95 at weight 1 / 0.95.  Below sc.rr_start (not a reference parameter;
// 0 by default) the path continues at weight 1; the draw is made either way.
template <typename R>
inline bool roulette(const Scene<R> &sc, Pcg32 &rng, int depth, V3<R> &beta) {
  bool go = rng.next_float() < 0.95f;
  if (depth < sc.rr_start) {
    if (slip() == kSlipRrStartWeight) beta = beta * R(1.0f / 0.95f);
    return true;
  }
  if (go) beta = beta * R(1.0f / 0.95f);
  return go;
}

template <typename R>
V3<R> path_mis_li(const Scene<R> &sc, Pcg32 &rng, V3<R> ro, V3<R> rd, R rtmin, R rtmax, PathStats *st) {
  using Isect = typename Scene<R>::Isect;
  using ERec = typename Scene<R>::ERec;
  PcgU1 u1(&rng);
  Isect is_next, is;
  bool found_next = false;
  V3<R> li(R(0)), beta(R(1));
  int depth = 0;
  R w_ems = 1, w_mats = 1;
  V3<R> to = ro, td = rd;
  R ttmin = rtmin, ttmax = rtmax;
  while (depth < sc.max_depth) {
    if (depth == 0) {
      if (st) st->ext++;
      if (!sc.intersect(Scene<R>::to_rayf(to, td, ttmin, ttmax), is)) {
        if (sc.has_env && !sc.force_background) {
          li = li + beta * sc.env_eval(td) / R(1);
          break;
        }
        return sc.background;
      }
    } else {
      if (found_next) is = is_next;
      else {
        if (sc.has_env) li = li + beta * sc.env_eval(td) / R(1);
        break;
      }
    }
    const auto &mesh = sc.meshes[is.mesh];
    const Material<R> &bsdf = sc.mats[mesh.material];
    if (mesh.emitter >= 0) {
      ERec er;
      er.ref = to; er.p = is.p; er.n = is.sh.n;
      er.wi = er.p - er.ref; er.dist = norm(er.wi); er.wi = er.wi / er.dist;
      V3<R> le = sc.area_eval(mesh, er);
      li = li + beta * w_mats * le;
    }
    for (const auto &em : sc.emitters) {
      ERec er;
      er.ref = is.p;
      if (em.kind == 1) er.dist = sc.bbox_radius;
      R sx = rng.next_float(), sy = rng.next_float(), s1 = rng.next_float();
      V3<R> ld = em.kind == 0 ? sc.area_sample(sc.meshes[em.mesh], er, sx, sy, s1) : sc.env_sample(er, sx, sy);
      R pl = er.pdf;
      if (!is_zero_fuzzy(ld)) {
        if (st) st->shadow++;
        RayF sr = Scene<R>::to_rayf(is.p, er.p - is.p, R(1e-4f), R(1.0f - 1e-4f));
        if (!sc.occluded(sr)) {
          BRec<R> br;
          br.wi = is.sh.to_local(R(-1) * td); br.wo = is.sh.to_local(er.wi); br.measure = kSolidAngle;
          R pb = bsdf.pdf(br);
          if (pl + pb != R(0)) w_ems = pl / (pl + pb);
          li = li + beta * bsdf.eval(br) * std::abs(cos_theta(br.wo)) * ld * w_ems;
        }
      }
    }
    BRec<R> br;
    br.wi = is.sh.to_local(R(-1) * td);
    R sx = rng.next_float(), sy = rng.next_float();
    V3<R> f = bsdf.sample(br, sx, sy, &u1);
    to = is.p; td = is.sh.to_world(br.wo); ttmin = R(1e-4f); ttmax = std::numeric_limits<R>::infinity();
    beta = beta * f;
    if (is_zero_fuzzy(beta)) break;
    if (st) st->ext++;
    found_next = sc.intersect(Scene<R>::to_rayf(to, td, ttmin, ttmax), is_next);
    if (found_next && sc.meshes[is_next.mesh].emitter >= 0) {
      ERec er;
      er.ref = is.p; er.p = is_next.p; er.n = is_next.sh.n;
      er.wi = er.p - er.ref; er.dist = norm(er.wi); er.wi = er.wi / er.dist;
      R plm = sc.area_pdf_eval(sc.meshes[is_next.mesh], er);
      R pbm = bsdf.pdf(br);
      if (pbm + plm != R(0)) w_mats = pbm / (pbm + plm);
    }
    if (br.measure == kDiscrete) { w_ems = 0; w_mats = 1; }
    if (!roulette(sc, rng, depth, beta)) break;
    depth++;
  }
  return li;
}

// src/integrator/PathMATSIntegrator.cpp:16-87: BSDF sampling only; every
// emitter hit adds its radiance at full weight.
template <typename R>
V3<R> path_mats_li(const Scene<R> &sc, Pcg32 &rng, V3<R> ro, V3<R> rd, R rtmin, R rtmax, PathStats *st) {
  using Isect = typename Scene<R>::Isect;
  using ERec = typename Scene<R>::ERec;
  PcgU1 u1(&rng);
  V3<R> li(R(0)), beta(R(1));
  int depth = 0;
  V3<R> to = ro, td = rd;
  R ttmin = rtmin, ttmax = rtmax;
  while (depth < sc.max_depth) {
    Isect is;
    if (st) st->ext++;
    if (!sc.intersect(Scene<R>::to_rayf(to, td, ttmin, ttmax), is)) {
      if (sc.has_env && !sc.force_background) li = li + beta * sc.env_eval(td) / R(1);
      else if (depth == 0) return sc.background;
      break;
    }
    const auto &mesh = sc.meshes[is.mesh];
    const Material<R> &bsdf = sc.mats[mesh.material];
    if (mesh.emitter >= 0) {
      ERec er;
      er.ref = to; er.p = is.p; er.n = is.sh.n; er.wi = td;
      li = li + sc.area_eval(mesh, er) * beta;
    }
    BRec<R> br;
    br.wi = is.sh.to_local(R(-1) * td);
    R sx = rng.next_float(), sy = rng.next_float();
    beta = beta * bsdf.sample(br, sx, sy, &u1);
    if (is_zero_fuzzy(beta)) break;
    if (!roulette(sc, rng, depth, beta)) break;
    to = is.p; td = is.sh.to_world(br.wo); ttmin = R(1e-4f); ttmax = std::numeric_limits<R>::infinity();
    depth++;
  }
  return li;
}

// src/integrator/PathEMSIntegrator.cpp:16-119: emitter sampling at diffuse
// surfaces (skipping the surface's own emitter), emission counted only on
// camera rays, after specular bounces and on consecutive hits of one emitter.
template <typename R>
V3<R> path_ems_li(const Scene<R> &sc, Pcg32 &rng, V3<R> ro, V3<R> rd, R rtmin, R rtmax, PathStats *st) {
  using Isect = typename Scene<R>::Isect;
  using ERec = typename Scene<R>::ERec;
  PcgU1 u1(&rng);
  V3<R> li(R(0)), beta(R(1));
  int depth = 0;
  bool last_specular = false;
  int last_emitter = -1;
  V3<R> to = ro, td = rd;
  R ttmin = rtmin, ttmax = rtmax;
  while (depth < sc.max_depth) {
    Isect is;
    if (st) st->ext++;
    if (!sc.intersect(Scene<R>::to_rayf(to, td, ttmin, ttmax), is)) {
      if (sc.has_env && !sc.force_background) li = li + beta * sc.env_eval(td) / R(1);
      else if (depth == 0) return sc.background;
      break;
    }
    const auto &mesh = sc.meshes[is.mesh];
    const Material<R> &bsdf = sc.mats[mesh.material];
    if (mesh.emitter >= 0 && (depth == 0 || last_specular || last_emitter == mesh.emitter)) {
      ERec er;
      er.ref = to; er.p = is.p; er.n = is.sh.n; er.wi = td;
      li = li + beta * sc.area_eval(mesh, er);
    }
    if (bsdf.is_diffuse()) {
      last_specular = false;
      for (int e = 0; e < int(sc.emitters.size()); e++) {
        const auto &em = sc.emitters[e];
        if (em.kind == 0 && em.mesh == int(is.mesh)) continue;  // the surface's own emitter
        ERec er;
        er.ref = is.p;
        if (em.kind == 1) er.dist = sc.bbox_radius;
        R sx = rng.next_float(), sy = rng.next_float(), s1 = rng.next_float();
        V3<R> ld = em.kind == 0 ? sc.area_sample(sc.meshes[em.mesh], er, sx, sy, s1) : sc.env_sample(er, sx, sy);
        if (!is_zero_fuzzy(ld)) {
          if (st) st->shadow++;
          RayF sr = Scene<R>::to_rayf(is.p, er.p - is.p, R(1e-4f), R(1.0f - 1e-4f));
          if (!sc.occluded(sr)) {
            BRec<R> br;
            br.wi = is.sh.to_local(R(-1) * td); br.wo = is.sh.to_local(er.wi); br.measure = kSolidAngle;
            li = li + beta * bsdf.eval(br) * std::abs(cos_theta(br.wo)) * ld;
          }
        }
      }
    } else {
      last_specular = true;
    }
    BRec<R> br;
    br.wi = is.sh.to_local(R(-1) * td);
    R sx = rng.next_float(), sy = rng.next_float();
    beta = beta * bsdf.sample(br, sx, sy, &u1);
    if (is_zero_fuzzy(beta)) break;
    if (!roulette(sc, rng, depth, beta)) break;
    last_emitter = mesh.emitter;
    to = is.p; td = is.sh.to_world(br.wo); ttmin = R(1e-4f); ttmax = std::numeric_limits<R>::infinity();
    depth++;
  }
  return li;
}

// src/integrator/WhittedIntegrator.cpp:26-116 (LiRecursive), unrolled: every
// vertex adds its emission; a diffuse vertex (BSDF::IsDiffuse) adds one
// emitter sample per emitter (its own emitter contributes
// Eval / Pdf |cos| Le along the shading normal, :66-76) and ends the path; any
// other vertex continues with probability 0.95 (drawn before the BSDF sample)
// while depth < the depth property.  The reference traces a shadow ray for
// every emitter sample; this restatement traces it only when the sample's
// radiance is not exactly zero (the contribution is then exactly zero).
template <typename R>
V3<R> whitted_li(const Scene<R> &sc, Pcg32 &rng, V3<R> ro, V3<R> rd, R rtmin, R rtmax, PathStats *st) {
  using Isect = typename Scene<R>::Isect;
  using ERec = typename Scene<R>::ERec;
  PcgU1 u1(&rng);
  V3<R> li(R(0)), beta(R(1));
  int depth = 0;
  V3<R> to = ro, td = rd;
  R ttmin = rtmin, ttmax = rtmax;
  for (;;) {
    Isect is;
    if (st) st->ext++;
    if (!sc.intersect(Scene<R>::to_rayf(to, td, ttmin, ttmax), is)) {  // :33-52
      if (sc.has_env && !sc.force_background) li = li + beta * sc.env_eval(td);
      else if (depth == 0) return sc.background;
      break;
    }
    const auto &mesh = sc.meshes[is.mesh];
    const Material<R> &bsdf = sc.mats[mesh.material];
    V3<R> le(R(0));
    if (mesh.emitter >= 0) {  // :54-62
      ERec er;
      er.ref = to; er.p = is.p; er.n = is.sh.n; er.wi = td;
      le = sc.area_eval(mesh, er);
      li = li + beta * le;
    }
    if (bsdf.is_diffuse()) {  // :64-98
      for (int e = 0; e < int(sc.emitters.size()); e++) {
        const auto &em = sc.emitters[e];
        BRec<R> br;
        br.wi = is.sh.to_local(R(-1) * td);
        br.measure = kSolidAngle;
        if (em.kind == 0 && em.mesh == int(is.mesh)) {
          br.wo = is.sh.to_local(is.sh.n);
          R p = bsdf.pdf(br);
          if (p != R(0)) li = li + beta * (bsdf.eval(br) / p * std::abs(cos_theta(br.wo)) * le);
          continue;
        }
        ERec er;
        er.ref = is.p;
        if (em.kind == 1) er.dist = sc.bbox_radius;
        R sx = rng.next_float(), sy = rng.next_float(), s1 = rng.next_float();
        V3<R> ld = em.kind == 0 ? sc.area_sample(sc.meshes[em.mesh], er, sx, sy, s1) : sc.env_sample(er, sx, sy);
        if (ld.x == R(0) && ld.y == R(0) && ld.z == R(0)) continue;
        if (st) st->shadow++;
        RayF sr = Scene<R>::to_rayf(is.p, er.p - is.p, R(1e-4f), R(1.0f - 1e-4f));
        if (!sc.occluded(sr)) {
          br.wo = is.sh.to_local(er.wi);
          li = li + beta * bsdf.eval(br) * std::abs(cos_theta(br.wo)) * ld;
        }
      }
      break;
    }
    bool go = rng.next_float() < 0.95f, rr = depth >= sc.rr_start;
    if (!((go || !rr) && depth < sc.max_depth)) break;  // :100-113
    BRec<R> br;
    br.wi = is.sh.to_local(R(-1) * td);
    R sx = rng.next_float(), sy = rng.next_float();
    V3<R> c = bsdf.sample(br, sx, sy, &u1);
    beta = beta * c * (rr ? R(1.0f / 0.95f) : R(1));
    if (beta.x == R(0) && beta.y == R(0) && beta.z == R(0)) break;
    to = is.p; td = is.sh.to_world(br.wo); ttmin = R(1e-4f); ttmax = std::numeric_limits<R>::infinity();
    depth++;
  }
  return li;
}

// ---------------------------------------------------------------- film + driver
// ImageBlock::Put (src/core/Block.cpp:95-134) on one image-sized block:
// pixel centres at +0.5, tabulated filter, invalid samples dropped.
template <typename R>
struct Film {
  int w, h;
  std::vector<double> acc;  // rgbw
  Film(int W, int H) : w(W), h(H), acc(size_t(W) * H * 4, 0.0) {}
  void put(const FilmFilter &ff, float px, float py, V3<R> val) {
    if (!color_valid(val)) return;
    float half = slip() == kSlipFilmCentre ? 0.0f : 0.5f;
    float cx = px - half, cy = py - half, r = ff.radius;
    int x0 = std::max(0, int(std::ceil(cx - r))), x1 = std::min(w - 1, int(std::floor(cx + r)));
    int y0 = std::max(0, int(std::ceil(cy - r))), y1 = std::min(h - 1, int(std::floor(cy + r)));
    for (int y = y0; y <= y1; y++) {
      float wy = ff.table[int(std::abs(float(y) - cy) * ff.lookup)];
      for (int x = x0; x <= x1; x++) {
        float wx = ff.table[int(std::abs(float(x) - cx) * ff.lookup)];
        double wgt = double(wx) * double(wy);
        double *a = &acc[(size_t(y) * w + x) * 4];
        a[0] += double(val.x) * wgt; a[1] += double(val.y) * wgt; a[2] += double(val.z) * wgt; a[3] += wgt;
      }
    }
  }
  void resolve(float *out) const {  // Color4f::DivideByFilterWeight (src/core/Color.cpp:111-121)
    for (size_t i = 0; i < size_t(w) * h; i++) {
      double wt = acc[i * 4 + 3];
      for (int c = 0; c < 3; c++) out[i * 3 + c] = wt != 0.0 ? float(acc[i * 4 + c] / wt) : 0.0f;
    }
  }
};

// One path sample with the per-path stream key (also used by the t-test mode).
template <typename R>
V3<R> render_sample(const Scene<R> &sc, uint32_t req_seed, uint32_t x, uint32_t y, uint32_t s, float &px, float &py,
                    PathStats *st) {
  Pcg32 rng;
  seed_path(rng, req_seed, y * uint32_t(sc.width) + x, s);
  float jx = rng.next_float(), jy = rng.next_float();
  rng.next_float(); rng.next_float();  // aperture sample (unused by the pinhole camera)
  px = float(x) + jx; py = float(y) + jy;
  V3<R> o, d;
  R tmin, tmax;
  sc.camera_ray(R(px), R(py), o, d, tmin, tmax);
  if (st) st->paths++;
  if (sc.integrator == 1) return path_ems_li(sc, rng, o, d, tmin, tmax, st);
  if (sc.integrator == 2) return path_mats_li(sc, rng, o, d, tmin, tmax, st);
  if (sc.integrator == 3) return whitted_li(sc, rng, o, d, tmin, tmax, st);
  return path_mis_li(sc, rng, o, d, tmin, tmax, st);
}

// Render rows [y0, y1) x samples [s0, s1) into a film.  Threads split rows.
template <typename R>
void render(const Scene<R> &sc, uint32_t req_seed, int s0, int s1, int threads, float *out_rgb, double *out_acc,
            PathStats *stats) {
  int W = sc.width, H = sc.height;
  threads = std::max(1, threads);
  std::vector<Film<R>> films(threads, Film<R>(W, H));
  std::vector<PathStats> sts(threads);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++) {
    pool.emplace_back([&, t]() {
      for (int y = t; y < H; y += threads)
        for (int x = 0; x < W; x++)
          for (int s = s0; s < s1; s++) {
            float px, py;
            V3<R> v = render_sample(sc, req_seed, uint32_t(x), uint32_t(y), uint32_t(s), px, py, &sts[t]);
            films[t].put(sc.film, px, py, v);
          }
    });
  }
  for (auto &th : pool) th.join();
  Film<R> total(W, H);
  for (auto &f : films)
    for (size_t i = 0; i < f.acc.size(); i++) total.acc[i] += f.acc[i];
  if (out_rgb) total.resolve(out_rgb);
  if (out_acc) std::memcpy(out_acc, total.acc.data(), total.acc.size() * sizeof(double));
  if (stats) {
    for (auto &s : sts) { stats->ext += s.ext; stats->shadow += s.shadow; stats->paths += s.paths; }
  }
}

}  // namespace hko

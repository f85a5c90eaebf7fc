// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY: extern "C" entry points of the
// CPU restatement (hk_oracle*.hpp) for the Python tests and the CPU-baseline
// leg of bench.py.  Nothing in the product loads this library.
#include <chrono>
#include <cstdio>
#include <memory>

#include "hk_oracle_scene.hpp"

using namespace hko;

extern "C" {

// Same field order as hk_scene_desc / hk_mesh_desc in include/hikari_amd.h.
struct ho_mesh {
  const float *v; uint32_t nv;
  const float *n; const float *uv;
  const uint32_t *f; uint32_t nt;
  int32_t material; int32_t is_emitter; float radiance[3];
};
struct ho_scene {
  const ho_mesh *meshes; int32_t n_meshes;
  const float *materials; int32_t n_materials;
  const float *env; int32_t env_w, env_h; float env_scale; float env_to_world[16];
  float cam_to_world[16]; float fov, near_clip, far_clip; int32_t width, height;
  int32_t filter; float filter_radius, filter_stddev, filter_b, filter_c;
  int32_t max_depth; float background[3]; int32_t force_background;
  const char *rft_dir;
  int32_t integrator;
  int32_t rr_start_depth;
};

static char g_err[512];
const char *ho_last_error() { return g_err; }

// ext/pcg32/pcg32-demo.cpp output format, seed(42, 54).
int ho_pcg32_demo(int rounds, char *buf, int buflen) {
  std::string s;
  char tmp[64];
  Pcg32 rng;
  rng.seed(42u, 54u);
  s += "pcg32_random_r:\n      -  result:      32-bit unsigned int (uint32_t)\n"
       "      -  period:      2^64   (* 2^63 streams)\n      -  state type:  pcg32_random_t (16 bytes)\n"
       "      -  output func: XSH-RR\n\n";
  static const char number[] = {'A', '2', '3', '4', '5', '6', '7', '8', '9', 'T', 'J', 'Q', 'K'};
  static const char suit[] = {'h', 'c', 'd', 's'};
  for (int round = 1; round <= rounds; ++round) {
    snprintf(tmp, sizeof tmp, "Round %d:\n", round); s += tmp;
    s += "  32bit:";
    for (int i = 0; i < 6; ++i) { snprintf(tmp, sizeof tmp, " 0x%08x", rng.next_uint()); s += tmp; }
    s += "\n  Coins: ";
    for (int i = 0; i < 65; ++i) s += rng.next_uint(2) ? 'H' : 'T';
    s += "\n  Rolls:";
    for (int i = 0; i < 33; ++i) { snprintf(tmp, sizeof tmp, " %d", (int)rng.next_uint(6) + 1); s += tmp; }
    s += "\n";
    char cards[52];
    for (int i = 0; i < 52; ++i) cards[i] = char(i);
    rng.shuffle(cards, cards + 52);
    s += "  Cards:";
    for (int i = 0; i < 52; ++i) {
      snprintf(tmp, sizeof tmp, " %c%c", number[cards[i] / 4], suit[cards[i] % 4]); s += tmp;
      if ((i + 1) % 22 == 0) s += "\n\t";
    }
    s += "\n\n";
  }
  if (int(s.size()) + 1 > buflen) return -1;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return int(s.size());
}

void ho_path_floats(uint32_t req_seed, uint32_t pixel, uint32_t sample, int n, float *out) {
  Pcg32 r;
  seed_path(r, req_seed, pixel, sample);
  for (int i = 0; i < n; i++) out[i] = r.next_float();
}

// Eigen-order vector kernels on n input pairs (a, b: n x 3): out rows of
// [dot, cross(3), normalized(a)(3), to_local(a in Frame(normalized b))(3),
//  to_world(3)] = 13 floats.
void ho_vec_ops(int n, const float *a, const float *b, float *out) {
  for (int i = 0; i < n; i++) {
    V3<float> x(a[3 * i], a[3 * i + 1], a[3 * i + 2]), y(b[3 * i], b[3 * i + 1], b[3 * i + 2]);
    float *o = out + 13 * i;
    o[0] = dot(x, y);
    V3<float> c = cross(x, y), nx = normalized(x);
    o[1] = c.x; o[2] = c.y; o[3] = c.z;
    o[4] = nx.x; o[5] = nx.y; o[6] = nx.z;
    Frame<float> fr(normalized(y));
    V3<float> l = fr.to_local(x), w = fr.to_world(x);
    o[7] = l.x; o[8] = l.y; o[9] = l.z;
    o[10] = w.x; o[11] = w.y; o[12] = w.z;
  }
}

// The scene transforms' FP32 point and vector products (M4::point / vec,
// src/core/Transform.cpp:36-50) of a row-major 4x4 matrix on n points:
// out rows of [point(3), vector(3)].
void ho_xform(const float *m16, int n, const float *p, float *out) {
  M4 m;
  for (int i = 0; i < 16; i++) m.m[i / 4][i % 4] = m16[i];
  for (int i = 0; i < n; i++) {
    V3<float> x(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    V3<float> a = m.point(x), b = m.vec(x);
    float *o = out + 6 * i;
    o[0] = a.x; o[1] = a.y; o[2] = a.z;
    o[3] = b.x; o[4] = b.y; o[5] = b.z;
  }
}

// Ray batch vs triangle soup.  tri: nt x 9 (p0, p1, p2); rays: n x 8
// (o.xyz, tmin, d.xyz, tmax).  prim = -1 on miss.
int ho_intersect(const float *tri, uint32_t nt, const float *rays, uint32_t n, int any, int use_bvh, float *out_t,
                 float *out_u, float *out_v, int64_t *out_prim) {
  TriSoup s;
  for (uint32_t i = 0; i < nt; i++) {
    const float *p = tri + 9 * i;
    s.p0.push_back({p[0], p[1], p[2]}); s.p1.push_back({p[3], p[4], p[5]}); s.p2.push_back({p[6], p[7], p[8]});
  }
  Bvh bvh;
  if (use_bvh) bvh.build(s);
  for (uint32_t i = 0; i < n; i++) {
    const float *r = rays + 8 * i;
    RayF ray;
    ray.o = {r[0], r[1], r[2]}; ray.tmin = r[3]; ray.d = {r[4], r[5], r[6]}; ray.tmax = r[7];
    Hit h;
    bool f = use_bvh ? bvh.intersect(ray, h, any != 0) : brute_intersect(s, ray, h, any != 0);
    out_t[i] = f ? h.t : 0.0f; out_u[i] = f ? h.u : 0.0f; out_v[i] = f ? h.v : 0.0f;
    out_prim[i] = f ? h.prim : -1;
  }
  return 0;
}

}  // extern "C"

struct ArrU1 : U1Source {
  const float *p; int k = 0, n;
  ArrU1(const float *q, int m) : p(q), n(m) {}
  float next() override { return k < n ? p[k++] : 0.5f; }
};

template <typename R>
static int bsdf_batch(const float *mats, int nmat, int idx, const char *rft_dir, int n, const float *wi,
                      const float *wo, const int *measure, const float *u, double *eval, double *pdf,
                      double *s_f, double *s_wo, int *s_measure, double *s_eta, double *s_pdf) {
  RftFile beck, ggx;
  bool need = false;
  for (int i = 0; i < nmat; i++) { int t = int(mats[i * 32]); if (t == kRoughPlastic || t == kRoughCoating) need = true; }
  if (need) {
    std::string d = rft_dir ? rft_dir : ".";
    beck.load(d + "/BeckmannRFTData.bin"); ggx.load(d + "/GGXRFTData.bin");
  }
  std::vector<Material<R>> m(nmat);
  std::vector<Rft<R> *> pool;
  for (int i = 0; i < nmat; i++) m[i].from_record(mats + 32 * i);
  link_materials(m, mats);
  for (int i = 0; i < nmat; i++) m[i].activate(&beck, &ggx, pool);
  const Material<R> &mat = m[idx];
  for (int i = 0; i < n; i++) {
    BRec<R> rec;
    rec.wi = V3<R>(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
    rec.wo = V3<R>(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
    rec.measure = measure[i];
    V3<R> f = mat.eval(rec);
    eval[3 * i] = f.x; eval[3 * i + 1] = f.y; eval[3 * i + 2] = f.z;
    pdf[i] = mat.pdf(rec);
    BRec<R> sr;
    sr.wi = rec.wi;
    ArrU1 extra(u + 4 * i + 2, 2);
    V3<R> sf = mat.sample(sr, R(u[4 * i]), R(u[4 * i + 1]), &extra);
    s_f[3 * i] = sf.x; s_f[3 * i + 1] = sf.y; s_f[3 * i + 2] = sf.z;
    s_wo[3 * i] = sr.wo.x; s_wo[3 * i + 1] = sr.wo.y; s_wo[3 * i + 2] = sr.wo.z;
    s_measure[i] = sr.measure;
    s_eta[i] = sr.eta;
    s_pdf[i] = mat.pdf(sr);
  }
  for (auto *p : pool) delete p;
  return 0;
}

extern "C" {

int ho_bsdf_batch(int fp64, const float *mats, int nmat, int idx, const char *rft_dir, int n, const float *wi,
                  const float *wo, const int *measure, const float *u, double *eval, double *pdf, double *s_f,
                  double *s_wo, int *s_measure, double *s_eta, double *s_pdf) {
  try {
    if (fp64) return bsdf_batch<double>(mats, nmat, idx, rft_dir, n, wi, wo, measure, u, eval, pdf, s_f, s_wo, s_measure, s_eta, s_pdf);
    return bsdf_batch<float>(mats, nmat, idx, rft_dir, n, wi, wo, measure, u, eval, pdf, s_f, s_wo, s_measure, s_eta, s_pdf);
  } catch (const std::exception &e) {
    snprintf(g_err, sizeof g_err, "%s", e.what());
    return -1;
  }
}

// Reduced RFT tables (RoughTransmittance::SetEta then SetAlpha) for tests.
int ho_rft_reduce(const char *path, float eta, float alpha, int stage, float *out_t, float *out_d) {
  try {
    RftFile f;
    f.load(path);
    Rft<float> r(f);
    if (stage >= 1) r.set_eta(eta);
    if (stage >= 2) r.set_alpha(alpha);
    std::memcpy(out_t, r.t.data(), r.t.size() * 4);
    std::memcpy(out_d, r.d.data(), r.d.size() * 4);
    return int(r.t.size());
  } catch (const std::exception &e) {
    snprintf(g_err, sizeof g_err, "%s", e.what());
    return -1;
  }
}
double ho_rft_eval(const char *path, int stage, float eta, float alpha, float cos_t) {
  RftFile f;
  f.load(path);
  Rft<double> r(f);
  if (stage >= 1) r.set_eta(eta);
  if (stage >= 2) r.set_alpha(alpha);
  return r.eval(cos_t, alpha, eta);
}

// RoughTransmittance::EvalDiffuse (src/core/RoughTransmitance.cpp:114-150) of
// the full table (stage 0), after SetEta (1) or after SetAlpha (2).
double ho_rft_eval_diffuse(const char *path, int stage, float eta, float alpha) {
  RftFile f;
  f.load(path);
  Rft<double> r(f);
  if (stage >= 1) r.set_eta(eta);
  if (stage >= 2) r.set_alpha(alpha);
  return r.eval_diffuse(alpha, eta);
}

// Selects a planted mistake of the restatement (hk_oracle.hpp, enum Slip);
// 0 restores the faithful one.
void ho_set_slip(int s) { slip() = s; }

// ImageBlock::Put (src/core/Block.cpp:95-134) of n samples (pos: n x 2 film
// coordinates, val: n x 3) into a W x H film with the tabulated filter of
// type ft (0 box, 1 gaussian, 2 tent, 3 mitchell); out: W x H x 4 sums.
int ho_film_splat(int ft, float radius, float stddev, float b, float c, int W, int H, int n, const float *pos,
                  const float *val, double *out) {
  if (W <= 0 || H <= 0 || n < 0) {
    snprintf(g_err, sizeof g_err, "bad film %d x %d or sample count %d", W, H, n);
    return -1;
  }
  FilmFilter ff;
  ff.setup(ft, radius, stddev, b, c);
  Film<float> film(W, H);
  for (int i = 0; i < n; i++)
    film.put(ff, pos[2 * i], pos[2 * i + 1], V3<float>(val[3 * i], val[3 * i + 1], val[3 * i + 2]));
  std::memcpy(out, film.acc.data(), film.acc.size() * sizeof(double));
  return 0;
}

double ho_fresnel_dielectric(double cos_i, double eta, double *cos_t) {
  return fresnel_dielectric<double>(cos_i, eta, 1.0 / eta, *cos_t);
}
void ho_fresnel_conductor(double cos_i, const double *eta, const double *k, double *out) {
  V3<double> r = fresnel_conductor<double>(cos_i, V3<double>(eta[0], eta[1], eta[2]), V3<double>(k[0], k[1], k[2]));
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
// Microfacet D / pdf / G1 / sample for quadrature tests (double).
void ho_microfacet(int type, double au, double av, int n, const double *m, const double *v, double *out_d,
                   double *out_g1, const double *u, double *out_sample, double *out_spdf) {
  Microfacet<double> md(type, au, av);
  for (int i = 0; i < n; i++) {
    V3<double> mm(m[3 * i], m[3 * i + 1], m[3 * i + 2]), vv(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    out_d[i] = md.eval(mm);
    out_g1[i] = md.smith_g1(vv, mm);
    double p;
    V3<double> s = md.sample(u[2 * i], u[2 * i + 1], p);
    out_sample[3 * i] = s.x; out_sample[3 * i + 1] = s.y; out_sample[3 * i + 2] = s.z;
    out_spdf[i] = p;
  }
}

static void to_in(const ho_scene *s, SceneIn &in) {
  for (int i = 0; i < s->n_meshes; i++) {
    const ho_mesh &m = s->meshes[i];
    MeshIn mi;
    mi.v = m.v; mi.nv = m.nv; mi.n = m.n; mi.uv = m.uv; mi.f = m.f; mi.nt = m.nt;
    mi.material = m.material; mi.is_emitter = m.is_emitter;
    for (int c = 0; c < 3; c++) mi.radiance[c] = m.radiance[c];
    in.meshes.push_back(mi);
  }
  in.materials = s->materials; in.n_materials = s->n_materials;
  in.env = s->env; in.env_w = s->env_w; in.env_h = s->env_h; in.env_scale = s->env_scale;
  for (int i = 0; i < 16; i++) { in.env_to_world[i] = s->env_to_world[i]; in.cam_to_world[i] = s->cam_to_world[i]; }
  in.fov = s->fov; in.near_clip = s->near_clip; in.far_clip = s->far_clip;
  in.width = s->width; in.height = s->height;
  in.filter = s->filter; in.filter_radius = s->filter_radius; in.filter_stddev = s->filter_stddev;
  in.filter_b = s->filter_b; in.filter_c = s->filter_c;
  in.max_depth = s->max_depth;
  for (int c = 0; c < 3; c++) in.background[c] = s->background[c];
  in.force_background = s->force_background;
  in.rft_dir = s->rft_dir;
  in.integrator = s->integrator;
  in.rr_start_depth = s->rr_start_depth;
}

// Render samples [s0, s1) of every pixel.  out_rgb: W*H*3 resolved image
// (nullable); out_acc: W*H*4 filter-weighted sums (nullable); stats: ext,
// shadow, paths, render microseconds excluding scene setup (nullable).
int ho_render(int fp64, const ho_scene *s, uint32_t req_seed, int s0, int s1, int threads, float *out_rgb,
              double *out_acc, uint64_t *stats) {
  try {
    SceneIn in;
    to_in(s, in);
    PathStats st;
    auto t0 = std::chrono::steady_clock::now(), t1 = t0;
    if (fp64) {
      Scene<double> sc; sc.load(in);
      t0 = std::chrono::steady_clock::now();
      render(sc, req_seed, s0, s1, threads, out_rgb, out_acc, &st);
      t1 = std::chrono::steady_clock::now();
    } else {
      Scene<float> sc; sc.load(in);
      t0 = std::chrono::steady_clock::now();
      render(sc, req_seed, s0, s1, threads, out_rgb, out_acc, &st);
      t1 = std::chrono::steady_clock::now();
    }
    if (stats) {
      stats[0] = st.ext; stats[1] = st.shadow; stats[2] = st.paths;
      stats[3] = uint64_t(std::chrono::duration<double, std::micro>(t1 - t0).count());  // render time, us
    }
    return 0;
  } catch (const std::exception &e) {
    snprintf(g_err, sizeof g_err, "%s", e.what());
    return -1;
  }
}

// Per-sample radiance for pixel (x, y), samples [s0, s0+n): out n x 5
// (r, g, b, px, py).  Used by the t-test harness and per-path comparisons.
int ho_pixel_samples(int fp64, const ho_scene *s, uint32_t req_seed, uint32_t x, uint32_t y, uint32_t s0, uint32_t n,
                     float *out) {
  try {
    SceneIn in;
    to_in(s, in);
    auto run = [&](auto &sc) {
      for (uint32_t k = 0; k < n; k++) {
        float px, py;
        auto v = render_sample(sc, req_seed, x, y, s0 + k, px, py, nullptr);
        out[5 * k] = float(v.x); out[5 * k + 1] = float(v.y); out[5 * k + 2] = float(v.z);
        out[5 * k + 3] = px; out[5 * k + 4] = py;
      }
    };
    if (fp64) { Scene<double> sc; sc.load(in); run(sc); }
    else { Scene<float> sc; sc.load(in); run(sc); }
    return 0;
  } catch (const std::exception &e) {
    snprintf(g_err, sizeof g_err, "%s", e.what());
    return -1;
  }
}

}  // extern "C"

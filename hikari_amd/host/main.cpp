// main.cpp -- hikari_render: the command-line driver (src/Main.cpp without
// the GUI).  Loads a scene or test description; a scene is rendered with its
// integrator and written as EXR or PFM, a test runs at load time and sets the
// exit status.
//
//   hikari_render FILE.xml [-o OUT.exr|OUT.pfm] [--spp N] [--dump] [--dump-mesh PREFIX]
//   hikari_render --convert IN.exr|IN.pfm OUT.exr|OUT.pfm
//   hikari_render --t-test MEAN VARIANCE REFERENCE N LEVEL N_TESTS
//   hikari_render --chi2-cdf X DOF
//   hikari_render --chi2-table FILE
//
// --dump prints the parsed description as JSON and touches no device;
// --dump-mesh also writes every mesh's world-space vertex positions and
// normals as raw FP32 (PREFIX.<mesh>.v.f32, PREFIX.<mesh>.n.f32);
// --convert rewrites an image (EXR written as uncompressed FLOAT RGB);
// --t-test runs the Student t-test of the test harness on given statistics
// and prints its verdict (exit status 0 accepted, 1 rejected); --chi2-cdf
// prints the chi-square distribution function of the chi2test harness;
// --chi2-table runs its pooled chi-square test on a text file holding
// "cells sample_count min_exp level n_tests", then the observed and the
// expected frequencies (exit status 0 accepted, 1 rejected).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "hikari.h"

using namespace hikari;

static void json_floats(const float *v, size_t n) {
  printf("[");
  for (size_t i = 0; i < n; i++) printf("%s%.9g", i ? ", " : "", double(v[i]));
  printf("]");
}

static std::string g_mesh_prefix;

static void write_f32(const std::string &path, const std::vector<float> &v) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) throw HikariException("cannot write \"%s\"", path.c_str());
  size_t n = v.empty() ? 0 : fwrite(v.data(), sizeof(float), v.size(), f);
  fclose(f);
  if (n != v.size()) throw HikariException("short write to \"%s\"", path.c_str());
}

static void dump_scene(const Scene *s) {
  printf("{\"meshes\": [");
  const auto &ms = s->GetMeshes();
  for (size_t i = 0; i < ms.size(); i++) {
    const Mesh *m = ms[i];
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (size_t k = 0; k < m->V.size(); k++) {
      lo[k % 3] = std::min(lo[k % 3], m->V[k]);
      hi[k % 3] = std::max(hi[k % 3], m->V[k]);
    }
    if (!g_mesh_prefix.empty()) {
      write_f32(g_mesh_prefix + "." + std::to_string(i) + ".v.f32", m->V);
      write_f32(g_mesh_prefix + "." + std::to_string(i) + ".n.f32", m->N);
    }
    std::vector<float> rec;
    m->bsdf->AppendRecords(rec);
    printf("%s{\"nv\": %zu, \"nt\": %zu, \"has_n\": %d, \"has_uv\": %d, \"lo\": ", i ? ", " : "", m->V.size() / 3,
           m->F.size() / 3, !m->N.empty(), !m->UV.empty());
    json_floats(lo, 3);
    printf(", \"hi\": ");
    json_floats(hi, 3);
    printf(", \"materials\": ");
    json_floats(rec.data(), rec.size());
    if (m->emitter) {
      float r[3] = {m->emitter->radiance.r, m->emitter->radiance.g, m->emitter->radiance.b};
      printf(", \"radiance\": ");
      json_floats(r, 3);
    }
    printf("}");
  }
  const Camera *c = s->GetCamera();
  float tw[16];
  c->to_world.ToFloat(tw);
  printf("], \"camera\": {\"width\": %d, \"height\": %d, \"fov\": %.9g, \"near\": %.9g, \"far\": %.9g, \"to_world\": ",
         c->width, c->height, double(c->fov), double(c->near_clip), double(c->far_clip));
  json_floats(tw, 16);
  printf(", \"filter\": %d, \"radius\": %.9g}", c->filter->kind, double(c->filter->radius));
  printf(", \"spp\": %d, \"integrator\": \"%s\"", s->GetSampler()->GetSampleCount(),
         s->GetIntegrator()->ToString().c_str());
  if (const EnvironmentLight *e = s->GetEnvironmentEmitter())
    printf(", \"env\": {\"width\": %d, \"height\": %d, \"scale\": %.9g}", e->width, e->height, double(e->scale));
  printf("}");
}

int main(int argc, char **argv) {
  if (argc == 4 && std::strcmp(argv[1], "--convert") == 0) {
    try {
      int w = 0, h = 0;
      std::vector<float> img = ReadImage(argv[2], w, h);
      std::string o = argv[3];
      std::string ext = o.size() >= 4 ? o.substr(o.size() - 4) : "";
      if (ext == ".pfm") WritePFM(o, img.data(), w, h);
      else WriteEXR(o, img.data(), w, h);
      printf("%dx%d\n", w, h);
      return 0;
    } catch (const std::exception &e) {
      fprintf(stderr, "error: %s\n", e.what());
      return 1;
    }
  }
  if (argc == 4 && std::strcmp(argv[1], "--chi2-cdf") == 0) {
    printf("%.17g\n", Chi2Cdf(std::atof(argv[2]), std::atoi(argv[3])));
    return 0;
  }
  if (argc == 3 && std::strcmp(argv[1], "--chi2-table") == 0) {
    FILE *fp = std::fopen(argv[2], "r");
    int n = 0, n_tests = 1;
    long count = 0;
    double min_exp = 0, level = 0;
    if (!fp || std::fscanf(fp, "%d %ld %lf %lf %d", &n, &count, &min_exp, &level, &n_tests) != 5 || n <= 0) {
      fprintf(stderr, "error: bad table %s\n", argv[2]);
      if (fp) std::fclose(fp);
      return 2;
    }
    std::vector<double> obs(static_cast<size_t>(n)), expv(static_cast<size_t>(n));
    bool ok = true;
    for (int i = 0; i < n && ok; i++) ok = std::fscanf(fp, "%lf", &obs[size_t(i)]) == 1;
    for (int i = 0; i < n && ok; i++) ok = std::fscanf(fp, "%lf", &expv[size_t(i)]) == 1;
    std::fclose(fp);
    if (!ok) {
      fprintf(stderr, "error: short table %s\n", argv[2]);
      return 2;
    }
    std::string msg;
    bool acc = PooledChi2Test(n, obs.data(), expv.data(), count, min_exp, level, n_tests, msg);
    printf("%s\n", msg.c_str());
    return acc ? 0 : 1;
  }
  if (argc == 8 && std::strcmp(argv[1], "--t-test") == 0) {
    std::string msg;
    bool ok = StudentsTTest(std::atof(argv[2]), std::atof(argv[3]), std::atof(argv[4]), std::atol(argv[5]),
                            std::atof(argv[6]), std::atoi(argv[7]), msg);
    printf("%s\n", msg.c_str());
    return ok ? 0 : 1;
  }
  std::string file, out;
  int spp = -1;
  bool dump = false;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a == "-o" && i + 1 < argc) out = argv[++i];
    else if (a == "--spp" && i + 1 < argc) spp = std::atoi(argv[++i]);
    else if (a == "--dump") dump = true;
    else if (a == "--dump-mesh" && i + 1 < argc) {
      dump = true;
      g_mesh_prefix = argv[++i];
    }
    else if (file.empty()) file = a;
    else {
      fprintf(stderr, "usage: hikari_render FILE.xml [-o OUT.exr|OUT.pfm] [--spp N] [--dump] [--dump-mesh PREFIX]\n");
      return 2;
    }
  }
  if (file.empty()) {
    fprintf(stderr, "usage: hikari_render FILE.xml [-o OUT.exr|OUT.pfm] [--spp N] [--dump] [--dump-mesh PREFIX]\n");
    return 2;
  }
  SetDryRun(dump);
  try {
    std::unique_ptr<Object> root(LoadFromXML(file));
    if (root->GetClassType() == Object::EClassType::ETest) {
      auto *base = static_cast<TestObject *>(root.get());
      if (std::strcmp(base->TestKind(), "ttest") != 0) {
        if (dump) {
          printf("{\"test\": \"%s\", \"description\": \"%s\"}\n", base->TestKind(), base->ToString().c_str());
          return 0;
        }
        return base->passed == base->total ? 0 : 1;
      }
      auto *t = static_cast<StudentTTest *>(root.get());
      if (dump) {
        printf("{\"test\": \"ttest\", \"references\": ");
        json_floats(t->GetReferences().data(), t->GetReferences().size());
        printf(", \"scenes\": [");
        for (size_t i = 0; i < t->GetScenes().size(); i++) {
          if (i) printf(", ");
          dump_scene(t->GetScenes()[i]);
        }
        printf("]}\n");
        return 0;
      }
      return t->passed == t->total ? 0 : 1;
    }
    if (root->GetClassType() != Object::EClassType::EScene) {
      fprintf(stderr, "%s: the root object is a %s, not a scene or test\n", file.c_str(),
              Object::ClassTypeName(root->GetClassType()).c_str());
      return 2;
    }
    auto *scene = static_cast<Scene *>(root.get());
    if (dump) {
      dump_scene(scene);
      printf("\n");
      return 0;
    }
    if (spp > 0) {
      PropertyList p;
      p.SetInteger("sampleCount", spp);
      const_cast<Sampler *>(scene->GetSampler())->sample_count = spp;
    }
    auto t0 = std::chrono::steady_clock::now();
    scene->GetIntegrator()->Preprocess(scene);
    std::vector<float> rgb;
    scene->GetIntegrator()->RenderImage(scene, rgb);
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const Camera *c = scene->GetCamera();
    if (out.empty()) {
      size_t dot = file.find_last_of('.');
      out = (dot == std::string::npos ? file : file.substr(0, dot)) + ".exr";
    }
    std::string ext = out.size() >= 4 ? out.substr(out.size() - 4) : "";
    if (ext == ".pfm") WritePFM(out, rgb.data(), c->width, c->height);
    else WriteEXR(out, rgb.data(), c->width, c->height);
    printf("Rendered %dx%d at %d spp in %.2f s -> %s\n", c->width, c->height, scene->GetSampler()->GetSampleCount(),
           secs, out.c_str());
    return 0;
  } catch (const std::exception &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}

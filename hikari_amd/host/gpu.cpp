// gpu.cpp -- the GPU plugins: integrator `gpuPathMIS` (registered also as
// `pathMIS`, the reference's name for the estimator it replaces; likewise
// pathEMS, pathMATS and whitted) and the accelerator `gpuBVH` (also `bvh`,
// `hlbvh`, `bruto`).  Both reach the device through the C ABI of
// include/hikari_amd.h only.
#include <cstring>
#include <mutex>

#include "hikari.h"

namespace hikari {

void CheckHk(int rc, const char *what) {
  if (rc != HK_OK) throw HikariException("%s: %s", what, hk_last_error(nullptr));
}

hk_ctx *GpuContext() {
  static std::mutex mu;
  static hk_ctx *ctx = nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!ctx) {
    if (hk_ctx_create(0, &ctx) != HK_OK) {
      ctx = nullptr;
      throw HikariException("no usable MI355X device: %s", hk_last_error(nullptr));
    }
  }
  return ctx;
}

static void check_ctx(int rc, const char *what) {
  if (rc != HK_OK) throw HikariException("%s: %s", what, hk_last_error(GpuContext()));
}

void Integrator::Preprocess(const Scene *) {}

namespace {

// PathMISIntegrator::Li (src/integrator/PathMISIntegrator.cpp:16-157) for
// whole images, on the wavefront kernels; the same class serves pathEMS
// (PathEMSIntegrator.cpp), pathMATS (PathMATSIntegrator.cpp) and whitted
// (WhittedIntegrator.cpp), which the bounce kernel selects by the scene's
// estimator id.  Per-ray Li() is not
// offered: the estimators run only on the device.
class GPUPathMISIntegrator : public Integrator {
 public:
  // depth: required by the path integrators (PathMISIntegrator.cpp:13),
  // optional for whitted (-1, no limit: WhittedIntegrator.cpp:13).
  // Two properties the reference does not have, both optional: queueSize,
  // the paths in flight (the context's "max_paths"; 0 keeps the context's
  // setting; no film depends on it), and rrStartDepth, the path depth from
  // which Russian roulette plays (0, the reference's schedule: every vertex).
  explicit GPUPathMISIntegrator(const PropertyList &p, int estimator = 0) : estimator_(estimator) {
    depth_ = estimator == 3 ? p.GetInteger("depth", -1) : p.GetInteger("depth");
    queue_size_ = p.GetInteger("queueSize", 0);
    rr_start_ = p.GetInteger("rrStartDepth", 0);
    if (queue_size_ < 0) throw HikariException("queueSize must be >= 0 (got %d)", queue_size_);
    if (rr_start_ < 0) throw HikariException("rrStartDepth must be >= 0 (got %d)", rr_start_);
  }
  int EstimatorId() const override { return estimator_; }
  ~GPUPathMISIntegrator() override {
    if (scene_) hk_scene_destroy(scene_);
  }
  void Preprocess(const Scene *scene) override {
    if (scene_) {
      hk_scene_destroy(scene_);
      scene_ = nullptr;
    }
    std::vector<float> records;
    std::vector<hk_mesh_desc> meshes;
    hk_scene_desc d = scene->Describe(depth_, estimator_, records, meshes);
    d.rr_start_depth = rr_start_;
    check_ctx(hk_scene_create(GpuContext(), &d, &scene_), "gpuPathMIS: scene upload");
    width_ = d.width;
    height_ = d.height;
  }
  void RenderImage(const Scene *scene, std::vector<float> &rgb) override {
    if (!scene_) Preprocess(scene);
    hk_request req;
    std::memset(&req, 0, sizeof req);
    req.use_scene_camera = 1;
    req.spp = uint32_t(scene->GetSampler()->GetSampleCount());
    rgb.assign(size_t(width_) * height_ * 3, 0.0f);
    if (queue_size_ > 0) check_ctx(hk_set_option(GpuContext(), "max_paths", queue_size_), "gpuPathMIS: queueSize");
    check_ctx(hk_render(scene_, &req, rgb.data(), nullptr), "gpuPathMIS: render");
  }
  // per-sample radiance of one pixel (the t-test harness)
  void PixelSamples(const Scene *scene, uint32_t seed, uint32_t x, uint32_t y, uint32_t n, std::vector<float> &out) {
    if (!scene_) Preprocess(scene);
    out.assign(size_t(n) * 5, 0.0f);
    check_ctx(hk_render_pixel_samples(scene_, seed, x, y, 0, n, out.data()), "gpuPathMIS: pixel samples");
  }
  std::string ToString() const override {
    static const char *names[] = {"pathMIS", "pathEMS", "pathMATS", "whitted"};
    std::string s = std::string("GPUPathIntegrator[") + names[estimator_] + ", depth = " + std::to_string(depth_);
    if (rr_start_) s += ", rrStartDepth = " + std::to_string(rr_start_);
    if (queue_size_) s += ", queueSize = " + std::to_string(queue_size_);
    return s + "]";
  }

 private:
  int estimator_;
  int depth_;
  int queue_size_ = 0, rr_start_ = 0;
  hk_scene *scene_ = nullptr;
  int width_ = 0, height_ = 0;
};

// Acceleration::RayIntersect over hk_accel (src/core/Acceleration.cpp:45-80
// semantics: closest hit, ties to the last tested triangle; shadow rays stop
// at the first hit).  The device BVH is built at the first query.
class GPUBVHAcceleration : public Acceleration {
 public:
  explicit GPUBVHAcceleration(const PropertyList &) {}
  ~GPUBVHAcceleration() override {
    if (accel_) hk_accel_destroy(accel_);
  }
  void Build() override { built_ = false; }
  bool RayIntersect(const float o[3], const float d[3], float tmin, float tmax, Hit &hit, bool shadow) const override {
    ensure();
    hk_ray r = {o[0], o[1], o[2], tmin, d[0], d[1], d[2], tmax};
    hk_hit h;
    check_ctx(hk_accel_intersect(accel_, &r, 1, &h, shadow ? 1 : 0), "gpuBVH: intersect");
    if (h.prim < 0) return false;
    hit.t = h.t;
    hit.u = h.u;
    hit.v = h.v;
    size_t k = 0;
    while (k + 1 < first_.size() && first_[k + 1] <= uint32_t(h.prim)) k++;
    hit.mesh = meshes_[k];
    hit.facet = uint32_t(h.prim) - first_[k];
    return true;
  }
  std::string ToString() const override { return "GPUBVHAcceleration[]"; }

 private:
  void ensure() const {
    if (built_) return;
    std::vector<float> v;
    std::vector<uint32_t> f;
    first_.clear();
    for (const Mesh *m : meshes_) {
      uint32_t base = uint32_t(v.size() / 3);
      first_.push_back(uint32_t(f.size() / 3));
      v.insert(v.end(), m->V.begin(), m->V.end());
      for (uint32_t i : m->F) f.push_back(base + i);
    }
    if (f.empty()) throw HikariException("gpuBVH: the scene has no triangles");
    if (accel_) hk_accel_destroy(accel_);
    accel_ = nullptr;
    check_ctx(hk_accel_build(GpuContext(), v.data(), uint32_t(v.size() / 3), f.data(), uint32_t(f.size() / 3), &accel_),
              "gpuBVH: build");
    built_ = true;
  }
  mutable hk_accel *accel_ = nullptr;
  mutable bool built_ = false;
  mutable std::vector<uint32_t> first_;
};

}  // namespace

struct GPUPathEMSIntegrator : GPUPathMISIntegrator {
  explicit GPUPathEMSIntegrator(const PropertyList &p) : GPUPathMISIntegrator(p, 1) {}
};
struct GPUPathMATSIntegrator : GPUPathMISIntegrator {
  explicit GPUPathMATSIntegrator(const PropertyList &p) : GPUPathMISIntegrator(p, 2) {}
};
struct GPUWhittedIntegrator : GPUPathMISIntegrator {
  explicit GPUWhittedIntegrator(const PropertyList &p) : GPUPathMISIntegrator(p, 3) {}
};

REGISTER_CLASS(GPUPathMISIntegrator, "gpuPathMIS")
REGISTER_CLASS(GPUPathMISIntegrator, "pathMIS")
REGISTER_CLASS(GPUPathEMSIntegrator, "gpuPathEMS")
REGISTER_CLASS(GPUPathEMSIntegrator, "pathEMS")
REGISTER_CLASS(GPUPathMATSIntegrator, "gpuPathMATS")
REGISTER_CLASS(GPUPathMATSIntegrator, "pathMATS")
REGISTER_CLASS(GPUWhittedIntegrator, "gpuWhitted")
REGISTER_CLASS(GPUWhittedIntegrator, "whitted")
REGISTER_CLASS(GPUBVHAcceleration, "gpuBVH")
REGISTER_CLASS(GPUBVHAcceleration, "bvh")
REGISTER_CLASS(GPUBVHAcceleration, "hlbvh")
REGISTER_CLASS(GPUBVHAcceleration, "bruto")

// pixel-sample access for the t-test (defined here, where the class is)
bool RenderPixelSamples(Integrator *integ, const Scene *scene, uint32_t seed, uint32_t n, std::vector<float> &out) {
  auto *g = dynamic_cast<GPUPathMISIntegrator *>(integ);
  if (!g) return false;
  g->PixelSamples(scene, seed, 0, 0, n, out);
  return true;
}

}  // namespace hikari

// objects.cpp -- scene object classes registered with the factory: BSDFs
// (material records), area / environment emitters, the OBJ mesh, camera,
// reconstruction filters, sampler and the scene itself.  Property names and
// defaults are the reference's (include/core/Common.hpp:66-312, :314-423).
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "hikari.h"

namespace hikari {

static std::string resolve(const std::string &name) {
  if (name.empty() || name[0] == '/') return name;
  std::string n = name;
  for (char &c : n)
    if (c == '\\') c = '/';  // the reference's scene files use Windows separators
  return CurrentSceneDirectory() + "/" + n;
}

// ------------------------------------------------------------------ BSDFs
namespace {

constexpr float kAir = 1.000277f;

enum RecType { kDiffuse = 0, kConductor, kDielectric, kPlastic, kRoughConductor, kRoughDielectric, kRoughPlastic,
               kCoating, kRoughCoating, kMirror, kMicrofacet, kRoughDiffuse, kTwoSided };

// One material record (HK_MATERIAL_STRIDE floats; layout in hikari_amd.h).
struct Record {
  float r[HK_MATERIAL_STRIDE];
  Record(int type) {
    std::memset(r, 0, sizeof r);
    r[0] = float(type);
    r[2] = -1.0f;
    r[12] = -1.0f;
  }
  void a(const Color3f &c) { r[4] = c.r; r[5] = c.g; r[6] = c.b; }
  void b(const Color3f &c) { r[8] = c.r; r[9] = c.g; r[10] = c.b; }
};

int dist_of(const PropertyList &p, const char *key) {
  std::string t = p.GetString(key, "beckmann");
  if (t == "beckmann") return 0;
  if (t == "ggx") return 1;
  throw HikariException("Unknown microfacet distribution \"%s\"", t.c_str());
}

class MaterialBSDF : public BSDF {
 public:
  MaterialBSDF(const std::string &type, const Record &rec, bool wants_nested)
      : type_(type), rec_(rec), wants_nested_(wants_nested) {}
  ~MaterialBSDF() override { delete nested_; }
  void AddChild(Object *child, const std::string &name) override {
    if (!wants_nested_ || child->GetClassType() != EClassType::EBSDF)
      throw HikariException("%s::AddChild(<%s>, <%s>) is not supported!", type_.c_str(),
                            ClassTypeName(child->GetClassType()).c_str(), name.c_str());
    if (nested_) throw HikariException("%s: tried to register multiple nested BSDFs!", type_.c_str());
    nested_ = static_cast<BSDF *>(child);
  }
  void Activate() override {
    if (wants_nested_ && !nested_) throw HikariException("%s: a nested BSDF is required", type_.c_str());
  }
  int AppendRecords(std::vector<float> &records) const override {
    Record r = rec_;
    if (nested_) r.r[2] = float(nested_->AppendRecords(records));
    int idx = int(records.size() / HK_MATERIAL_STRIDE);
    records.insert(records.end(), r.r, r.r + HK_MATERIAL_STRIDE);
    return idx;
  }
  std::string ToString() const override { return type_ + "[]"; }

 private:
  std::string type_;
  Record rec_;
  bool wants_nested_;
  BSDF *nested_ = nullptr;
};

Object *make_diffuse(const PropertyList &p) {
  Record r(kDiffuse);
  r.a(p.GetColor("albedo", Color3f(0.5f)));
  return new MaterialBSDF("diffuse", r, false);
}
Object *make_conductor(const PropertyList &p) {
  Record r(kConductor);
  r.a(p.GetColor("ks", Color3f(1.0f)));
  r.b(p.GetColor("k", Color3f(1.0f)));
  r.r[20] = p.GetFloat("intIOR", 1.5046f);
  r.r[21] = p.GetFloat("extIOR", kAir);
  return new MaterialBSDF("conductor", r, false);
}
Object *make_dielectric(const PropertyList &p) {
  Record r(kDielectric);
  r.a(p.GetColor("ksReflect", Color3f(1.0f)));
  r.b(p.GetColor("ksRefract", Color3f(1.0f)));
  r.r[20] = p.GetFloat("intIOR", 1.5046f);
  r.r[21] = p.GetFloat("extIOR", kAir);
  return new MaterialBSDF("dielectric", r, false);
}
Object *make_plastic(const PropertyList &p) {
  Record r(kPlastic);
  r.a(p.GetColor("ks", Color3f(1.0f)));
  r.b(p.GetColor("kd", Color3f(0.5f)));
  r.r[20] = p.GetFloat("intIOR", 1.5046f);
  r.r[21] = p.GetFloat("extIOR", kAir);
  r.r[3] = p.GetBoolean("nonlinear", false) ? 1.0f : 0.0f;
  return new MaterialBSDF("plastic", r, false);
}
// alpha / alphaU / alphaV with the "as" (anisotropic) switch
void rough_alpha(const PropertyList &p, Record &r) {
  bool aniso = p.GetBoolean("as", false);
  if (aniso) {
    r.r[22] = p.GetFloat("alphaU", 0.1f);
    r.r[23] = p.GetFloat("alphaV", 0.1f);
    r.r[3] = float(int(r.r[3]) | 2);
  } else {
    r.r[22] = r.r[23] = p.GetFloat("alpha", 0.1f);
  }
}
Object *make_rough_conductor(const PropertyList &p) {
  Record r(kRoughConductor);
  r.a(p.GetColor("ks", Color3f(1.0f)));
  r.b(p.GetColor("k", Color3f(1.0f)));
  r.r[20] = p.GetFloat("intIOR", 1.5046f);
  r.r[21] = p.GetFloat("extIOR", kAir);
  r.r[1] = float(dist_of(p, "type"));
  rough_alpha(p, r);
  return new MaterialBSDF("roughConductor", r, false);
}
Object *make_rough_dielectric(const PropertyList &p) {
  Record r(kRoughDielectric);
  r.a(p.GetColor("ksReflect", Color3f(1.0f)));
  r.b(p.GetColor("ksRefract", Color3f(1.0f)));
  r.r[20] = p.GetFloat("intIOR", 1.5046f);
  r.r[21] = p.GetFloat("extIOR", kAir);
  r.r[1] = float(dist_of(p, "type"));
  rough_alpha(p, r);
  return new MaterialBSDF("roughDielectric", r, false);
}
Object *make_rough_plastic(const PropertyList &p) {
  Record r(kRoughPlastic);
  r.a(p.GetColor("ks", Color3f(1.0f)));
  r.b(p.GetColor("kd", Color3f(0.5f)));
  r.r[20] = p.GetFloat("intIOR", 1.49f);
  r.r[21] = p.GetFloat("extIOR", kAir);
  r.r[3] = p.GetBoolean("nonlinear", false) ? 1.0f : 0.0f;
  r.r[1] = float(dist_of(p, "type"));
  r.r[22] = r.r[23] = p.GetFloat("alpha", 0.1f);
  return new MaterialBSDF("roughPlastic", r, false);
}
Object *make_coating(const PropertyList &p) {
  Record r(kCoating);
  r.a(p.GetColor("sigmaA", Color3f(0.0f)));
  r.b(p.GetColor("ks", Color3f(1.0f)));
  r.r[20] = p.GetFloat("intIOR", 1.49f);
  r.r[21] = p.GetFloat("extIOR", kAir);
  r.r[24] = p.GetFloat("thickness", 1.0f);
  return new MaterialBSDF("coating", r, true);
}
Object *make_rough_coating(const PropertyList &p) {
  Record r(kRoughCoating);
  r.a(p.GetColor("sigmaA", Color3f(0.0f)));
  r.b(p.GetColor("ks", Color3f(1.0f)));
  r.r[20] = p.GetFloat("intIOR", 1.49f);
  r.r[21] = p.GetFloat("extIOR", kAir);
  r.r[24] = p.GetFloat("thickness", 1.0f);
  r.r[1] = float(dist_of(p, "type"));
  r.r[22] = r.r[23] = p.GetFloat("alpha", 0.1f);
  return new MaterialBSDF("roughCoating", r, true);
}

Object *make_mirror(const PropertyList &) { return new MaterialBSDF("mirror", Record(kMirror), false); }
Object *make_microfacet(const PropertyList &p) {  // MicrofacetBSDF.cpp:10-28
  Record r(kMicrofacet);
  r.r[22] = r.r[23] = p.GetFloat("alpha", 0.1f);
  r.r[20] = p.GetFloat("intIOR", 1.5046f);
  r.r[21] = p.GetFloat("extIOR", kAir);
  r.a(p.GetColor("kd", Color3f(0.5f)));
  return new MaterialBSDF("microfacet", r, false);
}
Object *make_rough_diffuse(const PropertyList &p) {  // RoughDiffuseBSDF.cpp:10-21
  Record r(kRoughDiffuse);
  r.a(p.GetColor("albedo", Color3f(0.5f)));
  r.r[22] = r.r[23] = p.GetFloat("alpha", 0.2f);
  r.r[3] = p.GetBoolean("fastApprox", false) ? 4.0f : 0.0f;
  return new MaterialBSDF("roughDiffuse", r, false);
}

// TwoSidedBSDF (src/bsdf/TwoSidedBSDF.cpp): up to two nested BSDFs, upper side
// first; a missing lower side repeats the upper one.  Children that are not
// BSDFs are ignored, as in the reference's AddChild (:85-104).
class TwoSidedBSDF : public BSDF {
 public:
  explicit TwoSidedBSDF(const PropertyList &) {}
  ~TwoSidedBSDF() override {
    delete side_[0];
    delete side_[1];
  }
  void AddChild(Object *child, const std::string &) override {
    if (child->GetClassType() != EClassType::EBSDF) return;
    if (!side_[0]) side_[0] = static_cast<BSDF *>(child);
    else if (!side_[1]) side_[1] = static_cast<BSDF *>(child);
    else throw HikariException("TwoSidedBSDF: tried to specify the third nested BSDF");
  }
  void Activate() override {
    if (!side_[0]) throw HikariException("TwoSidedBSDF needs at least one nested BSDF!");
  }
  int AppendRecords(std::vector<float> &records) const override {
    Record r(kTwoSided);
    r.r[2] = float(side_[0]->AppendRecords(records));
    r.r[12] = side_[1] ? float(side_[1]->AppendRecords(records)) : -1.0f;
    for (int k : {int(r.r[2]), int(r.r[12])}) {  // :113-122
      if (k < 0) continue;
      int t = int(records[size_t(k) * HK_MATERIAL_STRIDE]);
      if (t == kCoating || t == kRoughCoating) t = int(records[size_t(records[size_t(k) * HK_MATERIAL_STRIDE + 2]) * HK_MATERIAL_STRIDE]);
      if (t == kDielectric || t == kRoughDielectric) throw HikariException("Only BSDF without transmission can be nested!");
    }
    int idx = int(records.size() / HK_MATERIAL_STRIDE);
    records.insert(records.end(), r.r, r.r + HK_MATERIAL_STRIDE);
    return idx;
  }
  std::string ToString() const override { return "TwoSidedBSDF[]"; }

 private:
  BSDF *side_[2] = {nullptr, nullptr};
};
Object *make_two_sided(const PropertyList &p) { return new TwoSidedBSDF(p); }

const bool kBsdfsRegistered = [] {
  ObjectFactory::RegisterClz("diffuse", make_diffuse);
  ObjectFactory::RegisterClz("conductor", make_conductor);
  ObjectFactory::RegisterClz("dielectric", make_dielectric);
  ObjectFactory::RegisterClz("plastic", make_plastic);
  ObjectFactory::RegisterClz("roughConductor", make_rough_conductor);
  ObjectFactory::RegisterClz("roughDielectric", make_rough_dielectric);
  ObjectFactory::RegisterClz("roughPlastic", make_rough_plastic);
  ObjectFactory::RegisterClz("coating", make_coating);
  ObjectFactory::RegisterClz("roughCoating", make_rough_coating);
  ObjectFactory::RegisterClz("mirror", make_mirror);
  ObjectFactory::RegisterClz("microfacet", make_microfacet);
  ObjectFactory::RegisterClz("roughDiffuse", make_rough_diffuse);
  ObjectFactory::RegisterClz("twoSided", make_two_sided);
  return true;
}();

}  // namespace

// ------------------------------------------------------------------ emitters
AreaLight::AreaLight(const PropertyList &p) { radiance = p.GetColor("radiance"); }
std::string AreaLight::ToString() const {
  char b[128];
  snprintf(b, sizeof b, "AreaLight[radiance = (%g, %g, %g)]", radiance.r, radiance.g, radiance.b);
  return b;
}
REGISTER_CLASS(AreaLight, "area")

EnvironmentLight::EnvironmentLight(const PropertyList &p) {
  filename = resolve(p.GetString("filename"));
  scale = p.GetFloat("scale", 1.0f);
  to_world = p.GetTransform("toWorld", Transform());
  texels = ReadImage(filename, width, height);
}
std::string EnvironmentLight::ToString() const {
  char b[256];
  snprintf(b, sizeof b, "EnvironmentLight[filename = \"%s\", %dx%d, scale = %g]", filename.c_str(), width, height,
           scale);
  return b;
}
REGISTER_CLASS(EnvironmentLight, "env")

// ------------------------------------------------------------------ OBJ mesh
namespace {

// src/mesh/WavefrontObjMesh.cpp:10-152: positions transformed by toWorld,
// normals by its inverse transpose and normalised, quads split as (0 1 2)
// (3 0 2), vertices deduplicated on their (position, uv, normal) index triple.
class WavefrontObjMesh : public Mesh {
 public:
  explicit WavefrontObjMesh(const PropertyList &p) {
    name = resolve(p.GetString("filename"));
    Transform T = p.GetTransform("toWorld", Transform());
    std::ifstream in(name);
    if (!in) throw HikariException("Unable to open OBJ file \"%s\"!", name.c_str());
    std::vector<Point3f> pos;
    std::vector<float> tex;
    std::vector<Vector3f> nrm;
    struct Key {
      uint32_t p, t, n;
      bool operator==(const Key &o) const { return p == o.p && t == o.t && n == o.n; }
    };
    struct KeyHash {
      size_t operator()(const Key &k) const { return (size_t(k.p) * 73856093u) ^ (size_t(k.t) * 19349663u) ^ (size_t(k.n) * 83492791u); }
    };
    std::unordered_map<Key, uint32_t, KeyHash> map;
    std::vector<Key> verts;
    auto parse_key = [&](const std::string &s) {
      Key k{0, 0, 0};
      size_t a = s.find('/');
      auto num = [&](const std::string &x) -> uint32_t {
        if (x.empty()) return 0;
        long v = std::stol(x);
        if (v <= 0) throw HikariException("Unsupported OBJ index \"%s\" in \"%s\"", x.c_str(), name.c_str());
        return uint32_t(v);
      };
      if (a == std::string::npos) {
        k.p = num(s);
      } else {
        k.p = num(s.substr(0, a));
        size_t b = s.find('/', a + 1);
        if (b == std::string::npos) {
          k.t = num(s.substr(a + 1));
        } else {
          k.t = num(s.substr(a + 1, b - a - 1));
          k.n = num(s.substr(b + 1));
        }
      }
      return k;
    };
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string pre;
      ls >> pre;
      if (pre == "v") {
        Point3f q;
        ls >> q.x >> q.y >> q.z;
        pos.push_back(T.ApplyPoint(q));
      } else if (pre == "vt") {
        float u = 0, v = 0;
        ls >> u >> v;
        tex.push_back(u);
        tex.push_back(v);
      } else if (pre == "vn") {
        Vector3f q;
        ls >> q.x >> q.y >> q.z;
        nrm.push_back(Normalized(T.ApplyNormal(q)));
      } else if (pre == "f") {
        std::string a, b, c, d;
        ls >> a >> b >> c >> d;
        Key k[6] = {parse_key(a), parse_key(b), parse_key(c)};
        int nv = 3;
        if (!d.empty()) {
          k[3] = parse_key(d);
          k[4] = k[0];
          k[5] = k[2];
          nv = 6;
        }
        for (int i = 0; i < nv; i++) {
          auto it = map.find(k[i]);
          if (it == map.end()) {
            map[k[i]] = uint32_t(verts.size());
            F.push_back(uint32_t(verts.size()));
            verts.push_back(k[i]);
          } else {
            F.push_back(it->second);
          }
        }
      }
    }
    for (const Key &k : verts) {
      if (k.p == 0 || k.p > pos.size()) throw HikariException("OBJ position index out of range in \"%s\"", name.c_str());
      V.push_back(pos[k.p - 1].x);
      V.push_back(pos[k.p - 1].y);
      V.push_back(pos[k.p - 1].z);
    }
    if (!nrm.empty())
      for (const Key &k : verts) {
        if (k.n == 0 || k.n > nrm.size()) throw HikariException("OBJ normal index out of range in \"%s\"", name.c_str());
        N.push_back(nrm[k.n - 1].x);
        N.push_back(nrm[k.n - 1].y);
        N.push_back(nrm[k.n - 1].z);
      }
    if (!tex.empty())
      for (const Key &k : verts) {
        if (k.t == 0 || 2 * k.t > tex.size()) throw HikariException("OBJ uv index out of range in \"%s\"", name.c_str());
        UV.push_back(tex[2 * (k.t - 1)]);
        UV.push_back(tex[2 * (k.t - 1) + 1]);
      }
  }
};

}  // namespace
REGISTER_CLASS(WavefrontObjMesh, "obj")

Mesh::~Mesh() {
  delete bsdf;
  delete emitter;
}
void Mesh::AddChild(Object *child, const std::string &cname) {
  switch (child->GetClassType()) {
    case EClassType::EBSDF:
      if (bsdf) throw HikariException("Mesh: tried to register multiple BSDF instances!");
      bsdf = static_cast<BSDF *>(child);
      break;
    case EClassType::EEmitter: {
      if (emitter) throw HikariException("Mesh: tried to register multiple Emitter instances!");
      Emitter *e = static_cast<Emitter *>(child);
      if (e->GetEmitterType() != Emitter::EEmitterType::EArea) {
        delete e;
        throw HikariException("Mesh: only area light can be attached!");
      }
      emitter = static_cast<AreaLight *>(e);
      break;
    }
    default:
      throw HikariException("Mesh::AddChild(<%s>, <%s>) is not supported!", ClassTypeName(child->GetClassType()).c_str(),
                            cname.c_str());
  }
}
void Mesh::Activate() {
  if (!bsdf) bsdf = static_cast<BSDF *>(ObjectFactory::CreateInstance("diffuse", PropertyList()));
}
std::string Mesh::ToString() const {
  char b[512];
  snprintf(b, sizeof b, "Mesh[name = \"%s\", vertexCount = %zu, triangleCount = %zu, emitter = %s]", name.c_str(),
           V.size() / 3, F.size() / 3, emitter ? emitter->ToString().c_str() : "null");
  return b;
}

// ------------------------------------------------------------------ filters
ReconstructionFilter::ReconstructionFilter(int k, const PropertyList &p) : kind(k) {
  if (kind == 1) {
    radius = p.GetFloat("radius", 2.0f);
    stddev = p.GetFloat("stddev", 0.5f);
  } else if (kind == 3) {
    radius = p.GetFloat("radius", 2.0f);
    b = p.GetFloat("B", 1.0f / 3.0f);
    c = p.GetFloat("C", 1.0f / 3.0f);
  } else if (kind == 0) {
    radius = 0.5f;
  } else {
    radius = 1.0f;
  }
}
std::string ReconstructionFilter::ToString() const {
  static const char *names[] = {"BoxFilter", "GaussianFilter", "TentFilter", "MitchellNetravaliFilter"};
  char buf[128];
  snprintf(buf, sizeof buf, "%s[radius = %g]", names[kind], radius);
  return buf;
}
namespace {
struct BoxFilterObj : ReconstructionFilter {
  explicit BoxFilterObj(const PropertyList &p) : ReconstructionFilter(0, p) {}
};
struct GaussianFilterObj : ReconstructionFilter {
  explicit GaussianFilterObj(const PropertyList &p) : ReconstructionFilter(1, p) {}
};
struct TentFilterObj : ReconstructionFilter {
  explicit TentFilterObj(const PropertyList &p) : ReconstructionFilter(2, p) {}
};
struct MitchellFilterObj : ReconstructionFilter {
  explicit MitchellFilterObj(const PropertyList &p) : ReconstructionFilter(3, p) {}
};
}  // namespace
REGISTER_CLASS(BoxFilterObj, "box")
REGISTER_CLASS(GaussianFilterObj, "gaussian")
REGISTER_CLASS(TentFilterObj, "tent")
REGISTER_CLASS(MitchellFilterObj, "mitchell")

// ------------------------------------------------------------------ camera
Camera::Camera(const PropertyList &p) {
  width = p.GetInteger("width", 1280);
  height = p.GetInteger("height", 720);
  fov = p.GetFloat("fov", 30.0f);
  near_clip = p.GetFloat("nearClip", 1e-4f);
  far_clip = p.GetFloat("farClip", 1e4f);
  to_world = p.GetTransform("toWorld", Transform());
  if (width <= 0 || height <= 0) throw HikariException("Camera: bad output size %dx%d", width, height);
}
Camera::~Camera() { delete filter; }
void Camera::AddChild(Object *child, const std::string &name) {
  if (child->GetClassType() != EClassType::EReconstructionFilter)
    throw HikariException("Camera::AddChild(<%s>, <%s>) is not supported!", ClassTypeName(child->GetClassType()).c_str(),
                          name.c_str());
  if (filter) throw HikariException("Camera: tried to register multiple reconstruction filters!");
  filter = static_cast<ReconstructionFilter *>(child);
}
void Camera::Activate() {
  if (!filter) filter = static_cast<ReconstructionFilter *>(ObjectFactory::CreateInstance("gaussian", PropertyList()));
}
std::string Camera::ToString() const {
  char b[256];
  snprintf(b, sizeof b, "PerspectiveCamera[outputSize = %dx%d, fov = %g, clip = [%g, %g]]", width, height, fov,
           near_clip, far_clip);
  return b;
}
REGISTER_CLASS(Camera, "perspective")

// ------------------------------------------------------------------ sampler
Sampler::Sampler(const PropertyList &p) { sample_count = p.GetInteger("sampleCount", 1); }
std::string Sampler::ToString() const { return "Independent[sampleCount = " + std::to_string(sample_count) + "]"; }
REGISTER_CLASS(Sampler, "independent")

// ------------------------------------------------------------------ scene
Scene::Scene(const PropertyList &p) {
  background_ = p.GetColor("background", Color3f(0.0f));
  force_background_ = p.GetBoolean("forceBackground", false);
}
Scene::~Scene() {
  for (Mesh *m : meshes_) delete m;
  delete env_;
  delete camera_;
  delete integrator_;
  delete sampler_;
  delete accel_;
}
void Scene::AddChild(Object *child, const std::string &name) {
  switch (child->GetClassType()) {
    case EClassType::EMesh: meshes_.push_back(static_cast<Mesh *>(child)); break;
    case EClassType::EEmitter: {
      Emitter *e = static_cast<Emitter *>(child);
      if (e->GetEmitterType() != Emitter::EEmitterType::EEnvironment) {
        delete e;
        throw HikariException("Scene: only the environment light can be a scene-level emitter in this build");
      }
      if (env_) throw HikariException("Scene: tried to register multiple environment emitters!");
      env_ = static_cast<EnvironmentLight *>(e);
      break;
    }
    case EClassType::EAcceleration:
      if (accel_) throw HikariException("There can only be one acceleration per scene!");
      accel_ = static_cast<Acceleration *>(child);
      break;
    case EClassType::ESampler:
      if (sampler_) throw HikariException("There can only be one sampler per scene!");
      sampler_ = static_cast<Sampler *>(child);
      break;
    case EClassType::ECamera:
      if (camera_) throw HikariException("There can only be one camera per scene!");
      camera_ = static_cast<Camera *>(child);
      break;
    case EClassType::EIntegrator:
      if (integrator_) throw HikariException("There can only be one integrator per scene!");
      integrator_ = static_cast<Integrator *>(child);
      break;
    default:
      throw HikariException("Scene::AddChild(<%s>, <%s>) is not supported!", ClassTypeName(child->GetClassType()).c_str(),
                            name.c_str());
  }
}
// src/core/Scene.cpp:128-165
void Scene::Activate() {
  if (!accel_) accel_ = static_cast<Acceleration *>(ObjectFactory::CreateInstance("bruto", PropertyList()));
  for (Mesh *m : meshes_) accel_->AddMesh(m);
  accel_->Build();
  if (!integrator_) throw HikariException("No integrator was specified!");
  if (!camera_) throw HikariException("No camera was specified!");
  if (!sampler_) sampler_ = static_cast<Sampler *>(ObjectFactory::CreateInstance("independent", PropertyList()));
}
std::string Scene::ToString() const {
  std::string s = "Scene[\n";
  for (Mesh *m : meshes_) s += "  " + m->ToString() + "\n";
  if (env_) s += "  " + env_->ToString() + "\n";
  if (camera_) s += "  " + camera_->ToString() + "\n";
  if (sampler_) s += "  " + sampler_->ToString() + "\n";
  return s + "]";
}

hk_scene_desc Scene::Describe(int max_depth, int estimator, std::vector<float> &records,
                              std::vector<hk_mesh_desc> &mdesc) const {
  hk_scene_desc d;
  std::memset(&d, 0, sizeof d);
  records.clear();
  mdesc.clear();
  for (Mesh *m : meshes_) {
    hk_mesh_desc md;
    std::memset(&md, 0, sizeof md);
    md.v = m->V.data();
    md.nv = uint32_t(m->V.size() / 3);
    md.n = m->N.empty() ? nullptr : m->N.data();
    md.uv = m->UV.empty() ? nullptr : m->UV.data();
    md.f = m->F.data();
    md.nt = uint32_t(m->F.size() / 3);
    md.material = m->bsdf->AppendRecords(records);
    if (m->emitter) {
      md.is_emitter = 1;
      md.radiance[0] = m->emitter->radiance.r;
      md.radiance[1] = m->emitter->radiance.g;
      md.radiance[2] = m->emitter->radiance.b;
    }
    mdesc.push_back(md);
  }
  d.meshes = mdesc.data();
  d.n_meshes = int32_t(mdesc.size());
  d.materials = records.data();
  d.n_materials = int32_t(records.size() / HK_MATERIAL_STRIDE);
  if (env_) {
    d.env = env_->texels.data();
    d.env_w = env_->width;
    d.env_h = env_->height;
    d.env_scale = env_->scale;
    env_->to_world.ToFloat(d.env_to_world);
  } else {
    Transform().ToFloat(d.env_to_world);
  }
  camera_->to_world.ToFloat(d.cam_to_world);
  d.fov = camera_->fov;
  d.near_clip = camera_->near_clip;
  d.far_clip = camera_->far_clip;
  d.width = camera_->width;
  d.height = camera_->height;
  const ReconstructionFilter *f = camera_->filter;
  d.filter = f->kind;
  d.filter_radius = f->radius;
  d.filter_stddev = f->stddev;
  d.filter_b = f->b;
  d.filter_c = f->c;
  d.max_depth = max_depth;
  d.background[0] = background_.r;
  d.background[1] = background_.g;
  d.background[2] = background_.b;
  d.force_background = force_background_ ? 1 : 0;
  d.rft_dir = nullptr;
  d.integrator = estimator;
  return d;
}
REGISTER_CLASS(Scene, "scene")

}  // namespace hikari

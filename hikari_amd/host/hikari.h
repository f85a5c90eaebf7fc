// hikari.h -- host-side plugin surface of hikari_amd.
//
// Mirrors the object model the reference's scene files are written against
// (include/core/Object.hpp, PropertyList.hpp, Parser.hpp, Scene.hpp,
// Integrator.hpp, Acceleration.hpp): objects are created by name through
// ObjectFactory from a PropertyList, receive their children through
// AddChild / SetParent and are finished by Activate().  The scene classes
// here hold plain data (triangles, material records, camera, emitters) and
// hand it to the GPU path through the C ABI of include/hikari_amd.h; the
// integrator `gpuPathMIS` (also registered as `pathMIS`) renders whole images
// on the device, and the accelerator `gpuBVH` (also `bvh`, `hlbvh`, `bruto`)
// answers ray queries on the device.
#pragma once
#include <cstdarg>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hikari_amd.h"

namespace hikari {

class HikariException : public std::runtime_error {
 public:
  explicit HikariException(const char *fmt, ...) __attribute__((format(printf, 2, 3)));
};

struct Color3f {
  float r = 0, g = 0, b = 0;
  Color3f() {}
  explicit Color3f(float v) : r(v), g(v), b(v) {}
  Color3f(float r_, float g_, float b_) : r(r_), g(g_), b(b_) {}
  float GetLuminance() const { return r * 0.212671f + g * 0.715160f + b * 0.072169f; }
};
struct Vector3f {
  float x = 0, y = 0, z = 0;
  Vector3f() {}
  Vector3f(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
};
using Point3f = Vector3f;

// 4x4 affine transform in FP32 (row-major, column vectors) with its inverse,
// the reference's Transform over Eigen::Matrix4f (include/core/
// Transform.hpp:16-57).  The XML ops premultiply the transform being built
// with Eigen::Affine3f's arithmetic (src/core/Parser.cpp:308-369); products
// and sums round in the order Eigen 3.3.5 evaluates them (a 3-term dot
// product as a0 + (a1 + a2), the 4x4 point product left to right).
struct Transform {
  float m[4][4];
  float inv[4][4];  // inverse: FP64 Gauss-Jordan of m, rounded to FP32
  Transform();      // identity
  // the XML ops (Parser.cpp:312-369), each premultiplied onto *this
  void PreTranslate(const Vector3f &v);
  void PreScale(const Vector3f &v);
  void PreRotate(float angle_deg, const Vector3f &axis);  // axis used as given (AngleAxis)
  void PreAffine(const float *rows16);                     // bottom row ignored, as Affine3f does
  void PreLookAt(const Point3f &origin, const Point3f &target, const Vector3f &up);
  Point3f ApplyPoint(const Point3f &p) const;     // Transform.cpp:46-50
  Vector3f ApplyVector(const Vector3f &v) const;  // Transform.cpp:36-39
  Vector3f ApplyNormal(const Vector3f &n) const;  // Transform.cpp:41-44 (inverse transpose, not normalised)
  void ToFloat(float *out16) const;
  void UpdateInverse();
};
// Eigen's normalized() in FP32 (Dot.h:124-134: divide by the square root of
// x*x + (y*y + z*z) when that is positive)
Vector3f Normalized(const Vector3f &v);

// Typed property bag (src/core/PropertyList.cpp): Get(name) throws when the
// property is missing or has another type; Get(name, default) returns the
// default when missing.
class PropertyList {
 public:
#define HIKARI_PROPERTY(CppType, Name)                                              \
  void Set##Name(const std::string &key, const CppType &value);                     \
  CppType Get##Name(const std::string &key) const;                                  \
  CppType Get##Name(const std::string &key, const CppType &default_value) const;
  HIKARI_PROPERTY(bool, Boolean)
  HIKARI_PROPERTY(int, Integer)
  HIKARI_PROPERTY(float, Float)
  HIKARI_PROPERTY(std::string, String)
  HIKARI_PROPERTY(Color3f, Color)
  HIKARI_PROPERTY(Point3f, Point)
  HIKARI_PROPERTY(Vector3f, Vector)
  HIKARI_PROPERTY(Transform, Transform)
#undef HIKARI_PROPERTY
  bool Has(const std::string &key) const { return props_.count(key) != 0; }

 private:
  enum Type { kBoolean, kInteger, kFloat, kString, kColor, kPoint, kVector, kTransform };
  struct Prop {
    Type type;
    bool b = false;
    int i = 0;
    float f = 0;
    std::string s;
    Color3f c;
    Vector3f v;
    Transform t;
  };
  std::map<std::string, Prop> props_;
  const Prop &get(const std::string &key, Type type, const char *type_name) const;
  const Prop *find(const std::string &key, Type type, const char *type_name) const;
  void set(const std::string &key, Prop p);
};

class Object {
 public:
  enum class EClassType {
    EScene = 0, EMesh, EBSDF, EPhaseFunction, EEmitter, ETexture, EMedium, ECamera, EIntegrator, ESampler, ETest,
    EReconstructionFilter, EAcceleration, EShape, EClassTypeCount
  };
  virtual ~Object();
  virtual EClassType GetClassType() const = 0;
  virtual void AddChild(Object *child, const std::string &name);
  virtual void SetParent(Object *parent, const std::string &name);
  virtual void Activate();
  virtual std::string ToString() const;
  static std::string ClassTypeName(EClassType t);
};

class ObjectFactory {
 public:
  using Constructor = std::function<Object *(const PropertyList &)>;
  static void RegisterClz(const std::string &name, const Constructor &ctor);
  static Object *CreateInstance(const std::string &name, const PropertyList &props);
  static bool Has(const std::string &name);
};

#define HIKARI_CAT2(a, b) a##b
#define HIKARI_CAT(a, b) HIKARI_CAT2(a, b)
// Registers `Class` (constructor Class(const PropertyList &)) under the XML
// type name `Name` at static-initialisation time.
#define REGISTER_CLASS(Class, Name)                                                                    \
  static const bool HIKARI_CAT(hikari_registered_, __LINE__) = [] {                                    \
    ::hikari::ObjectFactory::RegisterClz(Name, [](const ::hikari::PropertyList &p) -> ::hikari::Object * { \
      return new Class(p);                                                                             \
    });                                                                                                \
    return true;                                                                                       \
  }();

// Reads a scene / test description (src/core/Parser.cpp semantics): returns
// the root object, owned by the caller.  Relative file names in the document
// are resolved against the document's directory.
Object *LoadFromXML(const std::string &filename);
// Directory used to resolve relative file names while a document is loaded.
const std::string &CurrentSceneDirectory();

// ------------------------------------------------------------------ scene classes
class BSDF : public Object {
 public:
  EClassType GetClassType() const override { return EClassType::EBSDF; }
  // Appends this BSDF's material record(s) (HK_MATERIAL_STRIDE floats each,
  // nested BSDFs first) to `records`; returns the index of this BSDF's record.
  virtual int AppendRecords(std::vector<float> &records) const = 0;
};

class Emitter : public Object {
 public:
  enum class EEmitterType { EArea, EEnvironment };
  EClassType GetClassType() const override { return EClassType::EEmitter; }
  virtual EEmitterType GetEmitterType() const = 0;
};

class AreaLight : public Emitter {
 public:
  explicit AreaLight(const PropertyList &props);
  EEmitterType GetEmitterType() const override { return EEmitterType::EArea; }
  std::string ToString() const override;
  Color3f radiance;
};

class EnvironmentLight : public Emitter {
 public:
  explicit EnvironmentLight(const PropertyList &props);
  EEmitterType GetEmitterType() const override { return EEmitterType::EEnvironment; }
  std::string ToString() const override;
  std::vector<float> texels;  // height x width x 3 (RGB)
  int width = 0, height = 0;
  float scale = 1.0f;
  Transform to_world;
  std::string filename;
};

class Mesh : public Object {
 public:
  EClassType GetClassType() const override { return EClassType::EMesh; }
  void AddChild(Object *child, const std::string &name) override;
  void Activate() override;
  std::string ToString() const override;
  ~Mesh() override;
  std::vector<float> V, N, UV;  // 3 / 3 / 2 floats per vertex
  std::vector<uint32_t> F;      // 3 indices per triangle
  BSDF *bsdf = nullptr;
  AreaLight *emitter = nullptr;
  std::string name;
};

class ReconstructionFilter : public Object {
 public:
  explicit ReconstructionFilter(int kind, const PropertyList &props);
  EClassType GetClassType() const override { return EClassType::EReconstructionFilter; }
  std::string ToString() const override;
  int kind;  // HK filter ids: 0 box, 1 gaussian, 2 tent, 3 mitchell
  float radius = 2.0f, stddev = 0.5f, b = 1.0f / 3.0f, c = 1.0f / 3.0f;
};

class Camera : public Object {
 public:
  explicit Camera(const PropertyList &props);
  ~Camera() override;
  EClassType GetClassType() const override { return EClassType::ECamera; }
  void AddChild(Object *child, const std::string &name) override;
  void Activate() override;
  std::string ToString() const override;
  int width, height;
  float fov, near_clip, far_clip;
  Transform to_world;
  ReconstructionFilter *filter = nullptr;
};

class Sampler : public Object {
 public:
  explicit Sampler(const PropertyList &props);
  EClassType GetClassType() const override { return EClassType::ESampler; }
  std::string ToString() const override;
  int GetSampleCount() const { return sample_count; }
  int sample_count;
};

class Scene;

class Acceleration : public Object {
 public:
  EClassType GetClassType() const override { return EClassType::EAcceleration; }
  void AddMesh(Mesh *mesh) { meshes_.push_back(mesh); }
  virtual void Build() = 0;
  // Closest hit (or any hit for shadow rays) of one ray: t, barycentric
  // (u, v), the mesh and the facet index inside it.
  struct Hit {
    float t = 0, u = 0, v = 0;
    const Mesh *mesh = nullptr;
    uint32_t facet = 0;
  };
  virtual bool RayIntersect(const float o[3], const float d[3], float tmin, float tmax, Hit &hit,
                            bool shadow) const = 0;

 protected:
  std::vector<Mesh *> meshes_;
};

class Integrator : public Object {
 public:
  EClassType GetClassType() const override { return EClassType::EIntegrator; }
  // estimator id of hk_scene_desc.integrator (0 pathMIS, 1 pathEMS, 2 pathMATS)
  virtual int EstimatorId() const { return 0; }
  virtual void Preprocess(const Scene *scene);
  // Whole image (height x width x 3, filtered and normalised) at the scene
  // sampler's sample count.
  virtual void RenderImage(const Scene *scene, std::vector<float> &rgb) = 0;
};

class Scene : public Object {
 public:
  explicit Scene(const PropertyList &props);
  ~Scene() override;
  EClassType GetClassType() const override { return EClassType::EScene; }
  void AddChild(Object *child, const std::string &name) override;
  void Activate() override;
  std::string ToString() const override;
  const std::vector<Mesh *> &GetMeshes() const { return meshes_; }
  const EnvironmentLight *GetEnvironmentEmitter() const { return env_; }
  const Camera *GetCamera() const { return camera_; }
  Integrator *GetIntegrator() const { return integrator_; }
  const Sampler *GetSampler() const { return sampler_; }
  Acceleration *GetAccel() const { return accel_; }
  Color3f GetBackground() const { return background_; }
  bool GetForceBackground() const { return force_background_; }
  // Flat description for hk_scene_create; the arrays stay owned by the
  // scene objects and `records` (caller storage) and must outlive the call.
  hk_scene_desc Describe(int max_depth, int estimator, std::vector<float> &records,
                         std::vector<hk_mesh_desc> &meshes) const;

 private:
  std::vector<Mesh *> meshes_;
  EnvironmentLight *env_ = nullptr;
  Camera *camera_ = nullptr;
  Integrator *integrator_ = nullptr;
  Sampler *sampler_ = nullptr;
  Acceleration *accel_ = nullptr;
  Color3f background_;
  bool force_background_ = false;
};

// Process-wide device context of the host plugins (device 0), created on
// first use; throws when no device / library is usable.
hk_ctx *GpuContext();
void CheckHk(int rc, const char *what);

// Image files (PFM: little-endian RGB32F, bottom-to-top rows).
void WritePFM(const std::string &path, const float *rgb, int width, int height);
std::vector<float> ReadPFM(const std::string &path, int &width, int &height);
// OpenEXR (scanline, NONE / ZIPS / ZIP compression, HALF or FLOAT channels).
std::vector<float> ReadEXR(const std::string &path, int &width, int &height);
void WriteEXR(const std::string &path, const float *rgb, int width, int height);
// By extension: .exr or .pfm.
std::vector<float> ReadImage(const std::string &path, int &width, int &height);

// Student's t-test harness of the reference (src/core/StudentTTest.cpp):
// registered as "ttest"; Activate() runs the tests and logs the outcome.
// A test description (<test type="...">): runs at load time (Activate) and
// keeps its tally for the driver's exit status.
class TestObject : public Object {
 public:
  EClassType GetClassType() const override { return EClassType::ETest; }
  virtual const char *TestKind() const = 0;
  int passed = 0, total = 0;
};

class StudentTTest : public TestObject {
 public:
  explicit StudentTTest(const PropertyList &props);
  ~StudentTTest() override;
  const char *TestKind() const override { return "ttest"; }
  void AddChild(Object *child, const std::string &name) override;
  void Activate() override;
  std::string ToString() const override;
  const std::vector<Scene *> &GetScenes() const { return scenes_; }
  const std::vector<float> &GetReferences() const { return references_; }

 private:
  float significance_;
  std::vector<float> angles_, references_;
  int sample_count_;
  std::vector<BSDF *> bsdfs_;
  std::vector<Scene *> scenes_;
};

// Dry-run mode: tests parse and validate but do not run (scene dumps).
void SetDryRun(bool dry);
bool DryRun();

// Two-sided Student's t-test at `significance`, Sidak-corrected for
// `n_tests` simultaneous tests: true when the sample mean is compatible with
// the reference.  `msg` receives a one-line report.
bool StudentsTTest(double mean, double variance, double reference, long n, double significance, int n_tests,
                   std::string &msg);

// pcg32 (XSH-RR 64 -> 32, O'Neill's PCG family) with its default state and
// stream; NextFloat fills the mantissa of [1, 2) (ext/pcg32/pcg32.h).
struct HostPcg32 {
  uint64_t state = 0x853c49e6748fea9bull, inc = 0xda3e39cb94b95bdbull;
  uint32_t NextUint() {
    uint64_t old = state;
    state = old * 0x5851f42d4c957f2dull + inc;
    uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
  }
  float NextFloat() {
    uint32_t u = (NextUint() >> 9) | 0x3f800000u;
    float f;
    std::memcpy(&f, &u, 4);
    return f - 1.0f;
  }
};

// Chi-square distribution function with `dof` degrees of freedom at x.
double Chi2Cdf(double x, int dof);
// Pearson's chi^2 test of observed against expected cell frequencies, cells
// below min_exp pooled, at `significance` Sidak-corrected for n_tests
// (ext/hypothesis/hypothesis.h:140-240 semantics).
bool PooledChi2Test(int n_cells, const double *observed, const double *expected, long sample_count, double min_exp,
                    double significance, int n_tests, std::string &msg);

}  // namespace hikari

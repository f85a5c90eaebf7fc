// core.cpp -- exceptions, transforms, PropertyList, Object / ObjectFactory and
// the XML scene loader.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <set>
#include <sstream>

#include "hikari.h"
#include "xml.h"

namespace hikari {

static std::string vformat(const char *fmt, va_list ap) {
  char buf[2048];
  vsnprintf(buf, sizeof buf, fmt, ap);
  return buf;
}

HikariException::HikariException(const char *fmt, ...) : std::runtime_error("") {
  va_list ap;
  va_start(ap, fmt);
  std::string msg = vformat(fmt, ap);
  va_end(ap);
  static_cast<std::runtime_error &>(*this) = std::runtime_error(msg);
}

// ------------------------------------------------------------------ Transform
// FP32 arithmetic in Eigen 3.3.5's evaluation order, pinned bit for bit
// against the reference's vendored Eigen (oracle/_ref/eigen_xform): a dot
// product of three terms reduces as a0 + (a1 + a2) (Core/Redux.h
// redux_novec_unroller), Matrix4f * Vector4f accumulates column by column
// (((c0 x + c1 y) + c2 z) + c3 w, the packet product of
// Core/ProductEvaluators.h).  The host build turns FP contraction off.
namespace {
inline float dot3(float a0, float b0, float a1, float b1, float a2, float b2) { return a0 * b0 + (a1 * b1 + a2 * b2); }
Vector3f cross3(const Vector3f &a, const Vector3f &b) {  // Geometry/OrthoMethods.h
  return Vector3f(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
}  // namespace

Vector3f Normalized(const Vector3f &v) {
  float z = v.x * v.x + (v.y * v.y + v.z * v.z);
  if (!(z > 0.0f)) return v;
  float s = std::sqrt(z);
  return Vector3f(v.x / s, v.y / s, v.z / s);
}

Transform::Transform() {
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) m[i][j] = inv[i][j] = i == j ? 1.0f : 0.0f;
}
// Translation * Affine3f: the translation column gains v (Translation.h:125-129)
void Transform::PreTranslate(const Vector3f &v) {
  m[0][3] = m[0][3] + v.x;
  m[1][3] = m[1][3] + v.y;
  m[2][3] = m[2][3] + v.z;
  UpdateInverse();
}
// DiagonalMatrix * Affine3f: row i of the top three rows times v_i
// (Transform.h:478-486)
void Transform::PreScale(const Vector3f &v) {
  const float s[3] = {v.x, v.y, v.z};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 4; j++) m[i][j] = s[i] * m[i][j];
  UpdateInverse();
}
// AngleAxis<float>(DegToRad(angle), axis) * Affine3f: the rotation matrix of
// AngleAxis::toRotationMatrix (AngleAxis.h:218-243, the axis not normalised)
// times the top three rows (Transform.h:1459-1472)
void Transform::PreRotate(float angle_deg, const Vector3f &axis) {
  float ang = angle_deg * (float(M_PI) / 180.0f);  // DegToRad, include/core/Common.hpp:742
  const float a[3] = {axis.x, axis.y, axis.z};
  float s = std::sin(ang), c = std::cos(ang);
  float sa[3], ca[3];
  for (int i = 0; i < 3; i++) {
    sa[i] = s * a[i];
    ca[i] = (1.0f - c) * a[i];
  }
  float r[3][3], t;
  t = ca[0] * a[1]; r[0][1] = t - sa[2]; r[1][0] = t + sa[2];
  t = ca[0] * a[2]; r[0][2] = t + sa[1]; r[2][0] = t - sa[1];
  t = ca[1] * a[2]; r[1][2] = t - sa[0]; r[2][1] = t + sa[0];
  for (int i = 0; i < 3; i++) r[i][i] = ca[i] * a[i] + c;
  float o[3][4];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 4; j++) o[i][j] = dot3(r[i][0], m[0][j], r[i][1], m[1][j], r[i][2], m[2][j]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 4; j++) m[i][j] = o[i][j];
  UpdateInverse();
}
// Affine3f(A) * Affine3f: linear = A.linear * linear, translation =
// A.linear * translation + A.translation, bottom row (0, 0, 0, 1)
// (Transform.h:1481-1495)
void Transform::PreAffine(const float *a) {
  float o[3][4];
  for (int i = 0; i < 3; i++) {
    const float *ar = a + 4 * i;
    for (int j = 0; j < 3; j++) o[i][j] = dot3(ar[0], m[0][j], ar[1], m[1][j], ar[2], m[2][j]);
    o[i][3] = dot3(ar[0], m[0][3], ar[1], m[1][3], ar[2], m[2][3]) + ar[3];
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 4; j++) m[i][j] = o[i][j];
  m[3][0] = m[3][1] = m[3][2] = 0.0f;
  m[3][3] = 1.0f;
  UpdateInverse();
}
// Parser.cpp:356-367: columns left, new up, direction, origin
void Transform::PreLookAt(const Point3f &origin, const Point3f &target, const Vector3f &up) {
  Vector3f dir = Normalized(Vector3f(target.x - origin.x, target.y - origin.y, target.z - origin.z));
  Vector3f left = Normalized(cross3(Normalized(up), dir));
  Vector3f nup = Normalized(cross3(dir, left));
  const float a[16] = {left.x, nup.x, dir.x, origin.x, left.y, nup.y, dir.y, origin.y,
                       left.z, nup.z, dir.z, origin.z, 0.0f,   0.0f,  0.0f,  1.0f};
  PreAffine(a);
}
// The reference inverts with Eigen's FP32 4x4 routine (an SSE cofactor
// scheme); this one inverts in FP64 and rounds, which agrees with it to a
// few ulp.  Only normals use the inverse.
void Transform::UpdateInverse() {
  double a[4][8];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      a[i][j] = m[i][j];
      a[i][j + 4] = i == j;
    }
  for (int c = 0; c < 4; c++) {
    int p = c;
    for (int r = c + 1; r < 4; r++)
      if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
    if (a[p][c] == 0.0) throw HikariException("singular transform");
    for (int j = 0; j < 8; j++) std::swap(a[c][j], a[p][j]);
    double d = a[c][c];
    for (int j = 0; j < 8; j++) a[c][j] /= d;
    for (int r = 0; r < 4; r++)
      if (r != c) {
        double f = a[r][c];
        for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j];
      }
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) inv[i][j] = float(a[i][j + 4]);
}
Point3f Transform::ApplyPoint(const Point3f &p) const {
  float r[4];
  for (int i = 0; i < 4; i++) r[i] = ((m[i][0] * p.x + m[i][1] * p.y) + m[i][2] * p.z) + m[i][3] * 1.0f;
  return Point3f(r[0] / r[3], r[1] / r[3], r[2] / r[3]);
}
Vector3f Transform::ApplyVector(const Vector3f &v) const {
  return Vector3f(dot3(m[0][0], v.x, m[0][1], v.y, m[0][2], v.z), dot3(m[1][0], v.x, m[1][1], v.y, m[1][2], v.z),
                  dot3(m[2][0], v.x, m[2][1], v.y, m[2][2], v.z));
}
Vector3f Transform::ApplyNormal(const Vector3f &n) const {
  return Vector3f(dot3(inv[0][0], n.x, inv[1][0], n.y, inv[2][0], n.z),
                  dot3(inv[0][1], n.x, inv[1][1], n.y, inv[2][1], n.z),
                  dot3(inv[0][2], n.x, inv[1][2], n.y, inv[2][2], n.z));
}
void Transform::ToFloat(float *out) const {
  for (int i = 0; i < 16; i++) out[i] = m[i / 4][i % 4];
}

// ------------------------------------------------------------------ PropertyList
void PropertyList::set(const std::string &key, Prop p) {
  if (props_.count(key)) fprintf(stderr, "Property \"%s\" was specified multiple times!\n", key.c_str());
  props_[key] = p;
}
const PropertyList::Prop *PropertyList::find(const std::string &key, Type type, const char *type_name) const {
  auto it = props_.find(key);
  if (it == props_.end()) return nullptr;
  if (it->second.type != type)
    throw HikariException("Property '%s' has the wrong type! (expected <%s>)!", key.c_str(), type_name);
  return &it->second;
}
const PropertyList::Prop &PropertyList::get(const std::string &key, Type type, const char *type_name) const {
  const Prop *p = find(key, type, type_name);
  if (!p) throw HikariException("Property '%s' is missing!", key.c_str());
  return *p;
}

#define HIKARI_PROPERTY_IMPL(CppType, Name, T, field, tname)                                          \
  void PropertyList::Set##Name(const std::string &key, const CppType &value) {                        \
    Prop p;                                                                                           \
    p.type = T;                                                                                       \
    p.field = value;                                                                                  \
    set(key, p);                                                                                      \
  }                                                                                                   \
  CppType PropertyList::Get##Name(const std::string &key) const { return get(key, T, tname).field; } \
  CppType PropertyList::Get##Name(const std::string &key, const CppType &dv) const {                  \
    const Prop *p = find(key, T, tname);                                                              \
    return p ? p->field : dv;                                                                         \
  }
HIKARI_PROPERTY_IMPL(bool, Boolean, kBoolean, b, "boolean")
HIKARI_PROPERTY_IMPL(int, Integer, kInteger, i, "integer")
HIKARI_PROPERTY_IMPL(float, Float, kFloat, f, "float")
HIKARI_PROPERTY_IMPL(std::string, String, kString, s, "string")
HIKARI_PROPERTY_IMPL(Color3f, Color, kColor, c, "color")
HIKARI_PROPERTY_IMPL(Point3f, Point, kPoint, v, "point")
HIKARI_PROPERTY_IMPL(Vector3f, Vector, kVector, v, "vector")
HIKARI_PROPERTY_IMPL(Transform, Transform, kTransform, t, "transform")
#undef HIKARI_PROPERTY_IMPL

// ------------------------------------------------------------------ Object
Object::~Object() {}
void Object::AddChild(Object *, const std::string &) {
  throw HikariException("Object::AddChild() is not implemented for objects of type '%s'!",
                        ClassTypeName(GetClassType()).c_str());
}
void Object::SetParent(Object *, const std::string &) {}
void Object::Activate() {}
std::string Object::ToString() const { return ClassTypeName(GetClassType()) + "[]"; }
std::string Object::ClassTypeName(EClassType t) {
  static const char *names[] = {"Scene", "Mesh", "BSDF", "PhaseFunction", "Emitter", "Texture", "Medium",
                                "Camera", "Integrator", "Sampler", "Test", "ReconstructionFilter", "Acceleration",
                                "Shape"};
  int i = int(t);
  return i >= 0 && i < int(EClassType::EClassTypeCount) ? names[i] : "<Unknown>";
}

static std::map<std::string, ObjectFactory::Constructor> &registry() {
  static std::map<std::string, ObjectFactory::Constructor> r;
  return r;
}
void ObjectFactory::RegisterClz(const std::string &name, const Constructor &ctor) {
  if (registry().count(name)) {
    fprintf(stderr, "This registration was aborted because class '%s' has been registered before!\n", name.c_str());
    return;
  }
  registry()[name] = ctor;
}
Object *ObjectFactory::CreateInstance(const std::string &name, const PropertyList &props) {
  auto it = registry().find(name);
  if (it == registry().end()) throw HikariException("A constructor for class \"%s\" could not be found!", name.c_str());
  return it->second(props);
}
bool ObjectFactory::Has(const std::string &name) { return registry().count(name) != 0; }

// ------------------------------------------------------------------ XML loader
namespace {

enum class Tag {
  kObject,  // any object class (the class type is in `cls`)
  kBoolean, kInteger, kFloat, kString, kPoint, kVector, kColor, kTransform,
  kTranslate, kMatrix, kRotate, kScale, kLookAt
};

struct TagInfo {
  Tag tag;
  Object::EClassType cls;
};

const std::map<std::string, TagInfo> &tag_map() {
  using C = Object::EClassType;
  static const std::map<std::string, TagInfo> m = {
      {"scene", {Tag::kObject, C::EScene}},         {"mesh", {Tag::kObject, C::EMesh}},
      {"bsdf", {Tag::kObject, C::EBSDF}},           {"phase", {Tag::kObject, C::EPhaseFunction}},
      {"emitter", {Tag::kObject, C::EEmitter}},     {"texture", {Tag::kObject, C::ETexture}},
      {"medium", {Tag::kObject, C::EMedium}},       {"camera", {Tag::kObject, C::ECamera}},
      {"integrator", {Tag::kObject, C::EIntegrator}}, {"sampler", {Tag::kObject, C::ESampler}},
      {"test", {Tag::kObject, C::ETest}},           {"rfilter", {Tag::kObject, C::EReconstructionFilter}},
      {"acceleration", {Tag::kObject, C::EAcceleration}}, {"shape", {Tag::kObject, C::EShape}},
      {"boolean", {Tag::kBoolean, C::EClassTypeCount}}, {"integer", {Tag::kInteger, C::EClassTypeCount}},
      {"float", {Tag::kFloat, C::EClassTypeCount}},     {"string", {Tag::kString, C::EClassTypeCount}},
      {"point", {Tag::kPoint, C::EClassTypeCount}},     {"vector", {Tag::kVector, C::EClassTypeCount}},
      {"color", {Tag::kColor, C::EClassTypeCount}},     {"transform", {Tag::kTransform, C::EClassTypeCount}},
      {"translate", {Tag::kTranslate, C::EClassTypeCount}}, {"matrix", {Tag::kMatrix, C::EClassTypeCount}},
      {"rotate", {Tag::kRotate, C::EClassTypeCount}},   {"scale", {Tag::kScale, C::EClassTypeCount}},
      {"lookat", {Tag::kLookAt, C::EClassTypeCount}},
  };
  return m;
}

std::vector<std::string> tokenize(const std::string &s) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == ',' || ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur += ch;
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

float to_float(const std::string &s) {
  size_t pos = 0;
  float v;
  try {
    v = std::stof(s, &pos);
  } catch (...) {
    throw HikariException("Could not parse floating point value \"%s\"", s.c_str());
  }
  if (pos != s.size()) throw HikariException("Could not parse floating point value \"%s\"", s.c_str());
  return v;
}
int to_int(const std::string &s) {
  size_t pos = 0;
  long v;
  try {
    v = std::stol(s, &pos);
  } catch (...) {
    throw HikariException("Could not parse integer value \"%s\"", s.c_str());
  }
  if (pos != s.size()) throw HikariException("Could not parse integer value \"%s\"", s.c_str());
  return int(v);
}
bool to_bool(const std::string &s) {
  std::string l;
  for (char c : s) l += char(std::tolower((unsigned char)c));
  if (l == "true") return true;
  if (l == "false") return false;
  throw HikariException("Could not parse boolean value \"%s\"", s.c_str());
}
Vector3f to_vec3(const std::string &s) {
  std::vector<std::string> t = tokenize(s);
  if (t.size() != 3) throw HikariException("Expected 3 values");
  return Vector3f(to_float(t[0]), to_float(t[1]), to_float(t[2]));
}

struct Loader {
  std::string file;
  Transform trans;  // the transform being built (document order, each op premultiplied)

  std::string where(const XmlNode &n) const {
    char buf[64];
    snprintf(buf, sizeof buf, "line %d, col %d", n.line, n.col);
    return buf;
  }
  void check_attrs(const XmlNode &n, std::set<std::string> want, bool exact) const {
    for (auto &a : n.attrs) {
      auto it = want.find(a.first);
      if (it == want.end())
        throw HikariException("Unexpected attribute \"%s\" in \"%s\" at %s", a.first.c_str(), n.name.c_str(),
                              where(n).c_str());
      want.erase(it);
    }
    if (exact && !want.empty())
      throw HikariException("Missing attribute \"%s\" in \"%s\" at %s", want.begin()->c_str(), n.name.c_str(),
                            where(n).c_str());
  }
  static std::string attr(const XmlNode &n, const char *k) {
    const std::string *v = n.attr(k);
    return v ? *v : std::string();
  }

  Object *parse(const XmlNode &node, PropertyList &plist, const TagInfo *parent) {
    auto it = tag_map().find(node.name);
    if (it == tag_map().end())
      throw HikariException("Error while parsing \"%s\": unexpected tag \"%s\" at %s", file.c_str(),
                            node.name.c_str(), where(node).c_str());
    const TagInfo &ti = it->second;
    bool is_object = ti.tag == Tag::kObject;
    bool parent_is_transform = parent && parent->tag == Tag::kTransform;
    bool is_transform_op = ti.tag == Tag::kTranslate || ti.tag == Tag::kRotate || ti.tag == Tag::kScale ||
                           ti.tag == Tag::kLookAt || ti.tag == Tag::kMatrix;
    if (!parent && !is_object)
      throw HikariException("Error while parsing \"%s\": root element \"%s\" must be a Object (at %s)", file.c_str(),
                            node.name.c_str(), where(node).c_str());
    if (parent_is_transform != is_transform_op)
      throw HikariException("Error while parsing \"%s\": transform nodes can only contain transform operations (at %s)",
                            file.c_str(), where(node).c_str());
    if (parent && parent->tag != Tag::kObject && !(parent_is_transform && is_transform_op))
      throw HikariException("Error while parsing \"%s\": node \"%s\" requires a Object as parent (at %s)",
                            file.c_str(), node.name.c_str(), where(node).c_str());
    if (ti.tag == Tag::kTransform) trans = Transform();

    PropertyList child_props;
    std::vector<Object *> children;
    std::vector<std::string> names;
    try {
      for (auto &c : node.children) {
        Object *o = parse(*c, child_props, &ti);
        if (o) {
          children.push_back(o);
          names.push_back(attr(*c, "name"));
        }
      }
    } catch (...) {
      for (Object *o : children) delete o;
      throw;
    }
    Object *result = nullptr;
    try {
      if (is_object) {
        check_attrs(node, {"type", "name"}, false);
        std::string type = ti.cls == Object::EClassType::EScene ? std::string("scene") : attr(node, "type");
        result = ObjectFactory::CreateInstance(type, child_props);
        if (result->GetClassType() != ti.cls) {
          std::string got = Object::ClassTypeName(result->GetClassType());
          delete result;
          result = nullptr;
          throw HikariException("Unexpectedly constructed an object of type <%s> (expected type <%s>)", got.c_str(),
                                Object::ClassTypeName(ti.cls).c_str());
        }
        for (size_t i = 0; i < children.size(); i++) {
          Object *ch = children[i];
          children[i] = nullptr;
          result->AddChild(ch, names[i]);
          ch->SetParent(result, attr(node, "name"));
        }
        result->Activate();
      } else {
        std::string name = attr(node, "name"), value = attr(node, "value");
        switch (ti.tag) {
          case Tag::kString: check_attrs(node, {"name", "value"}, true); plist.SetString(name, value); break;
          case Tag::kFloat: check_attrs(node, {"name", "value"}, true); plist.SetFloat(name, to_float(value)); break;
          case Tag::kInteger: check_attrs(node, {"name", "value"}, true); plist.SetInteger(name, to_int(value)); break;
          case Tag::kBoolean: check_attrs(node, {"name", "value"}, true); plist.SetBoolean(name, to_bool(value)); break;
          case Tag::kPoint: check_attrs(node, {"name", "value"}, true); plist.SetPoint(name, to_vec3(value)); break;
          case Tag::kVector: check_attrs(node, {"name", "value"}, true); plist.SetVector(name, to_vec3(value)); break;
          case Tag::kColor: {
            check_attrs(node, {"name", "value"}, true);
            Vector3f v = to_vec3(value);
            plist.SetColor(name, Color3f(v.x, v.y, v.z));
            break;
          }
          case Tag::kTransform: check_attrs(node, {"name"}, true); plist.SetTransform(name, trans); break;
          case Tag::kTranslate: {
            check_attrs(node, {"value"}, true);
            trans.PreTranslate(to_vec3(value));
            break;
          }
          case Tag::kMatrix: {
            check_attrs(node, {"value"}, true);
            std::vector<std::string> tok = tokenize(value);
            if (tok.size() != 16) throw HikariException("Expected 16 values");
            float a[16];
            for (int i = 0; i < 16; i++) a[i] = to_float(tok[size_t(i)]);
            trans.PreAffine(a);
            break;
          }
          case Tag::kScale: {
            check_attrs(node, {"value"}, true);
            trans.PreScale(to_vec3(value));
            break;
          }
          case Tag::kRotate: {
            check_attrs(node, {"angle", "axis"}, true);
            trans.PreRotate(to_float(attr(node, "angle")), to_vec3(attr(node, "axis")));
            break;
          }
          case Tag::kLookAt: {
            check_attrs(node, {"origin", "target", "up"}, true);
            trans.PreLookAt(to_vec3(attr(node, "origin")), to_vec3(attr(node, "target")), to_vec3(attr(node, "up")));
            break;
          }
          default: throw HikariException("Unhandled element \"%s\"", node.name.c_str());
        }
      }
    } catch (const HikariException &e) {
      for (Object *o : children) delete o;
      delete result;
      std::string what = e.what();
      if (what.rfind("Error while parsing", 0) == 0) throw;
      throw HikariException("Error while parsing \"%s\": %s (at %s)", file.c_str(), what.c_str(), where(node).c_str());
    }
    return result;
  }
};

std::string g_scene_dir;

}  // namespace

const std::string &CurrentSceneDirectory() { return g_scene_dir; }

Object *LoadFromXML(const std::string &filename) {
  std::ifstream in(filename, std::ios::binary);
  if (!in) throw HikariException("Unable to open \"%s\"", filename.c_str());
  std::stringstream ss;
  ss << in.rdbuf();
  std::unique_ptr<XmlNode> root = ParseXml(ss.str(), filename);
  size_t slash = filename.find_last_of('/');
  std::string saved = g_scene_dir;
  g_scene_dir = slash == std::string::npos ? std::string(".") : filename.substr(0, slash);
  Loader L;
  L.file = filename;
  PropertyList top;
  try {
    Object *o = L.parse(*root, top, nullptr);
    g_scene_dir = saved;
    return o;
  } catch (...) {
    g_scene_dir = saved;
    throw;
  }
}

}  // namespace hikari

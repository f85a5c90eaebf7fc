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
Transform::Transform() {
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) m[i][j] = i == j ? 1.0 : 0.0;
}
Transform Transform::FromRows(const double *a) {
  Transform t;
  for (int i = 0; i < 16; i++) t.m[i / 4][i % 4] = a[i];
  return t;
}
Transform Transform::operator*(const Transform &b) const {
  Transform r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double s = 0;
      for (int k = 0; k < 4; k++) s += m[i][k] * b.m[k][j];
      r.m[i][j] = s;
    }
  return r;
}
Transform Transform::Inverse() const {
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
  Transform r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) r.m[i][j] = a[i][j + 4];
  return r;
}
Point3f Transform::ApplyPoint(const Point3f &p) const {
  double x = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
  double y = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
  double z = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
  double w = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
  return Point3f(float(x / w), float(y / w), float(z / w));
}
Vector3f Transform::ApplyVector(const Vector3f &v) const {
  return Vector3f(float(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z),
                  float(m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z),
                  float(m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z));
}
Vector3f Transform::ApplyNormal(const Vector3f &n) const {
  Transform inv = Inverse();
  return Vector3f(float(inv.m[0][0] * n.x + inv.m[1][0] * n.y + inv.m[2][0] * n.z),
                  float(inv.m[0][1] * n.x + inv.m[1][1] * n.y + inv.m[2][1] * n.z),
                  float(inv.m[0][2] * n.x + inv.m[1][2] * n.y + inv.m[2][2] * n.z));
}
void Transform::ToFloat(float *out) const {
  for (int i = 0; i < 16; i++) out[i] = float(m[i / 4][i % 4]);
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
            Vector3f v = to_vec3(value);
            Transform t;
            t.m[0][3] = v.x; t.m[1][3] = v.y; t.m[2][3] = v.z;
            trans = t * trans;
            break;
          }
          case Tag::kMatrix: {
            check_attrs(node, {"value"}, true);
            std::vector<std::string> tok = tokenize(value);
            if (tok.size() != 16) throw HikariException("Expected 16 values");
            double a[16];
            for (int i = 0; i < 16; i++) a[i] = to_float(tok[size_t(i)]);
            trans = Transform::FromRows(a) * trans;
            break;
          }
          case Tag::kScale: {
            check_attrs(node, {"value"}, true);
            Vector3f v = to_vec3(value);
            Transform t;
            t.m[0][0] = v.x; t.m[1][1] = v.y; t.m[2][2] = v.z;
            trans = t * trans;
            break;
          }
          case Tag::kRotate: {
            check_attrs(node, {"angle", "axis"}, true);
            double ang = double(to_float(attr(node, "angle"))) * M_PI / 180.0;
            Vector3f ax = to_vec3(attr(node, "axis"));
            double n = std::sqrt(double(ax.x) * ax.x + double(ax.y) * ax.y + double(ax.z) * ax.z);
            double x = ax.x / n, y = ax.y / n, z = ax.z / n, c = std::cos(ang), s = std::sin(ang), C1 = 1 - c;
            double r[16] = {c + x * x * C1, x * y * C1 - z * s, x * z * C1 + y * s, 0,
                            y * x * C1 + z * s, c + y * y * C1, y * z * C1 - x * s, 0,
                            z * x * C1 - y * s, z * y * C1 + x * s, c + z * z * C1, 0,
                            0, 0, 0, 1};
            trans = Transform::FromRows(r) * trans;
            break;
          }
          case Tag::kLookAt: {
            check_attrs(node, {"origin", "target", "up"}, true);
            Vector3f o = to_vec3(attr(node, "origin")), t = to_vec3(attr(node, "target")), u = to_vec3(attr(node, "up"));
            auto norm = [](double a[3]) {
              double l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
              a[0] /= l; a[1] /= l; a[2] /= l;
            };
            auto cross = [](const double a[3], const double b[3], double r[3]) {
              r[0] = a[1] * b[2] - a[2] * b[1];
              r[1] = a[2] * b[0] - a[0] * b[2];
              r[2] = a[0] * b[1] - a[1] * b[0];
            };
            double dir[3] = {double(t.x) - o.x, double(t.y) - o.y, double(t.z) - o.z};
            norm(dir);
            double up[3] = {u.x, u.y, u.z};
            norm(up);
            double left[3], nup[3];
            cross(up, dir, left);
            norm(left);
            cross(dir, left, nup);
            norm(nup);
            double r[16] = {left[0], nup[0], dir[0], o.x, left[1], nup[1], dir[1], o.y,
                            left[2], nup[2], dir[2], o.z, 0, 0, 0, 1};
            trans = Transform::FromRows(r) * trans;
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

// ref_eigen_xform.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Driver compiled against the reference's vendored Eigen 3.3.5 (ext/eigen,
// header-only) by `make -C oracle ref`; output goes to oracle/_ref/.  It
// composes a toWorld transform the way the reference's XML parser does
// (src/core/Parser.cpp:308-369: every op premultiplied onto an
// Eigen::Affine3f), builds the reference's Transform from it (matrix and
// Eigen inverse, src/core/Transform.cpp:9-11) and applies it to the "v" and
// "vn" lines of an OBJ file as src/mesh/WavefrontObjMesh.cpp:52-70 does
// (points through Matrix4f * Vector4f and division by w, normals through the
// inverse's transposed 3x3 block and normalized()), plus each "v" as a
// vector (the 3x3 block times it, src/core/Transform.cpp:36-39).  The host
// layer's FP32 transform code is compared with it bit for bit.
//
// Usage: ref_eigen_xform <ops.txt> <file.obj|-> <out.bin>
//   ops.txt: one op per line, in document order:
//     translate x y z | scale x y z | rotate angle ax ay az |
//     matrix m00 m01 ... m33 (row-major) | lookat ox oy oz tx ty tz ux uy uz
//   out.bin (float32 unless noted): matrix (16, row-major), inverse (16,
//     row-major), int32 nv, nv x 3 points, nv x 3 vectors, int32 nn,
//     nn x 3 normals
#include <Eigen/Core>
#include <Eigen/Geometry>
#include <Eigen/LU>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

typedef Eigen::Vector3f V3;

// the reference's DegToRad (include/core/Common.hpp:742)
static float deg_to_rad(float v) { return v * (float(M_PI) / 180.0f); }

int main(int argc, char **argv) {
  if (argc != 4) return 2;
  std::ifstream ops(argv[1]);
  if (!ops) return 3;
  Eigen::Affine3f trans;
  trans.setIdentity();
  std::string line;
  while (std::getline(ops, line)) {
    std::istringstream ls(line);
    std::string op;
    if (!(ls >> op)) continue;
    std::vector<float> a;
    std::string tok;
    while (ls >> tok) a.push_back(std::strtof(tok.c_str(), nullptr));
    if (op == "translate" && a.size() == 3) {
      trans = Eigen::Translation<float, 3>(a[0], a[1], a[2]) * trans;
    } else if (op == "scale" && a.size() == 3) {
      trans = Eigen::DiagonalMatrix<float, 3>(V3(a[0], a[1], a[2])) * trans;
    } else if (op == "rotate" && a.size() == 4) {
      trans = Eigen::AngleAxis<float>(deg_to_rad(a[0]), V3(a[1], a[2], a[3])) * trans;
    } else if (op == "matrix" && a.size() == 16) {
      Eigen::Matrix4f m;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) m(i, j) = a[size_t(i * 4 + j)];
      trans = Eigen::Affine3f(m) * trans;
    } else if (op == "lookat" && a.size() == 9) {
      V3 origin(a[0], a[1], a[2]), target(a[3], a[4], a[5]), up(a[6], a[7], a[8]);
      V3 dir = (target - origin).normalized();
      V3 left = up.normalized().cross(dir).normalized();
      V3 new_up = dir.cross(left).normalized();
      Eigen::Matrix4f m;
      m << left, new_up, dir, origin, 0, 0, 0, 1;
      trans = Eigen::Affine3f(m) * trans;
    } else {
      fprintf(stderr, "bad op line: %s\n", line.c_str());
      return 4;
    }
  }
  Eigen::Matrix4f M = trans.matrix();
  Eigen::Matrix4f Minv = M.inverse();
  std::vector<float> pts, vecs, nrms;
  if (std::string(argv[2]) != "-") {
    std::ifstream obj(argv[2]);
    if (!obj) return 5;
    while (std::getline(obj, line)) {
      std::stringstream ls;
      ls << line;
      std::string pre;
      ls >> pre;
      if (pre == "v") {
        V3 p;
        ls >> p.x() >> p.y() >> p.z();
        Eigen::Vector4f r = M * Eigen::Vector4f(p[0], p[1], p[2], 1.0f);
        V3 q = r.head<3>() / r.w();
        V3 v = M.topLeftCorner<3, 3>() * p;
        pts.insert(pts.end(), {q.x(), q.y(), q.z()});
        vecs.insert(vecs.end(), {v.x(), v.y(), v.z()});
      } else if (pre == "vn") {
        V3 n;
        ls >> n.x() >> n.y() >> n.z();
        V3 t = (Minv.topLeftCorner<3, 3>().transpose() * n).normalized();
        nrms.insert(nrms.end(), {t.x(), t.y(), t.z()});
      }
    }
  }
  FILE *f = fopen(argv[3], "wb");
  if (!f) return 6;
  float rm[16], ri[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      rm[i * 4 + j] = M(i, j);
      ri[i * 4 + j] = Minv(i, j);
    }
  fwrite(rm, 4, 16, f);
  fwrite(ri, 4, 16, f);
  int32_t nv = int32_t(pts.size() / 3), nn = int32_t(nrms.size() / 3);
  fwrite(&nv, 4, 1, f);
  fwrite(pts.data(), 4, pts.size(), f);
  fwrite(vecs.data(), 4, vecs.size(), f);
  fwrite(&nn, 4, 1, f);
  fwrite(nrms.data(), 4, nrms.size(), f);
  fclose(f);
  return 0;
}

// ref_eigen_ops.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Driver compiled against the reference's vendored Eigen 3.3.5
// (ext/eigen, header-only) by `make -C oracle ref`; output goes to
// oracle/_ref/.  It evaluates, with Eigen's own Vector3f arithmetic, the
// expressions the reference uses on the hot path -- dot, cross, normalized
// (src/core/Mesh.cpp:340-384, src/core/Frame.cpp:9-40) and the Frame built by
// CoordinateSystem (src/core/Common.cpp:591-604) -- so the oracle's
// restatement of the evaluation order can be compared bit for bit.
//
// Usage: ref_eigen_ops <in.bin> <out.bin>; in = n x 6 floats (a, b), out =
// n x 13 floats in the layout of ho_vec_ops().
#include <Eigen/Core>
#include <Eigen/Geometry>
#include <cmath>
#include <cstdio>
#include <vector>

typedef Eigen::Matrix<float, 3, 1> V;

static void coordinate_system(const V &a, V &b, V &c) {
  if (std::abs(a.x()) > std::abs(a.y())) {
    float inv = 1.0f / std::sqrt(a.x() * a.x() + a.z() * a.z());
    c = V(a.z() * inv, 0.0f, -a.x() * inv);
  } else {
    float inv = 1.0f / std::sqrt(a.y() * a.y() + a.z() * a.z());
    c = V(0.0f, a.z() * inv, -a.y() * inv);
  }
  b = c.cross(a);
}

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 3;
  std::vector<float> in;
  float buf[6];
  while (fread(buf, 4, 6, f) == 6) in.insert(in.end(), buf, buf + 6);
  fclose(f);
  size_t n = in.size() / 6;
  std::vector<float> out(n * 13);
  for (size_t i = 0; i < n; i++) {
    V a(in[6 * i], in[6 * i + 1], in[6 * i + 2]), b(in[6 * i + 3], in[6 * i + 4], in[6 * i + 5]);
    float *o = &out[13 * i];
    o[0] = a.dot(b);
    V c = a.cross(b), na = a.normalized();
    o[1] = c.x(); o[2] = c.y(); o[3] = c.z();
    o[4] = na.x(); o[5] = na.y(); o[6] = na.z();
    V nn = b.normalized(), s, t;
    coordinate_system(nn, s, t);
    V l(a.dot(s), a.dot(t), a.dot(nn));
    V w = s * a.x() + t * a.y() + nn * a.z();
    o[7] = l.x(); o[8] = l.y(); o[9] = l.z();
    o[10] = w.x(); o[11] = w.y(); o[12] = w.z();
  }
  f = fopen(argv[2], "wb");
  if (!f) return 4;
  fwrite(out.data(), 4, out.size(), f);
  fclose(f);
  return 0;
}

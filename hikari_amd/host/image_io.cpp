// image_io.cpp -- PFM and OpenEXR image files for environment maps and
// rendered films (the reference reads and writes EXR through OpenEXR in
// src/core/Bitmap.cpp).  EXR support: single-part scanline files, channels
// R, G, B (optionally A, ignored; a single Y channel is read as grey), HALF /
// FLOAT / UINT samples, compression NONE, ZIPS or ZIP (zlib + the format's
// byte predictor and interleave, as specified by the OpenEXR file layout).
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>

#include "hikari.h"

namespace hikari {

// ------------------------------------------------------------------ PFM
void WritePFM(const std::string &path, const float *rgb, int w, int h) {
  std::ofstream out(path, std::ios::binary);
  if (!out) throw HikariException("Unable to write \"%s\"", path.c_str());
  out << "PF\n" << w << " " << h << "\n-1\n";
  for (int y = h - 1; y >= 0; y--) out.write(reinterpret_cast<const char *>(rgb + size_t(y) * w * 3), size_t(w) * 12);
  if (!out) throw HikariException("Unable to write \"%s\"", path.c_str());
}

std::vector<float> ReadPFM(const std::string &path, int &w, int &h) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw HikariException("Unable to open \"%s\"", path.c_str());
  std::string magic;
  double scale;
  in >> magic >> w >> h >> scale;
  in.get();
  int ch = magic == "PF" ? 3 : (magic == "Pf" ? 1 : 0);
  if (!ch || w <= 0 || h <= 0) throw HikariException("\"%s\" is not a PFM file", path.c_str());
  bool big = scale > 0;
  std::vector<float> raw(size_t(w) * h * ch);
  in.read(reinterpret_cast<char *>(raw.data()), raw.size() * 4);
  if (!in) throw HikariException("\"%s\": truncated PFM data", path.c_str());
  if (big)
    for (float &f : raw) {
      uint32_t u;
      std::memcpy(&u, &f, 4);
      u = __builtin_bswap32(u);
      std::memcpy(&f, &u, 4);
    }
  std::vector<float> out(size_t(w) * h * 3);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++)
      for (int c = 0; c < 3; c++)
        out[(size_t(y) * w + x) * 3 + c] = raw[(size_t(h - 1 - y) * w + x) * ch + (ch == 3 ? c : 0)];
  return out;
}

// ------------------------------------------------------------------ EXR
namespace {

float half_to_float(uint16_t h) {
  uint32_t s = uint32_t(h >> 15) << 31, e = (h >> 10) & 0x1f, m = h & 0x3ff, u;
  if (e == 0) {
    if (m == 0) {
      u = s;
    } else {  // subnormal: normalise
      int k = -1;
      do { k++; m <<= 1; } while (!(m & 0x400));
      u = s | (uint32_t(127 - 15 - k) << 23) | ((m & 0x3ff) << 13);
    }
  } else if (e == 31) {
    u = s | 0x7f800000u | (m << 13);
  } else {
    u = s | ((e + 127 - 15) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

struct Reader {
  const std::vector<uint8_t> &b;
  size_t i = 0;
  const std::string &path;
  Reader(const std::vector<uint8_t> &buf, const std::string &p) : b(buf), path(p) {}
  void need(size_t n) const {
    if (i + n > b.size()) throw HikariException("\"%s\": truncated EXR file", path.c_str());
  }
  template <typename T> T get() {
    need(sizeof(T));
    T v;
    std::memcpy(&v, &b[i], sizeof(T));
    i += sizeof(T);
    return v;
  }
  std::string str() {
    size_t e = i;
    while (e < b.size() && b[e]) e++;
    if (e >= b.size()) throw HikariException("\"%s\": truncated EXR header", path.c_str());
    std::string s(reinterpret_cast<const char *>(&b[i]), e - i);
    i = e + 1;
    return s;
  }
};

struct Channel {
  std::string name;
  int type;  // 0 uint, 1 half, 2 float
};

}  // namespace

std::vector<float> ReadEXR(const std::string &path, int &w, int &h) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw HikariException("Unable to open \"%s\"", path.c_str());
  std::vector<uint8_t> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  Reader r(buf, path);
  if (r.get<uint32_t>() != 20000630u) throw HikariException("\"%s\" is not an OpenEXR file", path.c_str());
  uint32_t version = r.get<uint32_t>();
  if ((version & 0xff) != 2 || (version & 0x200) || (version & 0x1000))
    throw HikariException("\"%s\": only single-part scanline OpenEXR files are supported", path.c_str());
  std::vector<Channel> chans;
  int comp = -1, xmin = 0, ymin = 0, xmax = -1, ymax = -1;
  for (;;) {
    std::string name = r.str();
    if (name.empty()) break;
    std::string type = r.str();
    int32_t size = r.get<int32_t>();
    size_t end = r.i + size_t(size);
    r.need(size_t(size));
    if (name == "channels" && type == "chlist") {
      for (;;) {
        std::string cn = r.str();
        if (cn.empty()) break;
        int t = r.get<int32_t>();
        r.i += 4;  // pLinear + reserved
        int xs = r.get<int32_t>(), ys = r.get<int32_t>();
        if (xs != 1 || ys != 1) throw HikariException("\"%s\": subsampled channels are not supported", path.c_str());
        chans.push_back({cn, t});
      }
    } else if (name == "compression") {
      comp = r.get<uint8_t>();
    } else if (name == "dataWindow") {
      xmin = r.get<int32_t>(); ymin = r.get<int32_t>(); xmax = r.get<int32_t>(); ymax = r.get<int32_t>();
    }
    r.i = end;
  }
  if (chans.empty() || xmax < xmin || ymax < ymin) throw HikariException("\"%s\": incomplete EXR header", path.c_str());
  int lines;
  if (comp == 0 || comp == 2) lines = 1;          // NONE, ZIPS
  else if (comp == 3) lines = 16;                 // ZIP
  else throw HikariException("\"%s\": EXR compression %d is not supported (NONE, ZIPS, ZIP are)", path.c_str(), comp);
  w = xmax - xmin + 1;
  h = ymax - ymin + 1;
  int n_chunks = (h + lines - 1) / lines;
  std::vector<uint64_t> offsets(static_cast<size_t>(n_chunks));
  for (auto &o : offsets) o = r.get<uint64_t>();
  // channel byte sizes; channels are stored in the header's (alphabetical) order
  std::vector<size_t> csz;
  size_t pix_bytes = 0;
  for (auto &c : chans) {
    csz.push_back(c.type == 1 ? 2 : 4);
    pix_bytes += csz.back();
  }
  int idx_r = -1, idx_g = -1, idx_b = -1, idx_y = -1;
  for (size_t k = 0; k < chans.size(); k++) {
    if (chans[k].name == "R") idx_r = int(k);
    if (chans[k].name == "G") idx_g = int(k);
    if (chans[k].name == "B") idx_b = int(k);
    if (chans[k].name == "Y") idx_y = int(k);
  }
  if ((idx_r < 0 || idx_g < 0 || idx_b < 0) && idx_y < 0)
    throw HikariException("\"%s\": no R, G, B (or Y) channels", path.c_str());
  std::vector<float> out(size_t(w) * h * 3, 0.0f);
  std::vector<uint8_t> raw, tmp;
  for (int ci = 0; ci < n_chunks; ci++) {
    r.i = size_t(offsets[size_t(ci)]);
    int32_t y0 = r.get<int32_t>() - ymin;
    int32_t dsize = r.get<int32_t>();
    r.need(size_t(dsize));
    int nl = std::min(lines, h - y0);
    if (y0 < 0 || nl <= 0) throw HikariException("\"%s\": bad EXR chunk", path.c_str());
    size_t expect = size_t(nl) * w * pix_bytes;
    const uint8_t *src = &buf[r.i];
    if (comp == 0 || size_t(dsize) == expect) {
      raw.assign(src, src + dsize);
    } else {
      tmp.resize(expect);
      uLongf dl = uLongf(expect);
      if (uncompress(tmp.data(), &dl, src, uLong(dsize)) != Z_OK || dl != expect)
        throw HikariException("\"%s\": corrupt ZIP-compressed EXR chunk", path.c_str());
      // undo the predictor (running sum of byte differences), then the
      // split of even / odd bytes into two halves
      for (size_t k = 1; k < expect; k++) tmp[k] = uint8_t(int(tmp[k - 1]) + int(tmp[k]) - 128);
      raw.resize(expect);
      size_t half = (expect + 1) / 2;
      for (size_t k = 0; k < expect; k++) raw[k] = (k & 1) ? tmp[half + k / 2] : tmp[k / 2];
    }
    if (raw.size() < expect) throw HikariException("\"%s\": short EXR chunk", path.c_str());
    size_t off = 0;
    for (int l = 0; l < nl; l++) {
      int y = y0 + l;
      for (size_t k = 0; k < chans.size(); k++) {
        for (int x = 0; x < w; x++) {
          float v;
          if (chans[k].type == 1) {
            uint16_t hv;
            std::memcpy(&hv, &raw[off], 2);
            v = half_to_float(hv);
          } else if (chans[k].type == 2) {
            std::memcpy(&v, &raw[off], 4);
          } else {
            uint32_t u;
            std::memcpy(&u, &raw[off], 4);
            v = float(u);
          }
          off += csz[k];
          float *p = &out[(size_t(y) * w + x) * 3];
          if (int(k) == idx_r) p[0] = v;
          if (int(k) == idx_g) p[1] = v;
          if (int(k) == idx_b) p[2] = v;
          if (int(k) == idx_y && idx_r < 0) p[0] = p[1] = p[2] = v;
        }
      }
    }
  }
  return out;
}

void WriteEXR(const std::string &path, const float *rgb, int w, int h) {
  std::vector<uint8_t> o;
  auto put = [&](const void *p, size_t n) { o.insert(o.end(), (const uint8_t *)p, (const uint8_t *)p + n); };
  auto u32 = [&](uint32_t v) { put(&v, 4); };
  auto i32 = [&](int32_t v) { put(&v, 4); };
  auto f32 = [&](float v) { put(&v, 4); };
  auto str = [&](const char *s) { put(s, std::strlen(s) + 1); };
  auto attr = [&](const char *name, const char *type, int32_t size) {
    str(name);
    str(type);
    i32(size);
  };
  u32(20000630u);
  u32(2);
  attr("channels", "chlist", 3 * (2 + 16) + 1);
  for (const char *c : {"B", "G", "R"}) {
    str(c);
    i32(2);  // FLOAT
    u32(0);  // pLinear + reserved
    i32(1);
    i32(1);
  }
  o.push_back(0);
  attr("compression", "compression", 1);
  o.push_back(0);  // NONE
  attr("dataWindow", "box2i", 16);
  i32(0); i32(0); i32(w - 1); i32(h - 1);
  attr("displayWindow", "box2i", 16);
  i32(0); i32(0); i32(w - 1); i32(h - 1);
  attr("lineOrder", "lineOrder", 1);
  o.push_back(0);
  attr("pixelAspectRatio", "float", 4);
  f32(1.0f);
  attr("screenWindowCenter", "v2f", 8);
  f32(0.0f); f32(0.0f);
  attr("screenWindowWidth", "float", 4);
  f32(1.0f);
  o.push_back(0);  // end of header
  size_t table = o.size();
  o.resize(o.size() + size_t(h) * 8);
  for (int y = 0; y < h; y++) {
    uint64_t at = o.size();
    std::memcpy(&o[table + size_t(y) * 8], &at, 8);
    i32(y);
    i32(int32_t(size_t(w) * 12));
    for (int c = 2; c >= 0; c--)
      for (int x = 0; x < w; x++) f32(rgb[(size_t(y) * w + x) * 3 + c]);
  }
  std::ofstream out(path, std::ios::binary);
  if (!out) throw HikariException("Unable to write \"%s\"", path.c_str());
  out.write(reinterpret_cast<const char *>(o.data()), std::streamsize(o.size()));
}

std::vector<float> ReadImage(const std::string &path, int &w, int &h) {
  std::string ext = path.size() >= 4 ? path.substr(path.size() - 4) : "";
  for (char &c : ext) c = char(std::tolower((unsigned char)c));
  if (ext == ".exr") return ReadEXR(path, w, h);
  if (ext == ".pfm") return ReadPFM(path, w, h);
  throw HikariException("Unsupported image format \"%s\" (EXR or PFM expected)", path.c_str());
}

}  // namespace hikari

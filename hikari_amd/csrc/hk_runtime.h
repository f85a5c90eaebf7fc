// hk_runtime.h -- host-side objects behind the opaque C-ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hikari_amd.h"
#include "hk_cdf.h"
#include "hk_shared.h"

namespace hk {

struct HkError {
  int code;
  std::string msg;
  HkError(int c, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
};
void hip_check(hipError_t e, const char *what);
std::string default_data_dir();

// Owned device allocation.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DBuf() {}
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void alloc(size_t n);
  void release();
  void upload(const void *src, size_t n, hipStream_t st);
  template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
};

struct DevBvh {
  DBuf nodes, tris;
  uint32_t n_nodes = 0, n_tris = 0, n_leaves = 0, max_depth = 0;
};

// Device LBVH build (hk_bvh_gpu.hip) into the FP32 node format; tri9: n x 9
// floats on the host; cell_bits: Morton bits of the cells under the SAH upper
// tree (0: the radix tree all the way up).  Synchronizes `st`.
void build_bvh4_device(hipStream_t st, const float *tri9, uint32_t n, uint32_t cell_bits, DevBvh &out);
// 24 bits (256^3 cells): +4% trace speed over the reference's 12 on the
// 1.3 M / 2.1 M-triangle sphere and heightfield for +0.02 / +0.05 s of build
constexpr uint32_t kDefaultBvhCellBits = 24;

// path slots of the wavefront (~250-350 B of HBM each): 256 M by default for
// renders and batches (a larger wavefront spreads each launch's straggler
// tail over more work); a server's bulk wavefront 64 M (its persistent
// launches then leave the express lanes room sooner) and each express lane
// 4 M, unless max_paths is set (bulk max_paths, express 1/16 of it)
constexpr uint32_t kDefaultMaxPaths = 1u << 28;
constexpr uint32_t kDefaultServerBulkPaths = 1u << 26, kDefaultExpressPaths = 1u << 22;
constexpr uint32_t kDefaultMaxSamples = 1u << 28;   // samples per segment (16 B each in HBM)
// persistent traversal waves: idle lanes that trigger a refill, and queue
// entries (in 64s) reserved per atomic on the fetch counter
constexpr uint32_t kDefaultRefill = 40, kDefaultRefillShadow = 16, kDefaultFetchBatch = 4;
// HBM for the per-emitter shadow slots (48 B per path and emitter): scenes
// with many emitters run narrower wavefronts (paths <= this / (48 x emitters))
constexpr uint64_t kShadowSlotBytes = uint64_t(24) << 30;
uint32_t paths_for_emitters(uint32_t paths, uint32_t emitters);

// Owned pinned host allocation (staging for small per-bounce copies).
struct PinnedBuf {
  void *p = nullptr;
  size_t bytes = 0;
  PinnedBuf() {}
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  ~PinnedBuf() { release(); }
  void alloc(size_t n);
  void release();
  template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
};

// Wavefront buffers of one engine, grown on demand and reused across renders.
// The per-emitter shadow slots (and the emitter mask words beyond the first
// 32 emitters) are sized for paths x emitters of the current engine, apart
// from the queues, so a scene with many emitters and a narrow wavefront does
// not multiply with another scene's wide one.
struct Workspace {
  uint32_t cap_paths = 0, cap_samples = 0;
  size_t cap_slots = 0, cap_words = 0;
  bool tables = false;
  DBuf sample_L;
  DBuf q_o[2], q_d[2], q_beta[2], q_L[2], q_aux[2], q_meta[2], hit;
  DBuf s_ent, sh_o, sh_d, sh_c, sh_mask;
  DBuf counters, stats;
  DBuf reqs, items, plan, item_done, hdr;          // request / item tables, regen plan, per-item counts,
                                                   // bounce header
  DBuf req_invalid;                                // invalid samples dropped per request slot (u64)
  PinnedBuf h_reqs, h_items, h_plan, h_done, h_counters, h_hdr, h_req_invalid;  // host mirrors / staging
  void ensure(uint32_t paths, uint32_t emitters, uint32_t samples);
  uint64_t device_bytes() const;
  void release();
};

}  // namespace hk

struct hk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int n_cu = 0;
  std::string arch;
  uint32_t max_paths = hk::kDefaultMaxPaths;
  bool max_paths_set = false;  // max_paths given by hk_set_option (servers then use it, express 1/16)
  uint32_t max_samples = hk::kDefaultMaxSamples;
  uint32_t grid = 0;  // persistent grid for queue kernels (0: automatic)
  int bvh_builder = 0;  // 0 host binned SAH, 1 device LBVH
  uint32_t bvh_cell_bits = hk::kDefaultBvhCellBits;  // device LBVH: Morton bits of the cells under its SAH upper tree
  bool resolve_tiles = true;  // film resolve through LDS tiles (false: the plain gather)
  uint32_t tile = 8;          // camera-ray tile edge (1: scanlines)
  bool graph = true;          // replay each bounce as a captured HIP graph
  uint32_t refill_extend = hk::kDefaultRefill;       // idle lanes that make a persistent k_extend wave fetch rays
  uint32_t refill_shadow = hk::kDefaultRefillShadow; // the same for k_shadow
  uint32_t fetch_batch = hk::kDefaultFetchBatch;     // queue entries a persistent wave reserves per atomic, in 64s
  int schedule = 1;             // chunk admission / regeneration order: 0 FIFO, 1 shortest first
  bool express = true;          // server: small requests take an express lane beside the bulk lane
  bool profile = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  uint32_t *host_counters = nullptr;  // pinned
  hk::Workspace ws;
};

struct hk_accel {
  hk_ctx *ctx = nullptr;
  hk::DevBvh bvh;
  hk::DBuf fetch;  // the persistent intersect kernel's fetch counter (one word)
  double build_ms = 0;
};

struct hk_scene {
  hk_ctx *ctx = nullptr;
  hk::DevScene dev;
  hk::DevBvh bvh;
  hk::DBuf shade, meshes, mats, rft, emitters, cdf, cdf_guide;
  hk::DBuf env_tex, env_row_cdf, env_row_func, env_row_int, env_marg_cdf, env_marg_func, env_row_guide,
      env_marg_guide;
  hk_scene_desc desc;             // camera / film defaults (pointers not retained)
  uint32_t n_tris = 0;
  uint64_t bytes = 0;
};

namespace hk {
// 4x4 helpers in double (the camera / env transforms are set up once).
struct Mat4 {
  double m[4][4];
  static Mat4 from(const float *a) {
    Mat4 r;
    for (int i = 0; i < 16; i++) r.m[i / 4][i % 4] = a[i];
    return r;
  }
  static Mat4 identity() {
    Mat4 r;
    for (int i = 0; i < 16; i++) r.m[i / 4][i % 4] = (i / 4 == i % 4);
    return r;
  }
  Mat4 operator*(const Mat4 &b) const {
    Mat4 r;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        double s = 0;
        for (int k = 0; k < 4; k++) s += m[i][k] * b.m[k][j];
        r.m[i][j] = s;
      }
    return r;
  }
  Mat4 inverse() const {
    double a[4][8];
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) { a[i][j] = m[i][j]; a[i][j + 4] = i == j; }
    for (int c = 0; c < 4; c++) {
      int p = c;
      for (int r = c + 1; r < 4; r++)
        if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
      if (a[p][c] == 0.0) throw HkError(HK_ERR_INVALID, "singular transform");
      for (int j = 0; j < 8; j++) std::swap(a[c][j], a[p][j]);
      double d = a[c][c];
      for (int j = 0; j < 8; j++) a[c][j] /= d;
      for (int r = 0; r < 4; r++)
        if (r != c) {
          double f = a[r][c];
          for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j];
        }
    }
    Mat4 r;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) r.m[i][j] = a[i][j + 4];
    return r;
  }
  void to_float(float *out) const {
    for (int i = 0; i < 16; i++) out[i] = float(m[i / 4][i % 4]);
  }
};


void camera_setup(const float *c2w, float fov, int w, int h, float near_clip, float far_clip, DevRequest &r);
void filter_setup(const hk_scene_desc &d, DevRequest &r);
void upload_bvh(hk_ctx *ctx, const float *tri9, uint32_t nt, DevBvh &out);
void build_materials(const float *recs, int n, const char *rft_dir, std::vector<DevMaterial> &out,
                     std::vector<float> &rft_pool);
}  // namespace hk

// hk_shared.h -- plain-data records shared by the host runtime and the HIP
// kernels of libhikari_amd (no HIP types, so host .cpp files can include it).
#pragma once
#include <stdint.h>

namespace hk {

enum MatType : int32_t {
  kDiffuse = 0, kConductor = 1, kDielectric = 2, kPlastic = 3, kRoughConductor = 4,
  kRoughDielectric = 5, kRoughPlastic = 6, kCoating = 7, kRoughCoating = 8, kMirror = 9, kMicrofacet = 10,
  kRoughDiffuse = 11, kTwoSided = 12, kMatTypeCount = 13
};
enum Measure : int32_t { kUnknownMeasure = 0, kSolidAngle = 1, kDiscrete = 2 };
enum Dist : int32_t { kBeckmann = 0, kGGX = 1 };

constexpr int kRftThetaN = 100;  // reduced rough-transmittance table length (theta axis)

// Device material record (128 B).  Derived quantities are computed on the
// host exactly as the reference's constructors / Activate() do.
constexpr int32_t kMatFlagDiffuse = 4;  // BSDF::IsDiffuse (pathEMS / whitted sample emitters only there)
constexpr int32_t kMatFlagFast = 8;     // roughDiffuse fastApprox

struct alignas(16) DevMaterial {
  int32_t type, dist, nested, flags;    // flags: 1 nonlinear, 2 anisotropic, 4 diffuse, 8 fast
  float a[4];                           // colour A (see include/hikari_amd.h)
  float b[4];                           // colour B
  float eta, inv_eta, inv_eta2, spec_weight;
  float alpha_u, alpha_v, thickness, fdr_int;  // fdr_int: 1 - internal diffuse transmittance
  float eta3[4], eta_k[4];              // conductor eta, k / extIOR
  int32_t rft_off;                      // offset of the reduced theta table (kRftThetaN floats) or -1
  int32_t nested_back;                  // twoSided: lower-side material
  int32_t pad[2];
};
static_assert(sizeof(DevMaterial) == 128, "DevMaterial layout");

struct alignas(16) DevMesh {
  uint32_t first_prim, nt, material;
  int32_t emitter;                      // index in the emitter list or -1
  uint32_t has_n, has_uv, cdf_off, pad0;
  float radiance[4];
  float inv_area, pad1, pad2, pad3;
};
static_assert(sizeof(DevMesh) == 64, "DevMesh layout");

struct DevEmitter {
  int32_t kind;  // 0 area, 1 environment
  int32_t mesh;
};

// BVH4 node: 4 child boxes stored per axis (SoA inside the node) + 4 child
// references.  ref >= 0: internal node index; ref < 0: leaf, ~ref encodes
// (first triangle << 3) | (count - 1); kEmptyRef marks an unused slot.
constexpr int32_t kEmptyRef = int32_t(0x80000000);
constexpr int kLeafMax = 8;
struct alignas(64) Bvh4Node {
  float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
  int32_t child[4];
  int32_t pad[4];
};
static_assert(sizeof(Bvh4Node) == 128, "Bvh4Node layout");

// Node format the kernels walk: the 128-B FP32 node.
using TraceNode = Bvh4Node;
// The host builder lays out the first kTopNodesBfs nodes breadth-first (the
// root and the next three levels of a full tree); the traversal kernels copy
// the first kTopNodes (closest hit) or kTopNodesShadow (any hit) of them to
// LDS, since every ray's descent passes through them.
constexpr int kTopNodesBfs = 85;
constexpr int kTopNodes = 20;
constexpr int kTopNodesShadow = 84;

// Leaf-ordered triangle for traversal (48 B): p0 + prim id, e1 = p1 - p0,
// e2 = p2 - p0 (edges computed in FP32 on the host exactly as the reference
// computes them per test, src/core/Mesh.cpp:349).
struct alignas(16) TraceTri {
  float p0x, p0y, p0z; int32_t prim;
  float e1x, e1y, e1z, e2x;
  float e2y, e2z, pad0, pad1;
};
static_assert(sizeof(TraceTri) == 48, "TraceTri layout");

// Shading record of one triangle (128 B, one L2 line), in scene primitive
// order: the three vertex positions, normals and texture coordinates (zero
// where the mesh has none), its mesh, material and attribute flags.  A hit
// reads this one line instead of the face's index triple followed by nine
// scattered vertex attributes.
constexpr int kMaxEmitters = 1 << 16;  // emitters of one scene (the wavefront narrows as they grow)
constexpr uint32_t kMaxPathDepth = (1u << 18) - 1;  // vertices per path (18 bits of the path state)

constexpr uint32_t kShadeHasN = 1u, kShadeHasUV = 2u;
struct alignas(128) ShadeTri {
  float p[9];
  float n[9];
  float uv[6];
  uint32_t mesh, material, flags;
  uint32_t pad[5];
};
static_assert(sizeof(ShadeTri) == 128, "ShadeTri layout");

struct DevScene {
  const TraceNode *nodes;
  const TraceTri *tris;
  const ShadeTri *shade;   // nt shading records
  const DevMesh *meshes;
  const DevMaterial *mats;
  const float *rft;        // reduced table pool
  const DevEmitter *emitters;
  const float *cdf;        // area-light CDF pool (per mesh nt + 1 entries)
  const uint32_t *cdf_guide;  // guide tables of the CDFs (same offsets)
  const float *env_tex;    // h x w x 3
  const float *env_row_cdf;   // h x (w + 1)
  const float *env_row_func;  // h x w
  const float *env_row_int;   // h
  const float *env_marg_cdf;  // h + 1
  const float *env_marg_func; // h
  const uint32_t *env_row_guide;   // h x (w + 1), guide tables of the row CDFs
  const uint32_t *env_marg_guide;  // h + 1
  float env_marg_int;
  int32_t env_w, env_h, has_env;
  float env_scale;
  float env_to_world[9], env_to_local[9];
  int32_t n_emitters, n_meshes;
  float bbox_radius;
  int32_t max_depth;
  float background[3];
  int32_t force_background;
  int32_t root_ref;
  int32_t n_nodes;          // BVH nodes in `nodes`
  int32_t integrator;       // 0 pathMIS, 1 pathEMS, 2 pathMATS
  int32_t n_mats;           // records in mats
  int32_t rr_start;         // Russian roulette from this path depth on (0: every vertex)
};

// One render request as the kernels see it.
struct DevRequest {
  float s2c[16];           // sample -> camera, row-major
  float c2w[16];           // camera -> world, row-major
  float inv_w, inv_h, near_clip, far_clip;
  int32_t width, height;
  uint32_t seed;
  uint32_t pix0;           // first pixel of the film slice (0 for whole images)
  float filter[33];
  float filter_radius, lookup;
  uint32_t tile;           // whole images: edge of the square pixel tiles that order the
                           // sample slots (power of two, 1: scanlines); 0: a film slice at pix0
  float *accum;            // width x height x 4 (rgb * w, w)
};

// A chunk of one request: every pixel of its film slice, samples
// [s0, s0 + k).  It occupies a slot of the device item table while it runs;
// the radiance of its sample `local` (= sample offset * npix + pixel rank)
// goes to sample_L[buf_base + local].  The pixel rank is the pixel index, or
// for whole images the position in tile order (DevRequest::tile).
struct WorkItem {
  uint32_t req, npix, s0, k;
  uint32_t buf_base, pad0, pad1, pad2;
};
constexpr uint32_t kMaxItems = 4096;     // item slots (12 bits of the path state)
constexpr uint32_t kMaxRequests = 4096;  // request slots

// One entry of k_regen's plan: start `count` samples of the item in `slot`,
// local indices local0 .. local0 + count - 1; entries are consecutive in the
// launch (prefix = number of samples planned before this entry).
struct RegenEntry {
  uint32_t slot, local0, count, prefix;
};

}  // namespace hk

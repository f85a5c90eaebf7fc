// hk_kernels.h -- launch interface between the host runtime and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "hk_shared.h"

namespace hk {

// Per-bounce parameters of k_regen, read from device memory so that one
// captured graph serves every bounce: the new samples go to queue positions
// q_base .. q_base + n_new - 1, described by the first n_plan plan entries.
struct BounceHdr {
  uint32_t q_base, n_new, n_plan, pad;
};

// Everything the bounce kernels touch (passed by value to every kernel).
// Work arrives as chunks (WorkItem, one slot each); each bounce k_regen
// starts the samples listed in `plan` behind the continuing paths of the
// extension queue.  Path state lives in the extension queues (SoA,
// ping-pong): k_shade moves a continuing path from queue qi to the compacted
// queue qi^1; a finished path writes its radiance to the sample buffer and
// counts itself in item_done, so the host knows when a chunk is complete and
// can be resolved into its film.
struct RenderArgs {
  DevScene sc;
  const DevRequest *reqs;    // request table (by request slot)
  const WorkItem *items;     // item table (by item slot)
  const RegenEntry *plan;    // this bounce's new samples
  const BounceHdr *hdr;      // this bounce's k_regen parameters
  uint32_t n_paths;          // queue capacity (paths in flight)
  uint32_t refill_extend;    // persistent k_extend / k_shadow: idle lanes of a wave that
  uint32_t refill_shadow;    // trigger a fetch of new queue entries (1..64)
  uint32_t fetch_batch;      // persistent kernels: entries a wave reserves per atomic, in 64s
  uint32_t *item_done;       // finished samples per item slot, in kDoneWays partial counts:
                             // slot s, way w at w * kMaxItems + s (the host sums the ways)
  float4 *sample_L;          // final radiance per sample (rgb, unused)
  // per queue entry
  float4 *q_o[2];      // ray origin, tmin
  float4 *q_d[2];      // ray direction, tmax
  float4 *q_beta[2];   // throughput rgb, pdf of the last BSDF sample
  float4 *q_L[2];      // radiance rgb, MIS weight of BSDF sampling (W_MATS)
  uint4 *q_aux[2];     // pcg32 state (lo, hi), pixel index, local sample index
  uint32_t *q_meta[2]; // depth | flags << 16 | item slot << 20
  float4 *hit;         // closest hit of the current queue: t, u, v, prim bits
  // shadow queue: {source queue index, target (next-queue index, or sample
  // buffer index), mask of emitters 0..31, item slot | kShadowLast for a
  // finished path}; per (emitter, source index) rays at e * n_paths + i; the
  // mask words of emitters 32 and up at (w - 1) * n_paths + i
  uint4 *s_ent;
  float4 *sh_o, *sh_d, *sh_c;
  uint32_t *sh_mask;
  uint32_t *counters;            // at ctr(k): [0],[1] extension queues, [2] shadow queue, [4],[5] fetch
                                 // counters
  unsigned long long *stats;     // [0] extension rays, [1] shadow rays, [2..5] traversal counters
  unsigned long long *req_invalid;  // per request slot: samples the film resolve dropped
};

// Counter k of the bounce kernels lives at word ctr(k), 256 B from the
// next: atomics to one memory line serialize at the memory-side atomic unit
// even when they target different words, and k_shade bumps several counters
// per block.
constexpr uint32_t kCtrStride = 64;
__host__ __device__ constexpr uint32_t ctr(uint32_t k) { return k * kCtrStride; }
constexpr uint32_t kCounterWords = ctr(8);
// A chunk's finished-sample count is split over kDoneWays words 16 KiB
// apart (a block adds to way blockIdx & (kDoneWays - 1)), so that the waves
// of a large chunk do not all serialize on one memory line.
constexpr uint32_t kDoneWays = 8;

// grid_items: the most samples the launch may have to start (its grid)
hipError_t launch_regen(const RenderArgs &A, int qi, uint32_t grid_items, hipStream_t st);
hipError_t launch_extend(const RenderArgs &A, int qi, uint32_t grid, hipStream_t st);
hipError_t launch_shade(const RenderArgs &A, int qi, uint32_t grid, hipStream_t st);
hipError_t launch_shadow(const RenderArgs &A, int qn, uint32_t grid, hipStream_t st);
// tiles: allow the LDS-tiled kernel (whole images, filter halo <= 4 pixels);
// both kernels give the same sums bit for bit.
hipError_t launch_resolve(const RenderArgs &A, uint32_t slot, const DevRequest &rq, uint32_t npix, bool tiles,
                          hipStream_t st);
hipError_t launch_finalize(const void *accum, float *rgb, uint32_t n, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint32_t n, hipStream_t st);
// The library's fixed per-device scratch (one 256-B block) used by the
// entry point that allocates nothing: setup takes its device address once per
// device (call with that device current; repeated calls return at once);
// noalloc_scratch returns that address, or null before setup.
constexpr uint32_t kNoAllocScratchWords = 64;
constexpr int kNoAllocMaxDevices = 64;
hipError_t noalloc_scratch_setup(int device);
uint32_t *noalloc_scratch(int device);
// zeroes every way of item slot `slot`'s finished count
hipError_t launch_clear_done(uint32_t *item_done, uint32_t slot, hipStream_t st);
// Shading records of nt triangles from the global vertex arrays (N / UV may
// be null), face indices and per-triangle mesh ids.
hipError_t launch_shade_tris(const float *V, const float *N, const float *UV, const uint32_t *F,
                             const uint32_t *prim_mesh, const DevMesh *meshes, ShadeTri *out, uint32_t nt,
                             hipStream_t st);

// Closest / any hits of n rays (hk_ray / hk_hit layout) on persistent waves
// (counter: one device word, zeroed by the launch; refill / batch as the
// render's "refill" / "fetch_batch").
hipError_t launch_intersect_persistent(const TraceNode *nodes, const TraceTri *tris, int32_t root, const void *rays,
                                       void *hits, uint32_t n, int any, uint32_t *counter, uint32_t refill,
                                       uint32_t batch, hipStream_t st);
hipError_t launch_bsdf_batch(const DevMaterial *mats, const float *rft, int idx, uint32_t n, const float *wi,
                             const float *wo, const int32_t *meas, const float *u, float *out_eval, float *out_pdf,
                             float *out_f, float *out_wo, int32_t *out_meas, float *out_eta, float *out_spdf,
                             hipStream_t st);


}  // namespace hk

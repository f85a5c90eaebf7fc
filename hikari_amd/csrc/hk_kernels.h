// hk_kernels.h -- launch interface between the host runtime and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "hk_shared.h"

namespace hk {

// Everything one wavefront segment touches (passed by value to every kernel).
// A segment is a list of work items (fixed per-request sample chunks) whose
// samples are numbered 0..n_samples-1; n_paths path slots are recycled: a
// finished path writes its radiance to sample_L[sample id] and frees its
// slot, and k_regen starts the next samples in the freed slots.
struct RenderArgs {
  DevScene sc;
  const DevRequest *reqs;
  const WorkItem *items;
  uint32_t n_items;
  uint32_t n_paths;    // path slots
  uint32_t n_samples;  // samples of the segment
  float4 *sample_L;    // final radiance per sample (rgb, unused)
  uint32_t *free_slots;  // stack of free path slots, size in counters[3]
  // path state (indexed by path slot)
  float4 *ps_beta;   // throughput rgb, pdf of the last BSDF sample
  float4 *ps_L;      // radiance rgb, MIS weight of BSDF sampling (W_MATS)
  uint64_t *ps_rng;  // pcg32 state (stream = pixel)
  uint32_t *ps_pix;  // pixel index inside the request
  uint32_t *ps_sid;  // sample id inside the segment
  uint32_t *ps_meta; // depth | flags << 16
  // extension-ray queues (ping-pong, indexed by queue slot)
  uint32_t *q_path[2];
  float4 *q_o[2];    // origin, tmin
  float4 *q_d[2];    // direction, tmax
  float4 *q_hit[2];  // t, u, v, prim bits
  // shadow queue (slot -> path, emitter mask) + per (emitter, path) rays
  uint32_t *s_path, *s_mask;
  float4 *sh_o, *sh_d, *sh_c;
  uint32_t *counters;            // [0],[1] extension queues, [2] shadow queue, [3] free slots
  unsigned long long *stats;     // [0] extension rays, [1] shadow rays
};

hipError_t launch_intersect(const TraceNode *nodes, const TraceTri *tris, int32_t root, const void *rays, void *hits,
                            uint32_t n, int any, hipStream_t st);
hipError_t launch_bsdf_batch(const DevMaterial *mats, const float *rft, int idx, uint32_t n, const float *wi,
                             const float *wo, const int32_t *meas, const float *u, float *out_eval, float *out_pdf,
                             float *out_f, float *out_wo, int32_t *out_meas, float *out_eta, float *out_spdf,
                             hipStream_t st);
hipError_t launch_regen(const RenderArgs &A, int qi, uint32_t q_base, uint32_t free_top, uint32_t n_new,
                        uint32_t g0, hipStream_t st);
hipError_t launch_extend(const RenderArgs &A, int qi, uint32_t grid, hipStream_t st);
hipError_t launch_shade(const RenderArgs &A, int qi, uint32_t grid, hipStream_t st);
hipError_t launch_shadow(const RenderArgs &A, uint32_t grid, hipStream_t st);
hipError_t launch_resolve(const RenderArgs &A, uint32_t item, uint32_t npix, hipStream_t st);
hipError_t launch_finalize(const void *accum, float *rgb, uint32_t n, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint32_t n, hipStream_t st);
hipError_t launch_iota_u32(uint32_t *p, uint32_t n, hipStream_t st);

}  // namespace hk

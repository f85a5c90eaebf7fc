// hk_kernels.hip -- CDNA4 (gfx950) kernels of the path-tracing core.
//
//  * trace_ray<ANY>: 4-wide BVH traversal, stack of (ref, tnear) pairs with the
//    top kLdsStack entries in LDS and the rest in private memory, exact FP32
//    Moller-Trumbore in leaves (replaces Acceleration::RayIntersect).
//  * k_intersect: batched rayIntersect over hk_ray / hk_hit arrays.
//  * k_bsdf_batch: batched BSDF::Sample / Eval / Pdf for parity tests.
//  * wavefront pathMIS (PathMISIntegrator::Li, src/integrator/
//    PathMISIntegrator.cpp:16-157): [k_regen -> k_extend -> k_shade ->
//    k_shadow]* -> k_resolve.  Path state lives in the extension queues
//    (SoA, compacted with wave ballots, one atomic per block) so every kernel
//    reads and writes it coalesced; k_regen tops the queue up with new
//    samples, so every bounce launch stays full until the segment's samples
//    run out.
//    Film accumulation is owned per output pixel (gather in k_resolve), so
//    results are deterministic and independent of scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>

#include "hk_bsdf.h"
#include "hk_cdf.h"
#include "hk_kernels.h"

namespace hk {

constexpr int kBlock = 256;          // threads per block for every kernel here
#ifndef HK_LDS_STACK
#define HK_LDS_STACK 10
#endif
constexpr int kLdsStack = HK_LDS_STACK;  // stack entries per lane kept in LDS
constexpr int kPrivStack = 96 - kLdsStack;  // further entries in private memory
constexpr float kInfF = __builtin_huge_valf();
constexpr int32_t kDone = kEmptyRef;
// A lane whose next reference is on its stack (a leaf ref is ~code with code
// < 2^31 - 64, so it never takes this value).
constexpr int32_t kPop = kEmptyRef + 1;
__device__ __forceinline__ bool is_leaf(int32_t ref) { return ref < 0 && ref > kPop; }

// ------------------------------------------------------------------ traversal
// Traversal stack of one lane: the top kLdsStack entries in LDS, the rest in
// a private array.  The array is declared by the caller, apart from this
// struct, so that only the array lives in scratch memory and sp, the
// pointers and the ray state stay in registers.
// entry = (tnear bits << 32) | ref; the LDS pointer keeps its address space
typedef __attribute__((address_space(3))) uint64_t lds_u64;
struct TraceStack {
  lds_u64 *lds;     // kLdsStack x kBlock entries
  uint64_t *priv;   // kPrivStack entries
  int tid;
  int sp;
  __device__ __forceinline__ void push(int32_t ref, float t) {
    uint64_t e = ((uint64_t)__float_as_uint(t) << 32) | (uint32_t)ref;
    if (sp < kLdsStack) lds[sp * kBlock + tid] = e;
    else priv[sp - kLdsStack] = e;
    sp++;
  }
  // One pop per call: the top entry's ref if its tnear is within tlim; a
  // culled entry leaves kPop (more entries below) or kDone (stack empty), so
  // a lane skips culled entries one per traversal iteration instead of the
  // whole wave waiting in a loop for its slowest lane.  The empty asm
  // statement on the overflow branch keeps the two loads in separate branches
  // (a ds_read and a scratch load); without it the compiler merges them into
  // one generic-address load that is slow for LDS.
  __device__ __forceinline__ int32_t pop(float tlim) {
    if (sp == 0) return kDone;
    sp--;
    uint64_t e;
    if (sp < kLdsStack) {
      e = lds[sp * kBlock + tid];
    } else {
      e = priv[sp - kLdsStack];
      __asm__ volatile("" : "+v"(e));
    }
    if (__uint_as_float((uint32_t)(e >> 32)) <= tlim) return (int32_t)(uint32_t)e;
    return sp > 0 ? kPop : kDone;
  }
};

struct TraceCount {
  uint32_t nodes = 0, tris = 0;
};

// One lane's traversal state (closest or any hit over the BVH4).  Internal
// nodes are processed by node_step, leaves by leaf_tris; a lane is done when
// ref == kDone (kPop: its next reference is on the stack).  Slab tests are
// widened by a few ulp (tnear <= tfar + |tfar| 2^-21), which covers the
// rounding of (lo - o) * (1/d); the host pads triangle boxes for the
// inexactness of the triangle test.
//
// Near / far slab test (node_step<true>): per axis the ray's sign of 1/d
// picks which stored plane is the entry (near) and which the exit (far)
// plane, so each plane costs one FMA, plane * (1/d) - o * (1/d), and no
// min / max.  The product o * (1/d) is kept rounded up (cu*) for entry
// distances and rounded down (cd*) for exit distances: a computed entry is
// then never later, and an exit never earlier, than (plane - o) * (1/d) up to
// the FMA's own final rounding, which the widening covers.  Where o * (1/d)
// overflows, cu or cd is infinite or NaN and that axis stops constraining
// the box (fmaxf / fminf drop NaN), which is conservative.
// x + |x| 2^-21: for x >= 0 bit for bit x * (1 + 2^-21) (one rounding of the
// same exact value); for x < 0 it moves x up, where a product would move it
// down and reject boxes of rays whose t range is negative.
__device__ __forceinline__ float widen(float x) { return fmaf(fabsf(x), 4.76837158e-7f, x); }
struct RayTrav {
  vec3 o, d;
  float ix, iy, iz, tmin;
  float cux, cuy, cuz, cdx, cdy, cdz;  // o * (1/d) rounded up / down, per axis
  uint32_t offx, offy, offz;           // byte offset of the near plane array in a node
  float bt, bu, bv;
  int32_t bprim;
  int32_t ref;
  TraceStack st;
};

__device__ __forceinline__ void trav_begin(RayTrav &R, uint64_t *lds_stack, uint64_t *priv_stack, int32_t root, vec3 o,
                                           vec3 d, float tmin, float tmax) {
  R.o = o;
  R.d = d;
  R.ix = fabsf(d.x) > 1e-30f ? 1.0f / d.x : copysignf(1e30f, d.x);
  R.iy = fabsf(d.y) > 1e-30f ? 1.0f / d.y : copysignf(1e30f, d.y);
  R.iz = fabsf(d.z) > 1e-30f ? 1.0f / d.z : copysignf(1e30f, d.z);
  R.tmin = tmin;
  {
    // p +- |p| 2^-23 lies at least one ulp of p away from p, beyond the
    // rounding error (half an ulp) of p = o * (1/d)
    float px = o.x * R.ix, py = o.y * R.iy, pz = o.z * R.iz;
    const float k = 1.1920929e-07f;  // 2^-23
    R.cux = fmaf(fabsf(px), k, px); R.cdx = fmaf(-fabsf(px), k, px);
    R.cuy = fmaf(fabsf(py), k, py); R.cdy = fmaf(-fabsf(py), k, py);
    R.cuz = fmaf(fabsf(pz), k, pz); R.cdz = fmaf(-fabsf(pz), k, pz);
    R.offx = R.ix >= 0.0f ? 0u : 16u;
    R.offy = R.iy >= 0.0f ? 32u : 48u;
    R.offz = R.iz >= 0.0f ? 64u : 80u;
  }
  R.bt = tmax;
  R.bu = R.bv = 0.0f;
  R.bprim = -1;
  R.ref = root;
  R.st.lds = (lds_u64 *)lds_stack;
  R.st.priv = priv_stack;
  R.st.tid = threadIdx.x;
  R.st.sp = 0;
}

// Internal node R.ref: pushes the hit children far to near, continues with
// the nearest one (or the next stack entry that can still hold a hit).  Any-
// hit rays keep the near-to-far order too: taking the children in node order
// saves the sort but finds occluders later (2-3% slower overall).
// NF: the near / far slab test (fewer VALU, nine more live registers per
// lane); otherwise (lo - o) * (1/d) and (hi - o) * (1/d) with min / max.
// TOPN: nodes below TOPN are read from the LDS copy `top` of the first TOPN
// nodes (0: no copy).
template <bool NF, int TOPN = 0>
__device__ __forceinline__ void node_step(const TraceNode *__restrict__ nodes, const float4 *top, RayTrav &R,
                                          TraceCount &cnt) {
#ifdef HK_TRAVERSAL_STATS
  cnt.nodes++;
#endif
  const float ix = R.ix, iy = R.iy, iz = R.iz;
  float tlim = widen(R.bt);
  float tn[4];
  int32_t r[4];
  if constexpr (NF) {
    float4 nx4, fx4, ny4, fy4, nz4, fz4;
    int4 ch;
    if (TOPN > 0 && R.ref < TOPN) {
      const float4 *q = top + 8 * R.ref;  // an LDS array: the loads are ds_read
      nx4 = q[R.offx >> 4]; fx4 = q[(R.offx >> 4) ^ 1u];
      ny4 = q[R.offy >> 4]; fy4 = q[(R.offy >> 4) ^ 1u];
      nz4 = q[R.offz >> 4]; fz4 = q[(R.offz >> 4) ^ 1u];
      float4 c4 = q[6];
      ch = make_int4(__float_as_int(c4.x), __float_as_int(c4.y), __float_as_int(c4.z), __float_as_int(c4.w));
    } else {
      const char *nb = reinterpret_cast<const char *>(nodes);
      uint32_t base = uint32_t(R.ref) << 7;
      nx4 = *reinterpret_cast<const float4 *>(nb + (base + R.offx));
      fx4 = *reinterpret_cast<const float4 *>(nb + (base + (R.offx ^ 16u)));
      ny4 = *reinterpret_cast<const float4 *>(nb + (base + R.offy));
      fy4 = *reinterpret_cast<const float4 *>(nb + (base + (R.offy ^ 16u)));
      nz4 = *reinterpret_cast<const float4 *>(nb + (base + R.offz));
      fz4 = *reinterpret_cast<const float4 *>(nb + (base + (R.offz ^ 16u)));
      ch = *reinterpret_cast<const int4 *>(nb + (base + 96u));
    }
    float nxs[4] = {nx4.x, nx4.y, nx4.z, nx4.w}, fxs[4] = {fx4.x, fx4.y, fx4.z, fx4.w};
    float nys[4] = {ny4.x, ny4.y, ny4.z, ny4.w}, fys[4] = {fy4.x, fy4.y, fy4.z, fy4.w};
    float nzs[4] = {nz4.x, nz4.y, nz4.z, nz4.w}, fzs[4] = {fz4.x, fz4.y, fz4.z, fz4.w};
    r[0] = ch.x; r[1] = ch.y; r[2] = ch.z; r[3] = ch.w;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      float a = fmaxf(fmaxf(fmaf(nxs[c], ix, -R.cux), fmaf(nys[c], iy, -R.cuy)), fmaxf(fmaf(nzs[c], iz, -R.cuz), R.tmin));
      float b = fminf(fminf(fmaf(fxs[c], ix, -R.cdx), fmaf(fys[c], iy, -R.cdy)), fminf(fmaf(fzs[c], iz, -R.cdz), tlim));
      tn[c] = (r[c] != kEmptyRef && a <= widen(b)) ? a : kInfF;
    }
  } else {
  const vec3 o = R.o;
  float4 lx4, hx4, ly4, hy4, lz4, hz4;
  int4 ch;
  if (TOPN > 0 && R.ref < TOPN) {
    const float4 *q = top + 8 * R.ref;  // an LDS array: the loads are ds_read
    lx4 = q[0]; hx4 = q[1]; ly4 = q[2]; hy4 = q[3]; lz4 = q[4]; hz4 = q[5];
    float4 c4 = q[6];
    ch = make_int4(__float_as_int(c4.x), __float_as_int(c4.y), __float_as_int(c4.z), __float_as_int(c4.w));
  } else {
    const float4 *np = reinterpret_cast<const float4 *>(nodes + R.ref);
    lx4 = np[0]; hx4 = np[1]; ly4 = np[2]; hy4 = np[3]; lz4 = np[4]; hz4 = np[5];
    ch = *reinterpret_cast<const int4 *>(np + 6);
  }
  float lxs[4] = {lx4.x, lx4.y, lx4.z, lx4.w}, hxs[4] = {hx4.x, hx4.y, hx4.z, hx4.w};
  float lys[4] = {ly4.x, ly4.y, ly4.z, ly4.w}, hys[4] = {hy4.x, hy4.y, hy4.z, hy4.w};
  float lzs[4] = {lz4.x, lz4.y, lz4.z, lz4.w}, hzs[4] = {hz4.x, hz4.y, hz4.z, hz4.w};
  r[0] = ch.x; r[1] = ch.y; r[2] = ch.z; r[3] = ch.w;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    float t0x = (lxs[c] - o.x) * ix, t1x = (hxs[c] - o.x) * ix;
    float t0y = (lys[c] - o.y) * iy, t1y = (hys[c] - o.y) * iy;
    float t0z = (lzs[c] - o.z) * iz, t1z = (hzs[c] - o.z) * iz;
    float a = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), R.tmin));
    float b = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tlim));
    tn[c] = (r[c] != kEmptyRef && a <= widen(b)) ? a : kInfF;
  }
  }
#define HK_CSWAP(i, j)                                    \
  if (tn[j] < tn[i]) {                                    \
    float tt = tn[i]; tn[i] = tn[j]; tn[j] = tt;          \
    int32_t rr = r[i]; r[i] = r[j]; r[j] = rr;            \
  }
  HK_CSWAP(0, 1) HK_CSWAP(2, 3) HK_CSWAP(0, 2) HK_CSWAP(1, 3) HK_CSWAP(1, 2)
#undef HK_CSWAP
  if (tn[3] != kInfF) R.st.push(r[3], tn[3]);
  if (tn[2] != kInfF) R.st.push(r[2], tn[2]);
  if (tn[1] != kInfF) R.st.push(r[1], tn[1]);
  R.ref = tn[0] != kInfF ? r[0] : (R.st.sp > 0 ? kPop : kDone);
}

// The triangles of leaf `leaf`: exact FP32 Moller-Trumbore (no stack pop).
// Closest hit: among equal t the largest prim id wins (the order of the
// reference's brute-force loop); any hit: ref = kDone at the first hit.
template <bool ANY>
__device__ __forceinline__ void leaf_tris(const TraceTri *__restrict__ tris, RayTrav &R, int32_t leaf,
                                          TraceCount &cnt) {
  uint32_t code = uint32_t(~leaf);
  uint32_t first = code >> 3, ntri = (code & 7u) + 1u;
#ifdef HK_TRAVERSAL_STATS
  cnt.tris += ntri;
#endif
  for (uint32_t k = 0; k < ntri; k++) {
    const float4 *tp = reinterpret_cast<const float4 *>(tris + first + k);
    float4 a = tp[0], b = tp[1], c = tp[2];
    float u, v, t;
    if (mt_test(R.o, R.d, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(b.w, c.x, c.y), R.tmin, R.bt, u, v, t)) {
      int32_t prim = __float_as_int(a.w);
      if (ANY) {
        R.bt = t; R.bu = u; R.bv = v; R.bprim = prim;
        R.ref = kDone;
        return;
      }
      if (t < R.bt || prim > R.bprim) {
        R.bt = t; R.bu = u; R.bv = v; R.bprim = prim;
      }
    }
  }
}

constexpr int kLeafEarly = 8;
// One round of speculative while-while traversal for a wave (every lane
// must call it).  Node phase: lanes descend; a lane that reaches a leaf
// parks it in `pend` and keeps descending while enough lanes of the wave are
// still looking for their first leaf (closest-hit rays park up to two
// leaves; any-hit rays one, which keeps k_shadow's registers at five waves).
// Leaf phase: every parked leaf is tested.  The lane's ray is finished when
// ref == kDone and pend == kDone.
// Visiting nodes before a parked leaf has lowered bt only costs extra
// visits: the result is the minimum t (largest prim among equal t) in any
// order.
template <bool ANY, bool NF = true, int TOPN = 0>
__device__ __forceinline__ void trav_round(const TraceNode *__restrict__ nodes, const float4 *top,
                                           const TraceTri *__restrict__ tris, RayTrav &R, int32_t &pend,
                                           TraceCount &cnt) {
  int32_t pend2 = kDone;  // closest-hit rays park a second leaf rather than wait
  for (;;) {
    // park a leaf if a slot is free, then at most one stack read per lane
    // (one pop site: each site costs the wave its instructions whenever any
    // lane needs it)
    if (is_leaf(R.ref)) {
      if (pend == kDone) {
        pend = R.ref;
        R.ref = kPop;
      } else if (!ANY && pend2 == kDone) {
        pend2 = R.ref;
        R.ref = kPop;
      }
    }
    if (R.ref == kPop) R.ref = R.st.pop(widen(R.bt));
    // The leaf phase starts once fewer than kLeafEarly lanes still look for a
    // first leaf and some lane holds one: a few searching lanes no longer
    // keep the rest of the wave descending.  Every round makes progress
    // (node steps, stack pops, or at least one parked leaf tested), so the
    // loop ends.
    int searching = __popcll(__ballot((R.ref >= 0 || R.ref == kPop) && pend == kDone));
    if (searching == 0 || (searching < kLeafEarly && __ballot(pend != kDone) != 0ull)) break;
    if (R.ref >= 0) node_step<NF, TOPN>(nodes, top, R, cnt);
  }
  if (pend != kDone) {
    leaf_tris<ANY>(tris, R, pend, cnt);
    pend = kDone;
  }
  if (!ANY && pend2 != kDone) {
    leaf_tris<ANY>(tris, R, pend2, cnt);
    pend2 = kDone;
  }
}

// ------------------------------------------------------------------ BSDF batch
__global__ __launch_bounds__(kBlock) void k_bsdf_batch(const DevMaterial *__restrict__ mats,
                                                       const float *__restrict__ rft, int idx, uint32_t n,
                                                       const float *__restrict__ wi, const float *__restrict__ wo,
                                                       const int32_t *__restrict__ meas,
                                                       const float *__restrict__ u, float *out_eval, float *out_pdf,
                                                       float *out_f, float *out_wo, int32_t *out_meas,
                                                       float *out_eta, float *out_spdf) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    BRec r;
    r.wi = ld3(wi + 3 * i);
    r.wo = ld3(wo + 3 * i);
    r.measure = meas[i];
    r.eta = 1.0f;
    vec3 e = bsdf_eval(mats, rft, idx, r);
    out_eval[3 * i] = e.x; out_eval[3 * i + 1] = e.y; out_eval[3 * i + 2] = e.z;
    out_pdf[i] = bsdf_pdf(mats, rft, idx, r);
    BRec s;
    s.wi = r.wi;
    s.wo = v3(0.0f);
    s.eta = 1.0f;
    s.measure = kUnknownMeasure;
    U1 u1{nullptr, u + 4 * i + 2, 0, 2};
    vec3 f = bsdf_sample(mats, rft, idx, s, u[4 * i], u[4 * i + 1], u1);
    out_f[3 * i] = f.x; out_f[3 * i + 1] = f.y; out_f[3 * i + 2] = f.z;
    out_wo[3 * i] = s.wo.x; out_wo[3 * i + 1] = s.wo.y; out_wo[3 * i + 2] = s.wo.z;
    out_meas[i] = s.measure;
    out_eta[i] = s.eta;
    out_spdf[i] = bsdf_pdf(mats, rft, idx, s);
  }
}

// ------------------------------------------------------------------ wavefront helpers
__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Block-aggregated append of two predicates to two counters: one atomic per
// block and counter (the waves' counts are summed in LDS); slots in lane
// order within a wave and wave order within the block.  Every thread of the
// block must call it.  Returns the slots (0xffffffff where a predicate is
// false).
constexpr uint32_t kWaves = kBlock / 64;
struct BlockAppend {
  uint32_t cnt[2][kWaves];
  uint32_t base[2];
};
__device__ __forceinline__ void block_append2(BlockAppend &sh, uint32_t *c0, bool p0, uint32_t *c1, bool p1,
                                              uint32_t &s0, uint32_t &s1) {
  unsigned long long m0 = __ballot(p0), m1 = __ballot(p1);
  uint32_t w = threadIdx.x >> 6, lane = lane_id();
  if (lane == 0) {
    sh.cnt[0][w] = (uint32_t)__popcll(m0);
    sh.cnt[1][w] = (uint32_t)__popcll(m1);
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    uint32_t k = threadIdx.x, tot = 0;
    for (uint32_t v = 0; v < kWaves; v++) tot += sh.cnt[k][v];
    sh.base[k] = tot ? atomicAdd(k == 0 ? c0 : c1, tot) : 0u;
  }
  __syncthreads();
  uint32_t b0 = sh.base[0], b1 = sh.base[1];
  for (uint32_t v = 0; v < w; v++) {
    b0 += sh.cnt[0][v];
    b1 += sh.cnt[1][v];
  }
  unsigned long long below = lane == 0 ? 0ull : ((1ull << lane) - 1ull);
  s0 = p0 ? b0 + (uint32_t)__popcll(m0 & below) : 0xffffffffu;
  s1 = p1 ? b1 + (uint32_t)__popcll(m1 & below) : 0xffffffffu;
  __syncthreads();  // sh is reused by the next call
}

// Per-wave reservation of queue entries for the persistent kernels.  A wave
// takes batch x 64 entries per atomic on the shared fetch counter (fewer
// atomics on its memory line) and hands them to its idle lanes over several
// refills; when little of the queue is left it takes only what its idle
// lanes need, so that no wave holds a backlog while others run dry.  Entries
// are handed out exactly once whatever the batch: [next, end) is the wave's
// reserved, not yet started range.
struct WaveFetch {
  uint32_t next, end;
  uint32_t seen;  // the counter's value at the wave's last reservation
  bool dry;       // the counter has passed the queue length
};
__device__ __forceinline__ void fetch_init(WaveFetch &f) {
  f.next = f.end = f.seen = 0u;
  f.dry = false;
}
__device__ __forceinline__ bool fetch_exhausted(const WaveFetch &f) { return f.dry && f.next == f.end; }
// Gives idle lanes (bits set in `idle`, the same mask on all lanes) the next
// entries of the wave's range, in lane order, reserving a new range first
// when the current one is used up; lanes beyond the range's end stay idle
// until the next call.  Returns the lane's entry or 0xffffffff.  Every lane
// calls it.
__device__ __forceinline__ uint32_t wave_fetch(WaveFetch &f, uint32_t *counter, uint32_t count, uint32_t batch,
                                               unsigned long long idle) {
  uint32_t lane = lane_id();
  uint32_t n_idle = (uint32_t)__popcll(idle);
  if (f.next == f.end && !f.dry) {
    uint32_t want = n_idle;
    uint32_t left = count - min(f.seen, count);
    if (batch > 1u && left > gridDim.x * kWaves * batch * 64u) want = max(want, batch * 64u);
    uint32_t leader = (uint32_t)__ffsll((long long)idle) - 1u;
    uint32_t b = 0;
    if (lane == leader) b = atomicAdd(counter, want);
    // readlane (not a shuffle) keeps the reservation in scalar registers
    b = (uint32_t)__builtin_amdgcn_readlane((int)b, (int)leader);
    f.seen = b;
    f.next = min(b, count);
    f.end = min(b + want, count);
    if (b + want >= count) f.dry = true;
  }
  uint32_t take = min(f.end - f.next, n_idle);
  uint32_t idx = 0xffffffffu;
  if ((idle >> lane) & 1ull) {
    uint32_t r = (uint32_t)__popcll(idle & (lane == 0 ? 0ull : ((1ull << lane) - 1ull)));
    if (r < take) idx = f.next + r;
  }
  f.next += take;
  return idx;
}

// Adds a lane's traversal counters to stats[0] (node visits), stats[1]
// (triangle tests): one atomic per wave.
__device__ __forceinline__ void count_flush(unsigned long long *stats, TraceCount tc) {
  unsigned long long n = tc.nodes, t = tc.tris;
  for (int off = 32; off > 0; off >>= 1) {
    n += __shfl_down(n, off);
    t += __shfl_down(t, off);
  }
  if (lane_id() == 0) {
    atomicAdd(&stats[0], n);
    atomicAdd(&stats[1], t);
  }
}

// Plan entry holding launch index j: the last entry with prefix <= j.
__device__ __forceinline__ uint32_t find_plan(const RegenEntry *plan, uint32_t n, uint32_t j) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (plan[mid].prefix <= j) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Adds one to slot's finished count for every lane with pred set: one atomic
// per distinct slot in the wave, into the block's way of the count.  Every
// lane of the wave must call it.
__device__ __forceinline__ void count_done(uint32_t *item_done, bool pred, uint32_t slot) {
  unsigned long long m = __ballot(pred);
  if (m == 0ull) return;
  item_done += (blockIdx.x & (kDoneWays - 1u)) * kMaxItems;
  uint32_t lane = lane_id();
  while (m) {
    int leader = __ffsll((long long)m) - 1;
    uint32_t ls = __shfl(slot, leader);
    unsigned long long same = __ballot(pred && slot == ls);
    if ((int)lane == leader) atomicAdd(&item_done[ls], (uint32_t)__popcll(same));
    m &= ~same;
  }
}


// A 3x3 block (row stride rs) times v as the reference's Transform applies
// it to a vector (src/core/Transform.cpp:36-39): Eigen's coefficient product
// sums each row as a0 + (a1 + a2), every product and sum rounded on its own.
__device__ __forceinline__ vec3 mat3_mul(const float *m, vec3 v, int rs = 3) {
#pragma clang fp contract(off)
  return v3(m[0] * v.x + (m[1] * v.y + m[2] * v.z), m[rs] * v.x + (m[rs + 1] * v.y + m[rs + 2] * v.z),
            m[2 * rs] * v.x + (m[2 * rs + 1] * v.y + m[2 * rs + 2] * v.z));
}

// ---- emitters
struct ERec {
  vec3 p, n, wi;
  float pdf, dist;
};

// DiscretePDF1D::SampleContinuous / SampleDiscrete (src/core/DiscretePDF.cpp:32-78);
// the index search is cdf_find_guided (hk_cdf.h).
__device__ __forceinline__ float cdf_sample_cont(const float *cdf, const uint32_t *guide, const float *func, int n,
                                                 float integral, float s, float &pdf, int &idx) {
  int i = cdf_find_guided(cdf, guide, n, s);
  idx = i;
  float du = s - cdf[i];
  if (cdf[i + 1] - cdf[i] > 0.0f) du /= (cdf[i + 1] - cdf[i]);
  pdf = func[i] / integral;
  return ((float)i + du) / (float)n;
}

// AreaLight::Sample with Mesh::SamplePosition (src/emitter/AreaLight.cpp:14-35,
// src/core/Mesh.cpp:273-303)
__device__ __forceinline__ vec3 area_sample(const DevScene &sc, const DevMesh &m, vec3 ref, float sx, float sy,
                                            float s1, ERec &er) {
  int tri = cdf_find_guided(sc.cdf + m.cdf_off, sc.cdf_guide + m.cdf_off, (int)m.nt, s1);
  float sq = sqrtf(1.0f - sx);
  float al = 1.0f - sq, be = sy * sq, ga = 1.0f - al - be;
  const ShadeTri &st = sc.shade[m.first_prim + tri];
  vec3 p0 = ld3(st.p), p1 = ld3(st.p + 3), p2 = ld3(st.p + 6);
  er.p = (ga * p0 + al * p1) + be * p2;
  if (m.has_n) {
    er.n = (ga * ld3(st.n) + al * ld3(st.n + 3)) + be * ld3(st.n + 6);
  } else {
    er.n = normalized(cross(p1 - p0, p2 - p0));
  }
  vec3 wi = er.p - ref;
  er.dist = norm(wi);
  er.wi = normalized(wi);
  float gd = fabsf(dot(-er.wi, er.n));
  er.pdf = gd == 0.0f ? 0.0f : m.inv_area * ((er.dist * er.dist) / gd);
  if (er.pdf == 0.0f || isinf(er.pdf)) return v3(0.0f);
  vec3 le = dot(er.n, er.wi) < 0.0f ? ld3(m.radiance) : v3(0.0f);
  return le / er.pdf;
}

constexpr float kTwoPiF = 6.28318548202514648f;
constexpr float kPi2F = 9.8696044010893586f;  // float(pi * pi)

// EnvironmentLight::Sample (src/emitter/EnvironmentLight.cpp:40-76)
__device__ __forceinline__ vec3 env_sample(const DevScene &sc, vec3 ref, float dist, float sx, float sy, ERec &er) {
  float py, px;
  int iy, ix;
  float v = cdf_sample_cont(sc.env_marg_cdf, sc.env_marg_guide, sc.env_marg_func, sc.env_h, sc.env_marg_int, sy, py,
                            iy);
  float u = cdf_sample_cont(sc.env_row_cdf + size_t(iy) * (sc.env_w + 1), sc.env_row_guide + size_t(iy) * (sc.env_w + 1),
                            sc.env_row_func + size_t(iy) * sc.env_w, sc.env_w, sc.env_row_int[iy], sx, px, ix);
  float map_pdf = px * py;
  float theta = v * kPiF;
  float st = sinf(theta);
  if (map_pdf == 0.0f) return v3(0.0f);
  if (st == 0.0f) return v3(0.0f);
  er.pdf = map_pdf / (2.0f * kPi2F * st);
  float phi = u * kTwoPiF;
  vec3 dir = v3(st * cosf(phi), st * sinf(phi), cosf(theta));
  er.wi = mat3_mul(sc.env_to_world, dir);
  er.n = -er.wi;
  er.dist = dist;
  er.p = ref + er.wi * dist * 2.0f;
  vec3 rad = ld3(sc.env_tex + 3 * (size_t(iy) * sc.env_w + ix));
  if (!color_valid(rad)) return v3(0.0f);
  return rad * sc.env_scale / er.pdf;
}
// EnvironmentLight::Eval (src/emitter/EnvironmentLight.cpp:92-111)
__device__ __forceinline__ vec3 env_eval(const DevScene &sc, vec3 wi) {
  vec3 l = mat3_mul(sc.env_to_local, wi);
  float theta = acosf(l.z), phi = atan2f(l.y, l.x);
  if (phi < 0.0f) phi += 2.0f * kPiF;
  if (isnan(theta) || isnan(phi)) return v3(0.0f);
  float X = clampf(phi / kTwoPiF * (float)sc.env_w, 0.0f, (float)sc.env_w - 1.0f);
  float Y = clampf(theta / kPiF * (float)sc.env_h, 0.0f, (float)sc.env_h - 1.0f);
  vec3 rad = ld3(sc.env_tex + 3 * (size_t((int)Y) * sc.env_w + (int)X));
  if (!color_valid(rad)) return v3(0.0f);
  return rad * sc.env_scale;
}

// Triangle::PostIntersect (src/core/Mesh.cpp:67-172): position and shading frame.
__device__ __forceinline__ void post_intersect(const DevScene &sc, int32_t prim, float u, float v, vec3 &P, Frame &sh,
                                               uint32_t &mesh_id) {
  // the whole 128-B record in seven independent 16-B loads (one line)
  const float4 *rec = reinterpret_cast<const float4 *>(sc.shade + prim);
  float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3], r4 = rec[4], r5 = rec[5], r6 = rec[6];
  mesh_id = __float_as_uint(r6.x);
  uint32_t attr = __float_as_uint(r6.z);
  float b0 = 1.0f - (u + v), b1 = u, b2 = v;
  vec3 p0 = v3(r0.x, r0.y, r0.z), p1 = v3(r0.w, r1.x, r1.y), p2 = v3(r1.z, r1.w, r2.x);
  P = (b0 * p0 + b1 * p1) + b2 * p2;
  if (attr & kShadeHasN) {
    vec3 n0 = v3(r2.y, r2.z, r2.w), n1 = v3(r3.x, r3.y, r3.z), n2 = v3(r3.w, r4.x, r4.y);
    vec3 ns = (b0 * n0 + b1 * n1) + b2 * n2;
    if (attr & kShadeHasUV) {
      vec3 dp1 = p1 - p0, dp2 = p2 - p0;
      if (norm(cross(dp1, dp2)) == 0.0f) {
        sh = frame_from_n(normalized(cross(p1 - p0, p2 - p0)));
        return;
      }
      float du1x = r5.x - r4.z, du1y = r5.y - r4.w;
      float du2x = r5.z - r4.z, du2y = r5.w - r4.w;
      float det = du1x * du2y - du1y * du2x;
      if (det == 0.0f) {
        sh = frame_from_n(normalized(ns));
      } else {
        float idet = 1.0f / det;
        vec3 dpdu = (du2y * dp1 - du1y * dp2) * idet;
        float iln = 1.0f / norm(ns);
        sh = frame_from_n_dpdu(ns * iln, dpdu);
      }
    } else {
      sh = frame_from_n(normalized(ns));
    }
  } else {
    sh = frame_from_n(normalized(cross(p1 - p0, p2 - p0)));
  }
}

// q_meta: depth (18 bits) | flags (2 bits) << 18 | item slot (12 bits) << 20.
// Flags: bit 0 the last BSDF sample was discrete (pathMIS); bit 1 the last
// surface was not diffuse (pathEMS, which keeps that surface's emitter + 1 in
// the W_MATS slot of q_L).
constexpr uint32_t kDepthMask = 0x3ffffu;
constexpr uint32_t kFlagShift = 18;
constexpr uint32_t kFlagPrevDiscrete = 1u;
constexpr uint32_t kFlagLastSpecular = 2u;
constexpr uint32_t kShadowLast = 0x80000000u;  // shadow entry flag: the path ended at this bounce

// ------------------------------------------------------------------ k_regen
// Starts the n_new samples of this bounce's plan: their camera paths go to
// queue qi after the q_base continuing paths already there.  Per sample:
// RenderBlock's prologue (src/Main.cpp:22-31) + PerspectiveCamera::SampleRay
// (src/camera/PerspectiveCamera.cpp:77-102).  A sample's path depends only on
// (request, pixel, sample index), never on its queue position.
constexpr uint32_t kSlotShift = 20;

// Pixel order of a whole image's sample slots: T x T tiles (T a power of two),
// row-major inside a tile, tiles left to right in bands of T rows (the last
// band and the last tile of a band may be smaller), so that a wave's camera
// rays cover a compact patch; T = 1 is scanline order.  tile_pixel maps a
// rank to the pixel, tile_rank back.
__device__ __forceinline__ uint32_t tile_pixel(uint32_t r, uint32_t W, uint32_t H, uint32_t T) {
  uint32_t band = r / (T * W), y0 = band * T;
  uint32_t hb = min(T, H - y0);
  uint32_t rr = r - y0 * W;
  uint32_t x0 = (rr / (T * hb)) * T;
  uint32_t wb = min(T, W - x0);
  uint32_t r2 = rr - x0 * hb;
  return (y0 + r2 / wb) * W + x0 + r2 % wb;
}
__device__ __forceinline__ uint32_t tile_rank(uint32_t x, uint32_t y, uint32_t W, uint32_t H, uint32_t T) {
  uint32_t y0 = y & ~(T - 1u), x0 = x & ~(T - 1u);
  uint32_t hb = min(T, H - y0), wb = min(T, W - x0);
  return y0 * W + x0 * hb + (y - y0) * wb + (x - x0);
}

__global__ __launch_bounds__(kBlock) void k_regen(RenderArgs A, int qi) {
  const BounceHdr h = *A.hdr;
  const uint32_t q_base = h.q_base, n_new = h.n_new;
  uint32_t stride = gridDim.x * blockDim.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    A.counters[ctr(qi)] = q_base + n_new;
    A.counters[ctr(4)] = 0u;  // k_extend's fetch counter (persistent build)
  }
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n_new; j += stride) {
    const RegenEntry e = A.plan[find_plan(A.plan, h.n_plan, j)];
    const WorkItem it = A.items[e.slot];
    const DevRequest &rq = A.reqs[it.req];
    uint32_t local = e.local0 + (j - e.prefix);
    uint32_t rank = local % it.npix, s = it.s0 + local / it.npix;
    uint32_t q = rq.tile ? tile_pixel(rank, (uint32_t)rq.width, (uint32_t)rq.height, rq.tile) : rq.pix0 + rank;
    uint32_t x = q % (uint32_t)rq.width, y = q / (uint32_t)rq.width;
    Rng rng = rng_for_path(rq.seed, q, s);
    float jx = rng.next(), jy = rng.next();
    rng.next();
    rng.next();  // aperture sample, unused by the pinhole camera
    float px = (float)x + jx, py = (float)y + jy;
    float X = px * rq.inv_w, Y = py * rq.inv_h;
    const float *m = rq.s2c;
    float hx = m[0] * X + m[1] * Y + m[3], hy = m[4] * X + m[5] * Y + m[7];
    float hz = m[8] * X + m[9] * Y + m[11], hw = m[12] * X + m[13] * Y + m[15];
    vec3 dir = normalized(v3(hx / hw, hy / hw, hz / hw));
    float inv_z = 1.0f / dir.z;
    const float *c = rq.c2w;
    vec3 o = v3(c[3] / c[15], c[7] / c[15], c[11] / c[15]);
    vec3 d = mat3_mul(c, dir, 4);
    uint32_t pos = q_base + j;
    A.q_o[qi][pos] = make_float4(o.x, o.y, o.z, rq.near_clip * inv_z);
    A.q_d[qi][pos] = make_float4(d.x, d.y, d.z, rq.far_clip * inv_z);
    A.q_beta[qi][pos] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    A.q_L[qi][pos] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    A.q_aux[qi][pos] = make_uint4((uint32_t)rng.state, (uint32_t)(rng.state >> 32), q, local);
    A.q_meta[qi][pos] = e.slot << kSlotShift;
  }
}

// ------------------------------------------------------------------ k_extend
// Closest hit of every queued ray (Scene::RayIntersect, src/core/Scene.cpp:
// 117-120).  Persistent lanes: a wave keeps tracing and, whenever at least
// A.refill_extend of its lanes have finished their rays, gives them new queue
// entries from its reserved range (wave_fetch: A.fetch_batch x 64 entries per
// atomic on counters[4], cleared by k_regen / k_shade before), so a few long
// rays no longer hold a whole wave.  Both are run-time options; 40 idle lanes
// and 4 x 64 entries are fastest on the benchmark scenes (k_shadow: 16 idle
// lanes).
constexpr uint32_t kNoRay = 0xffffffffu;

// k_extend is held at 7 waves per SIMD (at most 72 VGPRs: it fits without
// spills); left alone the near / far slab test's per-ray constants take it to
// 80 VGPRs and 6 waves, which loses the gain.  k_shadow (95 VGPRs, 5 waves)
// keeps the min / max slab test: the constants would cost it a wave.
constexpr int kExtendMinWaves = 7;
__global__ __launch_bounds__(kBlock, kExtendMinWaves) void k_extend(RenderArgs A, int qi) {
  __shared__ uint64_t stack[kLdsStack * kBlock];
  // the top of the tree, read by every descent (2.5 KiB: with the stack the
  // block's LDS stays within a seventh of the CU's 160 KiB)
  __shared__ float4 s_top[kTopNodes * 8];
  {
    uint32_t n4 = uint32_t(min(A.sc.n_nodes, kTopNodes)) * 8u;
    const float4 *src = reinterpret_cast<const float4 *>(A.sc.nodes);
    for (uint32_t k = threadIdx.x; k < n4; k += kBlock) s_top[k] = src[k];
    __syncthreads();
  }
  uint32_t count = A.counters[ctr(qi)];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    A.counters[ctr(qi ^ 1)] = 0u;
    A.counters[ctr(2)] = 0u;
    A.counters[ctr(5)] = 0u;  // k_shadow's fetch counter
    atomicAdd(&A.stats[0], (unsigned long long)count);
  }
  const float4 *qo = A.q_o[qi], *qd = A.q_d[qi];
  float4 *qh = A.hit;
  TraceCount tc;
  RayTrav R;
  uint64_t priv[kPrivStack];
  R.ref = kDone;
  int32_t pend = kDone;  // parked leaf
  uint32_t rid = kNoRay;
  WaveFetch wf;
  fetch_init(wf);
  for (;;) {
    unsigned long long idle = __ballot(rid == kNoRay);
    int n_idle = __popcll(idle);
    if (!fetch_exhausted(wf) && n_idle >= (int)A.refill_extend) {
      uint32_t idx = wave_fetch(wf, &A.counters[ctr(4)], count, A.fetch_batch, idle);
      if (idx != kNoRay) {
        rid = idx;
        float4 a = qo[idx], b = qd[idx];
        trav_begin(R, stack, priv, A.sc.root_ref, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w);
        pend = kDone;
      }
    }
    if (__ballot(rid != kNoRay) == 0ull) {
      if (fetch_exhausted(wf)) break;
      continue;  // every lane idle: the next round fetches a full wave
    }
    trav_round<false, true, kTopNodes>(A.sc.nodes, s_top, A.sc.tris, R, pend, tc);
    if (rid != kNoRay && R.ref == kDone && pend == kDone) {
      qh[rid] = R.bprim >= 0 ? make_float4(R.bt, R.bu, R.bv, __int_as_float(R.bprim))
                             : make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
      rid = kNoRay;
    }
  }
#ifdef HK_TRAVERSAL_STATS
  count_flush(A.stats + 2, tc);
#endif
}

// ------------------------------------------------------------------ batched rayIntersect
// Acceleration::RayIntersect over an array of rays with k_extend's machinery:
// persistent waves that refill from reserved ranges of the ray array
// (counter: one zeroed word) and speculative traversal rounds.  Closest hits
// equal brute force bit for bit (the closest hit does not depend on the visit
// order); any-hit queries agree on whether a hit exists.  Held at 6 waves
// per SIMD (at most 80 VGPRs).
template <bool ANY>
__global__ __launch_bounds__(kBlock, 6) void k_intersect_persistent(const TraceNode *__restrict__ nodes,
                                                                 const TraceTri *__restrict__ tris, int32_t root,
                                                                 const float4 *__restrict__ rays,
                                                                 float4 *__restrict__ hits, uint32_t n,
                                                                 uint32_t *counter, uint32_t refill, uint32_t batch) {
  __shared__ uint64_t stack[kLdsStack * kBlock];
  TraceCount tc;
  RayTrav R;
  uint64_t priv[kPrivStack];
  R.ref = kDone;
  int32_t pend = kDone;
  uint32_t rid = kNoRay;
  WaveFetch wf;
  fetch_init(wf);
  for (;;) {
    unsigned long long idle = __ballot(rid == kNoRay);
    int n_idle = __popcll(idle);
    if (!fetch_exhausted(wf) && n_idle >= (int)refill) {
      uint32_t idx = wave_fetch(wf, counter, n, batch, idle);
      if (idx != kNoRay) {
        rid = idx;
        float4 a = rays[2 * size_t(idx)], b = rays[2 * size_t(idx) + 1];
        trav_begin(R, stack, priv, root, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w);
        pend = kDone;
      }
    }
    if (__ballot(rid != kNoRay) == 0ull) {
      if (fetch_exhausted(wf)) break;
      continue;
    }
    trav_round<ANY>(nodes, nullptr, tris, R, pend, tc);
    if (rid != kNoRay && R.ref == kDone && pend == kDone) {
      hits[rid] = R.bprim >= 0 ? make_float4(R.bt, R.bu, R.bv, __int_as_float(R.bprim))
                               : make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
      rid = kNoRay;
    }
  }
}

// ------------------------------------------------------------------ k_shade
constexpr uint32_t kSortKeys = 32;  // material keys of the block sort (last two: miss, idle lane)
// One bounce of PathMISIntegrator::Li for every queued path.
// The register allocator is asked for 4 waves per SIMD (128 VGPRs, a few
// bytes of spill per lane): with 3 waves (158 VGPRs, no spill) the kernel
// waits on memory more than it gains from the spare registers.
// HK_SHADE_MIN_WAVES overrides it in tuning builds.
#ifndef HK_SHADE_MIN_WAVES
#define HK_SHADE_MIN_WAVES 4
#endif
// LDS_TABLES: the scene's material records (at most kShadeMats) and mesh
// records (at most kShadeMeshes) are copied to LDS once per block; the BSDF
// code and the hit's emission read them there instead of from global memory.
constexpr int kShadeMats = 16, kShadeMeshes = 32;
// MANY: more than 32 emitters (shadow masks beyond the entry's one word);
// scenes with at most 32 use a build without that bookkeeping.
template <bool LDS_TABLES, bool MANY>
__global__ __launch_bounds__(kBlock, HK_SHADE_MIN_WAVES) void k_shade(RenderArgs A, int qi) {
  const DevScene &sc = A.sc;
  __shared__ DevMaterial s_mats[LDS_TABLES ? kShadeMats : 1];
  __shared__ DevMesh s_meshes[LDS_TABLES ? kShadeMeshes : 1];
  const DevMaterial *mats = sc.mats;
  const DevMesh *meshes = sc.meshes;
  if constexpr (LDS_TABLES) {
    constexpr int kVecM = int(sizeof(DevMaterial) / sizeof(float4)), kVecG = int(sizeof(DevMesh) / sizeof(float4));
    for (int k = (int)threadIdx.x; k < sc.n_mats * kVecM; k += kBlock)
      reinterpret_cast<float4 *>(s_mats)[k] = reinterpret_cast<const float4 *>(sc.mats)[k];
    for (int k = (int)threadIdx.x; k < sc.n_meshes * kVecG; k += kBlock)
      reinterpret_cast<float4 *>(s_meshes)[k] = reinterpret_cast<const float4 *>(sc.meshes)[k];
    __syncthreads();
    mats = s_mats;
    meshes = s_meshes;
  }
  uint32_t count = A.counters[ctr(qi)];
  if (blockIdx.x == 0 && threadIdx.x == 0) A.counters[ctr(4)] = 0u;  // k_extend's fetch counter
  int qn = qi ^ 1;
  uint32_t stride = gridDim.x * blockDim.x;
  // Block-local material sort: the block's 256 entries are counting-sorted by
  // the material they hit (misses and idle lanes last), and lane t shades
  // the entry of rank t, so waves see one material where they can.  Nothing
  // moves in memory; a path's result does not depend on its lane.
  __shared__ uint32_t s_cnt[kSortKeys];
  __shared__ uint32_t s_perm[kBlock];
  __shared__ BlockAppend s_app;
  for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
    uint32_t key = kSortKeys - 1;
    if (base + threadIdx.x < count) {
      int32_t hp = __float_as_int(A.hit[base + threadIdx.x].w);
      key = hp < 0 ? kSortKeys - 2 : min(sc.shade[hp].material, kSortKeys - 3);
    }
    if (threadIdx.x < kSortKeys) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    uint32_t rank = atomicAdd(&s_cnt[key], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t run = 0;
      for (uint32_t k = 0; k < kSortKeys; k++) {
        uint32_t c = s_cnt[k];
        s_cnt[k] = run;
        run += c;
      }
    }
    __syncthreads();
    s_perm[s_cnt[key] + rank] = threadIdx.x;
    __syncthreads();
    uint32_t i = base + s_perm[threadIdx.x];
    bool active = i < count;
    bool cont = false;
    // smask: emitters 0..31 with a shadow ray; sany (MANY): any emitter (the
    // words of emitters 32 and up go to A.sh_mask as they complete)
    uint32_t smask = 0, pix = 0, local = 0, islot = 0, meta_out = 0;
    bool sany = false;
    vec3 P = v3(0.0f), nd = v3(0.0f), L = v3(0.0f), beta = v3(0.0f);
    float pdf_b = 0.0f, w_mats = 0.0f;
    Rng rng;
    if (active) {
      float4 o4 = A.q_o[qi][i], d4 = A.q_d[qi][i], h4 = A.hit[i];
      float4 b4 = A.q_beta[qi][i], l4 = A.q_L[qi][i];
      uint4 aux = A.q_aux[qi][i];
      uint32_t meta = A.q_meta[qi][i];
      pix = aux.z;
      local = aux.w;
      islot = meta >> kSlotShift;
      rng.state = (uint64_t)aux.x | ((uint64_t)aux.y << 32);
      rng.inc = rng_inc(pix);
      uint32_t depth = meta & kDepthMask, flags = (meta >> kFlagShift) & 3u;
      const int mode = sc.integrator;  // 0 pathMIS, 1 pathEMS, 2 pathMATS, 3 whitted
      beta = v3(b4.x, b4.y, b4.z);
      L = v3(l4.x, l4.y, l4.z);
      pdf_b = b4.w;
      w_mats = l4.w;
      vec3 ro = v3(o4.x, o4.y, o4.z), rd = v3(d4.x, d4.y, d4.z);
      int32_t prim = __float_as_int(h4.w);
      if (prim < 0) {  // escaped (PathMISIntegrator.cpp:33-66)
        if (mode != 0) {  // PathEMSIntegrator.cpp:30-42, PathMATSIntegrator.cpp:28-40, WhittedIntegrator.cpp:33-52
          if (sc.has_env && !sc.force_background) L = L + beta * env_eval(sc, rd) / 1.0f;
          else if (depth == 0) L = ld3(sc.background);
        } else if (depth == 0) {
          if (sc.has_env && !sc.force_background) L = L + beta * env_eval(sc, rd) / 1.0f;
          else L = ld3(sc.background);
        } else if (sc.has_env) {
          L = L + beta * env_eval(sc, rd) / 1.0f;
        }
      } else {
        Frame fr;
        uint32_t mesh_id;
        post_intersect(sc, prim, h4.y, h4.z, P, fr, mesh_id);
        const DevMesh &mesh = meshes[mesh_id];
        int mat = (int)mesh.material;
        bool diffuse_mat = (mats[mat].flags & kMatFlagDiffuse) != 0;
        vec3 le = v3(0.0f);
        if (mesh.emitter >= 0 && mode != 0) {
          // PathMATSIntegrator.cpp:42-50 and WhittedIntegrator.cpp:54-62
          // (always); PathEMSIntegrator.cpp:44-60 (camera rays, after a
          // non-diffuse surface, or the same emitter again)
          // pathEMS keeps the last surface's emitter + 1 in the W_MATS slot
          // (unused by that estimator); a fresh path holds 1.0f there, which
          // no emitter index matches
          bool count = mode >= 2 || depth == 0 || (flags & kFlagLastSpecular) ||
                       (mode == 1 && __float_as_uint(w_mats) == (uint32_t)(mesh.emitter + 1));
          if (dot(fr.n, rd) < 0.0f) le = ld3(mesh.radiance);
          if (count) L = L + beta * le;
        } else if (mesh.emitter >= 0) {  // :73-81 (+ the MATS weight of :122-140)
          vec3 wi = P - ro;
          float dist = norm(wi);
          wi = wi / dist;
          if (depth > 0) {
            float gd = fabsf(dot(-wi, fr.n));
            float pl = gd == 0.0f ? 0.0f : mesh.inv_area * ((dist * dist) / gd);
            if (pdf_b + pl != 0.0f) w_mats = pdf_b / (pdf_b + pl);
            if (flags & kFlagPrevDiscrete) w_mats = 1.0f;
          }
          vec3 le = dot(fr.n, wi) < 0.0f ? ld3(mesh.radiance) : v3(0.0f);
          L = L + beta * w_mats * le;
        }
        vec3 wi_loc = to_local(fr, -1.0f * rd);
        // emitter sampling over every emitter (:83-109); pathEMS: only at
        // diffuse surfaces, skipping the surface's own emitter, no MIS weight
        // (PathEMSIntegrator.cpp:61-86); whitted: the same, where the
        // surface's own emitter adds Eval / Pdf |cos| Le along the shading
        // normal (WhittedIntegrator.cpp:64-98); pathMATS: none
        bool ems = mode == 0 || ((mode == 1 || mode == 3) && diffuse_mat);
        // One Eval / Pdf site serves both the emitter samples and whitted's
        // own-emitter term (each inlined BSDF copy costs registers).
        uint32_t word = 0;  // mask bits of the current 32 emitters
        for (int e = 0; ems && e < sc.n_emitters; e++) {
          if (MANY && e >= 32 && (e & 31) == 0) {  // a word of emitters 32.. is complete
            if (e > 32) A.sh_mask[size_t((e >> 5) - 2) * A.n_paths + i] = word;
            else smask = word;
            word = 0;
          }
          bool self = mode != 0 && e == mesh.emitter;
          if (self && mode != 3) continue;
          BRec br;
          br.wi = wi_loc;
          br.measure = kSolidAngle;
          br.eta = 1.0f;
          ERec er;
          vec3 ld = v3(0.0f);
          if (self) {
            br.wo = to_local(fr, fr.n);
          } else {
            float sx = rng.next(), sy = rng.next(), s1 = rng.next();
            DevEmitter em = sc.emitters[e];
            ld = em.kind == 0 ? area_sample(sc, sc.meshes[em.mesh], P, sx, sy, s1, er)
                              : env_sample(sc, P, sc.bbox_radius, sx, sy, er);
            // whitted traces every sample whose radiance is not exactly zero
            bool try_ems = mode == 3 ? !(ld.x == 0.0f && ld.y == 0.0f && ld.z == 0.0f) : !is_zero_fuzzy(ld);
            if (!try_ems) continue;
            br.wo = to_local(fr, er.wi);
          }
          float pb = (mode == 0 || self) ? bsdf_pdf(mats, sc.rft, mat, br) : 0.0f;
          vec3 fe = bsdf_eval(mats, sc.rft, mat, br);
          if (self) {
            if (pb != 0.0f) L = L + beta * (fe / pb * fabsf(br.wo.z) * le);
            continue;
          }
          // x * 1 is exact: the weight is 1 where no MIS applies
          float w_ems = (mode == 0 && er.pdf + pb != 0.0f) ? er.pdf / (er.pdf + pb) : 1.0f;
          vec3 c = beta * fe * fabsf(br.wo.z) * ld * w_ems;
          if (!(c.x == 0.0f && c.y == 0.0f && c.z == 0.0f)) {
            size_t slot = size_t(e) * A.n_paths + i;
            vec3 sd = er.p - P;
            A.sh_o[slot] = make_float4(P.x, P.y, P.z, 0.0f);
            A.sh_d[slot] = make_float4(sd.x, sd.y, sd.z, 0.0f);
            A.sh_c[slot] = make_float4(c.x, c.y, c.z, 0.0f);
            word |= 1u << (e & 31);
            if (MANY) sany = true;
          }
        }
        if (!MANY) smask = word;
        else if (ems) A.sh_mask[size_t(((sc.n_emitters - 1) >> 5) - 1) * A.n_paths + i] = word;
        // BSDF sampling (:111-120), MIS bookkeeping (:122-140), RR (:143-151).
        // whitted (WhittedIntegrator.cpp:100-113): a diffuse vertex ends the
        // path; any other continues with probability 0.95, drawn before the
        // BSDF sample, while depth < the depth property.  Below depth
        // sc.rr_start (0 unless set) the roulette's number is drawn but the
        // path continues at weight 1.
        const bool rr = depth >= (uint32_t)sc.rr_start;
        bool sample_bsdf =
            mode != 3 || (!diffuse_mat && (rng.next() < 0.95f || !rr) && depth < (uint32_t)sc.max_depth);
        if (sample_bsdf) {
          BRec br;
          br.wi = wi_loc;
          br.wo = v3(0.0f);
          br.eta = 1.0f;
          br.measure = kUnknownMeasure;
          float sx = rng.next(), sy = rng.next();
          U1 u1{&rng, nullptr, 0, 0};
          vec3 f = bsdf_sample(mats, sc.rft, mat, br, sx, sy, u1);
          nd = to_world(fr, br.wo);
          beta = beta * f;
          if (mode == 3) {
            if (rr) beta = beta * (1.0f / 0.95f);
            flags = 0u;
            cont = !(beta.x == 0.0f && beta.y == 0.0f && beta.z == 0.0f);
            if (cont) depth++;
          } else if (!is_zero_fuzzy(beta)) {
            pdf_b = bsdf_pdf(mats, sc.rft, mat, br);
            flags = 0u;
            if (br.measure == kDiscrete) {
              w_mats = 1.0f;
              flags = kFlagPrevDiscrete;
            }
            if (rng.next() < 0.95f || !rr) {
              if (rr) beta = beta * (1.0f / 0.95f);
              depth++;
              cont = depth < (uint32_t)sc.max_depth;
            }
            if (mode == 1) {
              flags = diffuse_mat ? 0u : kFlagLastSpecular;
              w_mats = __uint_as_float((uint32_t)(mesh.emitter + 1));
            }
          }
        }
      }
      meta_out = (depth & kDepthMask) | (flags << kFlagShift) | (islot << kSlotShift);
    }
    if (!MANY) sany = smask != 0u;
    uint32_t slot, ss;
    block_append2(s_app, &A.counters[ctr(qn)], cont, &A.counters[ctr(2)], sany, slot, ss);
    if (cont) {  // the path moves to the compacted next queue
      A.q_o[qn][slot] = make_float4(P.x, P.y, P.z, 1e-4f);
      A.q_d[qn][slot] = make_float4(nd.x, nd.y, nd.z, kInfF);
      A.q_beta[qn][slot] = make_float4(beta.x, beta.y, beta.z, pdf_b);
      A.q_L[qn][slot] = make_float4(L.x, L.y, L.z, w_mats);
      A.q_aux[qn][slot] = make_uint4((uint32_t)rng.state, (uint32_t)(rng.state >> 32), pix, local);
      A.q_meta[qn][slot] = meta_out;
    }
    // A finished path writes its radiance to the sample buffer: here, or in
    // k_shadow when emitter samples of this bounce are still pending (it is
    // counted as done where that final write happens).
    bool fin = active && !cont;
    uint32_t sbuf = A.items[islot].buf_base + local;
    if (fin) A.sample_L[sbuf] = make_float4(L.x, L.y, L.z, 0.0f);
    count_done(A.item_done, fin && !sany, islot);
    if (sany)
      A.s_ent[ss] = cont ? make_uint4(i, slot, smask, 0u) : make_uint4(i, sbuf, smask, islot | kShadowLast);
  }
}

// ------------------------------------------------------------------ k_shadow
// Scene::ShadowRayIntersect (src/core/Scene.cpp:122-126) for the queued
// emitter samples.  One lane owns a shadow entry (one path): it traces the
// entry's emitter rays in emitter order and adds the unoccluded
// contributions to the path's radiance, so no atomics are needed and the
// additions happen in the reference's order.  Lanes refill as in k_extend
// (fetch counter counters[5], cleared by k_extend).
template <bool MANY>
__global__ __launch_bounds__(kBlock) void k_shadow(RenderArgs A, int qn) {
  __shared__ uint64_t stack[kLdsStack * kBlock];
  // the top of the tree (10.6 KiB: at 5 waves per SIMD a block may use a
  // fifth of the CU's 160 KiB of LDS)
  __shared__ float4 s_top[kTopNodesShadow * 8];
  {
    uint32_t n4 = uint32_t(min(A.sc.n_nodes, kTopNodesShadow)) * 8u;
    const float4 *src = reinterpret_cast<const float4 *>(A.sc.nodes);
    for (uint32_t k = threadIdx.x; k < n4; k += kBlock) s_top[k] = src[k];
    __syncthreads();
  }
  uint32_t count = A.counters[ctr(2)];
  uint32_t traced = 0;
  TraceCount tc;
  RayTrav R;
  uint64_t priv[kPrivStack];
  R.ref = kDone;
  int32_t pend = kDone;  // parked leaf
  uint32_t ent = kNoRay, mask = 0, src = 0;
  // mask: the current word's emitters still to trace (word mword: emitters
  // 32 * mword ..); scenes with more than 32 emitters keep the further words
  // in A.sh_mask
  uint32_t mword = 0;
  const uint32_t last_word = MANY ? uint32_t(max(A.sc.n_emitters, 1) - 1) >> 5 : 0u;
  float4 *lp = nullptr;
  vec3 L = v3(0.0f), c = v3(0.0f);
  float lw = 0.0f;
  bool start = false, last = false;
  uint32_t islot = 0;
  uint32_t lane = lane_id();
  WaveFetch wf;
  fetch_init(wf);
  for (;;) {
    unsigned long long idle = __ballot(ent == kNoRay);
    int n_idle = __popcll(idle);
    if (!fetch_exhausted(wf) && n_idle >= (int)A.refill_shadow) {
      uint32_t idx = wave_fetch(wf, &A.counters[ctr(5)], count, A.fetch_batch, idle);
      if (idx != kNoRay) {
        ent = idx;
        uint4 e4 = A.s_ent[idx];
        src = e4.x;
        mask = e4.z;
        mword = 0;
        if (MANY)
          while (mask == 0u && mword < last_word) mask = A.sh_mask[size_t(mword++) * A.n_paths + src];
        last = (e4.w & kShadowLast) != 0u;
        islot = e4.w & ~kShadowLast;
        lp = last ? &A.sample_L[e4.y] : &A.q_L[qn][e4.y];
        float4 l4 = *lp;
        L = v3(l4.x, l4.y, l4.z);
        lw = l4.w;
        start = true;
      }
    }
    if (__ballot(ent != kNoRay) == 0ull) {
      if (fetch_exhausted(wf)) break;
      continue;
    }
    if (start) {  // the entry's next emitter ray
      start = false;
      int e = int(mword << 5) + __ffs((int)mask) - 1;
      mask &= mask - 1u;
      if (MANY)
        while (mask == 0u && mword < last_word) mask = A.sh_mask[size_t(mword++) * A.n_paths + src];
      size_t slot = size_t(e) * A.n_paths + src;
      float4 o = A.sh_o[slot], d = A.sh_d[slot], c4 = A.sh_c[slot];
      c = v3(c4.x, c4.y, c4.z);
      trav_begin(R, stack, priv, A.sc.root_ref, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), 1e-4f, 1.0f - 1e-4f);
      pend = kDone;
      traced++;
    }
    // min / max slab test (5 waves)
    trav_round<true, false, kTopNodesShadow>(A.sc.nodes, s_top, A.sc.tris, R, pend, tc);
    bool fin = false;
    if (ent != kNoRay && R.ref == kDone && pend == kDone) {
      if (R.bprim < 0) L = L + c;
      if (mask) {
        start = true;
      } else {
        *lp = make_float4(L.x, L.y, L.z, lw);
        ent = kNoRay;
        fin = last;
      }
    }
    count_done(A.item_done, fin, islot);
  }
  // one atomic per wave for the statistics
  for (int off = 32; off > 0; off >>= 1) traced += __shfl_down(traced, off);
  if (lane == 0 && traced) atomicAdd(&A.stats[1], (unsigned long long)traced);
#ifdef HK_TRAVERSAL_STATS
  count_flush(A.stats + 4, tc);
#endif
}

// ------------------------------------------------------------------ k_resolve
// ImageBlock::Put (src/core/Block.cpp:95-134) as a gather: each output pixel
// accumulates, in (sample, neighbour row, neighbour column) order, every
// sample of the chunk whose filter footprint covers it.  A sample's jitter is
// recomputed from its path stream.  This plain form serves any film.
__global__ __launch_bounds__(kBlock) void k_resolve_plain(RenderArgs A, uint32_t slot) {
  const WorkItem it = A.items[slot];
  const DevRequest &rq = A.reqs[it.req];
  int W = rq.width, H = rq.height;
  float R = rq.filter_radius;
  int rad = (int)floorf(R + 0.5f);
  uint32_t stride = gridDim.x * blockDim.x;
  uint32_t bad = 0;  // the pixel's own samples that the film drops
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < it.npix; p += stride) {
    int x = (int)(p % (uint32_t)W), y = (int)(p / (uint32_t)W);
    float4 *accum = reinterpret_cast<float4 *>(rq.accum);
    float4 acc = accum[p];
    uint32_t own = rq.tile ? tile_rank((uint32_t)x, (uint32_t)y, (uint32_t)W, (uint32_t)H, rq.tile) : p;
    for (uint32_t s = 0; s < it.k; s++) {
      float4 l4 = A.sample_L[it.buf_base + s * it.npix + own];
      if (!color_valid(v3(l4.x, l4.y, l4.z))) bad++;
    }
    for (uint32_t s = 0; s < it.k; s++) {
      for (int qy = max(0, y - rad); qy <= min(H - 1, y + rad); qy++) {
        for (int qx = max(0, x - rad); qx <= min(W - 1, x + rad); qx++) {
          uint32_t q = (uint32_t)(qy * W + qx);
          if (q >= it.npix) continue;
          Rng rng = rng_for_path(rq.seed, q, it.s0 + s);
          float jx = rng.next(), jy = rng.next();
          float cx = ((float)qx + jx) - 0.5f, cy = ((float)qy + jy) - 0.5f;
          if (x < (int)ceilf(cx - R) || x > (int)floorf(cx + R) || y < (int)ceilf(cy - R) || y > (int)floorf(cy + R))
            continue;
          uint32_t rank = rq.tile ? tile_rank((uint32_t)qx, (uint32_t)qy, (uint32_t)W, (uint32_t)H, rq.tile) : q;
          float4 l4 = A.sample_L[it.buf_base + s * it.npix + rank];
          vec3 L = v3(l4.x, l4.y, l4.z);
          if (!color_valid(L)) continue;  // invalid samples are dropped (Block.cpp:97-101)
          float wx = rq.filter[(int)(fabsf((float)x - cx) * rq.lookup)];
          float wy = rq.filter[(int)(fabsf((float)y - cy) * rq.lookup)];
          acc.x += L.x * wx * wy;
          acc.y += L.y * wx * wy;
          acc.z += L.z * wx * wy;
          acc.w += 1.0f * wx * wy;
        }
      }
    }
    accum[p] = acc;
  }
  if (bad) atomicAdd(&A.req_invalid[it.req], (unsigned long long)bad);
}

// The same accumulation for whole images, a 16x16 pixel tile per block: for
// each sample index the tile and its filter halo are staged in LDS once per
// halo pixel -- radiance, and the sample's filter weights toward each column
// and row its footprint can reach (jitter recomputed once, table lookups done
// once instead of once per neighbour) -- then every pixel sums its
// neighbours in the plain kernel's order with the same weights, so the sums
// are bit-identical.  Without branches: a dropped sample is staged as zero
// radiance and weight flag, a column or row outside a sample's footprint as
// weight 0, so what the plain kernel skips adds +0 here (products of finite
// non-negative values; the sums start at +0 and never become -0).
constexpr int kResTile = 16, kResHalo = 4;
// Dynamic LDS, sized for the filter's halo rad: radiance (w: 1 valid sample,
// 0 dropped or outside; all zero unless valid) per halo pixel, then its
// weights toward columns x - qx + rad and rows y - qy + rad (0 outside its
// footprint) (res_lds_bytes).
__host__ __device__ constexpr uint32_t res_lds_bytes(int rad) {
  return uint32_t((kResTile + 2 * rad) * (kResTile + 2 * rad)) * (16u + 8u * uint32_t(2 * rad + 1));
}
__global__ __launch_bounds__(kBlock) void k_resolve_tiles(RenderArgs A, uint32_t slot) {
  extern __shared__ float4 res_lds[];
  const WorkItem it = A.items[slot];
  const DevRequest &rq = A.reqs[it.req];
  int W = rq.width, H = rq.height;
  float R = rq.filter_radius;
  int rad = (int)floorf(R + 0.5f);  // <= kResHalo (checked on the host, which sized the LDS for it)
  int ew = kResTile + 2 * rad, taps = 2 * rad + 1;
  // the filter table, read ten times per halo pixel and sample
  __shared__ float s_filt[33];
  if (threadIdx.x < 33) s_filt[threadIdx.x] = rq.filter[threadIdx.x];
  float4 *s_l = res_lds;
  float *s_wx = reinterpret_cast<float *>(res_lds + ew * ew);
  float *s_wy = s_wx + ew * ew * taps;
  int tiles_x = (W + kResTile - 1) / kResTile, tiles_y = (H + kResTile - 1) / kResTile;
  float4 *accum = reinterpret_cast<float4 *>(rq.accum);
  const float lookup = rq.lookup;
  uint32_t bad = 0;  // dropped samples of the pixels inside this block's tiles (each counted once)
  for (int tile = blockIdx.x; tile < tiles_x * tiles_y; tile += gridDim.x) {
    int tx0 = (tile % tiles_x) * kResTile, ty0 = (tile / tiles_x) * kResTile;
    int x = tx0 + (int)threadIdx.x % kResTile, y = ty0 + (int)threadIdx.x / kResTile;
    bool own = x < W && y < H;
    float4 acc = own ? accum[y * W + x] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // the halo pixels this thread stages (at most kResPer), their sample
    // ranks, and their radiance for the next sample, loaded one sample ahead
    // so that the loads overlap the previous sample's sums
    constexpr int kResPer = ((kResTile + 2 * kResHalo) * (kResTile + 2 * kResHalo) + kBlock - 1) / kBlock;
    uint32_t rank[kResPer];
    float4 next[kResPer];
#pragma unroll
    for (int j = 0; j < kResPer; j++) {
      int e = (int)threadIdx.x + j * kBlock;
      int qx = tx0 - rad + e % ew, qy = ty0 - rad + e / ew;
      bool in = e < ew * ew && qx >= 0 && qx < W && qy >= 0 && qy < H;
      rank[j] = in ? tile_rank((uint32_t)qx, (uint32_t)qy, W, H, rq.tile) : 0xffffffffu;
      next[j] = (in && it.k > 0) ? A.sample_L[it.buf_base + rank[j]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    for (uint32_t s = 0; s < it.k; s++) {
      __syncthreads();  // the previous sample's tile has been read
#pragma unroll
      for (int j = 0; j < kResPer; j++) {
        int e = (int)threadIdx.x + j * kBlock;
        if (e >= ew * ew) break;
        int qx = tx0 - rad + e % ew, qy = ty0 - rad + e / ew;
        float4 l = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (rank[j] != 0xffffffffu) {
          uint32_t q = (uint32_t)(qy * W + qx);
          Rng rng = rng_for_path(rq.seed, q, it.s0 + s);
          float jx = rng.next(), jy = rng.next();
          float cx = ((float)qx + jx) - 0.5f, cy = ((float)qy + jy) - 0.5f;
          float4 l4 = next[j];
          if (s + 1 < it.k) next[j] = A.sample_L[it.buf_base + (s + 1) * it.npix + rank[j]];
          // invalid samples are dropped (Block.cpp:97-101)
          if (color_valid(v3(l4.x, l4.y, l4.z))) l = make_float4(l4.x, l4.y, l4.z, 1.0f);
          else if (qx >= tx0 && qx < tx0 + kResTile && qy >= ty0 && qy < ty0 + kResTile) bad++;
          int x0 = (int)ceilf(cx - R), x1 = (int)floorf(cx + R);
          int y0 = (int)ceilf(cy - R), y1 = (int)floorf(cy + R);
          for (int k = 0; k <= 2 * rad; k++) {
            int tx = qx - rad + k, ty = qy - rad + k;
            s_wx[e * taps + k] = (tx < x0 || tx > x1) ? 0.0f : s_filt[(int)(fabsf((float)tx - cx) * lookup)];
            s_wy[e * taps + k] = (ty < y0 || ty > y1) ? 0.0f : s_filt[(int)(fabsf((float)ty - cy) * lookup)];
          }
        }
        s_l[e] = l;
      }
      __syncthreads();
      if (!own) continue;
      for (int qy = max(0, y - rad); qy <= min(H - 1, y + rad); qy++) {
        for (int qx = max(0, x - rad); qx <= min(W - 1, x + rad); qx++) {
          int e = (qy - ty0 + rad) * ew + (qx - tx0 + rad);
          float wx = s_wx[e * taps + (x - qx + rad)];
          float wy = s_wy[e * taps + (y - qy + rad)];
          float4 l4 = s_l[e];
          acc.x += l4.x * wx * wy;
          acc.y += l4.y * wx * wy;
          acc.z += l4.z * wx * wy;
          acc.w += l4.w * wx * wy;
        }
      }
    }
    if (own) accum[y * W + x] = acc;
  }
  if (bad) atomicAdd(&A.req_invalid[it.req], (unsigned long long)bad);
}

__global__ __launch_bounds__(kBlock) void k_finalize(const float4 *accum, float *rgb, uint32_t n) {
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
    float4 a = accum[p];
    bool z = a.w == 0.0f;
    rgb[3 * p] = z ? 0.0f : a.x / a.w;
    rgb[3 * p + 1] = z ? 0.0f : a.y / a.w;
    rgb[3 * p + 2] = z ? 0.0f : a.z / a.w;
  }
}

// One thread per triangle: gathers the face's vertex attributes into its
// 128-B shading record, written as eight 16-B stores.
__global__ __launch_bounds__(kBlock) void k_shade_tris(const float *__restrict__ V, const float *__restrict__ N,
                                                       const float *__restrict__ UV, const uint32_t *__restrict__ F,
                                                       const uint32_t *__restrict__ prim_mesh,
                                                       const DevMesh *__restrict__ meshes, ShadeTri *__restrict__ out,
                                                       uint32_t nt) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  uint32_t mi = prim_mesh[t];
  const DevMesh &m = meshes[mi];
  bool has_n = m.has_n && N, has_uv = m.has_uv && UV;
  float r[32];
  for (int k = 0; k < 3; k++) {
    uint32_t i = F[3 * size_t(t) + k];
    for (int a = 0; a < 3; a++) {
      r[3 * k + a] = V[3 * size_t(i) + a];
      r[9 + 3 * k + a] = has_n ? N[3 * size_t(i) + a] : 0.0f;
    }
    for (int a = 0; a < 2; a++) r[18 + 2 * k + a] = has_uv ? UV[2 * size_t(i) + a] : 0.0f;
  }
  r[24] = __uint_as_float(mi);
  r[25] = __uint_as_float(m.material);
  r[26] = __uint_as_float((has_n ? kShadeHasN : 0u) | (has_uv ? kShadeHasUV : 0u));
  for (int k = 27; k < 32; k++) r[k] = 0.0f;
  float4 *o = reinterpret_cast<float4 *>(out + t);
  for (int k = 0; k < 8; k++) o[k] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
}

__global__ void k_clear_done(uint32_t *item_done, uint32_t slot) {
  if (threadIdx.x < kDoneWays) item_done[threadIdx.x * kMaxItems + slot] = 0u;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}


// ------------------------------------------------------------------ no-alloc scratch
// Fixed scratch of hk_accel_intersect_device_noalloc: the persistent
// intersect kernel's fetch counter in word 0, the rest of the 256-B block
// padding, so that no other data shares the counter's memory line.  Every
// device holds its own copy at its own address; host code uses only that
// address, taken once per device with the device current
// (noalloc_scratch_setup), never the array's name.
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];
static std::atomic<uint32_t *> g_scratch_dev[kNoAllocMaxDevices];
static std::mutex g_scratch_mu;

hipError_t noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device].load()) return hipSuccess;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_dev[device].store(static_cast<uint32_t *>(p));
  return hipSuccess;
}

uint32_t *noalloc_scratch(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
  return g_scratch_dev[device].load();
}

// ------------------------------------------------------------------ host launchers
static uint32_t grid_for(uint32_t n, uint32_t cap) {
  uint32_t g = (n + kBlock - 1) / kBlock;
  if (g > cap) g = cap;
  return g < 1 ? 1 : g;
}

hipError_t launch_bsdf_batch(const DevMaterial *mats, const float *rft, int idx, uint32_t n, const float *wi,
                             const float *wo, const int32_t *meas, const float *u, float *out_eval, float *out_pdf,
                             float *out_f, float *out_wo, int32_t *out_meas, float *out_eta, float *out_spdf,
                             hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bsdf_batch, dim3(grid_for(n, 1u << 16)), dim3(kBlock), 0, st, mats, rft, idx, n, wi, wo, meas,
                     u, out_eval, out_pdf, out_f, out_wo, out_meas, out_eta, out_spdf);
  return hipGetLastError();
}

hipError_t launch_regen(const RenderArgs &A, int qi, uint32_t grid_items, hipStream_t st) {
  hipLaunchKernelGGL(k_regen, dim3(grid_for(grid_items, 1u << 16)), dim3(kBlock), 0, st, A, qi);
  return hipGetLastError();
}
// Blocks of a kernel that are resident at once on the current device (the
// grid of the persistent kernels).
static uint32_t resident_blocks(const void *kernel) {
  int dev = 0, n_cu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1u << 16;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1u << 16;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess) return 1u << 16;
  return uint32_t(std::max(1, n_cu * per_cu));
}

hipError_t launch_intersect_persistent(const TraceNode *nodes, const TraceTri *tris, int32_t root, const void *rays,
                                       void *hits, uint32_t n, int any, uint32_t *counter, uint32_t refill,
                                       uint32_t batch, hipStream_t st) {
  if (n == 0) return hipSuccess;
  // the fetch counter may run past n by up to one reservation per wave
  if (refill < 1 || refill > 64 || batch < 1 || batch > 64 || n > (1u << 31)) return hipErrorInvalidValue;
  hipError_t e = launch_fill_u32(counter, 0u, 1u, st);
  if (e != hipSuccess) return e;
  uint32_t want = grid_for(n, 1u << 20);
  if (any) {
    static uint32_t cap = resident_blocks(reinterpret_cast<const void *>(k_intersect_persistent<true>));
    hipLaunchKernelGGL(k_intersect_persistent<true>, dim3(std::min(want, cap)), dim3(kBlock), 0, st, nodes, tris,
                       root, (const float4 *)rays, (float4 *)hits, n, counter, refill, batch);
  } else {
    static uint32_t cap = resident_blocks(reinterpret_cast<const void *>(k_intersect_persistent<false>));
    hipLaunchKernelGGL(k_intersect_persistent<false>, dim3(std::min(want, cap)), dim3(kBlock), 0, st, nodes, tris,
                       root, (const float4 *)rays, (float4 *)hits, n, counter, refill, batch);
  }
  return hipGetLastError();
}
hipError_t launch_extend(const RenderArgs &A, int qi, uint32_t grid, hipStream_t st) {
  static uint32_t cap = resident_blocks(reinterpret_cast<const void *>(k_extend));
  hipLaunchKernelGGL(k_extend, dim3(std::min(grid, cap)), dim3(kBlock), 0, st, A, qi);
  return hipGetLastError();
}
hipError_t launch_shade(const RenderArgs &A, int qi, uint32_t grid, hipStream_t st) {
  bool lds = A.sc.n_mats <= kShadeMats && A.sc.n_meshes <= kShadeMeshes, many = A.sc.n_emitters > 32;
  if (many) {
    if (lds) hipLaunchKernelGGL((k_shade<true, true>), dim3(grid), dim3(kBlock), 0, st, A, qi);
    else hipLaunchKernelGGL((k_shade<false, true>), dim3(grid), dim3(kBlock), 0, st, A, qi);
  } else {
    if (lds) hipLaunchKernelGGL((k_shade<true, false>), dim3(grid), dim3(kBlock), 0, st, A, qi);
    else hipLaunchKernelGGL((k_shade<false, false>), dim3(grid), dim3(kBlock), 0, st, A, qi);
  }
  return hipGetLastError();
}
hipError_t launch_shadow(const RenderArgs &A, int qn, uint32_t grid, hipStream_t st) {
  if (A.sc.n_emitters > 32) {
    static uint32_t cap = resident_blocks(reinterpret_cast<const void *>(k_shadow<true>));
    hipLaunchKernelGGL(k_shadow<true>, dim3(std::min(grid, cap)), dim3(kBlock), 0, st, A, qn);
  } else {
    static uint32_t cap = resident_blocks(reinterpret_cast<const void *>(k_shadow<false>));
    hipLaunchKernelGGL(k_shadow<false>, dim3(std::min(grid, cap)), dim3(kBlock), 0, st, A, qn);
  }
  return hipGetLastError();
}
hipError_t launch_resolve(const RenderArgs &A, uint32_t slot, const DevRequest &rq, uint32_t npix, bool tiles,
                          hipStream_t st) {
  int rad = (int)std::floor(rq.filter_radius + 0.5f);
  if (tiles && rq.tile && rad <= kResHalo && uint64_t(rq.width) * uint64_t(rq.height) == npix) {
    uint32_t tiles = uint32_t((rq.width + kResTile - 1) / kResTile) * uint32_t((rq.height + kResTile - 1) / kResTile);
    hipLaunchKernelGGL(k_resolve_tiles, dim3(std::min(tiles, 1u << 16)), dim3(kBlock), res_lds_bytes(rad), st, A,
                       slot);
  } else {
    hipLaunchKernelGGL(k_resolve_plain, dim3(grid_for(npix, 1u << 16)), dim3(kBlock), 0, st, A, slot);
  }
  return hipGetLastError();
}
hipError_t launch_finalize(const void *accum, float *rgb, uint32_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(grid_for(n, 1u << 16)), dim3(kBlock), 0, st, (const float4 *)accum, rgb, n);
  return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}
hipError_t launch_clear_done(uint32_t *item_done, uint32_t slot, hipStream_t st) {
  if (slot >= kMaxItems) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_clear_done, dim3(1), dim3(64), 0, st, item_done, slot);
  return hipGetLastError();
}
hipError_t launch_shade_tris(const float *V, const float *N, const float *UV, const uint32_t *F,
                             const uint32_t *prim_mesh, const DevMesh *meshes, ShadeTri *out, uint32_t nt,
                             hipStream_t st) {
  if (nt == 0) return hipSuccess;
  hipLaunchKernelGGL(k_shade_tris, dim3((nt + kBlock - 1) / kBlock), dim3(kBlock), 0, st, V, N, UV, F, prim_mesh,
                     meshes, out, nt);
  return hipGetLastError();
}

}  // namespace hk

// hk_bvh_gpu.hip -- device construction of the 4-wide BVH the traversal
// kernels walk, for callers that cannot wait for the host SAH build (scene
// admission while serving).  Replaces the role of HLBVHAcceleration::Build
// (src/acceleration/HLBVHAcceleration.cpp:81-204: Morton codes of triangle
// centroids, radix sort, a linear BVH over the sorted order), built here as:
//
//   1. vertex / centroid bounds (wave-reduced atomics on order-preserving
//      integer images of the floats);
//   2. 63-bit Morton codes (21 bits per axis) of the centroids, sorted with
//      their triangle ids (rocPRIM radix sort);
//   3. the binary radix tree of the sorted codes (Karras, "Maximizing
//      Parallelism in the Construction of BVHs, Octrees, and k-d Trees",
//      HPG 2012), equal codes ordered by position;
//   4. a bottom-up pass (the second child to arrive builds its parent) that
//      sets boxes and SAH costs and turns subtrees of at most kLeafMax
//      triangles into leaves where that is cheaper;
//   5. above the cells of the top `cell_bits` Morton bits (the reference's
//      treelets, HLBVHAcceleration.cpp:118-191, use 12: 4 per axis) the radix
//      tree's spatial-median nodes are replaced by a bucketed SAH tree over
//      the cells' subtree roots, built on the host into the same node slots
//      (those nodes cover no contiguous range and never become leaves);
//   6. a top-down, level-synchronous collapse into Bvh4Node: each 4-wide
//      node opens the internal child of largest area until it has four
//      children, as the host builder does.
//
// Every subtree covers a contiguous range of the sorted order, so a leaf is
// a range of the TraceTri array, written in that order.  Triangle boxes carry
// the same padding as the host build.  Closest-hit results do not depend on
// the tree (nearest t, ties to the larger triangle id), so both builders give
// the same hits bit for bit.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <future>
#include <vector>

#include "hk_build.h"
#include "hk_runtime.h"

namespace hk {

namespace {

constexpr int kB = 256;

__device__ __forceinline__ uint32_t ord(float f) {  // order-preserving float -> uint32
  uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ __forceinline__ float unord(uint32_t u) {
  uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}

__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// bounds[0..5]: ord() of vertex lo xyz / hi xyz; [6..11]: centroid lo / hi.
__global__ __launch_bounds__(kB) void k_bvh_bounds(const float *__restrict__ tri9, uint32_t n, uint32_t *bounds) {
  float vlo[3] = {INFINITY, INFINITY, INFINITY}, vhi[3] = {-INFINITY, -INFINITY, -INFINITY};
  float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float *p = tri9 + 9 * size_t(i);
    for (int a = 0; a < 3; a++) {
      float x = p[a], y = p[3 + a], z = p[6 + a];
      vlo[a] = fminf(vlo[a], fminf(x, fminf(y, z)));
      vhi[a] = fmaxf(vhi[a], fmaxf(x, fmaxf(y, z)));
      float c = (x + y + z) * (1.0f / 3.0f);
      clo[a] = fminf(clo[a], c);
      chi[a] = fmaxf(chi[a], c);
    }
  }
  for (int a = 0; a < 3; a++) {
    vlo[a] = wave_min(vlo[a]);
    vhi[a] = wave_max(vhi[a]);
    clo[a] = wave_min(clo[a]);
    chi[a] = wave_max(chi[a]);
  }
  if ((threadIdx.x & 63) == 0) {
    for (int a = 0; a < 3; a++) {
      atomicMin(&bounds[a], ord(vlo[a]));
      atomicMax(&bounds[3 + a], ord(vhi[a]));
      atomicMin(&bounds[6 + a], ord(clo[a]));
      atomicMax(&bounds[9 + a], ord(chi[a]));
    }
  }
}

__device__ __forceinline__ uint64_t spread21(uint32_t x) {  // bit k -> bit 3k
  uint64_t v = x & 0x1fffffu;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

struct MortonFrame {
  float lo[3], scale[3];  // q = (c - lo) * scale, clamped to [0, 2^21 - 1]
};

__global__ __launch_bounds__(kB) void k_bvh_morton(const float *__restrict__ tri9, uint32_t n, MortonFrame f,
                                                  uint64_t *keys, uint32_t *vals) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float *p = tri9 + 9 * size_t(i);
  uint64_t k = 0;
  for (int a = 0; a < 3; a++) {
    float c = (p[a] + p[3 + a] + p[6 + a]) * (1.0f / 3.0f);
    float q = (c - f.lo[a]) * f.scale[a];
    uint32_t qi = q <= 0.0f ? 0u : (q >= 2097151.0f ? 2097151u : uint32_t(q));
    k |= spread21(qi) << a;
  }
  keys[i] = k;
  vals[i] = i;
}

// Node numbering: internal nodes 0 .. n-2 (root 0), leaf i at n-1+i.
struct Tree {
  float4 *lo, *hi;        // boxes (2n-1)
  int2 *child;            // internal: (left, right)
  int32_t *prefix;        // internal: common-prefix length of the node's keys
  int32_t *parent;        // (2n-1); root: -1
  uint2 *range;           // (2n-1): first, last sorted position
  float *cost;            // SAH cost of the subtree as built
  uint32_t *leaf;         // 1: the node is a leaf of the final tree
  uint32_t *visit;        // internal: arrivals in the bottom-up pass
  uint32_t n;
};

// Padded triangle boxes and TraceTri records, in sorted order.  The padding
// is the host builder's (relative to the triangle's extent and coordinate
// magnitude and to the scene extent), so every ray the FP32 Moller-Trumbore
// test accepts passes the slab tests of the triangle's ancestors.
__global__ __launch_bounds__(kB) void k_bvh_leaves(const float *__restrict__ tri9, const uint32_t *__restrict__ order,
                                                  float scene_ext, Tree T, TraceTri *tris) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.n) return;
  uint32_t p = order[i];
  const float *q = tri9 + 9 * size_t(p);
  float lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    float l = fminf(q[a], fminf(q[3 + a], q[6 + a])), h = fmaxf(q[a], fmaxf(q[3 + a], q[6 + a]));
    float mag = fmaxf(fabsf(l), fabsf(h));
    float pad = (h - l) * 0x1p-12f + mag * 0x1p-18f + scene_ext * 0x1p-17f + 1e-30f;
    lo[a] = l - pad;
    hi[a] = h + pad;
  }
  uint32_t node = T.n - 1 + i;
  T.lo[node] = make_float4(lo[0], lo[1], lo[2], 0.0f);
  T.hi[node] = make_float4(hi[0], hi[1], hi[2], 0.0f);
  T.range[node] = make_uint2(i, i);
  float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  T.cost[node] = 2.0f * (dx * dy + dy * dz + dz * dx);
  T.leaf[node] = 1u;
  TraceTri t;
  t.p0x = q[0]; t.p0y = q[1]; t.p0z = q[2];
  t.prim = int32_t(p);
  t.e1x = q[3] - q[0]; t.e1y = q[4] - q[1]; t.e1z = q[5] - q[2];
  t.e2x = q[6] - q[0]; t.e2y = q[7] - q[1]; t.e2z = q[8] - q[2];
  t.pad0 = t.pad1 = 0.0f;
  tris[i] = t;
}

// Common-prefix length of sorted keys i and j (-1 outside [0, n)); equal keys
// extend the prefix with their positions.
__device__ __forceinline__ int delta(const uint64_t *__restrict__ keys, uint32_t n, int i, int j) {
  if (j < 0 || j >= int(n)) return -1;
  uint64_t a = keys[i], b = keys[j];
  if (a == b) return 64 + __clz(uint32_t(i) ^ uint32_t(j));
  return __clzll(a ^ b);
}

__global__ __launch_bounds__(kB) void k_bvh_karras(const uint64_t *__restrict__ keys, Tree T) {
  int i = int(blockIdx.x * blockDim.x + threadIdx.x);
  int n = int(T.n);
  if (i >= n - 1) return;
  int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
  int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta(keys, n, i, j);
  int s = 0;  // split: t = ceil(l / 2), ceil(l / 4), ..., 1
  for (int t = (l + 1) / 2;; t = (t + 1) / 2) {
    if (s + t <= l && delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    if (t == 1) break;
  }
  int gamma = i + s * d + min(d, 0);
  int first = min(i, j), last = max(i, j);
  int left = first == gamma ? n - 1 + gamma : gamma;
  int right = last == gamma + 1 ? n - 1 + gamma + 1 : gamma + 1;
  T.child[i] = make_int2(left, right);
  T.prefix[i] = dnode;
  T.range[i] = make_uint2(uint32_t(first), uint32_t(last));
  T.parent[left] = i;
  T.parent[right] = i;
  if (i == 0) T.parent[0] = -1;
}

// Bottom-up boxes and SAH costs: the second child to arrive at a node builds
// it (acquire / release on the arrival counter make the sibling's box
// visible).  Cost model: 1.2 x area per node step, area x count per leaf.
__global__ __launch_bounds__(kB) void k_bvh_refit(Tree T) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.n) return;
  int node = T.parent[T.n - 1 + i];
  while (node >= 0) {
    uint32_t prev = __hip_atomic_fetch_add(&T.visit[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == 0u) return;  // the sibling subtree is not done yet
    int2 c = T.child[node];
    float4 l0 = T.lo[c.x], h0 = T.hi[c.x], l1 = T.lo[c.y], h1 = T.hi[c.y];
    float4 lo = make_float4(fminf(l0.x, l1.x), fminf(l0.y, l1.y), fminf(l0.z, l1.z), 0.0f);
    float4 hi = make_float4(fmaxf(h0.x, h1.x), fmaxf(h0.y, h1.y), fmaxf(h0.z, h1.z), 0.0f);
    float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    float area = 2.0f * (dx * dy + dy * dz + dz * dx);
    uint2 r = T.range[node];
    uint32_t count = r.y - r.x + 1u;
    float split = 1.2f * area + T.cost[c.x] + T.cost[c.y];
    float as_leaf = area * float(count);
    bool leaf = count <= uint32_t(kLeafMax) && as_leaf <= split;
    T.lo[node] = lo;
    T.hi[node] = hi;
    T.cost[node] = leaf ? as_leaf : split;
    T.leaf[node] = leaf ? 1u : 0u;
    __threadfence();
    node = T.parent[node];
  }
}

struct Todo {
  int32_t bnode, slot;
};

// Boxes and ranges of the nodes idx[0 .. m).
__global__ __launch_bounds__(kB) void k_bvh_gather(const int32_t *__restrict__ idx, uint32_t m, Tree T, float4 *lo,
                                                  float4 *hi, uint2 *range) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  int32_t b = idx[k];
  lo[k] = T.lo[b];
  hi[k] = T.hi[b];
  range[k] = T.range[b];
}

// Internal nodes idx[j] become (child[j], box lo[j] / hi[j]).
__global__ __launch_bounds__(kB) void k_bvh_splice(const int32_t *__restrict__ idx, const int2 *__restrict__ child,
                                                  const float4 *__restrict__ lo, const float4 *__restrict__ hi,
                                                  uint32_t m, Tree T) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  int32_t b = idx[j];
  T.child[b] = child[j];
  T.lo[b] = lo[j];
  T.hi[b] = hi[j];
  T.leaf[b] = 0u;
}

// One level of the collapse.  counters[0]: 4-wide nodes allocated,
// counters[1]: entries written to `out`, counters[2]: leaves.
__global__ __launch_bounds__(kB) void k_bvh_collapse(Tree T, const Todo *__restrict__ in, uint32_t n_in, Todo *out,
                                                    Bvh4Node *nodes, uint32_t *counters) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_in) return;
  Todo t = in[k];
  int2 c0 = T.child[t.bnode];
  int32_t ch[4] = {c0.x, c0.y, -1, -1};
  int nc = 2;
  while (nc < 4) {
    int best = -1;
    float best_a = -1.0f;
    for (int c = 0; c < nc; c++) {
      int32_t b = ch[c];
      if (T.leaf[b]) continue;
      float4 lo = T.lo[b], hi = T.hi[b];
      float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
      float a = 2.0f * (dx * dy + dy * dz + dz * dx);
      if (a > best_a) { best_a = a; best = c; }
    }
    if (best < 0) break;
    int2 oc = T.child[ch[best]];
    ch[best] = oc.x;
    ch[nc++] = oc.y;
  }
  Bvh4Node nd;
  for (int c = 0; c < 4; c++) {
    if (c >= nc) {
      nd.lo_x[c] = nd.lo_y[c] = nd.lo_z[c] = INFINITY;
      nd.hi_x[c] = nd.hi_y[c] = nd.hi_z[c] = -INFINITY;
      nd.child[c] = kEmptyRef;
      continue;
    }
    int32_t b = ch[c];
    float4 lo = T.lo[b], hi = T.hi[b];
    nd.lo_x[c] = lo.x; nd.lo_y[c] = lo.y; nd.lo_z[c] = lo.z;
    nd.hi_x[c] = hi.x; nd.hi_y[c] = hi.y; nd.hi_z[c] = hi.z;
    if (T.leaf[b]) {
      uint2 r = T.range[b];
      nd.child[c] = ~int32_t((r.x << 3) | (r.y - r.x));
      atomicAdd(&counters[2], 1u);
    } else {
      int32_t slot = int32_t(atomicAdd(&counters[0], 1u));
      nd.child[c] = slot;
      out[atomicAdd(&counters[1], 1u)] = Todo{b, slot};
    }
  }
  for (int c = 0; c < 4; c++) nd.pad[c] = 0;
  nodes[t.slot] = nd;
}

// Bucketed SAH tree over the cell roots it[s .. e) (box, centroid, triangle
// count), built on the host.  Returns an upper-node number (>= 0; node 0 is
// the first call's) or ~item for a single root; kids / boxes receive the
// upper nodes.
struct UpperItem {
  float lo[3], hi[3], cen[3];
  double count;
  int32_t node;  // radix-tree node index
};
struct UpperBox {
  float lo[3], hi[3];
};
float box_area(const float *lo, const float *hi) {
  if (!(lo[0] <= hi[0])) return 0.0f;
  float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}
void box_grow(UpperBox &b, const float *lo, const float *hi) {
  for (int a = 0; a < 3; a++) {
    b.lo[a] = std::min(b.lo[a], lo[a]);
    b.hi[a] = std::max(b.hi[a], hi[a]);
  }
}
UpperBox box_empty() {
  UpperBox b;
  for (int a = 0; a < 3; a++) {
    b.lo[a] = INFINITY;
    b.hi[a] = -INFINITY;
  }
  return b;
}
// Below depth kUpperSahDepth the splits are medians, which bounds the
// recursion at kUpperSahDepth + log2(items) frames whatever the boxes.
constexpr int kUpperSahDepth = 48;
// The two halves of a node above depth kUpperParDepth with at least
// kUpperParItems roots are built on two host threads (at most 8 at once).
constexpr int kUpperParDepth = 3, kUpperParItems = 1 << 14;
int32_t upper_build(std::vector<UpperItem> &it, int s, int e, int depth, std::vector<int2> &kids,
                    std::vector<UpperBox> &boxes) {
  if (e - s == 1) return ~s;
  int32_t u = int32_t(kids.size());
  kids.push_back(make_int2(0, 0));
  boxes.push_back(box_empty());
  UpperBox cb = box_empty();
  for (int i = s; i < e; i++) box_grow(cb, it[size_t(i)].cen, it[size_t(i)].cen);
  int axis = 0;
  for (int a = 1; a < 3; a++)
    if (cb.hi[a] - cb.lo[a] > cb.hi[axis] - cb.lo[axis]) axis = a;
  float ext = cb.hi[axis] - cb.lo[axis], lo0 = cb.lo[axis];
  constexpr int kNB = 16;
  auto bucket = [&](const UpperItem &x) {
    return std::min(kNB - 1, std::max(0, int((x.cen[axis] - lo0) / ext * float(kNB))));
  };
  int mid = -1;
  if (ext > 0.0f && depth < kUpperSahDepth) {
    UpperBox bb[kNB];
    double cnt[kNB] = {0};
    int num[kNB] = {0};
    for (int b = 0; b < kNB; b++) bb[b] = box_empty();
    for (int i = s; i < e; i++) {
      int b = bucket(it[size_t(i)]);
      box_grow(bb[b], it[size_t(i)].lo, it[size_t(i)].hi);
      cnt[b] += it[size_t(i)].count;
      num[b]++;
    }
    // right-to-left sweep: the cost terms of buckets sp+1 .. kNB-1; then a
    // left-to-right sweep adds those of 0 .. sp
    double right_cost[kNB];
    int right_num[kNB];
    UpperBox r = box_empty();
    double cr = 0;
    int nr = 0;
    for (int b = kNB - 1; b > 0; b--) {
      box_grow(r, bb[b].lo, bb[b].hi);
      cr += cnt[b];
      nr += num[b];
      right_cost[b - 1] = double(box_area(r.lo, r.hi)) * cr;
      right_num[b - 1] = nr;
    }
    double best = INFINITY;
    int split = -1;
    UpperBox l = box_empty();
    double cl = 0;
    int nl = 0;
    for (int sp = 0; sp + 1 < kNB; sp++) {
      box_grow(l, bb[sp].lo, bb[sp].hi);
      cl += cnt[sp];
      nl += num[sp];
      if (nl == 0 || right_num[sp] == 0) continue;
      double c = double(box_area(l.lo, l.hi)) * cl + right_cost[sp];
      if (c < best) { best = c; split = sp; }
    }
    if (split >= 0)
      mid = int(std::partition(it.begin() + s, it.begin() + e, [&](const UpperItem &x) { return bucket(x) <= split; }) -
                it.begin());
  }
  if (mid <= s || mid >= e) {  // equal centroids, or deep in the tree: the median on the widest axis
    mid = s + (e - s) / 2;
    std::nth_element(it.begin() + s, it.begin() + mid, it.begin() + e,
                     [&](const UpperItem &x, const UpperItem &y) { return x.cen[axis] < y.cen[axis]; });
  }
  int32_t l, r;
  if (depth < kUpperParDepth && e - s >= kUpperParItems) {
    // each half into vectors of its own (its root at 0), then appended with
    // its upper-node numbers shifted; item numbers are global already
    std::vector<int2> kl, kr;
    std::vector<UpperBox> bl, br;
    auto left = std::async(std::launch::async, [&] { return upper_build(it, s, mid, depth + 1, kl, bl); });
    int32_t rr = upper_build(it, mid, e, depth + 1, kr, br);
    int32_t ll = left.get();
    auto append = [&](const std::vector<int2> &k, const std::vector<UpperBox> &b, int32_t root) {
      int32_t base = int32_t(kids.size());
      for (int2 c : k) kids.push_back(make_int2(c.x >= 0 ? c.x + base : c.x, c.y >= 0 ? c.y + base : c.y));
      boxes.insert(boxes.end(), b.begin(), b.end());
      return root >= 0 ? root + base : root;
    };
    l = append(kl, bl, ll);
    r = append(kr, br, rr);
  } else {
    l = upper_build(it, s, mid, depth + 1, kids, boxes);
    r = upper_build(it, mid, e, depth + 1, kids, boxes);
  }
  kids[size_t(u)] = make_int2(l, r);
  for (int32_t c : {l, r}) {
    if (c >= 0) box_grow(boxes[size_t(u)], boxes[size_t(c)].lo, boxes[size_t(c)].hi);
    else box_grow(boxes[size_t(u)], it[size_t(~c)].lo, it[size_t(~c)].hi);
  }
  return u;
}

__global__ void k_bvh_init(uint32_t *bounds, Todo *q, uint32_t *counters) {
  if (threadIdx.x < 12) bounds[threadIdx.x] = (threadIdx.x % 6) < 3 ? 0xffffffffu : 0u;
  if (threadIdx.x == 0) {
    q[0] = Todo{0, 0};
    counters[0] = 1u;  // the root's slot
    counters[1] = 0u;
    counters[2] = 0u;
  }
}

uint32_t blocks(uint32_t n) { return (n + kB - 1) / kB; }

}  // namespace

void build_bvh4_device(hipStream_t st, const float *tri9_host, uint32_t n, uint32_t cell_bits, DevBvh &out) {
  if (n < 2) throw HkError(HK_ERR_INVALID, "the device BVH builder needs at least two triangles");
  if (n >= (1u << 28)) throw HkError(HK_ERR_INVALID, "too many triangles for the leaf encoding");
  DBuf tri9, bounds, keys[2], vals[2], lo, hi, child, prefix, parent, range, cost, leaf, visit, q[2], counters, temp;
  tri9.upload(tri9_host, size_t(n) * 36, st);
  size_t nn = 2 * size_t(n) - 1;
  bounds.alloc(12 * 4);
  counters.alloc(4 * 4);
  q[0].alloc(size_t(n) * sizeof(Todo));
  q[1].alloc(size_t(n) * sizeof(Todo));
  hipLaunchKernelGGL(k_bvh_init, dim3(1), dim3(64), 0, st, bounds.as<uint32_t>(), q[0].as<Todo>(),
                     counters.as<uint32_t>());
  hipLaunchKernelGGL(k_bvh_bounds, dim3(std::min<uint32_t>(blocks(n), 4096)), dim3(kB), 0, st, tri9.as<float>(), n,
                     bounds.as<uint32_t>());
  uint32_t hb[12];
  hip_check(hipMemcpyAsync(hb, bounds.p, sizeof hb, hipMemcpyDeviceToHost, st), "BVH bounds");
  hip_check(hipStreamSynchronize(st), "BVH bounds");
  float vb[12];
  for (int k = 0; k < 12; k++) vb[k] = unord(hb[k]);
  float scene_ext = std::max(vb[3] - vb[0], std::max(vb[4] - vb[1], vb[5] - vb[2]));
  if (!std::isfinite(scene_ext)) throw HkError(HK_ERR_INVALID, "BVH build: non-finite vertex position");
  MortonFrame f;
  for (int a = 0; a < 3; a++) {
    float ext = vb[9 + a] - vb[6 + a];
    f.lo[a] = vb[6 + a];
    f.scale[a] = ext > 0.0f ? 2097152.0f / ext : 0.0f;
  }
  for (auto *b : {&keys[0], &keys[1]}) b->alloc(size_t(n) * 8);
  for (auto *b : {&vals[0], &vals[1]}) b->alloc(size_t(n) * 4);
  hipLaunchKernelGGL(k_bvh_morton, dim3(blocks(n)), dim3(kB), 0, st, tri9.as<float>(), n, f, keys[0].as<uint64_t>(),
                     vals[0].as<uint32_t>());
  size_t temp_bytes = 0;
  hip_check(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys[0].as<uint64_t>(), keys[1].as<uint64_t>(),
                                      vals[0].as<uint32_t>(), vals[1].as<uint32_t>(), n, 0, 63, st),
            "radix sort size");
  temp.alloc(std::max<size_t>(temp_bytes, 16));
  hip_check(rocprim::radix_sort_pairs(temp.p, temp_bytes, keys[0].as<uint64_t>(), keys[1].as<uint64_t>(),
                                      vals[0].as<uint32_t>(), vals[1].as<uint32_t>(), n, 0, 63, st),
            "radix sort");
  lo.alloc(nn * 16);
  hi.alloc(nn * 16);
  child.alloc(size_t(n) * 8);
  prefix.alloc(size_t(n) * 4);
  parent.alloc(nn * 4);
  range.alloc(nn * 8);
  cost.alloc(nn * 4);
  leaf.alloc(nn * 4);
  visit.alloc(size_t(n) * 4);
  hip_check(hipMemsetAsync(visit.p, 0, size_t(n) * 4, st), "BVH visit reset");
  hip_check(hipMemsetAsync(leaf.p, 0, nn * 4, st), "BVH leaf reset");
  Tree T{lo.as<float4>(), hi.as<float4>(), child.as<int2>(), prefix.as<int32_t>(), parent.as<int32_t>(),
         range.as<uint2>(), cost.as<float>(), leaf.as<uint32_t>(), visit.as<uint32_t>(), n};
  out.tris.alloc(size_t(n) * sizeof(TraceTri));
  hipLaunchKernelGGL(k_bvh_leaves, dim3(blocks(n)), dim3(kB), 0, st, tri9.as<float>(), vals[1].as<uint32_t>(),
                     scene_ext, T, out.tris.as<TraceTri>());
  hipLaunchKernelGGL(k_bvh_karras, dim3(blocks(n - 1)), dim3(kB), 0, st, keys[1].as<uint64_t>(), T);
  hipLaunchKernelGGL(k_bvh_refit, dim3(blocks(n)), dim3(kB), 0, st, T);
  // The SAH upper tree: the radix tree's nodes above the cells of the top
  // `cell_bits` Morton bits (common prefix <= cell_bits bits of the 64-bit
  // keys, whose top bit is always 0) are a binary tree over the cells' subtree
  // roots; a bucketed SAH tree over the same roots has as many internal nodes
  // and takes their slots (its root node 0).
  if (cell_bits > 0) {
    std::vector<int2> hchild(n - 1);
    std::vector<int32_t> hpre(n - 1);
    hip_check(hipMemcpyAsync(hchild.data(), child.p, sizeof(int2) * (n - 1), hipMemcpyDeviceToHost, st),
              "BVH children");
    hip_check(hipMemcpyAsync(hpre.data(), prefix.p, sizeof(int32_t) * (n - 1), hipMemcpyDeviceToHost, st),
              "BVH prefixes");
    hip_check(hipStreamSynchronize(st), "BVH tree");
    const int32_t kCellPrefix = int32_t(cell_bits) + 1;
    std::vector<int32_t> top, roots, stack{0};
    while (!stack.empty()) {
      int32_t b = stack.back();
      stack.pop_back();
      if (b >= int32_t(n - 1) || hpre[size_t(b)] >= kCellPrefix) {
        roots.push_back(b);
        continue;
      }
      top.push_back(b);
      stack.push_back(hchild[size_t(b)].y);
      stack.push_back(hchild[size_t(b)].x);
    }
    if (roots.size() >= 3) {
      uint32_t m = uint32_t(roots.size());
      DBuf d_idx, d_lo, d_hi, d_rng;
      d_idx.upload(roots.data(), size_t(m) * 4, st);
      d_lo.alloc(size_t(m) * 16);
      d_hi.alloc(size_t(m) * 16);
      d_rng.alloc(size_t(m) * 8);
      hipLaunchKernelGGL(k_bvh_gather, dim3(blocks(m)), dim3(kB), 0, st, d_idx.as<int32_t>(), m, T, d_lo.as<float4>(),
                         d_hi.as<float4>(), d_rng.as<uint2>());
      std::vector<float4> hlo(m), hhi(m);
      std::vector<uint2> hrng(m);
      hip_check(hipMemcpyAsync(hlo.data(), d_lo.p, size_t(m) * 16, hipMemcpyDeviceToHost, st), "BVH root boxes");
      hip_check(hipMemcpyAsync(hhi.data(), d_hi.p, size_t(m) * 16, hipMemcpyDeviceToHost, st), "BVH root boxes");
      hip_check(hipMemcpyAsync(hrng.data(), d_rng.p, size_t(m) * 8, hipMemcpyDeviceToHost, st), "BVH root ranges");
      hip_check(hipStreamSynchronize(st), "BVH roots");
      std::vector<UpperItem> items(m);
      for (uint32_t k = 0; k < m; k++) {
        UpperItem &x = items[k];
        float l4[3] = {hlo[k].x, hlo[k].y, hlo[k].z}, h4[3] = {hhi[k].x, hhi[k].y, hhi[k].z};
        for (int a = 0; a < 3; a++) {
          x.lo[a] = l4[a];
          x.hi[a] = h4[a];
          x.cen[a] = 0.5f * (l4[a] + h4[a]);
        }
        x.count = double(hrng[k].y - hrng[k].x + 1u);
        x.node = roots[k];
      }
      std::vector<int2> kids;
      std::vector<UpperBox> boxes;
      upper_build(items, 0, int(m), 0, kids, boxes);
      if (kids.size() != top.size()) throw HkError(HK_ERR_INVALID, "BVH upper tree: node count mismatch");
      std::vector<int2> rec_child(kids.size());
      std::vector<float4> rec_lo(kids.size()), rec_hi(kids.size());
      auto slot = [&](int32_t c) { return c >= 0 ? top[size_t(c)] : items[size_t(~c)].node; };
      for (size_t u = 0; u < kids.size(); u++) {
        rec_child[u] = make_int2(slot(kids[u].x), slot(kids[u].y));
        rec_lo[u] = make_float4(boxes[u].lo[0], boxes[u].lo[1], boxes[u].lo[2], 0.0f);
        rec_hi[u] = make_float4(boxes[u].hi[0], boxes[u].hi[1], boxes[u].hi[2], 0.0f);
      }
      DBuf d_top, d_child, d_ulo, d_uhi;
      d_top.upload(top.data(), top.size() * 4, st);
      d_child.upload(rec_child.data(), rec_child.size() * 8, st);
      d_ulo.upload(rec_lo.data(), rec_lo.size() * 16, st);
      d_uhi.upload(rec_hi.data(), rec_hi.size() * 16, st);
      hipLaunchKernelGGL(k_bvh_splice, dim3(blocks(uint32_t(top.size()))), dim3(kB), 0, st, d_top.as<int32_t>(),
                         d_child.as<int2>(), d_ulo.as<float4>(), d_uhi.as<float4>(), uint32_t(top.size()), T);
      hip_check(hipStreamSynchronize(st), "BVH upper tree");
    }
  }
  // The root's children always form a 4-wide node, also when the root could
  // be one leaf (n <= kLeafMax).
  hip_check(hipMemsetAsync(leaf.p, 0, 4, st), "BVH root");
  out.nodes.alloc(size_t(n) * sizeof(Bvh4Node));
  uint32_t n_in = 1, depth = 0, hc[3] = {1, 0, 0};
  int cur = 0;
  while (n_in > 0) {
    depth++;
    hip_check(hipMemsetAsync(counters.as<uint32_t>() + 1, 0, 4, st), "BVH queue reset");
    hipLaunchKernelGGL(k_bvh_collapse, dim3(blocks(n_in)), dim3(kB), 0, st, T, q[cur].as<Todo>(), n_in,
                       q[cur ^ 1].as<Todo>(), out.nodes.as<Bvh4Node>(), counters.as<uint32_t>());
    hip_check(hipMemcpyAsync(hc, counters.p, sizeof hc, hipMemcpyDeviceToHost, st), "BVH queue count");
    hip_check(hipStreamSynchronize(st), "BVH collapse");
    n_in = hc[1];
    cur ^= 1;
    if (depth > 64) throw HkError(HK_ERR_INVALID, "BVH collapse did not terminate");
  }
  // keep only the nodes in use (the build reserved one per triangle)
  DBuf used;
  used.alloc(size_t(hc[0]) * sizeof(Bvh4Node));
  hip_check(hipMemcpyAsync(used.p, out.nodes.p, size_t(hc[0]) * sizeof(Bvh4Node), hipMemcpyDeviceToDevice, st),
            "BVH nodes");
  hip_check(hipStreamSynchronize(st), "BVH nodes");
  std::swap(used.p, out.nodes.p);
  std::swap(used.bytes, out.nodes.bytes);
  out.n_nodes = hc[0];
  out.n_tris = n;
  out.n_leaves = hc[2];
  out.max_depth = depth;
}

}  // namespace hk

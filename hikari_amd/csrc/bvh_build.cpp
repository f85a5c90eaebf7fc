// bvh_build.cpp -- host construction of the 4-wide BVH the traversal kernels
// walk (replaces the reference's BVHAcceleration / HLBVHAcceleration::Build,
// src/acceleration/BVHAcceleration.cpp:46-248).
//
// Binned SAH (32 bins, centroid bounds) builds a binary tree whose leaves hold
// at most kLeafMax triangles; it is then collapsed into 4-wide nodes by
// repeatedly opening the child of largest surface area.  Triangle boxes are
// padded so that every ray the FP32 Moller-Trumbore test accepts also passes
// the (conservative) slab tests of its ancestors.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <vector>

#include "hk_build.h"

namespace hk {

namespace {

struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int a = 0; a < 3; a++) { lo[a] = INFINITY; hi[a] = -INFINITY; }
  }
  void grow(const Box &b) {
    for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); }
  }
  void grow(const float *p) {
    for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); }
  }
  float area() const {
    if (!(lo[0] <= hi[0])) return 0.0f;
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
  }
};

struct BNode {          // binary build node
  Box box;
  int32_t left = -1, right = -1;
  uint32_t start = 0, count = 0;
};

struct Builder {
  std::vector<Box> tb;            // triangle boxes (padded)
  std::vector<float> cen;         // centroids x3
  std::vector<uint32_t> idx;
  std::vector<BNode> nodes;
  std::atomic<int32_t> next{0};

  int32_t alloc2() { return next.fetch_add(2); }

  // Build the subtree of node `ni` over idx[s, e).
  void build(int32_t ni, uint32_t s, uint32_t e, int depth) {
    BNode &nd = nodes[ni];
    Box bb, cb;
    bb.reset(); cb.reset();
    for (uint32_t i = s; i < e; i++) {
      bb.grow(tb[idx[i]]);
      cb.grow(&cen[3 * idx[i]]);
    }
    nd.box = bb;
    uint32_t n = e - s;
    int axis = 0;
    float ext[3];
    for (int a = 0; a < 3; a++) ext[a] = cb.hi[a] - cb.lo[a];
    if (ext[1] > ext[axis]) axis = 1;
    if (ext[2] > ext[axis]) axis = 2;
    if (n <= 2 || ext[axis] <= 0.0f) {
      if (n <= (uint32_t)kLeafMax) { nd.start = s; nd.count = n; return; }
    }
    // binned SAH over the three axes
    const int B = 32;
    float best_cost = INFINITY;
    int best_axis = -1, best_split = -1;
    if (ext[axis] > 0.0f) {
      for (int a = 0; a < 3; a++) {
        if (!(ext[a] > 0.0f)) continue;
        Box bbox[B];
        uint32_t bcnt[B] = {0};
        for (int b = 0; b < B; b++) bbox[b].reset();
        float k = B / ext[a];
        for (uint32_t i = s; i < e; i++) {
          uint32_t p = idx[i];
          int b = std::min(B - 1, std::max(0, int((cen[3 * p + a] - cb.lo[a]) * k)));
          bcnt[b]++;
          bbox[b].grow(tb[p]);
        }
        float rarea[B];
        uint32_t rcnt[B];
        Box acc;
        acc.reset();
        uint32_t c = 0;
        for (int b = B - 1; b > 0; b--) {
          acc.grow(bbox[b]); c += bcnt[b];
          rarea[b] = acc.area(); rcnt[b] = c;
        }
        acc.reset(); c = 0;
        for (int b = 0; b < B - 1; b++) {
          acc.grow(bbox[b]); c += bcnt[b];
          if (c == 0 || rcnt[b + 1] == 0) continue;
          float cost = acc.area() * c + rarea[b + 1] * rcnt[b + 1];
          if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = b; }
        }
      }
    }
    float leaf_cost = bb.area() * n;
    float node_cost = 1.2f * bb.area();  // traversal step vs. triangle test
    if (best_axis < 0 || (n <= (uint32_t)kLeafMax && leaf_cost <= best_cost + node_cost)) {
      if (n <= (uint32_t)kLeafMax) { nd.start = s; nd.count = n; return; }
    }
    uint32_t mid;
    if (best_axis >= 0) {
      int a = best_axis;
      float k = B / ext[a];
      float lo = cb.lo[a];
      mid = uint32_t(std::partition(idx.begin() + s, idx.begin() + e, [&](uint32_t p) {
                       return std::min(B - 1, std::max(0, int((cen[3 * p + a] - lo) * k))) <= best_split;
                     }) - idx.begin());
    } else {
      mid = s;
    }
    if (mid == s || mid == e) {  // degenerate centroids: median split on the widest axis
      mid = s + n / 2;
      std::nth_element(idx.begin() + s, idx.begin() + mid, idx.begin() + e,
                       [&](uint32_t x, uint32_t y) { return cen[3 * x + axis] < cen[3 * y + axis]; });
    }
    int32_t l = alloc2();
    nd.left = l;
    nd.right = l + 1;
    if (n > 65536 && depth < 12) {
      auto f = std::async(std::launch::async, [=]() { build(l, s, mid, depth + 1); });
      build(l + 1, mid, e, depth + 1);
      f.get();
    } else {
      build(l, s, mid, depth + 1);
      build(l + 1, mid, e, depth + 1);
    }
  }
};

}  // namespace

void build_bvh4(const float *tri9, uint32_t nt, Bvh4Result &out) {
  out = Bvh4Result();
  if (nt == 0) throw std::runtime_error("cannot build a BVH over zero triangles");
  if (nt >= (1u << 28)) throw std::runtime_error("too many triangles for the leaf encoding");
  Builder b;
  b.tb.resize(nt);
  b.cen.resize(size_t(nt) * 3);
  b.idx.resize(nt);
  float slo[3] = {INFINITY, INFINITY, INFINITY}, shi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < size_t(nt) * 3; i++)
    for (int a = 0; a < 3; a++) {
      slo[a] = std::min(slo[a], tri9[3 * i + a]);
      shi[a] = std::max(shi[a], tri9[3 * i + a]);
    }
  float scene_ext = std::max(shi[0] - slo[0], std::max(shi[1] - slo[1], shi[2] - slo[2]));
  if (!std::isfinite(scene_ext)) throw std::runtime_error("non-finite vertex position");
  for (uint32_t i = 0; i < nt; i++) {
    const float *p = tri9 + 9 * size_t(i);
    Box bx;
    bx.reset();
    bx.grow(p); bx.grow(p + 3); bx.grow(p + 6);
    for (int a = 0; a < 3; a++) {
      // pad: relative to the triangle's extent and to its coordinate magnitude
      float mag = std::max(std::fabs(bx.lo[a]), std::fabs(bx.hi[a]));
      float pad = (bx.hi[a] - bx.lo[a]) * 0x1p-12f + mag * 0x1p-18f + scene_ext * 0x1p-17f + 1e-30f;
      bx.lo[a] -= pad;
      bx.hi[a] += pad;
      b.cen[3 * size_t(i) + a] = (p[a] + p[3 + a] + p[6 + a]) * (1.0f / 3.0f);
    }
    b.tb[i] = bx;
    b.idx[i] = i;
  }
  b.nodes.resize(size_t(2) * nt + 1);
  b.next = 1;
  b.build(0, 0, nt, 0);
  size_t nn = size_t(b.next.load());
  b.nodes.resize(nn);

  // Collapse to 4-wide nodes, depth-first.
  std::vector<Bvh4Node> &N4 = out.nodes;
  N4.reserve(nn / 2 + 1);
  uint32_t max_depth = 0, leaves = 0;
  struct Todo { int32_t bnode; int32_t slot; uint32_t depth; };
  std::vector<Todo> st;
  auto make_leaf_ref = [&](const BNode &n) -> int32_t {
    leaves++;
    return ~int32_t((n.start << 3) | (n.count - 1));
  };
  if (b.nodes[0].left < 0) {
    // single leaf: wrap it in a root with one used slot
    Bvh4Node root;
    std::memset(&root, 0, sizeof root);
    for (int c = 0; c < 4; c++) {
      root.lo_x[c] = root.lo_y[c] = root.lo_z[c] = INFINITY;
      root.hi_x[c] = root.hi_y[c] = root.hi_z[c] = -INFINITY;
      root.child[c] = kEmptyRef;
    }
    const BNode &n = b.nodes[0];
    root.lo_x[0] = n.box.lo[0]; root.lo_y[0] = n.box.lo[1]; root.lo_z[0] = n.box.lo[2];
    root.hi_x[0] = n.box.hi[0]; root.hi_y[0] = n.box.hi[1]; root.hi_z[0] = n.box.hi[2];
    root.child[0] = make_leaf_ref(n);
    N4.push_back(root);
    max_depth = 1;
  } else {
    N4.emplace_back();
    st.push_back({0, 0, 1});
    size_t head = 0;  // breadth-first while the top kTopNodesBfs slots fill, then depth-first
    while (head < st.size()) {
      Todo t;
      if (N4.size() < size_t(kTopNodesBfs)) {
        t = st[head++];
      } else {
        t = st.back();
        st.pop_back();
      }
      max_depth = std::max(max_depth, t.depth);
      int32_t ch[4] = {b.nodes[t.bnode].left, b.nodes[t.bnode].right, -1, -1};
      int nc = 2;
      while (nc < 4) {
        int best = -1;
        float best_a = -1.0f;
        for (int c = 0; c < nc; c++) {
          const BNode &cn = b.nodes[ch[c]];
          if (cn.left >= 0 && cn.box.area() > best_a) { best_a = cn.box.area(); best = c; }
        }
        if (best < 0) break;
        int32_t open = ch[best];
        ch[best] = b.nodes[open].left;
        ch[nc++] = b.nodes[open].right;
      }
      Bvh4Node nd;
      std::memset(&nd, 0, sizeof nd);
      for (int c = 0; c < 4; c++) {
        if (c >= nc) {
          nd.lo_x[c] = nd.lo_y[c] = nd.lo_z[c] = INFINITY;
          nd.hi_x[c] = nd.hi_y[c] = nd.hi_z[c] = -INFINITY;
          nd.child[c] = kEmptyRef;
          continue;
        }
        const BNode &cn = b.nodes[ch[c]];
        nd.lo_x[c] = cn.box.lo[0]; nd.lo_y[c] = cn.box.lo[1]; nd.lo_z[c] = cn.box.lo[2];
        nd.hi_x[c] = cn.box.hi[0]; nd.hi_y[c] = cn.box.hi[1]; nd.hi_z[c] = cn.box.hi[2];
        if (cn.left < 0) {
          nd.child[c] = make_leaf_ref(cn);
        } else {
          int32_t slot = int32_t(N4.size());
          N4.emplace_back();
          nd.child[c] = slot;
          st.push_back({ch[c], slot, t.depth + 1});
        }
      }
      N4[t.slot] = nd;
    }
  }
  out.tris.resize(nt);
  for (uint32_t i = 0; i < nt; i++) {
    uint32_t p = b.idx[i];
    const float *q = tri9 + 9 * size_t(p);
    TraceTri &t = out.tris[i];
    t.p0x = q[0]; t.p0y = q[1]; t.p0z = q[2];
    t.prim = int32_t(p);
    t.e1x = q[3] - q[0]; t.e1y = q[4] - q[1]; t.e1z = q[5] - q[2];
    t.e2x = q[6] - q[0]; t.e2y = q[7] - q[1]; t.e2z = q[8] - q[2];
    t.pad0 = t.pad1 = 0.0f;
  }
  out.max_depth = max_depth;
  out.n_leaves = leaves;
}

}  // namespace hk

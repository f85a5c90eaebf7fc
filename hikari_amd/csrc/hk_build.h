// hk_build.h -- host-side construction of device scene data.
#pragma once
#include <stdint.h>

#include <vector>

#include "hk_shared.h"

namespace hk {

struct Bvh4Result {
  std::vector<Bvh4Node> nodes;       // FP32 boxes, what the kernels walk
  std::vector<TraceTri> tris;
  uint32_t max_depth = 0, n_leaves = 0;
};

// tri9: nt x 9 floats (p0, p1, p2) in global prim order.
void build_bvh4(const float *tri9, uint32_t nt, Bvh4Result &out);

}  // namespace hk

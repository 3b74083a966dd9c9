#pragma once
#include <cstdint>
#include <vector>

#include "rtg_internal.h"

namespace rtg {

struct BuildPrim {
  float bmin[3], bmax[3], c[3];
};

struct BvhResult {
  std::vector<float> nodes;        // 16 floats (two 32-byte slots) per interior node
  uint32_t root_ref = 0;
  float root_min[3] = {0, 0, 0}, root_max[3] = {0, 0, 0};
  bool empty = true;
  int max_depth = 0;               // interior levels (traversal stack bound)
  std::vector<uint32_t> order_tri; // leaf order -> triangle index
  std::vector<uint32_t> order_sph; // leaf order -> sphere index
};

BvhResult build_bvh(const std::vector<BuildPrim>& tris, const std::vector<BuildPrim>& spheres);

}  // namespace rtg

#pragma once
#include <cstdint>
#include <vector>

#include "rtg_internal.h"

namespace rtg {

struct BuildPrim {
  float bmin[3], bmax[3], c[3];
};

struct BvhResult {
  std::vector<float> nodes;        // 32 floats (one 128-byte 4-wide node) per interior node
  uint32_t root_ref = 0;
  float root_min[3] = {0, 0, 0}, root_max[3] = {0, 0, 0};
  bool empty = true;
  int max_depth = 0;               // interior levels
  int max_stack = 0;               // traversal stack entries any ray can need
  std::vector<uint32_t> order_tri; // leaf order -> triangle index
  std::vector<uint32_t> order_sph; // leaf order -> sphere index
};

BvhResult build_bvh(const std::vector<BuildPrim>& tris, const std::vector<BuildPrim>& spheres);

}  // namespace rtg

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
  std::vector<uint32_t> qnodes;    // the same nodes compressed to 16 words (64 bytes) each, the
                                   // layout the device traverses (quantize_nodes)
  uint32_t root_ref = 0;
  float root_min[3] = {0, 0, 0}, root_max[3] = {0, 0, 0};
  bool empty = true;
  int max_depth = 0;               // interior levels
  int max_stack = 0;               // traversal stack entries any ray can need
  std::vector<uint32_t> order_tri; // leaf order -> triangle index
  std::vector<uint32_t> order_sph; // leaf order -> sphere index
};

BvhResult build_bvh(const std::vector<BuildPrim>& tris, const std::vector<BuildPrim>& spheres);

// Compressed 4-wide node (16 words): [0..2] origin (float bits), [3] biased exponents of the per-axis
// steps (bits 0-7 x, 8-15 y, 16-23 z), [4..9] 8-bit child bounds (lo x, lo y, lo z, hi x, hi y, hi z;
// byte j of each word for child j), [10..11] zero, [12..15] child refs. A bound decodes as
// origin + q * step (an exact product and one rounded float addition); every decoded child box
// contains the full-precision box.
void quantize_nodes(const std::vector<float>& nodes, std::vector<uint32_t>& qnodes);
// The decode of one bound, as the device computes it.
inline float dequantize_bound(float origin, uint32_t q, float step) { return origin + (float)q * step; }

}  // namespace rtg

// Internal data layout shared by the host library and the HIP kernels.
// Everything here is plain-old-data that is copied to device memory as is.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtg {

// ---------------------------------------------------------------- BVH layout
// Interior node = two 32-byte child slots (64 bytes, one node pair per load):
//   slot: float bmin[3], float bmax[3], uint32 ref, uint32 pad
// A child ref with bit 31 set is a leaf: bits 0..26 first primitive (leaf order),
// bits 27..30 count-1 (1..16 primitives). Otherwise the ref is an interior node index.
static const uint32_t kLeafBit = 0x80000000u;
static const int kMaxJobsDev = 256;  // job table entries (the flags word holds 8 bits of job slot)
static const int kMaxLeafPrims = 16;
__host__ __device__ inline uint32_t make_leaf_ref(uint32_t first, uint32_t count) {
  return kLeafBit | ((count - 1u) << 27) | first;
}
__host__ __device__ inline uint32_t leaf_first(uint32_t ref) { return ref & 0x07FFFFFFu; }
__host__ __device__ inline uint32_t leaf_count(uint32_t ref) { return ((ref >> 27) & 0xFu) + 1u; }

struct DevMaterial {
  int32_t type;
  int32_t dist;
  float alpha;     // Color(alpha).average(), as the reference evaluates the texture
  float eta_int;
  float eta_ext;
  float pad0, pad1, pad2;
  float albedo[4]; // Diffuse albedo or reflectance texture
  float eta[4];
  float k[4];
  float trans[4];
};

enum { LIGHT_MESH = 0, LIGHT_ENV = 1 };
struct DevLight {
  int32_t type;
  int32_t emit_offset;  // first emitter triangle in emit_tri
  int32_t emit_count;
  int32_t cdf_offset;   // offset of (emit_count + 1) CDF values in emit_cdf
  float radiance[4];
  float area;           // Mesh::surfaceArea (float sum in triangle order)
  int32_t has_normals;
  float pad0, pad1;
};

struct DevScene {
  const float4* nodes;       // 4 float4 per interior node
  uint32_t root_ref;         // ref of the root (interior index 0, or a leaf)
  int32_t empty;             // no primitives at all
  float root_min[3], root_max[3];
  const float* tri_v;        // 9 floats per triangle, leaf order
  const uint4* prim_info;    // x orig prim id, y material, z light (0xffffffff none), w mesh
  const int32_t* tri_vidx;   // 3 vertex ids per leaf-order triangle (vertex normals), or null
  const float* vnormals;     // global vertex normals (3 floats each) or null
  const uint8_t* tri_has_n;  // per leaf-order triangle: mesh has normals
  const float4* spheres;     // 4 float4 per sphere (leaf order): inverse 3x4 rows, (radius, ...)
  int32_t ntri, nsph;
  const DevMaterial* mats;
  int32_t nmat;
  int32_t nlights;
  const DevLight* lights;
  const float* light_cdf;    // nlights + 1
  int32_t env_light;         // index in lights, or -1
  int32_t strategy;
  const float* emit_tri;     // 9 floats per emitter triangle (mesh order)
  const float* emit_nrm;     // 9 floats per emitter triangle (vertex normals) when has_normals
  const float* emit_cdf;
  // environment map (light/EnvironmentLight.cpp)
  const float* env_map;      // H*W*3, unscaled
  const float* env_row_cdf;  // H * (W + 1)
  const float* env_marg_cdf; // H + 1
  const float* env_weight;   // H
  int32_t env_w, env_h;
  float env_factor, env_marg_sum, env_radius;
  int32_t pad;
};

// Per-render (per-request) state, one slot of the job table.
struct DevJob {
  float s2c[16];        // sampleToCamera (row-major)
  float c2w[16];        // camera-to-world (row-major)
  float inv_w, inv_h, near_clip, far_clip;
  int32_t width, height, spp, sqrt_spp;
  int32_t sampler, dimension, min_depth, max_depth;
  uint64_t seed;
  uint64_t total;       // width*height*spp
  uint64_t next;        // claim counter (atomic)
  uint64_t done;        // finished samples (atomic)
  float4* accum;        // width*height float4 (sum rgb, weight)
  int32_t scene;        // index into the scene table
  int32_t active;       // 1 while samples may be claimed
  int32_t tiles_x;      // 8x8 pixel tiles per row (sample ordering)
  int32_t pad;
};

// Path-pool state (structure of arrays). Slot flags word (rtg_kernels.hip):
//   bits 0..11 depth of the pending hit, bits 12..19 job slot, bit 20 LAST (only the emission
//   of the pending segment remains), bit 21 EMISSION (depth-0 emission enabled)
enum : uint32_t {
  ST_FREE = 0u, ST_ACTIVE = 1u, ST_DONE = 2u
};
struct DevPool {
  int32_t capacity;
  int32_t count_work;   // 1: accumulate traversal work counters (nodes, primitives)
  uint32_t* state;      // ST_*
  uint32_t* flags;
  uint32_t* pixel;
  uint32_t* sample;
  float4* ray_o;        // origin, minT
  float4* ray_d;        // direction, maxT
  float4* hit;          // t, b1, b2, prim (int bits; -1 miss)
  float4* T;            // throughput before the last BSDF sample, eta after it
  float4* F;            // last BSDF weight f*cos/pdf, its pdf
  float4* L;            // radiance, Russian-roulette divisor of the last step
  float4* sh_d;         // shadow direction, maxT
  float4* sh_c;         // shadow contribution if unoccluded
  // queues
  uint32_t* q_ext;      // extension queue
  uint32_t* q_shade;    // 4 queues of capacity each: diffuse, conductor, dielectric, miss
  uint32_t* q_shadow;
  uint8_t* cls_q;       // material class per extension-queue entry (0..2 BSDF type, 3 miss)
  uint8_t* shflag;      // per slot: the shade stage produced a shadow ray
  uint32_t* counters;   // C_* below; u64 counters occupy two words
};
enum {
  C_EXT = 0, C_SHADE = 1, C_SHADOW = 5, C_ACTIVE = 6, C_NAN = 7, C_EXTRAYS = 8, C_SHRAYS = 10, C_DONE = 12,
  C_NODES = 14, C_PRIMS = 16, C_SHNODES = 18, C_SHPRIMS = 20,  // u64 work counters (counting mode)
  C_COUNT = 24
};

// Launch wrappers implemented in rtg_kernels.hip
struct LaunchCfg {
  int grid;      // blocks for grid-stride kernels (traversal kernels use 2x grid blocks of 128)
  int block;     // threads per block
  int stack;     // traversal stack capacity (entries, >= tree height)
  uint2* gstack; // spill area: (stack - 16) entries x (2 * grid * 128) lanes
};
static const int kMaxStack = 64;
// ev (optional): 8 events recorded before regen and after each of the 7 stage launches
void launch_iteration(const DevPool& pool, const DevScene* scenes_dev, DevJob* jobs_dev, int njobs,
                      const LaunchCfg& cfg, hipStream_t stream, hipEvent_t* ev = nullptr);
void launch_reset_pool(const DevPool& pool, hipStream_t stream);
void launch_trace_tap(const DevScene* scene_dev, const float* rays, int n, int any_hit, int32_t* prim,
                      float* t, float* uv, float* p, float* nrm, const LaunchCfg& cfg, hipStream_t stream);
void launch_bsdf_tap(const DevMaterial* mat, int op, const float* wi, const float* wo_or_u, int n,
                     float* out_dir, float* f, float* pdf, float* rel_eta, hipStream_t stream);
void launch_fill_zero(void* ptr, size_t bytes, hipStream_t stream);

}  // namespace rtg

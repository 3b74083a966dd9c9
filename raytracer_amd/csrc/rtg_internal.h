// Internal data layout shared by the host library and the HIP kernels.
// Everything here is plain-old-data that is copied to device memory as is.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace rtg {

// ---------------------------------------------------------------- BVH layout
// 4-wide interior node as built (128 bytes, 8 float4): min x, min y, min z, max x, max y, max z of
// the four children (one float4 each, child j in component j), then the four child refs, then
// padding; the device reads it compressed to 64 bytes (rtg_bvh.h). An unused child slot holds
// kEmptyChild. A child ref with bit 31 set is a leaf: bits 0..26
// first primitive (leaf order), bits 27..30 count-1 (1..16 primitives). Otherwise the ref is an
// interior node index.
static const uint32_t kLeafBit = 0x80000000u;
static const int kMaxJobsDev = 256;  // job table entries (the flags word holds 8 bits of job slot)
static const int kMaxLeafPrims = 16;
static const uint32_t kEmptyChild = 0x7FFFFFFFu;
__host__ __device__ inline uint32_t make_leaf_ref(uint32_t first, uint32_t count) {
  return kLeafBit | ((count - 1u) << 27) | first;
}
__host__ __device__ inline uint32_t leaf_first(uint32_t ref) { return ref & 0x07FFFFFFu; }
__host__ __device__ inline uint32_t leaf_count(uint32_t ref) { return ((ref >> 27) & 0xFu) + 1u; }

// Integrators (the RTG_INTEGRATOR_* values of rtg.h)
enum { INTEGRATOR_PATH_MIS = 0, INTEGRATOR_VOLPATH = 1 };
// Material types (the RTG_MAT_* values of rtg.h)
enum { MAT_DIFFUSE = 0, MAT_ROUGH_CONDUCTOR = 1, MAT_ROUGH_DIELECTRIC = 2, MAT_CONDUCTOR = 3, MAT_DIELECTRIC = 4 };
// type: MAT_*; cls: the shade kernel and queue (0 diffuse, 1 conductors, 2 dielectrics)
struct DevMaterial {
  int32_t type;
  int32_t dist;
  float alpha;     // Color(alpha).average(), as the reference evaluates the texture
  float eta_int;
  float eta_ext;
  int32_t cls;
  float pad1, pad2;
  float albedo[4]; // Diffuse albedo or reflectance texture
  float eta[4];
  float k[4];
  float trans[4];
  int32_t tex[4];  // image textures (index into DevScene::textures, -1 none): albedo, alpha, transmitted
};

// An image texture (texture/ImageTexture.cpp, ScaleTexture.cpp): texels at `offset` floats into
// DevScene::texels, h rows of w RGB values.
struct DevTexture {
  int32_t w, h, invert_x, invert_y;
  float uv_scale, value_scale;
  uint32_t offset, pad;
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
  const float4* nodes;       // 4 x 16 bytes per compressed 4-wide interior node (rtg_bvh.h)
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
  const int32_t* env_marg_guide;  // kCdfGuide + 1 cut points of the marginal CDF (cdf_guide)
  const int32_t* env_row_guide;   // H * (kCdfGuide + 1) cut points of the row CDFs
  int32_t env_w, env_h;
  float env_factor, env_marg_sum, env_radius;
  int32_t pad;
  // image textures and per-vertex texture coordinates (2 floats per global vertex; tri_has_uv per
  // leaf-order triangle: its mesh has uvs, else the hit's barycentrics are its uv)
  const DevTexture* textures;
  const float* texels;
  const float* vuv;
  const uint8_t* tri_has_uv;
  // camera medium (medium/Heterogeneous.cpp, volume/GridVolume.cpp, medium/HGPhaseFunction.cpp):
  // density grid over the world box [0,1]^3, x fastest; sigma_max = max density x scale
  const float* grid;         // null: no heterogeneous medium
  int32_t gres[3];
  float gscale, sigma_max, hg_g;
  float albedo[3];
  // homogeneous camera medium (medium/Homogeneous.cpp), scaled: medium_kind 2
  int32_t medium_kind;       // 0 none, 1 heterogeneous grid, 2 homogeneous
  float sig_a[3], sig_s[3], sig_t[3];
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
  // reconstruction filter: 0 the box of radius 0.5 (one pixel, or two for a jitter of exactly 0),
  // 1 a tabulated filter (lut: 32x32 weights, radius rx/ry, sub-screen border ox/oy)
  int32_t filter;
  int32_t ox, oy;
  float rx, ry;
  const float* lut;
  int32_t integrator;   // RTG_INTEGRATOR_PATH_MIS or RTG_INTEGRATOR_VOLPATH
  int32_t exit_mis;     // 1: rough-dielectric exit vertices use the band proposals (dielectric_exit_mis)
  // 1: the whole render is in flight at once (a fresh pool with a slot for every sample) with the
  // box filter: paths keep their radiance in their slots and k_gather sums each pixel's samples in
  // sample order after the last iteration, instead of four float atomics per finished sample
  int32_t gather;
};
static const int kFilterLut = 32 * 32;
// Slots per claim chunk of the regeneration stage (one block, 256 threads x 8 slots).
static const int kClaimChunk = 2048;
// Cut points per environment CDF: entry k is the first index whose CDF value is >= k / kCdfGuide
// (the lower bound a full binary search would find for that value), entry kCdfGuide the CDF's size.
static const int kCdfGuide = 64;
// Host: cut points of a CDF of n values (core/CDF.h layout: c[0] = 0, non-decreasing): for
// k < kCdfGuide the lower bound of k / kCdfGuide (first index with c[i] >= k / kCdfGuide, by the
// same float comparison as the device's search), then n. A draw u in [k, k + 1) / kCdfGuide has its
// lower bound between entries k and k + 1, so the device searches only that range
// (cdf_sample_guided, rtg_device.h) and finds the index the full search would.
inline void cdf_guide(const float* c, int n, std::vector<int32_t>& out) {
  for (int k = 0; k < kCdfGuide; k++) {
    const float v = (float)k / (float)kCdfGuide;  // exact: kCdfGuide is a power of two
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (c[mid] < v) lo = mid + 1; else hi = mid;
    }
    out.push_back(lo);
  }
  out.push_back(n);
}

// ---------------------------------------------------------------- CDF (core/CDF.h:14-34)
// Largest path pool of a render with learned guiding: its training schedule counts iteration
// batches (64 warm-up steps, then one per batch up to 512), which at 2^21 slots completes within a
// 512x512 4096-spp render; larger pools cover more samples per batch and leave the network
// undertrained (about 200 steps at 2^23 slots).
#ifndef RTG_GUIDED_POOL_LOG2
#define RTG_GUIDED_POOL_LOG2 21
#endif
static const int kGuidedPoolMax = 1 << RTG_GUIDED_POOL_LOG2;
// Guided-render training schedule: at most kGuideMaxSteps Adam steps, at most kGuideStepsPerBatch
// per batch of iterations; on a context with a larger pool the training phase ends after those
// steps or 1/kGuidePhaseDiv of the render's samples, whichever comes first.
#ifndef RTG_GUIDE_MAX_STEPS
#define RTG_GUIDE_MAX_STEPS 128
#endif
#ifndef RTG_GUIDE_STEPS_PER_BATCH
#define RTG_GUIDE_STEPS_PER_BATCH 16
#endif
#ifndef RTG_GUIDE_PHASE_DIV
#define RTG_GUIDE_PHASE_DIV 16
#endif
static const uint64_t kGuideMaxSteps = RTG_GUIDE_MAX_STEPS;
static const int kGuideStepsPerBatch = RTG_GUIDE_STEPS_PER_BATCH;
static const uint64_t kGuidePhaseDiv = RTG_GUIDE_PHASE_DIV;
// lights per scene the all-lights strategy supports (shadow rays per vertex)
static const int kAllLights = 8;
static const uint32_t kAllLightDim = 1u << 22;  // 2D draws of lights 1.. at depth d: + d * kAllLights + light

// Path-pool state (structure of arrays). Slot flags word (rtg_kernels.hip):
//   bits 0..11 depth of the pending hit, bits 12..19 job slot, bit 20 LAST (only the emission
//   of the pending segment remains), bit 21 EMISSION (depth-0 emission enabled)
enum : uint32_t {
  ST_FREE = 0u, ST_ACTIVE = 1u, ST_DONE = 2u,
  ST_NONE = 0xffffffffu  // past the pool's end (never stored)
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
  // all-lights strategy (allocated on first use): up to kAllLights shadow rays per slot, entry
  // slot * kAllLights + k; shflag then holds their count
  int32_t all_lights;
  float4* sh_d2;
  float4* sh_c2;
  uint32_t* q_shadow2;
  float4* vol_a;        // volumetric paths: light-sample point of the last medium vertex, last hit t
  float4* vol_b;        // volumetric paths: light-sample normal of the last medium vertex
  // queues
  uint32_t* q_ext;      // extension queue
  uint32_t* q_shade;    // 5 queues of capacity each: diffuse, conductor, dielectric, miss, volumetric
  uint32_t* q_shadow;
  uint8_t* cls_q;       // material class per extension-queue entry (0..2 BSDF type, 3 miss)
  uint8_t* shflag;      // per slot: the shade stage produced a shadow ray
  uint32_t* counters;   // C_* below; u64 counters occupy two words
  uint32_t* chunk_of;   // gather renders: 1 + the slot chunk (kClaimChunk slots) that claimed each
                        // chunk of kClaimChunk consecutive samples in the first regeneration (0: none)
};
enum {
  C_EXT = 0, C_SHADE = 1, C_SHADOW = 5, C_ACTIVE = 6, C_NAN = 7, C_EXTRAYS = 8, C_SHRAYS = 10, C_DONE = 12,
  C_NODES = 14, C_PRIMS = 16, C_SHNODES = 18, C_SHPRIMS = 20,  // u64 work counters (counting mode)
  C_XFETCH = 22, C_SFETCH = 23,  // batch cursors of the persistent extension / shadow traversals
  C_VOLQ = 24,  // volumetric queue of a mixed iteration (entries of VolPathTracing jobs)
  C_SPILL = 25, // traversal-stack pushes past the LDS entries (counting mode)
  C_OVF = 26,   // 1 when a traversal stack overflowed (an error)
  C_GATHER = 27, // 1 when k_gather met a sample chunk outside the pool (an error)
  C_COUNT = 28
};

// ---------------------------------------------------------------- path guiding (rtg_guide.hip)
// Network: 12-frequency sin/cos encoding of (position in the scene box, shading normal) -> 144
// inputs, three hidden layers of 64 (ReLU), 40 outputs = 8 vMF lobes x (logit, mean xyz, log kappa).
// FP32 master parameters, row-major W[out][in] then bias, per layer:
static const int kGuideIn = 144, kGuideHid = 64, kGuideOut = 40, kGuideLobes = 8;
// The mixture as the renderer keeps it per guided vertex (rtg_device.h mixr_pdf): pi[8], pi times
// the vMF normaliser [8], mu[8][3], kappa[8].
static const int kGuideMixR = 48;
static const int kGOffW1 = 0, kGOffB1 = 9216, kGOffW2 = 9280, kGOffB2 = 13376, kGOffW3 = 13440, kGOffB3 = 17536,
                 kGOffW4 = 17600, kGOffB4 = 20160, kGuideParams = 20200;
// bf16 copies for the MFMA kernels: forward layout wf = W[out][in] per layer (L4 padded to 64 rows),
// backward layout wb = W^T[in][out] for layers 2..4 (L4 padded to 64 columns); FP32 biases padded.
static const int kGWf1 = 0, kGWf2 = 9216, kGWf3 = 13312, kGWf4 = 17408, kGWfSize = 21504;
static const int kGWb2 = 0, kGWb3 = 4096, kGWb4 = 8192, kGWbSize = 12288;
static const int kGBias1 = 0, kGBias2 = 64, kGBias3 = 128, kGBias4 = 192, kGBiasSize = 256;
static const int kGuideBatch = 65536;        // training samples per step
static const float kGuideLogKappaMin = -3.f, kGuideLogKappaMax = 8.f;
static const uint32_t kGuideSelDim = 1u << 20; // 1D sampler dimension of the strategy choice (+ depth)
// Guided vertices: diffuse ones up to depth kGuideRec-1 (their product f*cos*L_in does not depend
// on the incoming direction, so a network of position and normal can represent it).
#ifndef RTG_GUIDE_DEPTHS
#define RTG_GUIDE_DEPTHS 3
#endif
static const int kGuideRec = RTG_GUIDE_DEPTHS;
// training sample weights are clamped at this multiple of their batch mean
static const float kGuideWClamp = 16.f;

struct DevGuide {
  int32_t enabled;
  int32_t sample;            // 1: guided sampling; 0: plain sampling, vertices still recorded (warm-up)
  int32_t record;            // 1: record guided vertices and emit training samples
  int32_t job;               // job-table slot whose depth-0 vertices are guided
  float bmin[3], inv_ext[3]; // position normalisation to [0,1]^3 (scene box)
  const uint16_t* wf;        // bf16 forward weights
  const uint16_t* wb;        // bf16 backward weights
  const float* bias;         // FP32 biases
  float* mix;                // per pool slot: 8 lobes x (pi, mu xyz, kappa)
  float4* rec;               // per pool slot, kGuideRec records of 4 float4 (guided vertex at depth r):
                             //   pos(3) n(3) dir(3) L snapshot(3) throughput before the vertex(3) weight
  uint8_t* rstate;           // per slot and record: 0 none, 1 vertex recorded, 2 radiance snapshot taken
  uint32_t* gq;              // guide queue: pool slots of guided depth-0 vertices
  uint32_t* gq_count;
  float* ring;               // training samples: pos(3) n(3) dir(3) weight
  uint32_t ring_mask;        // capacity - 1 (power of two)
  unsigned long long* ring_count;
  unsigned long long* infer_count;  // inferences run (statistics)
};

// Launch wrappers implemented in rtg_kernels.hip
static const int kTraceRefill = 16;
static const int kTraceWaves = 6;  // waves per SIMD of a traversal launch (default)
static const int kMaxTraceWaves = 10;  // the spill area is sized for this many
struct LaunchCfg {
  int grid;      // blocks for grid-stride kernels
  int trace_blocks;  // blocks of 128 for the persistent traversal kernels
  int block;     // threads per block
  int stack;     // traversal stack capacity (entries, >= tree height)
  uint2* gstack; // spill area: (stack - kLdsStack) entries x (trace_blocks * 128) lanes
  int refill = kTraceRefill;  // persistent traversal: refill a wave once this many lanes are idle
  int regen_blocks = 1024;    // regeneration blocks: one resident round (4 per CU), each looping over chunks
  // blocks of 256 for the grid-stride stages: one resident round each (set by kernel_grids)
  int shade_grid[3] = {0, 0, 0};  // k_shade per material type
  int miss_grid = 0;              // k_shade_miss
  int vol_grid = 0;               // k_vol
  int infer_grid = 0;             // guide-network inference
};
// Fills the grid-stride stages' grids of cfg for a device of `cu_count` compute units: the blocks
// the device keeps resident at each kernel's register and LDS use (falls back to cfg.grid).
void kernel_grids(LaunchCfg& cfg, int cu_count);
int guide_infer_blocks_per_cu();
static const int kMaxStack = 96;
// traversal-stack entries of each lane kept in LDS (8 bytes each); the rest spill to global memory
#ifndef RTG_LDS_STACK
#define RTG_LDS_STACK 12
#endif
constexpr int kLdsStack = RTG_LDS_STACK;
// ev (optional): 8 events recorded before regen and after each of the 7 stage launches
// order_dev: [0] number of entries n, [1..n] job-table slots in claim order (free path slots go to
// the first job with samples left); read on the device so that a captured launch stays valid
void launch_iteration(const DevPool& pool, const DevScene* scenes_dev, DevJob* jobs_dev, const int32_t* order_dev,
                      const LaunchCfg& cfg, hipStream_t stream, hipEvent_t* ev = nullptr,
                      const DevGuide* guide = nullptr);
// One iteration of the volumetric integrator (regen, extension traversal, k_vol, shadow rays).
void launch_iteration_vol(const DevPool& pool, const DevScene* scenes_dev, DevJob* jobs_dev, const int32_t* order_dev,
                          const LaunchCfg& cfg, hipStream_t stream);
// One iteration over a pool holding paths of PathTracingMIS and VolPathTracing jobs: the plain
// shading stages for the former, k_vol over the volumetric queue for the latter.
void launch_iteration_mixed(const DevPool& pool, const DevScene* scenes_dev, DevJob* jobs_dev,
                            const int32_t* order_dev, const LaunchCfg& cfg, hipStream_t stream);
void launch_reset_pool(const DevPool& pool, hipStream_t stream);
// Gather renders (DevJob::gather): adds every pixel's samples of job `job` to its film.
void launch_gather(const DevJob* jobs_dev, int job, const DevPool& pool, int npix, hipStream_t stream);
// ctr: two device words, [0] += pushes into the global stack spill, [1] = 1 on a stack overflow
void launch_trace_tap(const DevScene* scene_dev, const float* rays, int n, int any_hit, int32_t* prim,
                      float* t, float* uv, float* p, float* nrm, uint32_t* ctr, const LaunchCfg& cfg,
                      hipStream_t stream);
void launch_tonemap(const float4* film, int n, int mode, float expo, unsigned int* maxbits, uint8_t* out,
                    hipStream_t stream);
// op 0 light sample (out n x 14), 1 light pdf (n), 2 environment radiance (n x 3), 3 light choice (n x 3)
void launch_light_tap(const DevScene* scene_dev, int op, int light, const float* in, const float* u, int n,
                      float* out, hipStream_t stream);
void launch_camera_tap(const DevJob& job, const float* pxy, int n, float* rays, hipStream_t stream);
void launch_medium_tap(const DevScene* scene_dev, int op, const float* in, int n, const float* u, int k, float* out,
                       hipStream_t stream);
void launch_bsdf_tap(const DevMaterial* mat, int op, const float* wi, const float* wo_or_u, int n,
                     float* out_dir, float* f, float* pdf, float* rel_eta, hipStream_t stream);
void launch_fill_zero(void* ptr, size_t bytes, hipStream_t stream);
// Allocation-free trace (rtg_trace_device_noalloc): a static device scratch array per GPU holds the
// trace counters (kScratchHeader bytes) and the stack spill area of up to kNoAllocBlocks blocks of
// 128 lanes, kNoAllocSpill entries each, so scenes whose stack bound is at most kNoAllocStack.
static const int kNoAllocBlocks = 512;
static const int kNoAllocSpill = 36;
static const int kNoAllocStack = kLdsStack + kNoAllocSpill;
static const size_t kScratchHeader = 256;
static const size_t kScratchBytes = kScratchHeader + (size_t)kNoAllocBlocks * 128 * kNoAllocSpill * 8;
static const int kMaxScratchDevices = 64;
// Device address of the current device's scratch array (hipGetSymbolAddress, once per device;
// the device must be current); null on failure.
void* scratch_address(int device);
void launch_trace_tap_noalloc(const DevScene* scene_dev, const float* rays, int n, int any_hit, int32_t* prim,
                              float* t, float* uv, float* p, float* nrm, void* scratch_dev, int stack, int refill,
                              hipStream_t stream);
// ps: n (pixel, sample) pairs -> out2 (n x n2 x 2) and out1 (n x n1) draws of the job's sampler
void launch_sampler_tap(const DevJob& J, const uint32_t* ps, int n, int n2, int n1, float* out2, float* out1,
                        hipStream_t stream);
void launch_guide_sample_tap(const float* mix, const float* u, int n, float* dirs, float* pdf, hipStream_t stream);

// Guiding kernels (rtg_guide.hip). Scratch for one training step (column-blocked bf16 tensors).
struct GuideTrainScratch {
  uint16_t* x0;   // B/8 x 160 x 8
  uint16_t* h[3]; // B/8 x 64 x 8 each (H1, H2, H3)
  uint16_t* g[4]; // B/8 x 64 x 8 each (dH1, dH2, dH3, dOut)
  float* grad;    // kGuideParams
  float* scal;    // [0] clamped weight sum, [1] loss sum, [2] weight clamp
};
void launch_guide_refresh(const float* master, uint16_t* wf, uint16_t* wb, float* bias, hipStream_t stream);
void launch_guide_infer_pool(const DevGuide& G, const DevPool& P, const DevScene* scenes, const DevJob* jobs,
                             int grid, hipStream_t stream);
void launch_guide_infer_array(const DevGuide& G, const float* in6, int n, float* mix_out, hipStream_t stream);
// One gradient evaluation over n samples (pos, n, dir, weight; 10 floats each) read from
// samples[(base + i) & mask]; leaves the gradient in S.grad and (weight sum, loss) in S.scal.
void launch_guide_grad(const DevGuide& G, const float* samples, uint64_t base, uint32_t mask, int n,
                       const GuideTrainScratch& S, hipStream_t stream);
void launch_guide_adam(float* master, float* m, float* v, const float* grad, int step, float lr, uint16_t* wf,
                       uint16_t* wb, float* bias, hipStream_t stream);

}  // namespace rtg

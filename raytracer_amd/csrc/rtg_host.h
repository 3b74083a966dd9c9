// Host-side objects behind the opaque handles of include/rtg.h.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rtg.h"
#include "rtg_internal.h"

namespace rtg {

static const int kMaxScenes = 64;
static const int kMaxJobs = 256;

struct RtgError : std::runtime_error {
  int code;
  RtgError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what);
void set_error(const char* msg);

// Path-guiding state of a context (rtg_guide_host.cpp).
struct GuideState {
  std::vector<void*> owned, pool_owned;
  float *master = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  uint16_t *wf = nullptr, *wb = nullptr;
  float* bias = nullptr;
  GuideTrainScratch S{};
  float* samples = nullptr;  // tap input (kGuideBatch x 10)
  int step = 0;              // Adam step count
  int capacity = 0;          // pool slots the per-slot buffers cover
  float* mix = nullptr;
  float4* rec = nullptr;
  uint8_t* rstate = nullptr;
  uint32_t *gq = nullptr, *gq_count = nullptr;
  float* ring = nullptr;
  uint32_t ring_cap = 0;
  unsigned long long* ring_count = nullptr;  // [0] ring count, [1] inference count
  // a trained network's inference on a larger pool: mixtures and guide queue per slot (only grow)
  std::vector<void*> infer_owned;
  int infer_capacity = 0;
  float* infer_mix = nullptr;
  uint32_t* infer_gq = nullptr;
  ~GuideState();
};

}  // namespace rtg

struct rtg_scene {
  rtg_ctx* ctx = nullptr;
  int slot = -1;
  rtg::DevScene dev;
  std::vector<void*> owned;
  uint64_t bytes = 0;
  int stack = 16;
  int ntri = 0, nsph = 0, ntex = 0;
  bool all_lights = false;  // RTG_LIGHTS_ALL
  std::atomic<int> users{0};  // requests of the serving loop that reference this scene
  std::vector<rtg_material> mats;  // the materials as uploaded (rtg_scene_material)
};

namespace rtg {
static const int kIterEvents = 9;  // events per profiled iteration: before regen, after each of 8 stages
// Kinds of wavefront iteration (run_iterations)
enum IterMode { IT_PLAIN = 0, IT_VOL = 1, IT_MIXED = 2 };
// A captured batch of iterations: everything its kernels receive by value is in the key.
struct GraphKey {
  int32_t mode, count, capacity, all_lights, count_work, grid, trace_blocks, stack, refill;
  uint64_t epoch;  // pool allocation generation
  bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof(GraphKey)) == 0; }
};
struct GraphEntry {
  GraphKey key;
  hipGraphExec_t exec;
};
}  // namespace rtg

struct rtg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string arch;
  int cu_count = 0;
  int grid = 1024;
  int pool_slots = 1 << 21;
  int profiling = 0;
  std::vector<hipEvent_t> events;
  rtg::DevPool pool{};
  int pool_alloc = 0;  // slots the pool's arrays hold (pool.capacity <= pool_alloc)
  std::vector<void*> pool_owned;
  std::vector<void*> scratch_owned;
  rtg::DevScene* scenes_dev = nullptr;
  rtg::DevJob* jobs_dev = nullptr;
  uint2* gstack = nullptr;
  uint32_t* tap_ctr = nullptr;        // trace tap: [0] spilled stack pushes, [1] stack overflow
  void* scratch = nullptr;            // this device's static scratch array (rtg_trace_device_noalloc)
  uint64_t last_trace_spills = 0;
  int refill = rtg::kTraceRefill;
  int trace_waves = rtg::kTraceWaves;  // waves per SIMD of the persistent traversal launches
  float* lut_dev = nullptr;  // kMaxJobs filter tables of kFilterLut weights
  int trace_blocks = 0;
  int32_t* order_dev = nullptr;    // claim order of job-table slots
  rtg_server* server = nullptr;    // the serving loop owning the pool, if any
  rtg::GuideState* guide = nullptr;  // path-guiding network and buffers, created on first use
  std::vector<rtg_scene*> scene_slots;
  // steady-state iteration batches captured as hipGraphs (one per launch configuration)
  std::vector<rtg::GraphEntry> graphs;
  uint64_t pool_epoch = 0;
  bool use_graphs = true;
  std::atomic<uint64_t> graph_launches{0}, graph_captures{0};
  void* pinned = nullptr;
  std::recursive_mutex mtx_r;
  std::mutex mtx;
};

namespace rtg {
rtg_scene* build_scene(rtg_ctx* ctx, const rtg_scene_desc* desc);
void release_scene(rtg_ctx* ctx, rtg_scene* s);
void ensure_pool(rtg_ctx* ctx, int capacity);
int stack_for_resident(rtg_ctx* ctx);
DevJob make_job(rtg_ctx* ctx, rtg_scene* scene, const rtg_render_desc& d, float4* accum);
void filter_setup(const rtg_render_desc& d, DevJob& J, float* lut);
void upload_job_filter(rtg_ctx* ctx, const rtg_render_desc& d, DevJob& J, int slot);
LaunchCfg launch_cfg(rtg_ctx* ctx);
// Enqueues `count` iterations of kind `mode` (IterMode) on the context's stream, replaying a cached
// hipGraph of the batch when graphs are on (captured on first use of a configuration).
// ev (IT_PLAIN only, or null): per-stage events of each iteration (kIterEvents per iteration); a
// profiled batch launches directly (its events are timed with hipEventElapsedTime).
void run_iterations(rtg_ctx* ctx, int mode, const LaunchCfg& cfg, int count, hipEvent_t* ev = nullptr);
void graphs_clear(rtg_ctx* ctx);
// rtg_render / rtg_render_device while a server runs on the context: the render becomes a request
// (rtg_serve.cpp); film_dev = the caller's device film (not cleared) or NULL.
int serve_render(rtg_server* s, rtg_scene* scene, const rtg_render_desc* desc, float* out_host, float* film_dev,
                 rtg_stats* st);
void camera_matrices(const rtg_render_desc& d, float* s2c, float* c2w);
void* dev_alloc(rtg_ctx* ctx, size_t bytes, std::vector<void*>& owned);
GuideState* guide_state(rtg_ctx* ctx);
void guide_destroy(rtg_ctx* ctx);
void guide_init_params(rtg_ctx* ctx, GuideState* g, uint64_t seed);
void guide_set_params(rtg_ctx* ctx, GuideState* g, const float* params);
DevGuide guide_dev(GuideState* g);
void guide_prepare_render(rtg_ctx* ctx, GuideState* g, rtg_scene* scene, uint64_t seed, DevGuide& G);
// The guide of a render's inference phase on a pool of `capacity` slots: the trained network with
// per-slot buffers of that size, recording (and so training) off; guided sampling stays as G has it
// (off only if the render's training never finished its warm-up steps).
DevGuide guide_infer_phase(rtg_ctx* ctx, GuideState* g, const DevGuide& G, int capacity);
void guide_train_step(rtg_ctx* ctx, GuideState* g, const DevGuide& G, const float* samples, uint64_t base,
                      uint32_t mask, int n, float lr);
}  // namespace rtg

// CPU stand-ins for librtg's kernel launchers (tests/test_serving_host.py): an iteration finishes
// every sample claimed before it, splats a deterministic value per (job, sample) into the job's film,
// then hands free path slots to the jobs in claim order, like the device regeneration stage.
#include <algorithm>
#include <cstdio>
#include "rtg_internal.h"
namespace rtg {
void launch_iteration(const DevPool& P, const DevScene*, DevJob* jobs, const int32_t* order, const LaunchCfg&,
                      hipStream_t, hipEvent_t*, const DevGuide*) {
  uint64_t free_slots = (uint64_t)P.capacity;
  for (int j = 0; j < kMaxJobsDev; j++) {  // finish everything claimed earlier
    DevJob& J = jobs[j];
    uint64_t fin = std::min(J.next, J.total);
    for (uint64_t s = J.done; s < fin; s++) {
      uint64_t npix = (uint64_t)J.width * J.height;
      float4& a = J.accum[s % npix];
      a.x += 1.f; a.y += (float)(s / npix); a.z += 0.5f; a.w += 1.f;
    }
    J.done = fin;
  }
  const int norder = order[0];
  for (int jj = 0; jj < norder; jj++) {
    int j = order[1 + jj];
    if (j < 0 || j >= kMaxJobsDev) { fprintf(stderr, "bad order entry\n"); abort(); }
    DevJob& J = jobs[j];
    if (!J.active || J.next >= J.total || free_slots == 0) continue;
    uint64_t take = std::min<uint64_t>(free_slots, J.total - J.next);
    J.next += take;
    free_slots -= take;
  }
}
// the volumetric and mixed iterations finish samples the same way
void launch_iteration_vol(const DevPool& P, const DevScene* S, DevJob* jobs, const int32_t* order,
                          const LaunchCfg& cfg, hipStream_t st) {
  launch_iteration(P, S, jobs, order, cfg, st, nullptr, nullptr);
}
void launch_iteration_mixed(const DevPool& P, const DevScene* S, DevJob* jobs, const int32_t* order,
                            const LaunchCfg& cfg, hipStream_t st) {
  launch_iteration(P, S, jobs, order, cfg, st, nullptr, nullptr);
}
void launch_medium_tap(const DevScene*, int, const float*, int, const float*, int, float*, hipStream_t) {}
void launch_light_tap(const DevScene*, int, int, const float*, const float*, int, float*, hipStream_t) {}
void launch_camera_tap(const DevJob&, const float*, int, float*, hipStream_t) {}
void launch_gather(const DevJob*, int, const DevPool&, int, hipStream_t) {}
void kernel_grids(LaunchCfg& cfg, int) {
  for (int m = 0; m < 3; m++) cfg.shade_grid[m] = cfg.grid;
  cfg.miss_grid = cfg.vol_grid = cfg.infer_grid = cfg.grid;
}
int guide_infer_blocks_per_cu() { return 0; }
void launch_tonemap(const float4*, int, int, float, unsigned int*, uint8_t*, hipStream_t) {}
// as k_reset_pool: every slot free, every counter zero
void launch_reset_pool(const DevPool& P, hipStream_t) {
  for (int i = 0; i < P.capacity; i++) P.state[i] = ST_FREE;
  memset(P.counters, 0, C_COUNT * sizeof(uint32_t));
}
void launch_trace_tap(const DevScene*, const float*, int, int, int32_t*, float*, float*, float*, float*, uint32_t*,
                      const LaunchCfg&, hipStream_t) {}
void launch_bsdf_tap(const DevMaterial*, int, const float*, const float*, int, float*, float*, float*, float*,
                     hipStream_t) {}
void launch_fill_zero(void* p, size_t b, hipStream_t) { memset(p, 0, b); }
void* scratch_address(int) {
  static unsigned char buf[256];
  return buf;
}
void launch_trace_tap_noalloc(const DevScene*, const float*, int, int, int32_t*, float*, float*, float*, float*,
                              void*, int, int, hipStream_t) {}
void launch_sampler_tap(const DevJob&, const uint32_t*, int, int, int, float*, float*, hipStream_t) {}
// guiding kernels are not exercised by the host-logic test
void launch_guide_refresh(const float*, uint16_t*, uint16_t*, float*, hipStream_t) {}
void launch_guide_infer_pool(const DevGuide&, const DevPool&, const DevScene*, const DevJob*, int, hipStream_t) {}
void launch_guide_infer_array(const DevGuide&, const float*, int, float*, hipStream_t) {}
void launch_guide_grad(const DevGuide&, const float*, uint64_t, uint32_t, int, const GuideTrainScratch&, hipStream_t) {}
void launch_guide_adam(float*, float*, float*, const float*, int, float, uint16_t*, uint16_t*, float*, hipStream_t) {}
void launch_guide_sample_tap(const float*, const float*, int, float*, float*, hipStream_t) {}
}  // namespace rtg

// Host side of the path-guiding network: device state, parameter initialisation, the training
// step, and the parity entry points rtg_guide_eval / rtg_guide_grad / rtg_guide_train.
// No counterpart in the reference renderer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "rtg_host.h"

namespace rtg {

static void* galloc(GuideState* g, size_t bytes) {
  void* p = nullptr;
  hip_check(hipMalloc(&p, bytes ? bytes : 16), "hipMalloc guide");
  g->owned.push_back(p);
  return p;
}

GuideState::~GuideState() {
  for (void* p : owned) (void)hipFree(p);
}

// Parameters, optimiser state, bf16 copies and training scratch (independent of the pool size).
static void guide_alloc_net(rtg_ctx* ctx, GuideState* g) {
  if (g->master) return;
  g->master = (float*)galloc(g, sizeof(float) * kGuideParams);
  g->adam_m = (float*)galloc(g, sizeof(float) * kGuideParams);
  g->adam_v = (float*)galloc(g, sizeof(float) * kGuideParams);
  g->wf = (uint16_t*)galloc(g, sizeof(uint16_t) * kGWfSize);
  g->wb = (uint16_t*)galloc(g, sizeof(uint16_t) * kGWbSize);
  g->bias = (float*)galloc(g, sizeof(float) * kGBiasSize);
  // padding rows/columns of the bf16 copies stay zero for ever
  launch_fill_zero(g->wf, sizeof(uint16_t) * kGWfSize, ctx->stream);
  launch_fill_zero(g->wb, sizeof(uint16_t) * kGWbSize, ctx->stream);
  launch_fill_zero(g->bias, sizeof(float) * kGBiasSize, ctx->stream);
  const size_t B = kGuideBatch;
  g->S.x0 = (uint16_t*)galloc(g, B * 160 * 2);
  for (int i = 0; i < 3; i++) g->S.h[i] = (uint16_t*)galloc(g, B * 64 * 2);
  for (int i = 0; i < 4; i++) g->S.g[i] = (uint16_t*)galloc(g, B * 64 * 2);
  g->S.grad = (float*)galloc(g, sizeof(float) * kGuideParams);
  g->S.scal = (float*)galloc(g, sizeof(float) * 4);
  g->samples = (float*)galloc(g, sizeof(float) * 10 * B);
}

// Per-path-slot buffers and the training-sample ring.
static void guide_alloc_pool(rtg_ctx* ctx, GuideState* g, int capacity) {
  if (g->capacity == capacity) return;
  if (g->capacity) {  // pool size changed: drop the per-slot buffers
    (void)hipStreamSynchronize(ctx->stream);
    for (void* p : g->pool_owned) (void)hipFree(p);
    g->pool_owned.clear();
  }
  auto A = [&](size_t bytes) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, bytes), "hipMalloc guide pool");
    g->pool_owned.push_back(p);
    return p;
  };
  const size_t n = (size_t)capacity;
  g->mix = (float*)A(n * kGuideMixR * sizeof(float));
  g->rec = (float4*)A(n * kGuideRec * 4 * sizeof(float4));
  g->rstate = (uint8_t*)A((n * kGuideRec + 3) & ~(size_t)3);
  g->gq = (uint32_t*)A(n * sizeof(uint32_t));
  g->gq_count = (uint32_t*)A(64);
  g->ring_cap = 1u << 21;
  g->ring = (float*)A((size_t)g->ring_cap * 10 * sizeof(float));
  g->ring_count = (unsigned long long*)A(64);
  g->capacity = capacity;
}

GuideState* guide_state(rtg_ctx* ctx) {
  if (!ctx->guide) ctx->guide = new GuideState();
  guide_alloc_net(ctx, ctx->guide);
  return ctx->guide;
}

void guide_destroy(rtg_ctx* ctx) {
  if (!ctx->guide) return;
  for (void* p : ctx->guide->pool_owned) (void)hipFree(p);
  for (void* p : ctx->guide->infer_owned) (void)hipFree(p);
  delete ctx->guide;
  ctx->guide = nullptr;
}

// Xavier-uniform weights, zero biases (SplitMix64 stream from the seed); the output layer's
// weights are scaled by 0.1 so the first mixtures are near-uniform lobe weights with broad lobes.
void guide_init_params(rtg_ctx* ctx, GuideState* g, uint64_t seed) {
  std::vector<float> p(kGuideParams, 0.f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  auto next = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  struct L { int off, nout, nin; double scale; };
  const L layers[4] = {{kGOffW1, 64, kGuideIn, 1.0}, {kGOffW2, 64, 64, 1.0}, {kGOffW3, 64, 64, 1.0},
                       {kGOffW4, kGuideOut, 64, 0.1}};
  for (const L& l : layers) {
    const double lim = l.scale * std::sqrt(6.0 / (l.nout + l.nin));
    for (int i = 0; i < l.nout * l.nin; i++) {
      double w = (2.0 * next() - 1.0) * lim;
      // first layer: features of frequency 2^f start with weight scale 0.7^f, so the initial
      // network is a smooth function of position and normal
      if (l.off == kGOffW1) w *= std::pow(0.7, (double)(((i % kGuideIn) % 24) / 2));
      p[l.off + i] = (float)w;
    }
  }
  guide_set_params(ctx, g, p.data());
}

void guide_set_params(rtg_ctx* ctx, GuideState* g, const float* params) {
  hip_check(hipMemcpy(g->master, params, sizeof(float) * kGuideParams, hipMemcpyHostToDevice), "guide params");
  launch_fill_zero(g->adam_m, sizeof(float) * kGuideParams, ctx->stream);
  launch_fill_zero(g->adam_v, sizeof(float) * kGuideParams, ctx->stream);
  launch_guide_refresh(g->master, g->wf, g->wb, g->bias, ctx->stream);
  g->step = 0;
  hip_check(hipGetLastError(), "guide refresh");
}

DevGuide guide_dev(GuideState* g) {
  DevGuide G;
  std::memset(&G, 0, sizeof(G));
  G.enabled = 1;
  G.sample = 1;
  G.record = 1;
  G.job = 0;
  for (int c = 0; c < 3; c++) {
    G.bmin[c] = 0.f;
    G.inv_ext[c] = 1.f;
  }
  G.wf = g->wf;
  G.wb = g->wb;
  G.bias = g->bias;
  G.mix = g->mix;
  G.rec = g->rec;
  G.rstate = g->rstate;
  G.gq = g->gq;
  G.gq_count = g->gq_count;
  G.ring = g->ring;
  G.ring_mask = g->ring_cap - 1u;
  G.ring_count = g->ring_count;
  G.infer_count = g->ring_count + 1;
  return G;
}

void guide_prepare_render(rtg_ctx* ctx, GuideState* g, rtg_scene* scene, uint64_t seed, DevGuide& G) {
  guide_alloc_pool(ctx, g, ctx->pool.capacity);
  guide_init_params(ctx, g, seed);
  launch_fill_zero(g->rstate, ((size_t)g->capacity * kGuideRec + 3) & ~(size_t)3, ctx->stream);
  launch_fill_zero(g->ring_count, 16, ctx->stream);
  G = guide_dev(g);
  for (int c = 0; c < 3; c++) {
    float lo = scene->dev.root_min[c], hi = scene->dev.root_max[c];
    float ext = hi - lo;
    G.bmin[c] = lo;
    G.inv_ext[c] = ext > 0.f ? 1.f / ext : 1.f;
  }
}

DevGuide guide_infer_phase(rtg_ctx* ctx, GuideState* g, const DevGuide& G, int capacity) {
  if (g->infer_capacity < capacity) {
    if (g->infer_capacity) {
      (void)hipStreamSynchronize(ctx->stream);
      for (void* p : g->infer_owned) (void)hipFree(p);
      g->infer_owned.clear();
      g->infer_capacity = 0;
    }
    auto A = [&](size_t bytes) {
      void* p = nullptr;
      hip_check(hipMalloc(&p, bytes), "hipMalloc guide inference buffers");
      g->infer_owned.push_back(p);
      return p;
    };
    g->infer_mix = (float*)A((size_t)capacity * kGuideMixR * sizeof(float));
    g->infer_gq = (uint32_t*)A((size_t)capacity * sizeof(uint32_t));
    g->infer_capacity = capacity;
  }
  DevGuide F = G;
  F.mix = g->infer_mix;
  F.gq = g->infer_gq;
  F.rec = nullptr;
  F.rstate = nullptr;
  F.record = 0;
  return F;
}

// One Adam step on n samples of `samples` starting at `base` (indices masked by `mask`).
void guide_train_step(rtg_ctx* ctx, GuideState* g, const DevGuide& G, const float* samples, uint64_t base,
                      uint32_t mask, int n, float lr) {
  launch_guide_grad(G, samples, base, mask, n, g->S, ctx->stream);
  g->step++;
  launch_guide_adam(g->master, g->adam_m, g->adam_v, g->S.grad, g->step, lr, g->wf, g->wb, g->bias, ctx->stream);
}

}  // namespace rtg

using namespace rtg;

extern "C" {

int rtg_guide_eval(rtg_ctx* ctx, const float* params, const float* in6, int32_t n, float* mix_out) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  try {
    if (!params || !in6 || !mix_out || n < 0) fail(RTG_ERR_INVALID, "bad guide_eval arguments");
    if (n == 0) return RTG_OK;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    GuideState* g = guide_state(ctx);
    guide_set_params(ctx, g, params);
    DevGuide G = guide_dev(g);
    std::vector<void*> own;
    float* din = (float*)dev_alloc(ctx, (size_t)n * 6 * sizeof(float), own);
    float* dout = (float*)dev_alloc(ctx, (size_t)n * kGuideOut * sizeof(float), own);
    try {
      hip_check(hipMemcpy(din, in6, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice), "inputs");
      launch_guide_infer_array(G, din, n, dout, ctx->stream);
      hip_check(hipGetLastError(), "guide infer");
      hip_check(hipStreamSynchronize(ctx->stream), "guide infer");
      hip_check(hipMemcpy(mix_out, dout, (size_t)n * kGuideOut * sizeof(float), hipMemcpyDeviceToHost), "mixture");
    } catch (...) {
      for (void* p : own) (void)hipFree(p);
      throw;
    }
    for (void* p : own) (void)hipFree(p);
    return RTG_OK;
  } catch (const RtgError& e) {
    set_error(e.what());
    return e.code;
  }
}

int rtg_guide_grad(rtg_ctx* ctx, const float* params, const float* samples, int32_t n, float* loss, float* grad) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  try {
    if (!params || !samples || n <= 0 || n > kGuideBatch) fail(RTG_ERR_INVALID, "bad guide_grad arguments");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    GuideState* g = guide_state(ctx);
    guide_set_params(ctx, g, params);
    DevGuide G = guide_dev(g);
    hip_check(hipMemcpy(g->samples, samples, (size_t)n * 10 * sizeof(float), hipMemcpyHostToDevice), "samples");
    launch_guide_grad(G, g->samples, 0, 0xffffffffu, n, g->S, ctx->stream);
    hip_check(hipGetLastError(), "guide grad");
    hip_check(hipStreamSynchronize(ctx->stream), "guide grad");
    float scal[2];
    hip_check(hipMemcpy(scal, g->S.scal, sizeof(scal), hipMemcpyDeviceToHost), "loss");
    if (loss) *loss = scal[1];
    if (grad) hip_check(hipMemcpy(grad, g->S.grad, sizeof(float) * kGuideParams, hipMemcpyDeviceToHost), "grad");
    return RTG_OK;
  } catch (const RtgError& e) {
    set_error(e.what());
    return e.code;
  }
}

int rtg_guide_train(rtg_ctx* ctx, const float* params, const float* samples, int32_t n, int32_t steps, float lr,
                    float* params_out, float* losses) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  try {
    if (!params || !samples || !params_out || n <= 0 || n > kGuideBatch || steps < 0 || !(lr > 0.f))
      fail(RTG_ERR_INVALID, "bad guide_train arguments");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    GuideState* g = guide_state(ctx);
    guide_set_params(ctx, g, params);
    DevGuide G = guide_dev(g);
    hip_check(hipMemcpy(g->samples, samples, (size_t)n * 10 * sizeof(float), hipMemcpyHostToDevice), "samples");
    for (int s = 0; s < steps; s++) {
      guide_train_step(ctx, g, G, g->samples, 0, 0xffffffffu, n, lr);
      if (losses) {
        float scal[2];
        hip_check(hipMemcpyAsync(scal, g->S.scal, sizeof(scal), hipMemcpyDeviceToHost, ctx->stream), "loss");
        hip_check(hipStreamSynchronize(ctx->stream), "guide train");
        losses[s] = scal[1];
      }
    }
    hip_check(hipGetLastError(), "guide train");
    hip_check(hipStreamSynchronize(ctx->stream), "guide train");
    hip_check(hipMemcpy(params_out, g->master, sizeof(float) * kGuideParams, hipMemcpyDeviceToHost), "params");
    return RTG_OK;
  } catch (const RtgError& e) {
    set_error(e.what());
    return e.code;
  }
}

int rtg_guide_params(rtg_ctx* ctx, float* params_out) {
  if (!ctx || !params_out) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  try {
    if (!ctx->guide || !ctx->guide->master) fail(RTG_ERR_INVALID, "no guided render has run on this context");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    hip_check(hipStreamSynchronize(ctx->stream), "guide params");
    hip_check(hipMemcpy(params_out, ctx->guide->master, sizeof(float) * kGuideParams, hipMemcpyDeviceToHost),
              "guide params");
    return RTG_OK;
  } catch (const RtgError& e) {
    set_error(e.what());
    return e.code;
  }
}

int rtg_guide_sample(rtg_ctx* ctx, const float* mix, const float* u, int32_t n, float* dirs, float* pdf) {
  if (!ctx) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  try {
    if (!mix || !u || !dirs || !pdf || n < 0) fail(RTG_ERR_INVALID, "bad guide_sample arguments");
    if (n == 0) return RTG_OK;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    std::vector<void*> own;
    float* dm = (float*)dev_alloc(ctx, (size_t)n * kGuideOut * sizeof(float), own);
    float* du = (float*)dev_alloc(ctx, (size_t)n * 3 * sizeof(float), own);
    float* dd = (float*)dev_alloc(ctx, (size_t)n * 3 * sizeof(float), own);
    float* dp = (float*)dev_alloc(ctx, (size_t)n * sizeof(float), own);
    try {
      hip_check(hipMemcpy(dm, mix, (size_t)n * kGuideOut * sizeof(float), hipMemcpyHostToDevice), "mix");
      hip_check(hipMemcpy(du, u, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice), "u");
      launch_guide_sample_tap(dm, du, n, dd, dp, ctx->stream);
      hip_check(hipGetLastError(), "guide sample");
      hip_check(hipStreamSynchronize(ctx->stream), "guide sample");
      hip_check(hipMemcpy(dirs, dd, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost), "dirs");
      hip_check(hipMemcpy(pdf, dp, (size_t)n * sizeof(float), hipMemcpyDeviceToHost), "pdf");
    } catch (...) {
      for (void* p : own) (void)hipFree(p);
      throw;
    }
    for (void* p : own) (void)hipFree(p);
    return RTG_OK;
  } catch (const RtgError& e) {
    set_error(e.what());
    return e.code;
  }
}

}  // extern "C"

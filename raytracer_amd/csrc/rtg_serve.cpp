// Continuous-batching serving loop (no counterpart in the reference renderer).
//
// Requests (scene + render description) queue on the host and are admitted into the device job
// table while the HBM budget for their FP32 films and a job slot allow. Admitted jobs share the
// context's path pool: every wavefront iteration hands free path slots to the jobs in the host's
// claim order (shortest remaining work first), so a request that arrives while others run starts
// as soon as slots free up, and the paths of several requests are traced in the same queues.
// A path's random numbers depend only on (seed, pixel, sample, dimension), so a request renders
// the same samples alone or in a mix; only the order of the film's float additions differs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "rtg_host.h"

using namespace rtg;

namespace {

using Clock = std::chrono::steady_clock;

enum ReqState { R_PENDING, R_RUNNING, R_DONE, R_CANCELLED };

struct Request {
  int64_t id = 0;
  rtg_scene* scene = nullptr;
  rtg_render_desc desc{};
  DevJob job{};
  int slot = -1;
  float4* film = nullptr;
  size_t film_bytes = 0;
  Clock::time_point t_submit, t_done;
  ReqState state = R_PENDING;
  bool cancel = false;
  uint64_t claimed = 0;       // samples claimed at the last job-table readback
  std::vector<float> result;  // host copy of the film once finished
};

}  // namespace

// Locking: ctx->mtx serialises device work (rtg_server_step holds it for the whole step);
// qm guards the request table and queue, so rtg_submit / rtg_poll / rtg_cancel from other
// threads do not wait for a running step.
struct rtg_server {
  rtg_ctx* ctx = nullptr;
  std::mutex qm;
  uint64_t budget = 0;
  uint64_t used = 0;
  int64_t next_id = 1;
  std::deque<Request*> pending;
  std::map<int64_t, std::unique_ptr<Request>> reqs;
  Request* slots[kMaxJobs] = {};
  DevJob* pinned_jobs = nullptr;  // job table read back after every iteration batch
  // Host sources of asynchronous uploads live in pinned memory owned by the server: a slot's
  // entry is rewritten only after the stream has been synchronised (a slot is freed by the
  // completion check, which follows a synchronisation), so no copy reads memory that changed.
  DevJob* up_jobs = nullptr;
  int32_t* up_order = nullptr;
  rtg_server_stats st{};
};

namespace {

void free_request_device(rtg_server* s, Request* r) {
  if (r->film) {
    (void)hipFree(r->film);
    r->film = nullptr;
    s->used -= r->film_bytes;
    s->st.film_bytes = s->used;
  }
  // The device entry of a released slot stays as it is: it is absent from the claim order and
  // no path references it any more (all of the request's paths have been finalised).
  if (r->slot >= 0) {
    s->slots[r->slot] = nullptr;
    r->slot = -1;
  }
  if (r->scene) {
    r->scene->users--;
    r->scene = nullptr;
  }
}

// Admit queued requests in arrival order while a job slot and the HBM budget allow.
void admit(rtg_server* s) {
  rtg_ctx* ctx = s->ctx;
  std::lock_guard<std::mutex> lk(s->qm);
  while (!s->pending.empty()) {
    Request* r = s->pending.front();
    if (r->cancel) {
      s->pending.pop_front();
      s->st.cancelled++;
      r->state = R_CANCELLED;
      r->t_done = Clock::now();
      free_request_device(s, r);
      continue;
    }
    int slot = -1;
    for (int i = 0; i < kMaxJobs; i++)
      if (!s->slots[i]) {
        slot = i;
        break;
      }
    if (slot < 0) return;
    size_t bytes = (size_t)r->desc.width * r->desc.height * sizeof(float4);
    if (s->used + bytes > s->budget) {
      bool idle = true;
      for (auto* q : s->slots)
        if (q) idle = false;
      if (!idle) return;  // wait for films to be released
      if (bytes > s->budget) {  // can never fit: fail the request
        s->pending.pop_front();
        s->st.cancelled++;
        r->state = R_CANCELLED;
        r->t_done = Clock::now();
        free_request_device(s, r);
        continue;
      }
    }
    hip_check(hipMalloc(&r->film, bytes), "hipMalloc film");
    r->film_bytes = bytes;
    s->used += bytes;
    s->st.film_bytes = s->used;
    s->st.film_bytes_peak = std::max<uint64_t>(s->st.film_bytes_peak, s->used);
    s->st.admitted++;
    launch_fill_zero(r->film, bytes, ctx->stream);
    r->job = make_job(ctx, r->scene, r->desc, r->film);
    upload_job_filter(ctx, r->desc, r->job, slot);
    r->claimed = 0;
    r->slot = slot;
    s->slots[slot] = r;
    s->up_jobs[slot] = r->job;
    hip_check(hipMemcpyAsync(ctx->jobs_dev + slot, &s->up_jobs[slot], sizeof(DevJob), hipMemcpyHostToDevice,
                             ctx->stream),
              "job upload");
    r->state = R_RUNNING;
    s->pending.pop_front();
  }
}

}  // namespace

extern "C" {

int rtg_server_create(rtg_ctx* ctx, uint64_t hbm_budget_bytes, rtg_server** out) {
  if (!ctx || !out) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  try {
    if (ctx->server) fail(RTG_ERR_INVALID, "the context already has a server");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    std::unique_ptr<rtg_server> s(new rtg_server());
    s->ctx = ctx;
    s->budget = hbm_budget_bytes ? hbm_budget_bytes : (8ull << 30);
    hip_check(hipHostMalloc((void**)&s->pinned_jobs, sizeof(DevJob) * kMaxJobs, hipHostMallocDefault),
              "hipHostMalloc");
    hip_check(hipHostMalloc((void**)&s->up_jobs, sizeof(DevJob) * kMaxJobs, hipHostMallocDefault),
              "hipHostMalloc");
    hip_check(hipHostMalloc((void**)&s->up_order, sizeof(int32_t) * kMaxJobs, hipHostMallocDefault),
              "hipHostMalloc");
    std::memset(s->pinned_jobs, 0, sizeof(DevJob) * kMaxJobs);
    ensure_pool(ctx, ctx->pool_slots);
    ctx->pool.count_work = 0;
    launch_reset_pool(ctx->pool, ctx->stream);
    launch_fill_zero(ctx->jobs_dev, sizeof(DevJob) * kMaxJobs, ctx->stream);
    hip_check(hipStreamSynchronize(ctx->stream), "server init");
    ctx->server = s.get();
    *out = s.release();
    return RTG_OK;
  } catch (const RtgError& e) {
    set_error(e.what());
    return e.code;
  }
}

int rtg_server_destroy(rtg_server* s) {
  if (!s) return RTG_ERR_INVALID;
  rtg_ctx* ctx = s->ctx;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  {
    // the request-table lock is released before the server (which owns the mutex) is deleted
    std::lock_guard<std::mutex> lq(s->qm);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : s->reqs) {
      Request* r = kv.second.get();
      if (r->film) {
        (void)hipFree(r->film);
        r->film = nullptr;
      }
      if (r->scene) r->scene->users--;
      r->scene = nullptr;
    }
    if (s->pinned_jobs) (void)hipHostFree(s->pinned_jobs);
    if (s->up_jobs) (void)hipHostFree(s->up_jobs);
    if (s->up_order) (void)hipHostFree(s->up_order);
  }
  ctx->server = nullptr;
  delete s;
  return RTG_OK;
}

int rtg_submit(rtg_server* s, rtg_scene* scene, const rtg_render_desc* desc, int64_t* job_id) {
  if (!s || !scene || !desc || !job_id) return RTG_ERR_INVALID;
  rtg_ctx* ctx = s->ctx;
  std::lock_guard<std::mutex> lk(s->qm);
  try {
    // validate now so a bad request fails at submission
    if (desc->guiding) fail(RTG_ERR_INVALID, "guided renders are not served (use rtg_render)");
    if (desc->integrator != RTG_INTEGRATOR_PATH_MIS)
      fail(RTG_ERR_INVALID, "volumetric renders are not served (use rtg_render)");
    if (scene->all_lights) fail(RTG_ERR_INVALID, "all_lights scenes are not served (use rtg_render)");
    (void)make_job(ctx, scene, *desc, nullptr);
    std::unique_ptr<Request> r(new Request());
    r->id = s->next_id++;
    r->scene = scene;
    scene->users++;
    r->desc = *desc;
    r->t_submit = Clock::now();
    s->pending.push_back(r.get());
    s->st.submitted++;
    *job_id = r->id;
    s->reqs[r->id] = std::move(r);
    return RTG_OK;
  } catch (const RtgError& e) {
    set_error(e.what());
    return e.code;
  }
}

int rtg_server_step(rtg_server* s, double max_ms) {
  if (!s) return RTG_ERR_INVALID;
  rtg_ctx* ctx = s->ctx;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  try {
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    auto t0 = Clock::now();
    LaunchCfg cfg = launch_cfg(ctx);
    ctx->pool.all_lights = 0;  // served scenes sample one light per vertex
    const int batch = 2;
    while (true) {
      admit(s);
      // claim order: shortest remaining work first (unclaimed samples of each running job)
      std::vector<std::pair<uint64_t, int>> run;
      for (int i = 0; i < kMaxJobs; i++)
        if (s->slots[i]) run.push_back({s->slots[i]->job.total - s->slots[i]->claimed, i});
      if (run.empty()) break;
      s->st.max_running = std::max<int32_t>(s->st.max_running, (int32_t)run.size());
      std::sort(run.begin(), run.end());
      int norder = 0;
      for (auto& p : run) s->up_order[norder++] = p.second;
      hip_check(hipMemcpyAsync(ctx->order_dev, s->up_order, sizeof(int32_t) * norder, hipMemcpyHostToDevice,
                               ctx->stream),
                "order");
      for (int b = 0; b < batch; b++)
        launch_iteration(ctx->pool, ctx->scenes_dev, ctx->jobs_dev, ctx->order_dev, norder, cfg, ctx->stream);
      s->st.iterations += batch;
      hip_check(hipGetLastError(), "kernel launch");
      hip_check(hipMemcpyAsync(s->pinned_jobs, ctx->jobs_dev, sizeof(DevJob) * kMaxJobs, hipMemcpyDeviceToHost,
                               ctx->stream),
                "job readback");
      hip_check(hipStreamSynchronize(ctx->stream), "server iteration");
      // completions
      std::lock_guard<std::mutex> lq(s->qm);
      for (int i = 0; i < kMaxJobs; i++) {
        Request* r = s->slots[i];
        if (!r) continue;
        const DevJob& dj = s->pinned_jobs[i];
        r->claimed = std::min<uint64_t>(dj.next, r->job.total);
        uint64_t target = r->job.total;
        if (r->cancel) {
          if (dj.active) {
            // stop further claims; paths already claimed finish normally
            s->up_jobs[i].active = 0;
            hip_check(hipMemcpyAsync(&ctx->jobs_dev[i].active, &s->up_jobs[i].active, sizeof(int32_t),
                                     hipMemcpyHostToDevice, ctx->stream),
                      "cancel");
            hip_check(hipStreamSynchronize(ctx->stream), "cancel");
            s->pinned_jobs[i].active = 0;
            continue;  // re-read the claim count after the stop is visible
          }
          target = std::min<uint64_t>(dj.next, r->job.total);
        }
        if (dj.done >= target) {
          r->t_done = Clock::now();
          if (!r->cancel) {
            r->result.resize(r->film_bytes / sizeof(float));
            hip_check(hipMemcpy(r->result.data(), r->film, r->film_bytes, hipMemcpyDeviceToHost), "film readback");
          }
          r->state = r->cancel ? R_CANCELLED : R_DONE;
          if (r->cancel) {
            s->st.cancelled++;
          } else {
            s->st.completed++;
            s->st.samples += r->job.total;
          }
          free_request_device(s, r);
        }
      }
      double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
      if (max_ms > 0 && ms >= max_ms) break;
    }
    return RTG_OK;
  } catch (const RtgError& e) {
    set_error(e.what());
    return e.code;
  }
}

int rtg_poll(rtg_server* s, int64_t job_id, float* out_rgbw, double* latency_ms) {
  if (!s) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->qm);
  auto it = s->reqs.find(job_id);
  if (it == s->reqs.end()) {
    set_error("unknown job id");
    return RTG_ERR_INVALID;
  }
  Request* r = it->second.get();
  if (r->state == R_PENDING || r->state == R_RUNNING) return RTG_ERR_NOT_READY;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  if (r->state == R_CANCELLED) return RTG_ERR_CANCELLED;
  if (out_rgbw) std::memcpy(out_rgbw, r->result.data(), r->result.size() * sizeof(float));
  return RTG_OK;
}

int rtg_cancel(rtg_server* s, int64_t job_id) {
  if (!s) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->qm);
  auto it = s->reqs.find(job_id);
  if (it == s->reqs.end()) {
    set_error("unknown job id");
    return RTG_ERR_INVALID;
  }
  it->second->cancel = true;
  return RTG_OK;
}

int rtg_server_query(rtg_server* s, rtg_server_stats* out) {
  if (!s || !out) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->qm);
  rtg_server_stats st = s->st;
  st.pending = (int32_t)s->pending.size();
  st.running = 0;
  for (auto* r : s->slots)
    if (r) st.running++;
  *out = st;
  return RTG_OK;
}

}  // extern "C"

// Continuous-batching serving loop (no counterpart in the reference renderer; it replaces the
// job loop of JobManager::addJobs / jobRun, core/JobManager.cpp:103-232, for many renders).
//
// Requests (scene + render description) queue on the host and are admitted into the device job
// table while the HBM budget for their FP32 films and a job slot allow. Admitted jobs share the
// context's path pool: every wavefront iteration hands free path slots to the jobs in the host's
// claim order (shortest remaining work first), so a request that arrives while others run starts
// as soon as slots free up, and the paths of several requests — PathTracingMIS and VolPathTracing
// jobs, any resident scene — are traced in the same queues. A path's random numbers depend only
// on (seed, pixel, sample, dimension), so a request renders the same samples alone or in a mix;
// only the order of the film's float additions differs.
//
// The device side of a step is a replayed hipGraph of a batch of iterations; the host between
// batches only uploads the claim order, reads the job table back and retires finished requests.
// The loop runs on the caller's thread (rtg_server_step, rtg_wait) or on the server's own thread
// (rtg_server_start).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
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
  float4* film = nullptr;     // device film (the server's, or the caller's for an external film)
  bool external = false;      // film owned by the caller (rtg_render_device through the server)
  size_t film_bytes = 0;      // bytes the budget accounts for (0 for an external film)
  Clock::time_point t_submit, t_done;
  ReqState state = R_PENDING;
  bool cancel = false;
  uint64_t claimed = 0;       // samples claimed at the last job-table readback
  std::vector<float> result;  // host copy of the film once finished
};

}  // namespace

// Locking: ctx->mtx serialises device work (a step holds it for one slice of iteration batches);
// qm guards the request table and queue, so rtg_submit / rtg_poll / rtg_wait / rtg_cancel from
// other threads do not wait for a running step. cv_done is signalled (under qm) whenever requests
// finish; cv_work whenever a request arrives or the server thread is asked to stop.
struct rtg_server {
  rtg_ctx* ctx = nullptr;
  std::mutex qm;
  std::condition_variable cv_done, cv_work;
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
  int32_t* up_order = nullptr;    // [0] count, [1..] job slots in claim order
  uint32_t* pinned_ovf = nullptr; // the pool's traversal-stack overflow flag, read after each batch
  rtg_server_stats st{};
  std::thread loop;
  bool loop_on = false;           // the server thread runs the loop
  bool stop = false;
  int loop_error = RTG_OK;        // first error the server thread met
  std::string loop_error_msg;
};

namespace {

void free_request_device(rtg_server* s, Request* r) {
  if (r->film && !r->external) {
    (void)hipFree(r->film);
    s->used -= r->film_bytes;
    s->st.film_bytes = s->used;
  }
  r->film = nullptr;
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

void finish_request(rtg_server* s, Request* r, ReqState state) {
  r->state = state;
  r->t_done = Clock::now();
  if (state == R_CANCELLED) s->st.cancelled++;
  free_request_device(s, r);
}

bool any_running(const rtg_server* s) {
  for (auto* q : s->slots)
    if (q) return true;
  return false;
}

// Admit queued requests in arrival order while a job slot and the HBM budget allow. Caller holds qm.
void admit(rtg_server* s) {
  rtg_ctx* ctx = s->ctx;
  while (!s->pending.empty()) {
    Request* r = s->pending.front();
    if (r->cancel) {
      s->pending.pop_front();
      finish_request(s, r, R_CANCELLED);
      continue;
    }
    int slot = -1;
    for (int i = 0; i < kMaxJobs; i++)
      if (!s->slots[i]) {
        slot = i;
        break;
      }
    if (slot < 0) return;
    const size_t bytes = r->external ? 0 : (size_t)r->desc.width * r->desc.height * sizeof(float4);
    if (s->used + bytes > s->budget) {
      if (any_running(s)) return;  // wait for films to be released
      if (bytes > s->budget) {     // can never fit: fail the request
        s->pending.pop_front();
        finish_request(s, r, R_CANCELLED);
        continue;
      }
    }
    if (!r->external) {
      hip_check(hipMalloc(&r->film, bytes), "hipMalloc film");
      r->film_bytes = bytes;
      s->used += bytes;
      s->st.film_bytes = s->used;
      s->st.film_bytes_peak = std::max<uint64_t>(s->st.film_bytes_peak, s->used);
      launch_fill_zero(r->film, bytes, ctx->stream);
    }
    s->st.admitted++;
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

// One slice of the loop: iteration batches until nothing is pending or running, or max_ms passed
// (0: until idle). Caller holds ctx->mtx.
void step_impl(rtg_server* s, double max_ms) {
  rtg_ctx* ctx = s->ctx;
  hip_check(hipSetDevice(ctx->device), "hipSetDevice");
  const auto t0 = Clock::now();
  const LaunchCfg cfg = launch_cfg(ctx);
  ctx->pool.all_lights = 0;  // served scenes sample one light per vertex
  ctx->pool.count_work = 0;
  const int batch = 2;
  while (true) {
    int norder = 0;
    bool vol = false;
    {
      std::lock_guard<std::mutex> lq(s->qm);
      admit(s);
      // claim order: shortest remaining work first (unclaimed samples of each running job)
      std::vector<std::pair<uint64_t, int>> run;
      for (int i = 0; i < kMaxJobs; i++)
        if (s->slots[i]) {
          run.push_back({s->slots[i]->job.total - s->slots[i]->claimed, i});
          vol = vol || s->slots[i]->desc.integrator == RTG_INTEGRATOR_VOLPATH;
        }
      if (run.empty()) break;
      s->st.max_running = std::max<int32_t>(s->st.max_running, (int32_t)run.size());
      std::sort(run.begin(), run.end());
      for (auto& p : run) s->up_order[1 + norder++] = p.second;
      s->up_order[0] = norder;
    }
    hip_check(hipMemcpyAsync(ctx->order_dev, s->up_order, sizeof(int32_t) * (1 + norder), hipMemcpyHostToDevice,
                             ctx->stream),
              "order");
    run_iterations(ctx, vol ? IT_MIXED : IT_PLAIN, cfg, batch);
    s->st.iterations += batch;
    hip_check(hipGetLastError(), "kernel launch");
    hip_check(hipMemcpyAsync(s->pinned_jobs, ctx->jobs_dev, sizeof(DevJob) * kMaxJobs, hipMemcpyDeviceToHost,
                             ctx->stream),
              "job readback");
    hip_check(hipMemcpyAsync(s->pinned_ovf, ctx->pool.counters + C_OVF, sizeof(uint32_t), hipMemcpyDeviceToHost,
                             ctx->stream),
              "overflow flag readback");
    hip_check(hipStreamSynchronize(ctx->stream), "server iteration");
    if (*s->pinned_ovf) fail(RTG_ERR_HIP, "BVH traversal stack overflow (the builder's depth bound was exceeded)");
    // completions
    bool finished = false;
    {
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
          if (!r->cancel && !r->external) {
            r->result.resize(r->film_bytes / sizeof(float));
            hip_check(hipMemcpy(r->result.data(), r->film, r->film_bytes, hipMemcpyDeviceToHost), "film readback");
          }
          if (!r->cancel) {
            s->st.completed++;
            s->st.samples += r->job.total;
          }
          finish_request(s, r, r->cancel ? R_CANCELLED : R_DONE);
          finished = true;
        }
      }
    }
    if (finished) s->cv_done.notify_all();
    const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    if (max_ms > 0 && ms >= max_ms) break;
  }
}

int step_locked(rtg_server* s, double max_ms) {
  std::lock_guard<std::mutex> lk(s->ctx->mtx);
  try {
    step_impl(s, max_ms);
    return RTG_OK;
  } catch (const RtgError& e) {
    set_error(e.what());
    return e.code;
  }
}

// The server thread: runs slices of the loop while requests are pending or running, sleeps on
// cv_work otherwise. A slice ends after 20 ms so that other calls on the context get the device.
void loop_main(rtg_server* s) {
  while (true) {
    {
      std::unique_lock<std::mutex> lq(s->qm);
      s->cv_work.wait(lq, [&] { return s->stop || !s->pending.empty() || any_running(s); });
      if (s->stop) return;
    }
    const int rc = step_locked(s, 20.0);
    if (rc != RTG_OK) {
      std::lock_guard<std::mutex> lq(s->qm);
      s->loop_error = rc;
      s->loop_error_msg = rtg_last_error(s->ctx);
      // fail everything still waiting so that no rtg_wait blocks forever
      for (auto& kv : s->reqs) {
        Request* r = kv.second.get();
        if (r->state == R_PENDING || r->state == R_RUNNING) finish_request(s, r, R_CANCELLED);
      }
      s->pending.clear();
      s->cv_done.notify_all();
      return;
    }
  }
}

int result_of(rtg_server* s, Request* r, float* out_rgbw, double* latency_ms) {
  if (r->state == R_PENDING || r->state == R_RUNNING) return RTG_ERR_NOT_READY;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  if (r->state == R_CANCELLED) {
    if (s->loop_error != RTG_OK) {
      set_error(("serving loop failed: " + s->loop_error_msg).c_str());
      return s->loop_error;
    }
    return RTG_ERR_CANCELLED;
  }
  if (out_rgbw && !r->external) std::memcpy(out_rgbw, r->result.data(), r->result.size() * sizeof(float));
  return RTG_OK;
}

int submit_impl(rtg_server* s, rtg_scene* scene, const rtg_render_desc* desc, float* external, int64_t* job_id) {
  rtg_ctx* ctx = s->ctx;
  std::lock_guard<std::mutex> lk(s->qm);
  try {
    // validate now so a bad request fails at submission
    if (desc->guiding == RTG_GUIDING_LEARNED)
      fail(RTG_ERR_INVALID, "renders with learned guiding are not served (use rtg_render without a server)");
    if (desc->integrator != RTG_INTEGRATOR_PATH_MIS && desc->integrator != RTG_INTEGRATOR_VOLPATH)
      fail(RTG_ERR_INVALID, "unknown integrator");
    // all_lights: PathTracingMIS samples every light; VolPathTracing samples one whatever the strategy
    if (scene->all_lights && desc->integrator == RTG_INTEGRATOR_PATH_MIS)
      fail(RTG_ERR_INVALID, "all_lights scenes are not served (use rtg_render without a server)");
    if (s->loop_error != RTG_OK) fail(s->loop_error, "the serving loop failed: " + s->loop_error_msg);
    (void)make_job(ctx, scene, *desc, nullptr);
    std::unique_ptr<Request> r(new Request());
    r->id = s->next_id++;
    r->scene = scene;
    scene->users++;
    r->desc = *desc;
    r->film = (float4*)external;
    r->external = external != nullptr;
    r->t_submit = Clock::now();
    s->pending.push_back(r.get());
    s->st.submitted++;
    *job_id = r->id;
    s->reqs[r->id] = std::move(r);
  } catch (const RtgError& e) {
    set_error(e.what());
    return e.code;
  }
  s->cv_work.notify_one();
  return RTG_OK;
}

int wait_impl(rtg_server* s, int64_t job_id, double timeout_ms, float* out_rgbw, double* latency_ms) {
  const auto t0 = Clock::now();
  auto left_ms = [&]() {
    return timeout_ms - std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
  };
  while (true) {
    bool self_drive = false;
    {
      std::unique_lock<std::mutex> lq(s->qm);
      auto it = s->reqs.find(job_id);
      if (it == s->reqs.end()) {
        set_error("unknown job id");
        return RTG_ERR_INVALID;
      }
      Request* r = it->second.get();
      auto ready = [&] { return r->state == R_DONE || r->state == R_CANCELLED; };
      if (ready()) return result_of(s, r, out_rgbw, latency_ms);
      if (timeout_ms > 0 && left_ms() <= 0) return RTG_ERR_NOT_READY;
      if (s->loop_on) {
        if (timeout_ms > 0)
          s->cv_done.wait_for(lq, std::chrono::duration<double, std::milli>(left_ms()), ready);
        else
          s->cv_done.wait(lq, ready);
        if (ready()) return result_of(s, r, out_rgbw, latency_ms);
        continue;
      }
      self_drive = true;
    }
    // no server thread: this caller runs the loop until its request is done
    if (self_drive) {
      const int rc = step_locked(s, 5.0);
      if (rc != RTG_OK) return rc;
    }
  }
}

}  // namespace

namespace rtg {
// rtg_render / rtg_render_device on a context with a server: the render is submitted as a request
// (into `film_dev` when given: the caller's device buffer, not cleared) and waited for.
int serve_render(rtg_server* s, rtg_scene* scene, const rtg_render_desc* desc, float* out_host, float* film_dev,
                 rtg_stats* st) {
  int64_t id = 0;
  int rc = submit_impl(s, scene, desc, film_dev, &id);
  if (rc != RTG_OK) return rc;
  double lat = 0.0;
  rc = wait_impl(s, id, 0.0, out_host, &lat);
  if (rc != RTG_OK) return rc;
  if (st) {
    std::memset(st, 0, sizeof(*st));
    std::lock_guard<std::mutex> lq(s->qm);
    const Request* r = s->reqs[id].get();
    st->samples = r->job.total;
    st->milliseconds = lat;
  }
  std::lock_guard<std::mutex> lq(s->qm);
  s->reqs.erase(id);
  return RTG_OK;
}
}  // namespace rtg

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
    hip_check(hipHostMalloc((void**)&s->up_order, sizeof(int32_t) * (kMaxJobs + 1), hipHostMallocDefault),
              "hipHostMalloc");
    hip_check(hipHostMalloc((void**)&s->pinned_ovf, 64, hipHostMallocDefault), "hipHostMalloc");
    *s->pinned_ovf = 0;
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

int rtg_server_start(rtg_server* s) {
  if (!s) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lq(s->qm);
  if (s->loop_on) {
    set_error("the server thread already runs");
    return RTG_ERR_INVALID;
  }
  s->stop = false;
  s->loop_on = true;
  s->loop = std::thread(loop_main, s);
  return RTG_OK;
}

int rtg_server_destroy(rtg_server* s) {
  if (!s) return RTG_ERR_INVALID;
  rtg_ctx* ctx = s->ctx;
  {
    std::lock_guard<std::mutex> lq(s->qm);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->loop.joinable()) s->loop.join();
  std::lock_guard<std::mutex> lk(ctx->mtx);
  {
    // the request-table lock is released before the server (which owns the mutex) is deleted
    std::lock_guard<std::mutex> lq(s->qm);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : s->reqs) {
      Request* r = kv.second.get();
      if (r->film && !r->external) (void)hipFree(r->film);
      r->film = nullptr;
      if (r->scene) r->scene->users--;
      r->scene = nullptr;
    }
    if (s->pinned_jobs) (void)hipHostFree(s->pinned_jobs);
    if (s->up_jobs) (void)hipHostFree(s->up_jobs);
    if (s->up_order) (void)hipHostFree(s->up_order);
    if (s->pinned_ovf) (void)hipHostFree(s->pinned_ovf);
  }
  ctx->server = nullptr;
  delete s;
  return RTG_OK;
}

int rtg_submit(rtg_server* s, rtg_scene* scene, const rtg_render_desc* desc, int64_t* job_id) {
  if (!s || !scene || !desc || !job_id) return RTG_ERR_INVALID;
  return submit_impl(s, scene, desc, nullptr, job_id);
}

int rtg_server_step(rtg_server* s, double max_ms) {
  if (!s) return RTG_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lq(s->qm);
    if (s->loop_on) {
      set_error("the server thread runs the loop (rtg_server_start); use rtg_wait");
      return RTG_ERR_INVALID;
    }
  }
  return step_locked(s, max_ms);
}

int rtg_poll(rtg_server* s, int64_t job_id, float* out_rgbw, double* latency_ms) {
  if (!s) return RTG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->qm);
  auto it = s->reqs.find(job_id);
  if (it == s->reqs.end()) {
    set_error("unknown job id");
    return RTG_ERR_INVALID;
  }
  return result_of(s, it->second.get(), out_rgbw, latency_ms);
}

int rtg_wait(rtg_server* s, int64_t job_id, double timeout_ms, float* out_rgbw, double* latency_ms) {
  if (!s) return RTG_ERR_INVALID;
  return wait_impl(s, job_id, timeout_ms, out_rgbw, latency_ms);
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
  st.graph_launches = s->ctx->graph_launches;
  *out = st;
  return RTG_OK;
}

}  // extern "C"

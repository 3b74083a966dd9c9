// Wavefront MIS path tracer kernels for gfx950 (wave64).
//
// One iteration of the pool:
//   regen -> extend -> classify -> shade (diffuse, conductor, dielectric, miss) -> compact -> shadow.
// Slots keep per-path state in structure-of-arrays buffers. Every stage consumes a compacted
// queue, so lanes of a wave run the same BSDF code. Queues are built by block-wide prefix sums
// over 2048-entry chunks with ONE global atomic per chunk: same-address atomics are resolved
// outside the eight XCDs' L2 caches and serialise, so they are kept to ~n/2048 per launch.
// The control flow restates integrator/PathTracingMIS.cpp:27-263 and
// integrator/Integrator.cpp:176-238 split at the two ray casts of each bounce.
#include <cstring>
#include <mutex>

#include "rtg_device.h"

namespace rtg {
using namespace dev;


// ---------------------------------------------------------------- camera (NoriPerspective.cpp:80-118)
__device__ __forceinline__ void camera_ray(const DevJob& J, float px, float py, f3& o, f3& d, float& tmin,
                                           float& tmax) {
  const float4* S2C = reinterpret_cast<const float4*>(J.s2c);
  const float4* C2W = reinterpret_cast<const float4*>(J.c2w);
  f3 nearP = xform_point(S2C, mk(px * J.inv_w, py * J.inv_h, 0.0f));
  f3 dd = normalized(nearP);
  float invZ = 1.0f / dd.z;
  o = xform_point(C2W, mk(0.f, 0.f, 0.f));
  d = xform_vector(C2W, dd);
  tmin = J.near_clip * invZ;
  tmax = J.far_clip * invZ;
}

// Sample number s of a job -> (pixel, sample index). Pixels are walked in 8-row strips,
// column by column, so a wave covers an 8x8 block (coherent primary rays).
__device__ __forceinline__ void decode_sample(const DevJob& J, uint64_t s, uint32_t& pixel, uint32_t& si,
                                              int& x, int& y) {
  uint64_t npix = (uint64_t)J.width * (uint64_t)J.height;
  si = (uint32_t)(s / npix);
  uint32_t r = (uint32_t)(s - (uint64_t)si * npix);
  uint32_t strip = r / (8u * (uint32_t)J.width);
  uint32_t hs = min(8u, (uint32_t)J.height - strip * 8u);
  uint32_t within = r - strip * 8u * (uint32_t)J.width;
  x = (int)(within / hs);
  y = (int)(strip * 8u + within % hs);
  pixel = (uint32_t)y * (uint32_t)J.width + (uint32_t)x;
}

__device__ __forceinline__ void film_add(const DevJob& J, int x, int y, f3 L, float w) {
  float4* a = J.accum + (size_t)y * J.width + x;
  // hardware float atomics (the film is ordinary coarse-grained device memory)
  unsafeAtomicAdd(&a->x, L.x * w);
  unsafeAtomicAdd(&a->y, L.y * w);
  unsafeAtomicAdd(&a->z, L.z * w);
  unsafeAtomicAdd(&a->w, w);
}

// Splat into the film (core/Screen.cpp:44-111, SampleJob.cpp:35-71): the sample goes to the
// pixels of its 32x32 block's sub-screen (the block plus the filter's overlap border) within the
// filter radius of (x + sx - 0.5, y + sy - 0.5), weighted by the filter table; sub-screen pixels
// outside the image are dropped. Box filter of radius 0.5 (filter 0): its own pixel, and for a
// jitter of exactly 0 also the previous pixel inside the same block.
__device__ __forceinline__ void splat(const DevJob& J, uint32_t pixel, float sx, float sy, f3 L) {
  const int x = (int)(pixel % (uint32_t)J.width), y = (int)(pixel / (uint32_t)J.width);
  if (J.filter == 0) {
    int x0 = x, y0 = y;
    if (sx == 0.f && (x & 31) != 0) x0 = x - 1;
    if (sy == 0.f && (y & 31) != 0) y0 = y - 1;
    for (int yy = y0; yy <= y; yy++)
      for (int xx = x0; xx <= x; xx++)
        if (!J.gather || xx != x || yy != y) film_add(J, xx, yy, L, 1.0f);  // gather: own pixel later
    return;
  }
  const int bx0 = x & ~31, by0 = y & ~31;
  const int pw = min(32, J.width - bx0) + 2 * J.ox, ph = min(32, J.height - by0) + 2 * J.oy;
  const float fx = ((float)x + sx) - 0.5f - (float)(bx0 - J.ox);
  const float fy = ((float)y + sy) - 0.5f - (float)(by0 - J.oy);
  const int x0 = max(0, (int)ceilf(fx - J.rx)), x1 = min(pw - 1, (int)floorf(fx + J.rx));
  const int y0 = max(0, (int)ceilf(fy - J.ry)), y1 = min(ph - 1, (int)floorf(fy + J.ry));
  for (int py = y0; py <= y1; py++) {
    const int gy = by0 - J.oy + py;
    if (gy < 0 || gy >= J.height) continue;
    // ReconstructionFilter.h:53-62: table index by truncation
    const int iy = min((int)((32.f * fabsf((float)py - fy)) / J.ry), 31);
    for (int px = x0; px <= x1; px++) {
      const int gx = bx0 - J.ox + px;
      if (gx < 0 || gx >= J.width) continue;
      const int ix = min((int)((32.f * fabsf((float)px - fx)) / J.rx), 31);
      film_add(J, gx, gy, L, J.lut[ix + iy * 32]);
    }
  }
}

// ---------------------------------------------------------------- block primitives
// Exclusive prefix sum over the block (NW waves); total over the block in `total`.
template <int NW, typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* s_w, T& total) {
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    T y = __shfl_up(x, off);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) s_w[wave] = x;
  __syncthreads();
  T woff = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; w++) {
    T c = s_w[w];
    if ((uint32_t)w < wave) woff += c;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return woff + x - v;
}
// One value per block: thread 0 reserves `total` entries of a queue and every thread learns the base.
__device__ __forceinline__ uint32_t block_reserve(uint32_t* counter, uint32_t total, uint32_t* s_base) {
  if (threadIdx.x == 0) *s_base = total ? atomicAdd(counter, total) : 0u;
  __syncthreads();
  uint32_t b = *s_base;
  __syncthreads();
  return b;
}

// ---------------------------------------------------------------- regen
// Each block owns chunks of kRegenSlots consecutive slots (kRegenK per thread, strided by the block size).
// It finalises finished paths into their job's film, claims new samples for its free slots (one
// 64-bit atomic per chunk and job) and appends its active slots to the extension queue (one atomic
// per chunk).
constexpr int kRegenK = 8;
constexpr int kRegenSlots = 256 * kRegenK;
static_assert(kRegenSlots == kClaimChunk, "the gather's slot map assumes one claim chunk per regeneration block");

// Vertices the network guides: diffuse surfaces at the first kGuideRec depths. (Rough conductors
// are left to their own sampler: their GGX lobes are narrow, so the product the network would
// learn is the BSDF's own shape, and guiding them cost more time than it saved.)
__device__ __forceinline__ bool guided_vertex(int mt, uint32_t depth) {
  return depth < (uint32_t)kGuideRec && mt == 0;
}

// GUIDED: every guided vertex of a finished path whose radiance snapshot was taken appends one
// training sample (position, normal, direction, luminance of the radiance gathered after the
// vertex divided by the throughput before it) to the guide's ring; one 64-bit atomic per chunk.
// While the guide does not record, the per-slot records are neither read nor written (an inference
// phase runs without them).
template <bool GUIDED>
__global__ void __launch_bounds__(256, 4) k_regen(DevPool P, DevJob* jobs, const int32_t* order, DevGuide G) {
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_base;
  __shared__ unsigned long long s_base64;
  __shared__ uint32_t s_kc[kRegenK * 4];
  __shared__ uint32_t s_done[kMaxJobsDev];
  __shared__ uint32_t s_nan;
  __shared__ uint32_t s_claim;
  uint32_t* C = P.counters;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // shade queues are filled by the classify stage of this iteration
    C[C_SHADE + 0] = 0; C[C_SHADE + 1] = 0; C[C_SHADE + 2] = 0; C[C_SHADE + 3] = 0; C[C_VOLQ] = 0;
    C[C_XFETCH] = 0;  // extension-ray batches of this iteration
    if (GUIDED) *G.gq_count = 0;
  }
  for (int j = threadIdx.x; j < kMaxJobsDev; j += blockDim.x) s_done[j] = 0;
  if (threadIdx.x == 0) {
    s_nan = 0;
    // samples left to claim in any job of the claim order (claim counters only grow, so a block
    // that sees none now would find none later either)
    uint32_t left = 0;
    for (int jj = 0; jj < order[0] && !left; jj++) {
      const DevJob& J = jobs[order[1 + jj]];
      left = J.active && *(volatile unsigned long long*)&J.next < J.total;
    }
    s_claim = left;
  }
  __syncthreads();
  const bool claimable = s_claim != 0;
  const int cap = P.capacity;
  const int nchunks = (cap + kRegenSlots - 1) / kRegenSlots;
  for (int chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    // slot base + k * 256 is this thread's k-th: every pool access of a wave is coalesced
    const int base = chunk * kRegenSlots + threadIdx.x;
    uint32_t need = 0;  // bit k: slot base + k * 256 is free
    uint32_t act = 0;   // bit k: slot is active after this kernel
    uint32_t trn = 0;   // bit k: slot yields a guide training sample
    // the loads of all kRegenK slots are issued before any of them is used
    uint32_t stv[kRegenK], flv[kRegenK];
    float4 Lv[kRegenK];
#pragma unroll
    for (int k = 0; k < kRegenK; k++) {
      const int i = base + k * 256;
      stv[k] = i < cap ? P.state[i] : ST_NONE;
    }
    // a chunk with no finished or active path, and either no free slot or nothing to claim, has
    // nothing to do (late iterations of a render whose whole sample set is in flight)
    {
      bool busy = false;
#pragma unroll
      for (int k = 0; k < kRegenK; k++)
        busy = busy || stv[k] == ST_DONE || stv[k] == ST_ACTIVE || (stv[k] == ST_FREE && claimable);
      if (!__syncthreads_or(busy)) continue;
    }
#pragma unroll
    for (int k = 0; k < kRegenK; k++) {
      const int i = base + k * 256;
      if (stv[k] == ST_DONE) {
        flv[k] = P.flags[i];
        Lv[k] = P.L[i];
      }
    }
#pragma unroll
    for (int k = 0; k < kRegenK; k++) {
      int i = base + k * 256;
      uint32_t st = stv[k];
      if (st == ST_DONE) {
        uint32_t f = flv[k];
        uint32_t jb = fl_job(f);
        const DevJob& J = jobs[jb];
        float4 L = Lv[k];
        f3 l = mk(L.x, L.y, L.z);
        const bool bad = isnan(l.x) || isnan(l.y) || isnan(l.z);
        if (bad) {
          atomicAdd(&s_nan, 1u);
        } else if (J.filter == 0) {
          // box filter: only whether the jitter was 0 matters (flags, set at the claim)
          splat(J, P.pixel[i], (f & FL_XEDGE) ? 0.f : 0.5f, (f & FL_YEDGE) ? 0.f : 0.5f, l);
        } else {
          uint32_t pix = P.pixel[i], si = P.sample[i];
          PathRng rng = make_rng(J.seed, pix, si, J);
          float sx, sy;
          sample2d(rng, 0, sx, sy);
          splat(J, pix, sx, sy, l);
        }
        if (GUIDED && G.record) {
#pragma unroll
          for (int r = 0; r < kGuideRec; r++) {
            const uint8_t rs = G.rstate[(size_t)i * kGuideRec + r];
            if (rs == 2 && !bad) {
              // f*cos*L_in / p of the vertex = radiance gathered after it / throughput before it
              float4* q = G.rec + ((size_t)i * kGuideRec + r) * 4;
              const float4 q2 = q[2];
              float4 q3 = q[3];
              const f3 d = mk(l.x - q2.y, l.y - q2.z, l.z - q2.w);
              const f3 t = mk(q3.x > 0.f ? d.x / q3.x : 0.f, q3.y > 0.f ? d.y / q3.y : 0.f,
                              q3.z > 0.f ? d.z / q3.z : 0.f);
              const float w = luminance(t);
              if (w > 0.f && w < 3.0e38f) {
                q3.w = w;
                q[3] = q3;
                trn |= 1u << (k * kGuideRec + r);
              }
            }
            if (rs) G.rstate[(size_t)i * kGuideRec + r] = 0;
          }
        }
        atomicAdd(&s_done[jb], 1u);
        P.state[i] = ST_FREE;
        st = ST_FREE;
      }
      if (st == ST_FREE) need |= 1u << k;
      if (st == ST_ACTIVE) act |= 1u << k;
    }
    if (GUIDED) {
      uint32_t tot;
      uint32_t toff = block_excl_scan<4, uint32_t>((uint32_t)__popc(trn), s_w, tot);
      if (threadIdx.x == 0) s_base64 = tot ? atomicAdd(G.ring_count, (unsigned long long)tot) : 0ull;
      __syncthreads();
      unsigned long long rb = s_base64 + toff;
      __syncthreads();
#pragma unroll 1
      for (int b = 0; b < kRegenK * kGuideRec; b++) {
        if (!(trn & (1u << b))) continue;
        const int k = b / kGuideRec, r = b - k * kGuideRec;
        const float4* q = G.rec + ((size_t)(base + k * 256) * kGuideRec + r) * 4;
        const float4 r0 = q[0], r1 = q[1], r2 = q[2], r3 = q[3];
        float* d = G.ring + 10 * (size_t)(rb & G.ring_mask);
        d[0] = r0.x; d[1] = r0.y; d[2] = r0.z; d[3] = r0.w; d[4] = r1.x;
        d[5] = r1.y; d[6] = r1.z; d[7] = r1.w; d[8] = r2.x; d[9] = r3.w;
        rb++;
      }
    }
    // claims, jobs in the host's claim order
    const int norder = order[0];
    for (int jj = 0; jj < norder; jj++) {
      const int j = order[1 + jj];
      uint32_t total;
      uint32_t my = (uint32_t)__popc(need);
      uint32_t off = block_excl_scan<4, uint32_t>(my, s_w, total);
      if (total == 0) break;
      DevJob& J = jobs[j];
      if (!J.active) continue;
      if (threadIdx.x == 0) {
        unsigned long long nx = *(volatile unsigned long long*)&J.next;
        s_base64 = (nx >= J.total) ? ~0ull : atomicAdd((unsigned long long*)&J.next, (unsigned long long)total);
      }
      __syncthreads();
      unsigned long long b = s_base64;
      __syncthreads();
      if (b == ~0ull || b >= J.total) continue;
      // gather renders claim every sample in their first regeneration, each block kClaimChunk of
      // them into its (fully free) chunk of slots: remember which slot chunk holds which samples
      if (J.gather && threadIdx.x == 0) P.chunk_of[b / kClaimChunk] = (uint32_t)chunk + 1u;  // 0: unclaimed
      uint64_t got = J.total - b;
      if (got > total) got = total;
      uint32_t idx = off;
#pragma unroll 1
      for (int k = 0; k < kRegenK; k++) {
        if (!(need & (1u << k))) continue;
        if ((uint64_t)idx < got) {
          int i = base + k * 256;
          uint64_t s = b + idx;
          uint32_t pix, si;
          int x, y;
          decode_sample(J, s, pix, si, x, y);
          PathRng rng = make_rng(J.seed, pix, si, J);
          float sx, sy;
          sample2d(rng, 0, sx, sy);  // SampleJob.cpp:49-53 pixel jitter
          f3 o, d;
          float tmin, tmax;
          camera_ray(J, (float)x + sx, (float)y + sy, o, d, tmin, tmax);
          P.ray_o[i] = make_float4(o.x, o.y, o.z, tmin);
          P.ray_d[i] = make_float4(d.x, d.y, d.z, tmax);
          P.T[i] = make_float4(1.f, 1.f, 1.f, 1.f);
          P.F[i] = make_float4(1.f, 1.f, 1.f, 0.f);
          P.L[i] = make_float4(0.f, 0.f, 0.f, 1.f);
          P.pixel[i] = pix;
          P.sample[i] = si;
          P.flags[i] = ((uint32_t)j << JOB_SHIFT) | FL_EMIT | (J.integrator == INTEGRATOR_VOLPATH ? FL_VOL : 0u) |
                       (sx == 0.f ? FL_XEDGE : 0u) | (sy == 0.f ? FL_YEDGE : 0u);
          P.state[i] = ST_ACTIVE;
          need &= ~(1u << k);
          act |= 1u << k;
        }
        idx++;
      }
    }
    // extension queue, in slot order: entry (k, thread) for slot base + k * 256, so that the
    // stages reading the queue (kRegenK consecutive entries per thread) access the pool coalesced
    {
      const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
      const unsigned long long below = (1ull << lane) - 1ull;
      uint32_t rank[kRegenK];
#pragma unroll
      for (int k = 0; k < kRegenK; k++) {
        const unsigned long long m = __ballot((act >> k) & 1u);
        rank[k] = (uint32_t)__popcll(m & below);
        if (lane == 0) s_kc[k * 4 + wave] = (uint32_t)__popcll(m);
      }
      __syncthreads();
      uint32_t pre[kRegenK], run = 0;
#pragma unroll
      for (int k = 0; k < kRegenK; k++)
#pragma unroll
        for (uint32_t w = 0; w < 4; w++) {
          if (w == wave) pre[k] = run;
          run += s_kc[k * 4 + w];
        }
      const uint32_t qb = block_reserve(&C[C_EXT], run, &s_base);
#pragma unroll
      for (int k = 0; k < kRegenK; k++)
        if (act & (1u << k)) P.q_ext[qb + pre[k] + rank[k]] = (uint32_t)(base + k * 256);
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kMaxJobsDev; j += blockDim.x)
    if (s_done[j]) atomicAdd((unsigned long long*)&jobs[j].done, (unsigned long long)s_done[j]);
  if (threadIdx.x == 0 && s_nan) atomicAdd(&C[C_NAN], s_nan);
}

// ---------------------------------------------------------------- extend: closest hit
__device__ __forceinline__ TravStack make_stack(uint2* smem, uint2* gstack, int cap, uint32_t* spill,
                                                uint32_t* ovf) {
  TravStack s;
  s.spill = spill;
  s.ovf = ovf;
  s.l = smem + threadIdx.x;
  s.lstride = blockDim.x;
  s.gstride = gridDim.x * blockDim.x;
  s.g = gstack + blockIdx.x * blockDim.x + threadIdx.x;
  s.cap = cap;
  return s;
}

// Work counters (counting mode): block sums, one atomic per block.
__device__ __forceinline__ void add_work(uint32_t* C, int idx, uint32_t nv, uint32_t np, uint32_t* s_w) {
  uint32_t a = wave_sum(nv), b = wave_sum(np);
  uint32_t wave = threadIdx.x >> 6;
  if (__lane_id() == 0) { s_w[2 * wave] = a; s_w[2 * wave + 1] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long ta = 0, tb = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; w++) { ta += s_w[2 * w]; tb += s_w[2 * w + 1]; }
    atomicAdd((unsigned long long*)&C[idx], ta);
    atomicAdd((unsigned long long*)&C[idx + 2], tb);
  }
  __syncthreads();
}

// Persistent traversal over a ray queue. A lane that finishes its ray takes the next entry of its
// wave's current batch of 64 queue entries, and a wave takes a new batch (Src::batch) when it runs
// out, so a wave keeps all its lanes traversing while rays are left instead of waiting for its
// slowest lane before the next 64. Idle lanes are refilled once at least `refill` of them are idle
// (or all are). Src: batch() -> first entry of a new batch or kNoBatch (wave-uniform);
// load(k, ctx, state, geo) -> loads entry k and starts its traversal, false when it finished at
// once (its result already stored); finish(ctx, state) stores a result.
constexpr uint32_t kNoBatch = 0xffffffffu;
constexpr uint32_t kTraceBatch = 64;

template <bool ANY, class Src>
__device__ __forceinline__ void trace_persistent(Src& src, uint32_t n, const TravStack& stk, int refill,
                                                 uint32_t& nv, uint32_t& np) {
  const uint32_t lane = __lane_id();
  const uint64_t below = (1ull << lane) - 1ull;
  const uint64_t wave_mask = __ballot(true);
  bool busy = false;
  TravState s;
  SceneGeo geo;
  typename Src::Ctx c;
  uint32_t wnext = 0, wend = 0;  // the wave's current batch [wnext, wend)
  bool more = true;              // batches may remain
  while (true) {
    const uint64_t idle = __ballot(!busy);
    if (more && (idle == wave_mask || __popcll(idle) >= refill)) {
      uint64_t pending = idle;
      while (pending) {
        if (wnext >= wend) {
          const uint32_t b = src.batch();
          if (b == kNoBatch || b >= n) {
            more = false;
            break;
          }
          wnext = b;
          wend = min(b + kTraceBatch, n);
        }
        const uint32_t rank = (uint32_t)__popcll(pending & below);
        const bool mine = ((pending >> lane) & 1ull) && rank < wend - wnext;
        const uint64_t took = __ballot(mine);
        if (mine) busy = src.load(wnext + rank, c, s, geo);
        wnext += (uint32_t)__popcll(took);
        pending &= ~took;
      }
    }
    if (!__any(busy)) {
      if (!more) break;
      continue;
    }
    if (busy && trav_step<ANY>(geo, s, stk, nv, np)) {
      src.finish(c, s);
      busy = false;
    }
  }
}

// One wave-uniform atomic per batch on a queue's fetch counter.
__device__ __forceinline__ uint32_t fetch_batch(uint32_t* ctr) {
  uint32_t b = 0;
  if (__lane_id() == (uint32_t)__builtin_amdgcn_readfirstlane(__lane_id())) b = atomicAdd(ctr, kTraceBatch);
  return (uint32_t)__builtin_amdgcn_readfirstlane(b);
}

// Extension rays: closest hit, hit record and material class per queue entry.
struct ExtSrc {
  DevPool P;
  const DevScene* scenes;
  const DevJob* jobs;
  struct Ctx {
    uint32_t k, slot;
    const DevScene* S;
  };
  __device__ __forceinline__ uint32_t batch() { return fetch_batch(&P.counters[C_XFETCH]); }
  __device__ __forceinline__ bool load(uint32_t k, Ctx& c, TravState& s, SceneGeo& geo) {
    c.k = k;
    c.slot = P.q_ext[k];
    const uint32_t f = P.flags[c.slot];
    c.S = &scenes[jobs[fl_job(f)].scene];
    const float4 o = P.ray_o[c.slot], d = P.ray_d[c.slot];
    s.r.o = mk(o.x, o.y, o.z);
    s.r.d = mk(d.x, d.y, d.z);
    s.r.inv = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
    s.r.tmin = o.w;
    s.r.tmax = d.w;
    if (!trav_begin(*c.S, s)) {
      finish(c, s);
      return false;
    }
    geo = scene_geo(*c.S);
    return true;
  }
  __device__ __forceinline__ void finish(const Ctx& c, const TravState& s) {
    P.hit[c.slot] = make_float4(s.h.t, s.h.b1, s.h.b2, __int_as_float(s.h.prim));
    // material class for the shade queues: 0..2 BSDF type, 3 miss or no BSDF
    uint8_t cls = 3;
    if (s.h.prim >= 0) {
      const uint32_t m = c.S->prim_info[s.h.prim].y;
      if (m != 0xffffffffu) cls = (uint8_t)c.S->mats[m].cls;
    }
    P.cls_q[c.k] = cls;
  }
};

// Occupancy of the persistent traversal kernels: waves per SIMD the register allocation must allow
// (the LDS stack, kLdsStack x 8 bytes per lane, bounds it as well).
#ifndef RTG_TRAV_WAVES_PER_EU
#define RTG_TRAV_WAVES_PER_EU 6
#endif
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(RTG_TRAV_WAVES_PER_EU)))
k_extend(DevPool P, const DevScene* scenes, const DevJob* jobs, uint2* gstack, int stack_depth, int refill) {
  extern __shared__ uint2 smem[];
  __shared__ uint32_t s_w[4];
  uint32_t* C = P.counters;
  const TravStack stk = make_stack(smem, gstack, stack_depth, P.count_work ? &C[C_SPILL] : nullptr, &C[C_OVF]);
  const uint32_t n = C[C_EXT];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    C[C_SHADOW] = 0;   // compact stage of this iteration fills it
    C[C_SFETCH] = 0;   // shadow-ray batches of this iteration
    atomicAdd((unsigned long long*)&C[C_EXTRAYS], (unsigned long long)n);
  }
  uint32_t nv = 0, np = 0;
  ExtSrc src{P, scenes, jobs};
  trace_persistent<false>(src, n, stk, refill, nv, np);
  if (P.count_work) add_work(C, C_NODES, nv, np, s_w);
}

// ---------------------------------------------------------------- classify: shade queues
// Chunks of 2048 extension-queue entries per block; the five per-class counts (at most 2048 each)
// are packed into 12-bit fields of one 64-bit word so a single block scan serves all queues.
// MIXED: entries of VolPathTracing paths (flag FL_VOL) go to the volumetric queue (class 4), hit
// or miss, for k_vol; the others keep their material class.
// GUIDED: also queues the guided vertices of the guided job (guided_vertex) for the guide
// network's inference.
constexpr int kClassBits = 12;
constexpr unsigned long long kClassMask = (1ull << kClassBits) - 1ull;
static_assert(kRegenSlots < (1 << kClassBits), "class counts of a chunk must fit their field");

template <bool GUIDED, bool MIXED>
__global__ void __launch_bounds__(256) k_classify(DevPool P, DevGuide G) {
  __shared__ unsigned long long s_w[4];
  __shared__ uint32_t s_base[5];
  __shared__ uint32_t s_w32[4];
  __shared__ uint32_t s_gbase;
  uint32_t* C = P.counters;
  const uint32_t n = C[C_EXT];
  const uint32_t nchunks = (n + kRegenSlots - 1) / kRegenSlots;
  for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint32_t base = chunk * kRegenSlots + threadIdx.x * kRegenK;  // kRegenK consecutive entries
    uint32_t cls[kRegenK];
    unsigned long long cnt = 0;
#pragma unroll
    for (int k = 0; k < kRegenK; k++) {
      uint32_t e = base + k;
      cls[k] = e < n ? (uint32_t)P.cls_q[e] : 5u;
      if (MIXED && e < n && (P.flags[P.q_ext[e]] & FL_VOL)) {
        cls[k] = 4u;
        P.cls_q[e] = 4;  // the shadow compaction takes the entry's shadow flag
      }
      if (cls[k] < 5) cnt += 1ull << (kClassBits * cls[k]);
    }
    unsigned long long total;
    unsigned long long off = block_excl_scan<4, unsigned long long>(cnt, s_w, total);
    if (threadIdx.x < 5) {
      uint32_t t = (uint32_t)((total >> (kClassBits * threadIdx.x)) & kClassMask);
      s_base[threadIdx.x] = t ? atomicAdd(&C[threadIdx.x < 4 ? C_SHADE + threadIdx.x : C_VOLQ], t) : 0u;
    }
    __syncthreads();
    uint32_t pos[5];
#pragma unroll
    for (int q = 0; q < 5; q++) pos[q] = s_base[q] + (uint32_t)((off >> (kClassBits * q)) & kClassMask);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRegenK; k++) {
      uint32_t c = cls[k];
      if (c < 5) {
        uint32_t e = base + k;
        P.q_shade[(size_t)c * P.capacity + pos[c]] = P.q_ext[e];
        pos[c]++;
      }
    }
    if (GUIDED) {
      uint32_t gmask = 0;
#pragma unroll
      for (int k = 0; k < kRegenK; k++) {
        if (G.sample && cls[k] < 2) {
          const uint32_t f = P.flags[P.q_ext[base + k]];
          if (guided_vertex((int)cls[k], fl_depth(f)) && (int)fl_job(f) == G.job) gmask |= 1u << k;
        }
      }
      uint32_t gtot;
      uint32_t goff = block_excl_scan<4, uint32_t>((uint32_t)__popc(gmask), s_w32, gtot);
      uint32_t gb = block_reserve(G.gq_count, gtot, &s_gbase);
      uint32_t gp = gb + goff;
#pragma unroll
      for (int k = 0; k < kRegenK; k++)
        if (gmask & (1u << k)) G.gq[gp++] = P.q_ext[base + k];
    }
  }
}

// ---------------------------------------------------------------- shade helpers
__device__ __forceinline__ float power_heuristic(float a, float b) {  // Integrator.h:78-83
  float l = a * a, r = b * b;
  return l / (l + r);
}

// Emission picked up by the BSDF-sampled ray of the previous vertex (PathTracingMIS.cpp:191-220)
// and the throughput update that follows it (:222-232).
__device__ __forceinline__ void prev_segment(const DevScene& S, f3 Le, float pdfLight, bool caught, float4& T,
                                             const float4& F, float4& L) {
  if (caught) {
    // a delta bounce (F.w < 0) cannot be reached by light sampling: light pdf 0, weight 1
    // (PathTracingMIS.cpp:205)
    float w = F.w < 0.f ? power_heuristic(-F.w, 0.f) : power_heuristic(F.w, pdfLight);
    f3 c = cmul(cmul(mk(T.x, T.y, T.z), Le), mk(F.x, F.y, F.z)) * w;
    L.x += c.x; L.y += c.y; L.z += c.z;
  }
  f3 t = cmul(mk(T.x, T.y, T.z), mk(F.x, F.y, F.z));
  t = t / L.w;  // Russian-roulette divisor (1 when no roulette was played)
  T.x = t.x; T.y = t.y; T.z = t.z;
}

// Radiance snapshot for guide training, taken at the first event after a guided depth-0 vertex
// (before that event adds anything): what the path gathers afterwards is the radiance arriving
// along the guided direction, weighted by the vertex's sampling weight.
// (depth: the event's depth; the vertex recorded at depth - 1 is the one awaiting it.)
__device__ __forceinline__ void guide_snapshot(const DevGuide& G, uint32_t slot, uint32_t depth, const float4& L) {
  if (!G.record || depth < 1 || depth > (uint32_t)kGuideRec) return;
  const size_t r = (size_t)slot * kGuideRec + (depth - 1);
  if (G.rstate[r] == 1) {
    float4* q = G.rec + r * 4;
    float4 q2 = q[2];
    q2.y = L.x; q2.z = L.y; q2.w = L.z;
    q[2] = q2;
    G.rstate[r] = 2;
  }
}

template <int MT, bool GUIDED>
__global__ void __launch_bounds__(256) k_shade(DevPool P, const DevScene* scenes, const DevJob* jobs, DevGuide G) {
  uint32_t* C = P.counters;
  const uint32_t n = C[C_SHADE + MT];
  const uint32_t* q = P.q_shade + (size_t)MT * P.capacity;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    uint32_t nshadow = 0;
    uint32_t slot = q[k];
    {
      uint32_t f = P.flags[slot];
      uint32_t depth = fl_depth(f);
      const DevJob& J = jobs[fl_job(f)];
      const DevScene& S = scenes[J.scene];
      float4 ro = P.ray_o[slot], rd = P.ray_d[slot], hv = P.hit[slot];
      float4 T = P.T[slot], F = P.F[slot], L = P.L[slot];
      f3 o = mk(ro.x, ro.y, ro.z), d = mk(rd.x, rd.y, rd.z);
      HitRec h;
      h.t = hv.x; h.b1 = hv.y; h.b2 = hv.z; h.prim = __float_as_int(hv.w);
      Surface sf = surface_at(S, h, o, d);
      if (GUIDED) guide_snapshot(G, slot, depth, L);
      if (depth > 0) {
        bool caught = sf.light != 0xffffffffu;
        f3 Le = mk(0.f, 0.f, 0.f);
        float pl = 0.f;
        if (caught) {
          const DevLight& LL = S.lights[sf.light];
          if (dot(-d, sf.sh.n) > 0.f) Le = mat_col(LL.radiance);  // MeshLight.h:21-26
          pl = mesh_light_pdf(LL, o, sf.p, sf.sh.n);
        }
        prev_segment(S, Le, pl, caught, T, F, L);
      }
      bool done = (f & FL_LAST) || (int)depth >= J.max_depth;
      uint32_t nf = f;
      if (!done) {
        // emission seen directly from the camera (PathTracingMIS.cpp:73-74)
        if ((f & FL_EMIT) && sf.light != 0xffffffffu) {
          const DevLight& LL = S.lights[sf.light];
          if (dot(-d, sf.sh.n) > 0.f) {
            f3 c = cmul(mk(T.x, T.y, T.z), mat_col(LL.radiance));
            L.x += c.x; L.y += c.y; L.z += c.z;
          }
        }
        DevMaterial M = S.mats[sf.material];
        resolve_material(S, M, h, o, d);
        PathRng rng = make_rng(J.seed, P.pixel[slot], P.sample[slot], J);
        f3 wo_l = to_local(sf.sh, -d);
        // one-light next-event estimation (Integrator.cpp:176-238, PathTracingMIS.cpp:119-153);
        // from a smooth conductor or dielectric it contributes 0 (eval is 0), so it is skipped
        if (P.all_lights && S.nlights > 0 && !is_delta(M)) {
          // all_lights (PathTracingMIS.cpp:119-137): every light sampled once, without the
          // selection probability; one shadow ray per contributing light
          uint32_t cnt = 0;
          for (int li = 0; li < S.nlights; li++) {
            float ux, uy;
            sample2d(rng, li == 0 ? 1u + 2u * depth : kAllLightDim + depth * (uint32_t)kAllLights + (uint32_t)li, ux, uy);
            const DevLight& LL = S.lights[li];
            const LightSample ls = (LL.type == LIGHT_ENV) ? env_sample(S, sf.p, ux, uy)
                                                          : mesh_light_sample(S, LL, sf.p, ux, uy);
            if (ls.pdf > 0.f && luminance(ls.intensity) > 0.f) {
              const f3 value = ls.intensity / ls.pdf;
              if (!is_zero(value)) {
                const f3 wi_l = to_local(sf.sh, ls.dir);
                const float cosT = fabsf(wi_l.z);
                const float w = power_heuristic(ls.pdf, bsdf_pdf<MT>(M, wi_l, wo_l));
                const f3 c = cmul(cmul(mk(T.x, T.y, T.z), value), bsdf_eval<MT>(M, wi_l, wo_l)) * cosT * w;
                if (!exactly_zero(c)) {  // a zero contribution adds nothing, visible or not: no shadow ray
                  const size_t e = (size_t)slot * kAllLights + cnt;
                  P.sh_d2[e] = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, ls.dist - EPS);
                  P.sh_c2[e] = make_float4(c.x, c.y, c.z, 0.f);
                  cnt++;
                }
              }
            }
          }
          nshadow = cnt;
        } else if (S.nlights > 0 && !is_delta(M)) {
          float ux, uy;
          sample2d(rng, 1u + 2u * depth, ux, uy);
          float lw;
          int li = cdf_sample_reuse(S.light_cdf, S.nlights + 1, ux, lw);
          const DevLight& LL = S.lights[li];
          LightSample ls = (LL.type == LIGHT_ENV) ? env_sample(S, sf.p, ux, uy)
                                                  : mesh_light_sample(S, LL, sf.p, ux, uy);
          if (ls.pdf > 0.f && luminance(ls.intensity) > 0.f) {
            f3 value = ls.intensity / ls.pdf;
            float lpdf = ls.pdf * lw;
            value = value / lw;
            if (!is_zero(value)) {
              f3 wi_l = to_local(sf.sh, ls.dir);
              float cosT = fabsf(wi_l.z);
              float bp = bsdf_pdf<MT>(M, wi_l, wo_l);
              float w = power_heuristic(lpdf, bp);
              f3 fe = bsdf_eval<MT>(M, wi_l, wo_l);
              f3 c = cmul(cmul(mk(T.x, T.y, T.z), value), fe) * cosT * w;
              // a zero contribution (the light below the surface for this BSDF, or a zero MIS
              // weight) adds nothing whether the light is visible or not: no shadow ray
              if (!exactly_zero(c)) {
                P.sh_d[slot] = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, ls.dist - EPS);
                P.sh_c[slot] = make_float4(c.x, c.y, c.z, 0.f);
                nshadow = 1;
              }
            }
          }
        }
        // BSDF sampling (PathTracingMIS.cpp:156-164) and roulette (:225-253)
        float bx, by;
        sample2d(rng, 2u + 2u * depth, bx, by);
        BsdfSample bs;
        bool ok;
        if (MT == 2 && J.exit_mis && M.type == MAT_ROUGH_DIELECTRIC) {
          ok = dielectric_exit_mis(M, wo_l, sample1d(rng, kGuideSelDim + depth), bx, by, bs);
        } else if (GUIDED && guided_vertex(MT, depth) && (int)fl_job(f) == G.job && !is_delta(M)) {
          // one-sample MIS of BSDF sampling and the guide, each chosen with probability 1/2:
          // weight = (sampler weight * sampler density) / mixture density, so the estimator keeps
          // the plain tracer's expectation; the MIS weights of the next hit use the BSDF's pdf.
          const float* mx = G.mix + (size_t)slot * kGuideMixR;
          const float us = G.sample ? sample1d(rng, kGuideSelDim + depth) : 0.f;
          f3 wl;
          if (us < 0.5f) {
            ok = bsdf_sample<MT>(M, wo_l, bx, by, bs);
            wl = bs.wo;
          } else {
            wl = to_local(sf.sh, mixr_sample(mx, (us - 0.5f) * 2.f, bx, by));
            ok = true;
          }
          f3 wref;
          float ps = 0.f, pref = 0.f, rel = 1.f;
          if (ok) ok = bsdf_density<MT>(M, wo_l, wl, wref, ps, pref, rel);
          const f3 wd = to_world(sf.sh, wl);
          if (ok) {
            const float pm = G.sample ? 0.5f * ps + 0.5f * mixr_pdf(mx, wd) : ps;
            if (us < 0.5f) {
              wref = bs.f;
              pref = bs.pdf;
              rel = bs.rel_eta;
            }
            bs.wo = wl;
            bs.f = wref * (ps / pm);
            bs.pdf = pref;
            bs.rel_eta = rel;
            ok = pm > 0.f;
          }
          if (G.record) {
            const f3 np = mk((sf.p.x - G.bmin[0]) * G.inv_ext[0], (sf.p.y - G.bmin[1]) * G.inv_ext[1],
                             (sf.p.z - G.bmin[2]) * G.inv_ext[2]);
            float4* q = G.rec + ((size_t)slot * kGuideRec + depth) * 4;
            q[0] = make_float4(np.x, np.y, np.z, sf.sh.n.x);
            q[1] = make_float4(sf.sh.n.y, sf.sh.n.z, wd.x, wd.y);
            q[2] = make_float4(wd.z, 0.f, 0.f, 0.f);
            q[3] = make_float4(T.x, T.y, T.z, 0.f);
            G.rstate[(size_t)slot * kGuideRec + depth] = ok ? 1 : 0;
          }
        } else {
          ok = bsdf_sample<MT>(M, wo_l, bx, by, bs);
        }
        if (!ok || is_zero(bs.f)) {
          done = true;
        } else {
          f3 nd = to_world(sf.sh, bs.wo);
          P.ray_d[slot] = make_float4(nd.x, nd.y, nd.z, MAX_REAL);
          F = make_float4(bs.f.x, bs.f.y, bs.f.z, is_delta(M) ? -bs.pdf : bs.pdf);
          float eta = T.w * bs.rel_eta;
          float q = 1.f;
          nf = (f & ~(DEPTH_MASK | FL_EMIT)) | (depth + 1u);
          if ((int)depth >= J.min_depth) {
            f3 tn = cmul(mk(T.x, T.y, T.z), bs.f);
            float stop = smin(0.95f, luminance(tn) * eta * eta);
            float u1 = sample1d(rng, depth - (uint32_t)J.min_depth);
            if (u1 < stop) q = stop;
            else nf |= FL_LAST;
          }
          T.w = eta;
          L.w = q;
        }
      }
      P.ray_o[slot] = make_float4(sf.p.x, sf.p.y, sf.p.z, EPS);
      P.T[slot] = T;
      P.F[slot] = F;
      P.L[slot] = L;
      P.flags[slot] = nf;
      P.state[slot] = done ? ST_DONE : ST_ACTIVE;
    }
    P.shflag[slot] = (uint8_t)nshadow;
  }
}

// Misses and hits without a BSDF (PathTracingMIS.cpp:38-51,67-70,182-189,217-219)
template <bool GUIDED>
__global__ void __launch_bounds__(256) k_shade_miss(DevPool P, const DevScene* scenes, const DevJob* jobs,
                                                    DevGuide G) {
  uint32_t* C = P.counters;
  const uint32_t n = C[C_SHADE + 3];
  const uint32_t* q = P.q_shade + (size_t)3 * P.capacity;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    uint32_t slot = q[k];
    uint32_t f = P.flags[slot];
    uint32_t depth = fl_depth(f);
    const DevJob& J = jobs[fl_job(f)];
    const DevScene& S = scenes[J.scene];
    float4 ro = P.ray_o[slot], rd = P.ray_d[slot], hv = P.hit[slot];
    float4 T = P.T[slot], F = P.F[slot], L = P.L[slot];
    f3 o = mk(ro.x, ro.y, ro.z), d = mk(rd.x, rd.y, rd.z);
    int prim = __float_as_int(hv.w);
    if (GUIDED) guide_snapshot(G, slot, depth, L);
    if (prim >= 0) {
      // hit without a BSDF: the emission of the previous segment still counts, then the path ends
      if (depth > 0) {
        HitRec h;
        h.t = hv.x; h.b1 = hv.y; h.b2 = hv.z; h.prim = prim;
        Surface sf = surface_at(S, h, o, d);
        bool caught = sf.light != 0xffffffffu;
        f3 Le = mk(0.f, 0.f, 0.f);
        float pl = 0.f;
        if (caught) {
          const DevLight& LL = S.lights[sf.light];
          if (dot(-d, sf.sh.n) > 0.f) Le = mat_col(LL.radiance);
          pl = mesh_light_pdf(LL, o, sf.p, sf.sh.n);
        }
        prev_segment(S, Le, pl, caught, T, F, L);
      }
    } else if (S.env_light >= 0) {
      f3 Le = env_le(S, d);
      if (depth == 0) {
        L.x += Le.x; L.y += Le.y; L.z += Le.z;
      } else {
        float pl = env_pdf(S, d);
        prev_segment(S, Le, pl, true, T, F, L);
      }
    }
    P.L[slot] = L;
    P.state[slot] = ST_DONE;
  }
}

// ---------------------------------------------------------------- volumetric path tracing
// VolPathTracing::li (integrator/VolPathTracing.cpp:25-346) for scenes whose only medium is the
// camera medium and whose shapes all carry a BSDF (no medium transitions, no null-BSDF
// boundaries, no emissive media). One kernel per iteration after the shared closest-hit
// traversal; per extension-queue entry it
//   1. adds the emission met by the direction sampled at the previous vertex, weighted by the
//      ratio-tracking transmittance of that segment and the MIS weight (:112-157 after a medium
//      vertex, :257-325 after a surface vertex), and applies that vertex's throughput update and
//      roulette divisor;
//   2. runs delta tracking along the segment (:60): a collision is a medium vertex (albedo,
//      light sampling attenuated by ratio tracking and weighted against the phase function,
//      phase sampling), otherwise the segment ends at the surface (depth-0 emission, attenuated
//      light sampling with the BSDF MIS weight, BSDF sampling) or escapes (depth-0 environment);
//   3. plays Russian roulette on max(throughput) * eta^2 (:331-343).
// A segment that missed the scene is tracked up to the distance of the path's last hit, as the
// reference's Intersection keeps its previous t when computeIntersection fails.
static const uint32_t FL_PEND = 1u << 22;  // the sampled direction's emission is still to be added
static const uint32_t FL_PMED = 1u << 23;  // ... and that direction was sampled at a medium vertex
constexpr uint32_t kVolDim = 1u << 24;      // raw 1D draws of depth k: kVolDim + k * kVolDimDepth + j
constexpr uint32_t kVolDimDepth = 1u << 16;

struct VolDraws {
  const PathRng* r;
  uint32_t k;
  __device__ __forceinline__ float operator()() { return sample1d(*r, k++); }
};

__device__ __forceinline__ f3 bsdf_eval_any(const DevMaterial& m, f3 wi, f3 wo) {
  return m.cls == 0 ? bsdf_eval<0>(m, wi, wo) : m.cls == 1 ? bsdf_eval<1>(m, wi, wo) : bsdf_eval<2>(m, wi, wo);
}
__device__ __forceinline__ float bsdf_pdf_any(const DevMaterial& m, f3 wi, f3 wo) {
  return m.cls == 0 ? bsdf_pdf<0>(m, wi, wo) : m.cls == 1 ? bsdf_pdf<1>(m, wi, wo) : bsdf_pdf<2>(m, wi, wo);
}
__device__ __forceinline__ bool bsdf_sample_any(const DevMaterial& m, f3 wi, float ux, float uy, BsdfSample& bs) {
  return m.cls == 0   ? bsdf_sample<0>(m, wi, ux, uy, bs)
         : m.cls == 1 ? bsdf_sample<1>(m, wi, ux, uy, bs)
                      : bsdf_sample<2>(m, wi, ux, uy, bs);
}
__device__ __forceinline__ float max3(f3 v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }

// MIXED = false: every extension-queue entry is volumetric (a render of one VolPathTracing job).
// MIXED = true: the volumetric queue of a mixed iteration (classify wrote the entries' classes).
template <bool MIXED>
__global__ void __launch_bounds__(256, 3) k_vol(DevPool P, const DevScene* scenes, const DevJob* jobs) {
  uint32_t* C = P.counters;
  const uint32_t n = MIXED ? C[C_VOLQ] : C[C_EXT];
  const uint32_t* queue = MIXED ? P.q_shade + (size_t)4 * P.capacity : P.q_ext;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
    const uint32_t slot = queue[e];
    const uint32_t f = P.flags[slot];
    const uint32_t depth = fl_depth(f);
    const DevJob& J = jobs[fl_job(f)];
    const DevScene& S = scenes[J.scene];
    const float4 ro = P.ray_o[slot], rd = P.ray_d[slot], hv = P.hit[slot];
    float4 T = P.T[slot], L = P.L[slot];
    const float4 F = P.F[slot];
    float4 va = P.vol_a[slot];
    const f3 o = mk(ro.x, ro.y, ro.z), d = mk(rd.x, rd.y, rd.z);
    const bool medium = S.medium_kind != 0;
    HitRec h;
    h.t = hv.x; h.b1 = hv.y; h.b2 = hv.z; h.prim = __float_as_int(hv.w);
    const bool hit = h.prim >= 0;
    const float tEnd = hit ? h.t : (depth == 0 ? MAX_REAL : va.w);
    if (hit || depth == 0) va.w = tEnd;
    const PathRng rng = make_rng(J.seed, P.pixel[slot], P.sample[slot], J);
    VolDraws u{&rng, kVolDim + depth * kVolDimDepth};
    f3 tp = mk(T.x, T.y, T.z);
    // 1. emission along the previous vertex's sampled direction
    if (f & FL_PEND) {
      const bool after_medium = (f & FL_PMED) != 0;
      bool caught = false;
      f3 Le = mk(0.f, 0.f, 0.f);
      float pl = 0.f;
      if (hit) {
        // the surface is built again for the vertex below: not keeping it live across the
        // tracking loops saves registers
        const Surface sf = surface_at(S, h, o, d);
        if (sf.light != 0xffffffffu) {
          caught = true;
          const DevLight& LL = S.lights[sf.light];
          if (dot(-d, sf.sh.n) > 0.f) Le = mat_col(LL.radiance);
          if (F.w >= 0.f) {
            if (after_medium) {
              const float4 vb = P.vol_b[slot];
              pl = medium_hit_light_pdf(LL, sf.p, mk(va.x, va.y, va.z), mk(vb.x, vb.y, vb.z));
            } else {
              pl = mesh_light_pdf(LL, o, sf.p, sf.sh.n);
            }
          }
        }
      } else if (!after_medium && S.env_light >= 0) {
        caught = true;
        Le = env_le(S, d);
        if (F.w >= 0.f) pl = env_pdf(S, d);
      }
      if (caught) {
        // evalTransmittanceFromSampledBSDF (:372-423): ratio tracking over [minT, hit or maxT]
        const f3 tr = medium ? medium_tr(S, o, d, ro.w, hit ? h.t : rd.w, u) : mk(1.f, 1.f, 1.f);
        const float w = power_heuristic(fabsf(F.w), pl);
        const f3 c = cmul(cmul(cmul(tp, Le), mk(F.x, F.y, F.z)), tr) * w;
        L.x += c.x; L.y += c.y; L.z += c.z;
      }
      tp = cmul(tp, mk(F.x, F.y, F.z)) / L.w;
    }
    bool done = (f & FL_LAST) || (int)depth >= J.max_depth;
    uint32_t nf = f & ~(FL_PEND | FL_PMED | FL_EMIT);
    bool shadow = false;
    f3 vp = o, nd = d;
    float nmin = 0.f;
    float4 nF = F;
    bool pass = false;
    if (!done) {
      const bool emission = depth == 0;  // RadianceType::EMISSION in the first iteration only
      float ts = 0.f;
      f3 wm = mk(1.f, 1.f, 1.f);
      const bool scat = medium && medium_distance(S, o, d, ro.w, tEnd, u, ts, wm);
      tp = cmul(tp, wm);  // the medium's weight, with or without a collision (:64-66, :160-165)
      if (scat) {
        // medium vertex (:60-103)
        vp = o + ts * d;
        if (S.nlights > 0) {
          float ux, uy;
          sample2d(rng, 1u + 2u * depth, ux, uy);
          float lw;
          const int li = cdf_sample_reuse(S.light_cdf, S.nlights + 1, ux, lw);
          const DevLight& LL = S.lights[li];
          const LightSample ls =
              (LL.type == LIGHT_ENV) ? env_sample(S, vp, ux, uy) : mesh_light_sample(S, LL, vp, ux, uy);
          va.x = ls.point.x; va.y = ls.point.y; va.z = ls.point.z;
          P.vol_b[slot] = make_float4(ls.normal.x, ls.normal.y, ls.normal.z, 0.f);
          if (ls.pdf > 0.f && luminance(ls.intensity) > 0.f) {
            const f3 tr = medium_tr(S, vp, ls.dir, EPS, ls.dist - EPS, u);
            const float lpdf = ls.pdf * lw;
            const f3 value = cmul(tr, ls.intensity) / lpdf;
            if (!is_zero(value)) {
              const float pf = hg_eval(S.hg_g, ls.dir, -d);
              const float w = power_heuristic(lpdf, pf);
              const f3 c = (cmul(tp, value) * pf) * w;
              if (!exactly_zero(c)) {  // a zero contribution needs no shadow ray
                P.sh_d[slot] = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, ls.dist - EPS);
                P.sh_c[slot] = make_float4(c.x, c.y, c.z, 0.f);
                shadow = true;
              }
            }
          }
        }
        float bx, by, ppdf;
        sample2d(rng, 2u + 2u * depth, bx, by);
        nd = hg_sample(S.hg_g, -d, bx, by, ppdf);
        nmin = 0.f;
        nF = make_float4(1.f, 1.f, 1.f, ppdf);
        nf |= FL_PEND | FL_PMED;
      } else if (!hit) {
        // escaped (:170-183)
        if (emission && S.env_light >= 0) {
          const f3 Le = env_le(S, d);
          L.x += tp.x * Le.x; L.y += tp.y * Le.y; L.z += tp.z * Le.z;
        }
        done = true;
      } else if (S.prim_info[h.prim].y == 0xffffffffu) {
        // a shape without a BSDF is passed through (:187-195): same direction, same depth, no roulette
        vp = surface_at(S, h, o, d).p;
        nmin = EPS;
        nF = make_float4(1.f, 1.f, 1.f, 0.f);
        pass = true;
      } else {
        // surface vertex (:185-330)
        const Surface sf = surface_at(S, h, o, d);
        vp = sf.p;
        if (emission && sf.light != 0xffffffffu && dot(-d, sf.sh.n) > 0.f) {
          const f3 c = cmul(tp, mat_col(S.lights[sf.light].radiance));
          L.x += c.x; L.y += c.y; L.z += c.z;
        }
        DevMaterial M = S.mats[sf.material];
        resolve_material(S, M, h, o, d);
        const f3 wo_l = to_local(sf.sh, -d);
        if (S.nlights > 0 && !is_delta(M)) {
          float ux, uy;
          sample2d(rng, 1u + 2u * depth, ux, uy);
          float lw;
          const int li = cdf_sample_reuse(S.light_cdf, S.nlights + 1, ux, lw);
          const DevLight& LL = S.lights[li];
          const LightSample ls =
              (LL.type == LIGHT_ENV) ? env_sample(S, sf.p, ux, uy) : mesh_light_sample(S, LL, sf.p, ux, uy);
          if (ls.pdf > 0.f && luminance(ls.intensity) > 0.f) {
            const f3 tr = medium ? medium_tr(S, sf.p, ls.dir, EPS, ls.dist - EPS, u) : mk(1.f, 1.f, 1.f);
            const float lpdf = ls.pdf * lw;
            const f3 value = cmul(tr, ls.intensity) / lpdf;
            if (!is_zero(value)) {
              const f3 wi_l = to_local(sf.sh, ls.dir);
              const float cosT = fabsf(wi_l.z);
              const float w = power_heuristic(lpdf, bsdf_pdf_any(M, wi_l, wo_l));
              const f3 c = (cmul(cmul(tp, value), bsdf_eval_any(M, wi_l, wo_l)) * cosT) * w;
              if (!exactly_zero(c)) {  // a zero contribution needs no shadow ray
                P.sh_d[slot] = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, ls.dist - EPS);
                P.sh_c[slot] = make_float4(c.x, c.y, c.z, 0.f);
                shadow = true;
              }
            }
          }
        }
        float bx, by;
        sample2d(rng, 2u + 2u * depth, bx, by);
        BsdfSample bs;
        const bool ok = bsdf_sample_any(M, wo_l, bx, by, bs);
        if (!ok || is_zero(bs.f)) {
          done = true;
        } else {
          nd = to_world(sf.sh, bs.wo);
          nmin = EPS;
          nF = make_float4(bs.f.x, bs.f.y, bs.f.z, is_delta(M) ? -bs.pdf : bs.pdf);
          T.w *= bs.rel_eta;
          nf |= FL_PEND;
        }
      }
      if (!done && pass) {
        L.w = 1.f;
        P.ray_d[slot] = make_float4(nd.x, nd.y, nd.z, MAX_REAL);
      } else if (!done) {
        float q = 1.f;
        if ((int)depth >= J.min_depth) {
          const float stop = smin(0.95f, max3(cmul(tp, mk(nF.x, nF.y, nF.z))) * T.w * T.w);
          if (u() < stop) q = stop;
          else nf |= FL_LAST;
        }
        L.w = q;
        nf = (nf & ~DEPTH_MASK) | (depth + 1u);
        P.ray_d[slot] = make_float4(nd.x, nd.y, nd.z, MAX_REAL);
      }
      P.ray_o[slot] = make_float4(vp.x, vp.y, vp.z, nmin);
    }
    P.T[slot] = make_float4(tp.x, tp.y, tp.z, T.w);
    P.F[slot] = nF;
    P.L[slot] = L;
    P.vol_a[slot] = va;
    P.flags[slot] = nf;
    P.state[slot] = done ? ST_DONE : ST_ACTIVE;
    P.shflag[slot] = shadow ? 1 : 0;
    if (!MIXED) P.cls_q[e] = 0;  // the shadow compaction reads every entry's flag
  }
}

// ---------------------------------------------------------------- compact: shadow queue
__global__ void __launch_bounds__(256) k_compact_shadow(DevPool P) {
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_base;
  uint32_t* C = P.counters;
  const uint32_t n = C[C_EXT];
  const uint32_t nchunks = (n + kRegenSlots - 1) / kRegenSlots;
  for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint32_t base = chunk * kRegenSlots + threadIdx.x * kRegenK;  // kRegenK consecutive entries
    uint32_t cnt[kRegenK];
    uint32_t slots[kRegenK];
    uint32_t mine = 0;
#pragma unroll
    for (int k = 0; k < kRegenK; k++) {
      uint32_t e = base + k;
      slots[k] = 0;
      cnt[k] = 0;
      if (e < n && P.cls_q[e] != 3) {  // misses of plain paths produce no shadow ray
        slots[k] = P.q_ext[e];
        cnt[k] = P.shflag[slots[k]];
      }
      mine += cnt[k];
    }
    uint32_t total;
    uint32_t off = block_excl_scan<4, uint32_t>(mine, s_w, total);
    uint32_t pos = block_reserve(&C[C_SHADOW], total, &s_base) + off;
    if (!P.all_lights) {
#pragma unroll
      for (int k = 0; k < kRegenK; k++)
        if (cnt[k]) P.q_shadow[pos++] = slots[k];
    } else {
      // entry slot * kAllLights + j for each of the slot's shadow rays
#pragma unroll
      for (int k = 0; k < kRegenK; k++)
        for (uint32_t j = 0; j < cnt[k]; j++) P.q_shadow2[pos++] = slots[k] * (uint32_t)kAllLights + j;
    }
  }
}

// ---------------------------------------------------------------- shadow rays (Integrator.cpp:228-235)
// Shadow rays: any hit; an unoccluded ray adds its stored NEE contribution to the path.
struct ShadowSrc {
  DevPool P;
  const DevScene* scenes;
  const DevJob* jobs;
  struct Ctx {
    uint32_t slot, entry;
  };
  __device__ __forceinline__ uint32_t batch() { return fetch_batch(&P.counters[C_SFETCH]); }
  __device__ __forceinline__ bool load(uint32_t k, Ctx& c, TravState& s, SceneGeo& geo) {
    if (P.all_lights) {
      c.entry = P.q_shadow2[k];
      c.slot = c.entry / (uint32_t)kAllLights;
    } else {
      c.slot = P.q_shadow[k];
      c.entry = c.slot;
    }
    const uint32_t f = P.flags[c.slot];
    const DevScene& S = scenes[jobs[fl_job(f)].scene];
    const float4 o = P.ray_o[c.slot], d = P.all_lights ? P.sh_d2[c.entry] : P.sh_d[c.slot];
    s.r.o = mk(o.x, o.y, o.z);
    s.r.d = mk(d.x, d.y, d.z);
    s.r.inv = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
    s.r.tmin = EPS;
    s.r.tmax = d.w;
    if (!trav_begin(S, s)) {
      finish(c, s);
      return false;
    }
    geo = scene_geo(S);
    return true;
  }
  __device__ __forceinline__ void finish(const Ctx& c, const TravState& s) {
    if (!s.hit) {
      if (P.all_lights) {
        // a slot's shadow rays may finish together: add atomically
        const float4 cc = P.sh_c2[c.entry];
        unsafeAtomicAdd(&P.L[c.slot].x, cc.x);
        unsafeAtomicAdd(&P.L[c.slot].y, cc.y);
        unsafeAtomicAdd(&P.L[c.slot].z, cc.z);
      } else {
        const float4 cc = P.sh_c[c.slot];
        float4 L = P.L[c.slot];
        L.x += cc.x; L.y += cc.y; L.z += cc.z;
        P.L[c.slot] = L;
      }
    }
  }
};

__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(RTG_TRAV_WAVES_PER_EU)))
k_shadow(DevPool P, const DevScene* scenes, const DevJob* jobs, uint2* gstack, int stack_depth, int refill) {
  extern __shared__ uint2 smem[];
  __shared__ uint32_t s_w[4];
  uint32_t* C = P.counters;
  const TravStack stk = make_stack(smem, gstack, stack_depth, P.count_work ? &C[C_SPILL] : nullptr, &C[C_OVF]);
  const uint32_t n = C[C_SHADOW];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    C[C_EXT] = 0;  // regen of the next iteration fills it
    atomicAdd((unsigned long long*)&C[C_SHRAYS], (unsigned long long)n);
  }
  uint32_t nv = 0, np = 0;
  ShadowSrc src{P, scenes, jobs};
  trace_persistent<true>(src, n, stk, refill, nv, np);
  if (P.count_work) add_work(C, C_SHNODES, nv, np, s_w);
}

__global__ void k_reset_pool(DevPool P) {
  int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.capacity; i += stride) P.state[i] = ST_FREE;
  if (blockIdx.x == 0 && threadIdx.x < C_COUNT) P.counters[threadIdx.x] = 0;
}

__global__ void k_fill_zero(uint32_t* p, size_t n) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// ---------------------------------------------------------------- parity taps
// Rays from a host array (8 floats each), one scene; batches are taken in a fixed grid-stride order
// (no counter), so the tap runs the same persistent traversal as the render kernels.
struct TapSrc {
  const DevScene* scene;
  const float* rays;
  int32_t *prim;
  float *t, *uv, *p, *nrm;
  bool any;
  uint32_t next, stride;  // wave-uniform batch cursor
  struct Ctx {
    uint32_t i;
  };
  __device__ __forceinline__ uint32_t batch() {
    const uint32_t b = next;
    next += stride;
    return b;
  }
  __device__ __forceinline__ bool load(uint32_t i, Ctx& c, TravState& s, SceneGeo& geo) {
    c.i = i;
    const float* q = rays + 8 * (size_t)i;
    s.r.o = mk(q[0], q[1], q[2]);
    s.r.d = mk(q[3], q[4], q[5]);
    s.r.inv = mk(1.f / q[3], 1.f / q[4], 1.f / q[5]);
    s.r.tmin = q[6];
    s.r.tmax = q[7];
    if (!trav_begin(*scene, s)) {
      finish(c, s);
      return false;
    }
    geo = scene_geo(*scene);
    return true;
  }
  __device__ __forceinline__ void finish(const Ctx& c, const TravState& s) {
    const uint32_t i = c.i;
    const DevScene& S = *scene;
    const bool hit = s.hit;
    prim[i] = hit ? (int32_t)S.prim_info[s.h.prim].x : -1;
    if (t) t[i] = hit ? s.h.t : 0.f;
    if (uv) {
      uv[2 * i] = hit ? s.h.b1 : 0.f;
      uv[2 * i + 1] = hit ? s.h.b2 : 0.f;
    }
    if (p || nrm) {
      f3 pp = mk(0.f, 0.f, 0.f), nn = mk(0.f, 0.f, 0.f);
      if (hit && !any) {
        // surface_at takes the ray as given (adapt_eps only changes tmin)
        Surface sf = surface_at(S, s.h, s.r.o, s.r.d);
        pp = sf.p;
        nn = sf.sh.n;
      }
      if (p) { p[3 * i] = pp.x; p[3 * i + 1] = pp.y; p[3 * i + 2] = pp.z; }
      if (nrm) { nrm[3 * i] = nn.x; nrm[3 * i + 1] = nn.y; nrm[3 * i + 2] = nn.z; }
    }
  }
};

__global__ void __launch_bounds__(128) k_trace_tap(const DevScene* scene, const float* rays, int n, int any_hit,
                                                   int32_t* prim, float* t, float* uv, float* p, float* nrm,
                                                   uint32_t* ctr, uint2* gstack, int stack_depth, int refill) {
  extern __shared__ uint2 smem[];
  const TravStack stk = make_stack(smem, gstack, stack_depth, ctr, ctr + 1);
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64u;
  const uint32_t nwaves = (gridDim.x * blockDim.x) / 64u;
  TapSrc src{scene, rays, prim, t, uv, p, nrm, any_hit != 0, wave * kTraceBatch, nwaves * kTraceBatch};
  uint32_t nv = 0, np = 0;
  if (any_hit)
    trace_persistent<true>(src, (uint32_t)n, stk, refill, nv, np);
  else
    trace_persistent<false>(src, (uint32_t)n, stk, refill, nv, np);
}

// Sampler tap: the 2D and 1D draws the renderer's sampler gives (pixel, sample) pairs of a job.
__global__ void k_sampler_tap(DevJob J, const uint32_t* ps, int n, int n2, int n1, float* out2, float* out1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PathRng r = make_rng(J.seed, ps[2 * i], ps[2 * i + 1], J);
  for (int k = 0; k < n2; k++) {
    float x, y;
    sample2d(r, (uint32_t)k, x, y);
    out2[((size_t)i * n2 + k) * 2] = x;
    out2[((size_t)i * n2 + k) * 2 + 1] = y;
  }
  for (int k = 0; k < n1; k++) out1[(size_t)i * n1 + k] = sample1d(r, (uint32_t)k);
}

void launch_sampler_tap(const DevJob& J, const uint32_t* ps, int n, int n2, int n1, float* out2, float* out1,
                        hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sampler_tap, dim3((n + 255) / 256), dim3(256), 0, stream, J, ps, n, n2, n1, out2, out1);
}

__global__ void k_bsdf_tap(const DevMaterial* mat, int op, const float* wi, const float* wo_or_u, int n,
                           float* out_dir, float* f, float* pdf, float* rel_eta) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevMaterial& M = *mat;
  f3 a = ld3(wi + 3 * (size_t)i);
  if (op == 0) {
    f3 b = ld3(wo_or_u + 3 * (size_t)i);
    f3 e;
    float pd;
    if (M.cls == 0) { e = bsdf_eval<0>(M, a, b); pd = bsdf_pdf<0>(M, a, b); }
    else if (M.cls == 1) { e = bsdf_eval<1>(M, a, b); pd = bsdf_pdf<1>(M, a, b); }
    else { e = bsdf_eval<2>(M, a, b); pd = bsdf_pdf<2>(M, a, b); }
    f[3 * i] = e.x; f[3 * i + 1] = e.y; f[3 * i + 2] = e.z;
    pdf[i] = pd;
  } else if (op == 2) {
    // sampling density of bsdf_sample at wo, the weight it returns there and the pdf it reports
    f3 b = ld3(wo_or_u + 3 * (size_t)i);
    f3 w = mk(0.f, 0.f, 0.f);
    float ps = 0.f, pref = 0.f, rel = 1.f;
    bool ok = M.cls == 0 ? bsdf_density<0>(M, a, b, w, ps, pref, rel) : bsdf_density<1>(M, a, b, w, ps, pref, rel);
    if (!ok) { w = mk(0.f, 0.f, 0.f); ps = 0.f; pref = 0.f; }
    f[3 * i] = w.x; f[3 * i + 1] = w.y; f[3 * i + 2] = w.z;
    pdf[i] = ps;
    rel_eta[i] = pref;
  } else if (op == 3) {
    // rough-dielectric exit proposals (dielectric_exit_mis) with u = (technique, ux, uy): the
    // MIS-weighted estimate -> f, the pdf the sampler reports -> pdf, the MIS factor -> rel_eta
    const float us = wo_or_u[3 * i], ux = wo_or_u[3 * i + 1], uy = wo_or_u[3 * i + 2];
    BsdfSample bs;
    bs.wo = mk(0.f, 0.f, 0.f);
    bs.f = mk(0.f, 0.f, 0.f);
    bs.pdf = -1.f;
    bs.rel_eta = 1.f;
    float k = 0.f;
    if (!dielectric_exit_mis(M, a, us, ux, uy, bs, &k)) { bs.f = mk(0.f, 0.f, 0.f); k = 0.f; }
    out_dir[3 * i] = bs.wo.x; out_dir[3 * i + 1] = bs.wo.y; out_dir[3 * i + 2] = bs.wo.z;
    f[3 * i] = bs.f.x; f[3 * i + 1] = bs.f.y; f[3 * i + 2] = bs.f.z;
    pdf[i] = bs.pdf;
    rel_eta[i] = k;
  } else {
    float ux = wo_or_u[2 * i], uy = wo_or_u[2 * i + 1];
    BsdfSample bs;
    bs.wo = mk(0.f, 0.f, 0.f);
    bs.f = mk(0.f, 0.f, 0.f);
    bs.pdf = -1.f;
    bs.rel_eta = 1.f;
    bool ok;
    if (M.cls == 0) ok = bsdf_sample<0>(M, a, ux, uy, bs);
    else if (M.cls == 1) ok = bsdf_sample<1>(M, a, ux, uy, bs);
    else ok = bsdf_sample<2>(M, a, ux, uy, bs);
    if (!ok) bs.f = mk(0.f, 0.f, 0.f);
    out_dir[3 * i] = bs.wo.x; out_dir[3 * i + 1] = bs.wo.y; out_dir[3 * i + 2] = bs.wo.z;
    f[3 * i] = bs.f.x; f[3 * i + 1] = bs.f.y; f[3 * i + 2] = bs.f.z;
    pdf[i] = bs.pdf;
    rel_eta[i] = bs.rel_eta;
  }
}

// Camera-medium parity tap (the C ABI's rtg_medium_batch): one item per thread.
struct TapDraws {
  const float* u;
  int k, i;
  __device__ __forceinline__ float operator()() {
    const float v = i < k ? u[i] : 0.5f;
    i++;
    return v;
  }
};
__global__ void k_medium_tap(const DevScene* scene, int op, const float* in, int n, const float* u, int k, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevScene& S = *scene;
  if (op == 0) {
    out[i] = medium_sigma(S, mk(in[3 * i], in[3 * i + 1], in[3 * i + 2]));
  } else if (op == 1 || op == 2) {
    const float* r = in + 8 * i;
    const f3 o = mk(r[0], r[1], r[2]), d = mk(r[3], r[4], r[5]);
    TapDraws dr{u + (size_t)i * k, k, 0};
    if (op == 1) {
      float t = r[7];
      const bool sc = medium_sample_distance(S, o, d, r[6], r[7], dr, t);
      out[3 * i] = sc ? 1.f : 0.f;
      out[3 * i + 1] = sc ? t : fminf(t, r[7]);
      out[3 * i + 2] = (float)dr.i;
    } else {
      out[3 * i] = medium_transmittance(S, o, d, r[6], r[7], dr);
      out[3 * i + 1] = (float)dr.i;
      out[3 * i + 2] = 0.f;
    }
  } else if (op == 3) {
    const float* r = in + 6 * i;
    out[i] = hg_eval(S.hg_g, mk(r[0], r[1], r[2]), mk(r[3], r[4], r[5]));
  } else {
    float pdf;
    const f3 wo = hg_sample(S.hg_g, mk(in[3 * i], in[3 * i + 1], in[3 * i + 2]), u[2 * i], u[2 * i + 1], pdf);
    out[4 * i] = wo.x; out[4 * i + 1] = wo.y; out[4 * i + 2] = wo.z; out[4 * i + 3] = pdf;
  }
}

// ---------------------------------------------------------------- gather (box-filter renders in one pass)
// Sample s of a gather job sits in slot (chunk_of[s / 2048] - 1) * 2048 + t + 256 k with s % 2048 = 8 t + k
// (the regeneration's claim order inside a fully free chunk). One thread per pixel, pixels in the
// claim order's strip walk; a pixel's samples are summed in sample order (as the reference adds them,
// core/Screen.cpp:44-97) and added to the film, NaN samples dropped with their weight.
__global__ void __launch_bounds__(256) k_gather(const DevJob* jobs, int job, DevPool P, int npix) {
  const DevJob& J = jobs[job];
  const uint32_t W = (uint32_t)J.width, Hh = (uint32_t)J.height;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < (uint32_t)npix; r += gridDim.x * blockDim.x) {
    const uint32_t strip = r / (8u * W);
    const uint32_t hs = min(8u, Hh - strip * 8u);
    const uint32_t within = r - strip * 8u * W;
    const uint32_t x = within / hs, y = strip * 8u + within % hs;
    float sr = 0.f, sg = 0.f, sb = 0.f, w = 0.f;
    const uint32_t nchunks = (uint32_t)(P.capacity / kClaimChunk);
    for (int si = 0; si < J.spp; si++) {
      const uint64_t s = (uint64_t)si * (uint64_t)npix + r;
      const uint32_t rr = (uint32_t)(s % kClaimChunk);
      const uint32_t c = P.chunk_of[s / kClaimChunk] - 1u;  // unclaimed (0) wraps past nchunks
      if (c >= nchunks) {  // never expected: the host gathers only renders the pool holds whole
        P.counters[C_GATHER] = 1u;
        continue;
      }
      const size_t slot = (size_t)c * kClaimChunk + (rr >> 3) + (rr & 7u) * 256u;
      const float4 L = P.L[slot];
      if (!(isnan(L.x) || isnan(L.y) || isnan(L.z))) {
        sr += L.x; sg += L.y; sb += L.z; w += 1.f;
      }
    }
    float4* a = J.accum + (size_t)y * W + x;
    const float4 v = *a;
    *a = make_float4(v.x + sr, v.y + sg, v.z + sb, v.w + w);
  }
}

// ---------------------------------------------------------------- light and camera taps
// Light tap (rtg_light_batch): the functions the shading kernels call for next-event estimation
// and for the MIS weight of an emitter hit, one item per thread.
__global__ void k_light_tap(const DevScene* scene, int op, int li, const float* in, const float* u, int n,
                            float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevScene& S = *scene;
  if (op == 0) {  // Light::sample from a point
    const DevLight& L = S.lights[li];
    const f3 p = mk(in[3 * i], in[3 * i + 1], in[3 * i + 2]);
    const LightSample ls =
        L.type == LIGHT_ENV ? env_sample(S, p, u[2 * i], u[2 * i + 1]) : mesh_light_sample(S, L, p, u[2 * i], u[2 * i + 1]);
    float* o = out + 14 * (size_t)i;
    o[0] = ls.dir.x; o[1] = ls.dir.y; o[2] = ls.dir.z; o[3] = ls.dist; o[4] = ls.pdf;
    o[5] = ls.intensity.x; o[6] = ls.intensity.y; o[7] = ls.intensity.z;
    o[8] = ls.point.x; o[9] = ls.point.y; o[10] = ls.point.z;
    o[11] = ls.normal.x; o[12] = ls.normal.y; o[13] = ls.normal.z;
  } else if (op == 1) {  // Light::pdf (from point, light point, light normal, direction)
    const DevLight& L = S.lights[li];
    const float* a = in + 12 * (size_t)i;
    out[i] = L.type == LIGHT_ENV ? env_pdf(S, mk(a[9], a[10], a[11]))
                                 : mesh_light_pdf(L, mk(a[0], a[1], a[2]), mk(a[3], a[4], a[5]), mk(a[6], a[7], a[8]));
  } else if (op == 2) {  // EnvironmentLight::le
    const f3 c = env_le(S, mk(in[3 * i], in[3 * i + 1], in[3 * i + 2]));
    out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
  } else {  // one-light selection: CDF::sample with the draw reused (Integrator.cpp:176-190)
    float s = in[i], w;
    const int k = cdf_sample_reuse(S.light_cdf, S.nlights + 1, s, w);
    out[3 * i] = (float)k; out[3 * i + 1] = s; out[3 * i + 2] = w;
  }
}

// Camera tap (rtg_camera_rays): NoriPerspective::getRay through the renderer's camera_ray.
__global__ void k_camera_tap(DevJob J, const float* pxy, int n, float* rays) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f3 o, d;
  float tmin, tmax;
  camera_ray(J, pxy[2 * i], pxy[2 * i + 1], o, d, tmin, tmax);
  float* r = rays + 8 * (size_t)i;
  r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = d.x; r[4] = d.y; r[5] = d.z; r[6] = tmin; r[7] = tmax;
}

// ---------------------------------------------------------------- tone mapping (core/main.cpp:185-290)
__device__ __forceinline__ f3 resolved(const float4 a) {  // Pixel::applyColor (core/Screen.cpp:8-15)
  return a.w != 0.f ? mk(a.x / a.w, a.y / a.w, a.z / a.w) : mk(a.x, a.y, a.z);
}
__device__ __forceinline__ float to_srgb(float v) {  // Color.h:450-461
  if (v < 0.0031308f) return 12.92f * v;
  return 1.055f * powf(v, 0.41666f) - 0.055f;
}
// math::thresholding(v * 255, 0, 255) converted to sf::Uint8 (truncation; NaN gives 0 here)
__device__ __forceinline__ uint8_t to_byte(float v) {
  const float x = v * 255.f;
  const float c = x < 0.f ? 0.f : (x > 255.f ? 255.f : x);
  return (c == c) ? (uint8_t)(int)c : (uint8_t)0;
}
// maximum luminance, starting from the smallest normal float (numeric_limits<real>::min()):
// every candidate is then positive, so the unsigned bit patterns order like the values
__global__ void k_tonemap_max(const float4* film, int n, unsigned int* maxbits) {
  float m = 1.17549435e-38f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    m = fmaxf(m, luminance(resolved(film[i])));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(maxbits, __float_as_uint(m));
}
__global__ void k_tonemap(const float4* film, int n, int mode, float expo, const unsigned int* maxbits,
                          uint8_t* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const f3 c = resolved(film[i]);
  float r, g, b;
  if (mode == 0) {  // EXRToClampedRGB
    r = c.x; g = c.y; b = c.z;
  } else if (mode == 1) {  // gammaToneMapper with gamma -1 (sRGB)
    const f3 e = c * expo;
    r = to_srgb(e.x); g = to_srgb(e.y); b = to_srgb(e.z);
  } else {  // globalReinhardToneMapper with gamma -1
    const float mx = __uint_as_float(*maxbits);
    const float oldLum = luminance(c);
    const float numerator = (float)((double)oldLum * (1.0 + (double)(oldLum / (mx * mx))));
    const float lum = numerator / (1.0f + oldLum);
    const f3 f = (c * lum) / oldLum;
    r = to_srgb(f.x); g = to_srgb(f.y); b = to_srgb(f.z);
  }
  out[3 * i] = to_byte(r);
  out[3 * i + 1] = to_byte(g);
  out[3 * i + 2] = to_byte(b);
}

// ---------------------------------------------------------------- launchers
void launch_tonemap(const float4* film, int n, int mode, float expo, unsigned int* maxbits, uint8_t* out,
                    hipStream_t stream) {
  if (mode == 2) hipLaunchKernelGGL(k_tonemap_max, dim3(1024), dim3(256), 0, stream, film, n, maxbits);
  hipLaunchKernelGGL(k_tonemap, dim3((n + 255) / 256), dim3(256), 0, stream, film, n, mode, expo, maxbits, out);
}
void launch_medium_tap(const DevScene* scene_dev, int op, const float* in, int n, const float* u, int k, float* out,
                       hipStream_t stream) {
  hipLaunchKernelGGL(k_medium_tap, dim3((n + 127) / 128), dim3(128), 0, stream, scene_dev, op, in, n, u, k, out);
}
void launch_light_tap(const DevScene* scene_dev, int op, int light, const float* in, const float* u, int n,
                      float* out, hipStream_t stream) {
  hipLaunchKernelGGL(k_light_tap, dim3((n + 127) / 128), dim3(128), 0, stream, scene_dev, op, light, in, u, n, out);
}
void launch_camera_tap(const DevJob& job, const float* pxy, int n, float* rays, hipStream_t stream) {
  hipLaunchKernelGGL(k_camera_tap, dim3((n + 127) / 128), dim3(128), 0, stream, job, pxy, n, rays);
}
static size_t stack_bytes(int block) { return (size_t)block * kLdsStack * 8u; }

// A grid-stride kernel launched with more blocks than the device keeps resident runs a partial
// second round of blocks after the first (each block has the same share of the queue), so the
// stages get exactly one resident round: blocks per CU at the kernel's register and LDS use.
template <class K>
static int resident_blocks_per_cu(K kernel) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(kernel), 256, 0) !=
      hipSuccess)
    return 0;
  return nb;
}
void kernel_grids(LaunchCfg& cfg, int cu_count) {
  static std::mutex mu;
  static int per_cu[6] = {-1, -1, -1, -1, -1, -1};  // shade 0..2, miss, vol, infer
  {
    std::lock_guard<std::mutex> lk(mu);
    if (per_cu[0] < 0) {
      // the guided and plain instantiations of a stage use the same registers to within a few
      per_cu[0] = std::min(resident_blocks_per_cu(k_shade<0, true>), resident_blocks_per_cu(k_shade<0, false>));
      per_cu[1] = std::min(resident_blocks_per_cu(k_shade<1, true>), resident_blocks_per_cu(k_shade<1, false>));
      per_cu[2] = std::min(resident_blocks_per_cu(k_shade<2, true>), resident_blocks_per_cu(k_shade<2, false>));
      per_cu[3] = std::min(resident_blocks_per_cu(k_shade_miss<true>), resident_blocks_per_cu(k_shade_miss<false>));
      per_cu[4] = std::min(resident_blocks_per_cu(k_vol<true>), resident_blocks_per_cu(k_vol<false>));
      per_cu[5] = guide_infer_blocks_per_cu();
    }
  }
  auto grid = [&](int i) { return per_cu[i] > 0 ? per_cu[i] * cu_count : cfg.grid; };
  for (int m = 0; m < 3; m++) cfg.shade_grid[m] = grid(m);
  cfg.miss_grid = grid(3);
  cfg.vol_grid = grid(4);
  cfg.infer_grid = grid(5);
}

void launch_iteration(const DevPool& pool, const DevScene* scenes_dev, DevJob* jobs_dev, const int32_t* order_dev,
                      const LaunchCfg& cfg, hipStream_t stream, hipEvent_t* ev, const DevGuide* guide) {
  const DevJob* cj = jobs_dev;
  const int chunks = (pool.capacity + kRegenSlots - 1) / kRegenSlots;
  const bool gd = guide && guide->enabled;
  DevGuide G;
  if (gd) {
    G = *guide;
  } else {
    memset(&G, 0, sizeof(G));
    G.job = -1;
  }
  if (ev) (void)hipEventRecord(ev[0], stream);
  if (gd)
    hipLaunchKernelGGL(k_regen<true>, dim3(std::min(chunks, cfg.regen_blocks)), dim3(256), 0, stream, pool, jobs_dev,
                       order_dev, G);
  else
    hipLaunchKernelGGL(k_regen<false>, dim3(std::min(chunks, cfg.regen_blocks)), dim3(256), 0, stream, pool, jobs_dev,
                     order_dev, G);
  if (ev) (void)hipEventRecord(ev[1], stream);
  hipLaunchKernelGGL(k_extend, dim3(cfg.trace_blocks), dim3(128), stack_bytes(128), stream, pool, scenes_dev, cj,
                     cfg.gstack, cfg.stack, cfg.refill);
  if (gd)
    hipLaunchKernelGGL((k_classify<true, false>), dim3(std::min(chunks, 2 * cfg.regen_blocks)), dim3(256), 0, stream, pool, G);
  else
    hipLaunchKernelGGL((k_classify<false, false>), dim3(std::min(chunks, 2 * cfg.regen_blocks)), dim3(256), 0, stream, pool, G);
  if (ev) (void)hipEventRecord(ev[2], stream);
  if (gd) {
    if (G.sample) launch_guide_infer_pool(G, pool, scenes_dev, cj, cfg.infer_grid, stream);
    if (ev) (void)hipEventRecord(ev[3], stream);
    hipLaunchKernelGGL((k_shade<0, true>), dim3(cfg.shade_grid[0]), dim3(256), 0, stream, pool, scenes_dev, cj, G);
    if (ev) (void)hipEventRecord(ev[4], stream);
    hipLaunchKernelGGL((k_shade<1, true>), dim3(cfg.shade_grid[1]), dim3(256), 0, stream, pool, scenes_dev, cj, G);
    if (ev) (void)hipEventRecord(ev[5], stream);
    hipLaunchKernelGGL((k_shade<2, true>), dim3(cfg.shade_grid[2]), dim3(256), 0, stream, pool, scenes_dev, cj, G);
    if (ev) (void)hipEventRecord(ev[6], stream);
    hipLaunchKernelGGL(k_shade_miss<true>, dim3(cfg.miss_grid), dim3(256), 0, stream, pool, scenes_dev, cj, G);
  } else {
    if (ev) (void)hipEventRecord(ev[3], stream);
    hipLaunchKernelGGL((k_shade<0, false>), dim3(cfg.shade_grid[0]), dim3(256), 0, stream, pool, scenes_dev, cj, G);
    if (ev) (void)hipEventRecord(ev[4], stream);
    hipLaunchKernelGGL((k_shade<1, false>), dim3(cfg.shade_grid[1]), dim3(256), 0, stream, pool, scenes_dev, cj, G);
    if (ev) (void)hipEventRecord(ev[5], stream);
    hipLaunchKernelGGL((k_shade<2, false>), dim3(cfg.shade_grid[2]), dim3(256), 0, stream, pool, scenes_dev, cj, G);
    if (ev) (void)hipEventRecord(ev[6], stream);
    hipLaunchKernelGGL(k_shade_miss<false>, dim3(cfg.miss_grid), dim3(256), 0, stream, pool, scenes_dev, cj, G);
  }
  if (ev) (void)hipEventRecord(ev[7], stream);
  hipLaunchKernelGGL(k_compact_shadow, dim3(std::min(chunks, 2 * cfg.regen_blocks)), dim3(256), 0, stream, pool);
  hipLaunchKernelGGL(k_shadow, dim3(cfg.trace_blocks), dim3(128), stack_bytes(128), stream, pool, scenes_dev, cj,
                     cfg.gstack, cfg.stack, cfg.refill);
  if (ev) (void)hipEventRecord(ev[8], stream);
}

void launch_iteration_vol(const DevPool& pool, const DevScene* scenes_dev, DevJob* jobs_dev, const int32_t* order_dev,
                          const LaunchCfg& cfg, hipStream_t stream) {
  const int chunks = (pool.capacity + kRegenSlots - 1) / kRegenSlots;
  DevGuide G;
  memset(&G, 0, sizeof(G));
  G.job = -1;
  hipLaunchKernelGGL(k_regen<false>, dim3(std::min(chunks, cfg.regen_blocks)), dim3(256), 0, stream, pool, jobs_dev,
                     order_dev, G);
  hipLaunchKernelGGL(k_extend, dim3(cfg.trace_blocks), dim3(128), stack_bytes(128), stream, pool, scenes_dev, jobs_dev,
                     cfg.gstack, cfg.stack, cfg.refill);
  hipLaunchKernelGGL(k_vol<false>, dim3(cfg.vol_grid), dim3(256), 0, stream, pool, scenes_dev, jobs_dev);
  hipLaunchKernelGGL(k_compact_shadow, dim3(std::min(chunks, 2 * cfg.regen_blocks)), dim3(256), 0, stream, pool);
  hipLaunchKernelGGL(k_shadow, dim3(cfg.trace_blocks), dim3(128), stack_bytes(128), stream, pool, scenes_dev,
                     jobs_dev, cfg.gstack, cfg.stack, cfg.refill);
}

void launch_iteration_mixed(const DevPool& pool, const DevScene* scenes_dev, DevJob* jobs_dev,
                            const int32_t* order_dev, const LaunchCfg& cfg, hipStream_t stream) {
  const int chunks = (pool.capacity + kRegenSlots - 1) / kRegenSlots;
  DevGuide G;
  memset(&G, 0, sizeof(G));
  G.job = -1;
  hipLaunchKernelGGL(k_regen<false>, dim3(std::min(chunks, cfg.regen_blocks)), dim3(256), 0, stream, pool, jobs_dev,
                     order_dev, G);
  hipLaunchKernelGGL(k_extend, dim3(cfg.trace_blocks), dim3(128), stack_bytes(128), stream, pool, scenes_dev, jobs_dev,
                     cfg.gstack, cfg.stack, cfg.refill);
  hipLaunchKernelGGL((k_classify<false, true>), dim3(std::min(chunks, 2 * cfg.regen_blocks)), dim3(256), 0, stream, pool, G);
  hipLaunchKernelGGL((k_shade<0, false>), dim3(cfg.shade_grid[0]), dim3(256), 0, stream, pool, scenes_dev, jobs_dev, G);
  hipLaunchKernelGGL((k_shade<1, false>), dim3(cfg.shade_grid[1]), dim3(256), 0, stream, pool, scenes_dev, jobs_dev, G);
  hipLaunchKernelGGL((k_shade<2, false>), dim3(cfg.shade_grid[2]), dim3(256), 0, stream, pool, scenes_dev, jobs_dev, G);
  hipLaunchKernelGGL(k_shade_miss<false>, dim3(cfg.miss_grid), dim3(256), 0, stream, pool, scenes_dev, jobs_dev, G);
  hipLaunchKernelGGL(k_vol<true>, dim3(cfg.vol_grid), dim3(256), 0, stream, pool, scenes_dev, jobs_dev);
  hipLaunchKernelGGL(k_compact_shadow, dim3(std::min(chunks, 2 * cfg.regen_blocks)), dim3(256), 0, stream, pool);
  hipLaunchKernelGGL(k_shadow, dim3(cfg.trace_blocks), dim3(128), stack_bytes(128), stream, pool, scenes_dev,
                     jobs_dev, cfg.gstack, cfg.stack, cfg.refill);
}

void launch_gather(const DevJob* jobs_dev, int job, const DevPool& pool, int npix, hipStream_t stream) {
  const int blocks = std::min((npix + 255) / 256, 8192);
  if (blocks > 0) hipLaunchKernelGGL(k_gather, dim3(blocks), dim3(256), 0, stream, jobs_dev, job, pool, npix);
}

void launch_reset_pool(const DevPool& pool, hipStream_t stream) {
  hipLaunchKernelGGL(k_reset_pool, dim3(1024), dim3(256), 0, stream, pool);
}

void launch_trace_tap(const DevScene* scene_dev, const float* rays, int n, int any_hit, int32_t* prim, float* t,
                      float* uv, float* p, float* nrm, uint32_t* ctr, const LaunchCfg& cfg, hipStream_t stream) {
  if (n == 0) return;
  int blocks = std::min((n + 127) / 128, cfg.trace_blocks);
  hipLaunchKernelGGL(k_trace_tap, dim3(blocks), dim3(128), stack_bytes(128), stream, scene_dev, rays, n, any_hit,
                     prim, t, uv, p, nrm, ctr, cfg.gstack, cfg.stack, cfg.refill);
}

void launch_bsdf_tap(const DevMaterial* mat, int op, const float* wi, const float* wo_or_u, int n, float* out_dir,
                     float* f, float* pdf, float* rel_eta, hipStream_t stream) {
  int blocks = (n + 255) / 256;
  if (blocks == 0) return;
  hipLaunchKernelGGL(k_bsdf_tap, dim3(blocks), dim3(256), 0, stream, mat, op, wi, wo_or_u, n, out_dir, f, pdf,
                     rel_eta);
}

// Mixture sampling tap: the (pi, mu, kappa) x 8 mixture put into the renderer's layout as the
// network kernel stores it, then directions with u = (lobe, u1, u2) and the mixture density at each
// sampled direction, by the functions guided shading calls.
__global__ void k_guide_sample_tap(const float* mix, const float* u, int n, float* dirs, float* pdf) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* mx = mix + (size_t)kGuideOut * i;
  float4 mr4[kGuideMixR / 4];
  float* mr = reinterpret_cast<float*>(mr4);
  mixr_store_half(mr, 0, mx);
  mixr_store_half(mr, 1, mx + kGuideOut / 2);
  f3 d = mixr_sample(mr, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
  dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
  pdf[i] = mixr_pdf(mr, d);
}

void launch_guide_sample_tap(const float* mix, const float* u, int n, float* dirs, float* pdf, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_guide_sample_tap, dim3((n + 255) / 256), dim3(256), 0, stream, mix, u, n, dirs, pdf);
}

// ---------------------------------------------------------------- allocation-free trace
// rtg_trace_device_noalloc's scratch: the trace counters (first 256 bytes) and the traversal-stack
// spill area of a launch of up to kNoAllocBlocks blocks of 128 lanes with kNoAllocSpill entries
// per lane beyond the LDS part. One copy per GPU, kept for the life of the process; host code
// reaches it only through the address scratch_address() returns.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

void* scratch_address(int device) {
  static std::mutex mu;
  static void* addr[kMaxScratchDevices] = {};
  if (device < 0 || device >= kMaxScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    addr[device] = p;
  }
  return addr[device];
}

void launch_trace_tap_noalloc(const DevScene* scene_dev, const float* rays, int n, int any_hit, int32_t* prim,
                              float* t, float* uv, float* p, float* nrm, void* scratch_dev, int stack, int refill,
                              hipStream_t stream) {
  if (n == 0) return;
  uint32_t* ctr = (uint32_t*)scratch_dev;
  uint2* spill = (uint2*)((char*)scratch_dev + kScratchHeader);
  hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(64), 0, stream, ctr, (size_t)2);
  const int blocks = std::min((n + 127) / 128, kNoAllocBlocks);
  hipLaunchKernelGGL(k_trace_tap, dim3(blocks), dim3(128), stack_bytes(128), stream, scene_dev, rays, n, any_hit,
                     prim, t, uv, p, nrm, ctr, spill, stack, refill);
}

void launch_fill_zero(void* ptr, size_t bytes, hipStream_t stream) {
  size_t n = bytes / 4;
  if (n == 0) return;
  int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, stream, (uint32_t*)ptr, n);
}

}  // namespace rtg

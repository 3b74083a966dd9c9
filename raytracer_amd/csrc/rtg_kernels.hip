// Wavefront MIS path tracer kernels for gfx950 (wave64).
//
// One iteration of the pool = regen -> extend(+classify) -> shade per material type -> shadow.
// Slots keep per-path state in structure-of-arrays buffers; every stage consumes a compacted
// queue built with wave ballots + one atomic per wave, so lanes of a wave run the same BSDF
// code. The control flow restates integrator/PathTracingMIS.cpp:27-263 and
// integrator/Integrator.cpp:176-238 split at the two ray casts of each bounce.
#include "rtg_device.h"

namespace rtg {
using namespace dev;

static const uint32_t DEPTH_MASK = 0xFFFu;
static const uint32_t JOB_SHIFT = 12;
static const uint32_t JOB_MASK = 0xFFu;
static const uint32_t FL_LAST = 1u << 20;
static const uint32_t FL_EMIT = 1u << 21;

__device__ __forceinline__ uint32_t fl_depth(uint32_t f) { return f & DEPTH_MASK; }
__device__ __forceinline__ uint32_t fl_job(uint32_t f) { return (f >> JOB_SHIFT) & JOB_MASK; }

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

// Box-filter splat (core/Screen.cpp:44-97 with 32x32 sub-screens, r = 0.5): the sample lands in
// its own pixel; a jitter of exactly 0 also reaches the previous pixel inside the same block.
__device__ __forceinline__ void splat(const DevJob& J, uint32_t pixel, float sx, float sy, f3 L) {
  int x = (int)(pixel % (uint32_t)J.width), y = (int)(pixel / (uint32_t)J.width);
  int x0 = x, y0 = y;
  if (sx == 0.f && (x & 31) != 0) x0 = x - 1;
  if (sy == 0.f && (y & 31) != 0) y0 = y - 1;
  for (int yy = y0; yy <= y; yy++)
    for (int xx = x0; xx <= x; xx++) {
      float4* a = J.accum + (size_t)yy * J.width + xx;
      // hardware float atomics (the film is ordinary coarse-grained device memory)
      unsafeAtomicAdd(&a->x, L.x);
      unsafeAtomicAdd(&a->y, L.y);
      unsafeAtomicAdd(&a->z, L.z);
      unsafeAtomicAdd(&a->w, 1.0f);
    }
}

// ---------------------------------------------------------------- regen
// Finalises finished paths into their job's film, claims new samples for free slots and builds the
// extension queue. Claims are aggregated per block (one 64-bit atomic per block and job), the
// finished-sample count per wave.
__global__ void __launch_bounds__(256) k_regen(DevPool P, DevJob* jobs, int njobs) {
  __shared__ uint32_t s_app[8];
  __shared__ uint32_t s_cnt[4];
  __shared__ unsigned long long s_base;
  uint32_t* C = P.counters;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // shade queues are filled by the extend stage of this iteration
    C[C_SHADE + 0] = 0; C[C_SHADE + 1] = 0; C[C_SHADE + 2] = 0; C[C_SHADE + 3] = 0;
  }
  const int cap = P.capacity;
  const int stride = gridDim.x * blockDim.x;
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  // iterate whole blocks so ballots and barriers see every lane
  int nIter = (cap + stride - 1) / stride;
  for (int it = 0; it < nIter; it++) {
    int i = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = i < cap;
    uint32_t st = valid ? P.state[i] : ST_ACTIVE;
    bool fin = valid && st == ST_DONE;
    uint32_t fjob = 0;
    if (fin) {
      uint32_t f = P.flags[i];
      fjob = fl_job(f);
      const DevJob& J = jobs[fjob];
      float4 L = P.L[i];
      f3 l = mk(L.x, L.y, L.z);
      bool nan = isnan(l.x) || isnan(l.y) || isnan(l.z);
      if (nan) {
        atomicAdd(&C[C_NAN], 1u);
      } else {
        uint32_t pix = P.pixel[i], si = P.sample[i];
        PathRng rng = make_rng(J.seed, pix, si, J);
        float sx, sy;
        sample2d(rng, 0, sx, sy);
        splat(J, pix, sx, sy, l);
      }
      st = ST_FREE;
      P.state[i] = ST_FREE;
    }
    // finished-sample counts: one atomic per (wave, job) present in the wave
    {
      unsigned long long m = __ballot(fin);
      while (m) {
        uint32_t leader = (uint32_t)__builtin_ctzll(m);
        uint32_t lj = (uint32_t)__shfl((int)fjob, (int)leader);
        unsigned long long same = __ballot(fin && fjob == lj);
        if (lane == leader) atomicAdd((unsigned long long*)&jobs[lj].done, (unsigned long long)__popcll(same));
        m &= ~same;
      }
    }
    bool need = valid && st == ST_FREE;
    for (int j = 0; j < njobs; j++) {
      if (!__syncthreads_or(need)) break;
      DevJob& J = jobs[j];
      if (!J.active) continue;
      unsigned long long mask = __ballot(need);
      if (lane == 0) s_cnt[wave] = (uint32_t)__popcll(mask);
      __syncthreads();
      if (threadIdx.x == 0) {
        uint32_t tot = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
        s_base = tot ? atomicAdd((unsigned long long*)&J.next, (unsigned long long)tot) : 0ull;
      }
      __syncthreads();
      uint32_t woff = 0;
      for (uint32_t w = 0; w < wave; w++) woff += s_cnt[w];
      uint64_t s = s_base + woff + (uint64_t)__popcll(mask & ((1ull << lane) - 1ull));
      if (need && s < J.total) {
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
        P.flags[i] = ((uint32_t)j << JOB_SHIFT) | FL_EMIT;
        P.state[i] = ST_ACTIVE;
        st = ST_ACTIVE;
        need = false;
      }
      __syncthreads();  // s_cnt / s_base are rewritten by the next job's round
    }
    bool act = valid && st == ST_ACTIVE;
    uint32_t* qc[1] = {&C[C_EXT]};
    uint32_t pos = block_append<4, 1>(qc, act ? 0 : -1, s_app);
    if (act) P.q_ext[pos] = (uint32_t)i;
  }
}

// ---------------------------------------------------------------- extend: closest hit + classify
__device__ __forceinline__ TravStack make_stack(uint32_t* smem, uint2* gstack, int cap) {
  TravStack s;
  s.lref = smem + threadIdx.x;
  s.lt = reinterpret_cast<float*>(smem + (size_t)kLdsStack * blockDim.x) + threadIdx.x;
  s.lstride = blockDim.x;
  s.gstride = gridDim.x * blockDim.x;
  s.g = gstack + blockIdx.x * blockDim.x + threadIdx.x;
  s.cap = cap;
  return s;
}

__global__ void __launch_bounds__(128) k_extend(DevPool P, const DevScene* scenes, const DevJob* jobs,
                                                uint2* gstack, int stack_depth) {
  extern __shared__ uint32_t smem[];
  __shared__ uint32_t s_app[12];
  const TravStack stk = make_stack(smem, gstack, stack_depth);
  uint32_t* C = P.counters;
  if (blockIdx.x == 0 && threadIdx.x == 0) C[C_SHADOW] = 0;  // shade stage of this iteration fills it
  const uint32_t n = C[C_EXT];
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t nIter = (n + stride - 1) / stride;
  for (uint32_t it = 0; it < nIter; it++) {
    uint32_t k = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = k < n;
    int cls = -1;
    uint32_t slot = 0;
    uint32_t nv = 0, np = 0;
    if (valid) {
      slot = P.q_ext[k];
      uint32_t f = P.flags[slot];
      const DevScene& S = scenes[jobs[fl_job(f)].scene];
      float4 o = P.ray_o[slot], d = P.ray_d[slot];
      Ray r;
      r.o = mk(o.x, o.y, o.z);
      r.d = mk(d.x, d.y, d.z);
      r.inv = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
      r.tmin = o.w;
      r.tmax = d.w;
      HitRec h;
      traverse<false>(S, r, stk, h, nv, np);
      P.hit[slot] = make_float4(h.t, h.b1, h.b2, __int_as_float(h.prim));
      if (h.prim < 0) {
        cls = 3;
      } else {
        uint32_t m = S.prim_info[h.prim].y;
        cls = (m == 0xffffffffu) ? 3 : S.mats[m].type;
        if (m == 0xffffffffu) cls = 4;  // no BSDF: handled by the miss kernel as a terminator
      }
    }
    wave_count64((unsigned long long*)&C[C_EXTRAYS], valid);
    if (P.count_work) {
      uint32_t a = wave_sum(nv), b = wave_sum(np);
      if (__lane_id() == 0) {
        atomicAdd((unsigned long long*)&C[C_NODES], (unsigned long long)a);
        atomicAdd((unsigned long long*)&C[C_PRIMS], (unsigned long long)b);
      }
    }
    // classify into the four shade queues (no-BSDF hits go with misses)
    int q = valid ? (cls == 4 ? 3 : cls) : -1;
    uint32_t* qc[4] = {&C[C_SHADE + 0], &C[C_SHADE + 1], &C[C_SHADE + 2], &C[C_SHADE + 3]};
    uint32_t pos = block_append<2, 4>(qc, q, s_app);
    if (q >= 0) P.q_shade[(size_t)q * P.capacity + pos] = slot;
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
    float w = power_heuristic(F.w, pdfLight);
    f3 c = cmul(cmul(mk(T.x, T.y, T.z), Le), mk(F.x, F.y, F.z)) * w;
    L.x += c.x; L.y += c.y; L.z += c.z;
  }
  f3 t = cmul(mk(T.x, T.y, T.z), mk(F.x, F.y, F.z));
  t = t / L.w;  // Russian-roulette divisor (1 when no roulette was played)
  T.x = t.x; T.y = t.y; T.z = t.z;
}

template <int MT>
__global__ void __launch_bounds__(256) k_shade(DevPool P, const DevScene* scenes, const DevJob* jobs) {
  __shared__ uint32_t s_app[8];
  uint32_t* C = P.counters;
  const uint32_t n = C[C_SHADE + MT];
  const uint32_t* q = P.q_shade + (size_t)MT * P.capacity;
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t nIter = (n + stride - 1) / stride;
  for (uint32_t it = 0; it < nIter; it++) {
    uint32_t k = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = k < n;
    bool shadow = false;
    uint32_t slot = 0;
    if (valid) {
      slot = q[k];
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
        const DevMaterial& M = S.mats[sf.material];
        PathRng rng = make_rng(J.seed, P.pixel[slot], P.sample[slot], J);
        f3 wo_l = to_local(sf.sh, -d);
        // one-light next-event estimation (Integrator.cpp:176-238, PathTracingMIS.cpp:119-153)
        if (S.nlights > 0) {
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
              P.sh_d[slot] = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, ls.dist - EPS);
              P.sh_c[slot] = make_float4(c.x, c.y, c.z, 0.f);
              shadow = true;
            }
          }
        }
        // BSDF sampling (PathTracingMIS.cpp:156-164) and roulette (:225-253)
        float bx, by;
        sample2d(rng, 2u + 2u * depth, bx, by);
        BsdfSample bs;
        bool ok = bsdf_sample<MT>(M, wo_l, bx, by, bs);
        if (!ok || is_zero(bs.f)) {
          done = true;
        } else {
          f3 nd = to_world(sf.sh, bs.wo);
          P.ray_d[slot] = make_float4(nd.x, nd.y, nd.z, MAX_REAL);
          F = make_float4(bs.f.x, bs.f.y, bs.f.z, bs.pdf);
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
    uint32_t* qc[1] = {&C[C_SHADOW]};
    uint32_t pos = block_append<4, 1>(qc, shadow ? 0 : -1, s_app);
    if (shadow) P.q_shadow[pos] = slot;
  }
}

// Misses and hits without a BSDF (PathTracingMIS.cpp:38-51,67-70,182-189,217-219)
__global__ void __launch_bounds__(256) k_shade_miss(DevPool P, const DevScene* scenes, const DevJob* jobs) {
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

// ---------------------------------------------------------------- shadow rays (Integrator.cpp:228-235)
__global__ void __launch_bounds__(128) k_shadow(DevPool P, const DevScene* scenes, const DevJob* jobs,
                                                uint2* gstack, int stack_depth) {
  extern __shared__ uint32_t smem[];
  const TravStack stk = make_stack(smem, gstack, stack_depth);
  uint32_t* C = P.counters;
  if (blockIdx.x == 0 && threadIdx.x == 0) C[C_EXT] = 0;  // regen of the next iteration fills it
  const uint32_t n = C[C_SHADOW];
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t nIter = (n + stride - 1) / stride;
  for (uint32_t it = 0; it < nIter; it++) {
    uint32_t k = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = k < n;
    uint32_t nv = 0, np = 0;
    if (valid) {
      uint32_t slot = P.q_shadow[k];
      uint32_t f = P.flags[slot];
      const DevScene& S = scenes[jobs[fl_job(f)].scene];
      float4 o = P.ray_o[slot], d = P.sh_d[slot];
      Ray r;
      r.o = mk(o.x, o.y, o.z);
      r.d = mk(d.x, d.y, d.z);
      r.inv = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
      r.tmin = EPS;
      r.tmax = d.w;
      HitRec h;
      bool occ = traverse<true>(S, r, stk, h, nv, np);
      if (!occ) {
        float4 c = P.sh_c[slot];
        float4 L = P.L[slot];
        L.x += c.x; L.y += c.y; L.z += c.z;
        P.L[slot] = L;
      }
    }
    wave_count64((unsigned long long*)&C[C_SHRAYS], valid);
    if (P.count_work) {
      uint32_t a = wave_sum(nv), b = wave_sum(np);
      if (__lane_id() == 0) {
        atomicAdd((unsigned long long*)&C[C_SHNODES], (unsigned long long)a);
        atomicAdd((unsigned long long*)&C[C_SHPRIMS], (unsigned long long)b);
      }
    }
  }
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
__global__ void __launch_bounds__(128) k_trace_tap(const DevScene* scene, const float* rays, int n, int any_hit,
                                                   int32_t* prim, float* t, float* uv, float* p, float* nrm,
                                                   uint2* gstack, int stack_depth) {
  extern __shared__ uint32_t smem[];
  const TravStack stk = make_stack(smem, gstack, stack_depth);
  const DevScene& S = *scene;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float* q = rays + 8 * (size_t)i;
    Ray r;
    r.o = mk(q[0], q[1], q[2]);
    r.d = mk(q[3], q[4], q[5]);
    r.inv = mk(1.f / q[3], 1.f / q[4], 1.f / q[5]);
    r.tmin = q[6];
    r.tmax = q[7];
    HitRec h;
    uint32_t nv = 0, np = 0;
    bool hit = any_hit ? traverse<true>(S, r, stk, h, nv, np) : traverse<false>(S, r, stk, h, nv, np);
    prim[i] = hit ? (int32_t)S.prim_info[h.prim].x : -1;
    if (t) t[i] = hit ? h.t : 0.f;
    if (uv) {
      uv[2 * i] = hit ? h.b1 : 0.f;
      uv[2 * i + 1] = hit ? h.b2 : 0.f;
    }
    if (p || nrm) {
      f3 pp = mk(0.f, 0.f, 0.f), nn = mk(0.f, 0.f, 0.f);
      if (hit && !any_hit) {
        Surface sf = surface_at(S, h, r.o, r.d);
        pp = sf.p;
        nn = sf.sh.n;
      }
      if (p) { p[3 * i] = pp.x; p[3 * i + 1] = pp.y; p[3 * i + 2] = pp.z; }
      if (nrm) { nrm[3 * i] = nn.x; nrm[3 * i + 1] = nn.y; nrm[3 * i + 2] = nn.z; }
    }
  }
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
    if (M.type == 0) { e = bsdf_eval<0>(M, a, b); pd = bsdf_pdf<0>(M, a, b); }
    else if (M.type == 1) { e = bsdf_eval<1>(M, a, b); pd = bsdf_pdf<1>(M, a, b); }
    else { e = bsdf_eval<2>(M, a, b); pd = bsdf_pdf<2>(M, a, b); }
    f[3 * i] = e.x; f[3 * i + 1] = e.y; f[3 * i + 2] = e.z;
    pdf[i] = pd;
  } else {
    float ux = wo_or_u[2 * i], uy = wo_or_u[2 * i + 1];
    BsdfSample bs;
    bs.wo = mk(0.f, 0.f, 0.f);
    bs.f = mk(0.f, 0.f, 0.f);
    bs.pdf = -1.f;
    bs.rel_eta = 1.f;
    bool ok;
    if (M.type == 0) ok = bsdf_sample<0>(M, a, ux, uy, bs);
    else if (M.type == 1) ok = bsdf_sample<1>(M, a, ux, uy, bs);
    else ok = bsdf_sample<2>(M, a, ux, uy, bs);
    if (!ok) bs.f = mk(0.f, 0.f, 0.f);
    out_dir[3 * i] = bs.wo.x; out_dir[3 * i + 1] = bs.wo.y; out_dir[3 * i + 2] = bs.wo.z;
    f[3 * i] = bs.f.x; f[3 * i + 1] = bs.f.y; f[3 * i + 2] = bs.f.z;
    pdf[i] = bs.pdf;
    rel_eta[i] = bs.rel_eta;
  }
}

// ---------------------------------------------------------------- launchers
static size_t stack_bytes(int block) { return (size_t)block * kLdsStack * 8u; }

void launch_iteration(const DevPool& pool, const DevScene* scenes_dev, DevJob* jobs_dev, int njobs,
                      const LaunchCfg& cfg, hipStream_t stream, hipEvent_t* ev) {
  int g = cfg.grid;
  const DevJob* cj = jobs_dev;
  if (ev) (void)hipEventRecord(ev[0], stream);
  hipLaunchKernelGGL(k_regen, dim3(g), dim3(256), 0, stream, pool, jobs_dev, njobs);
  if (ev) (void)hipEventRecord(ev[1], stream);
  hipLaunchKernelGGL(k_extend, dim3(g * 2), dim3(128), stack_bytes(128), stream, pool, scenes_dev, cj,
                     cfg.gstack, cfg.stack);
  if (ev) (void)hipEventRecord(ev[2], stream);
  hipLaunchKernelGGL(k_shade<0>, dim3(g), dim3(256), 0, stream, pool, scenes_dev, cj);
  if (ev) (void)hipEventRecord(ev[3], stream);
  hipLaunchKernelGGL(k_shade<1>, dim3(g), dim3(256), 0, stream, pool, scenes_dev, cj);
  if (ev) (void)hipEventRecord(ev[4], stream);
  hipLaunchKernelGGL(k_shade<2>, dim3(g), dim3(256), 0, stream, pool, scenes_dev, cj);
  if (ev) (void)hipEventRecord(ev[5], stream);
  hipLaunchKernelGGL(k_shade_miss, dim3(g), dim3(256), 0, stream, pool, scenes_dev, cj);
  if (ev) (void)hipEventRecord(ev[6], stream);
  hipLaunchKernelGGL(k_shadow, dim3(g * 2), dim3(128), stack_bytes(128), stream, pool, scenes_dev, cj,
                     cfg.gstack, cfg.stack);
  if (ev) (void)hipEventRecord(ev[7], stream);
}

void launch_reset_pool(const DevPool& pool, hipStream_t stream) {
  hipLaunchKernelGGL(k_reset_pool, dim3(1024), dim3(256), 0, stream, pool);
}

void launch_trace_tap(const DevScene* scene_dev, const float* rays, int n, int any_hit, int32_t* prim, float* t,
                      float* uv, float* p, float* nrm, const LaunchCfg& cfg, hipStream_t stream) {
  if (n == 0) return;
  int blocks = std::min((n + 127) / 128, cfg.grid * 2);
  hipLaunchKernelGGL(k_trace_tap, dim3(blocks), dim3(128), stack_bytes(128), stream, scene_dev, rays, n, any_hit,
                     prim, t, uv, p, nrm, cfg.gstack, cfg.stack);
}

void launch_bsdf_tap(const DevMaterial* mat, int op, const float* wi, const float* wo_or_u, int n, float* out_dir,
                     float* f, float* pdf, float* rel_eta, hipStream_t stream) {
  int blocks = (n + 255) / 256;
  if (blocks == 0) return;
  hipLaunchKernelGGL(k_bsdf_tap, dim3(blocks), dim3(256), 0, stream, mat, op, wi, wo_or_u, n, out_dir, f, pdf,
                     rel_eta);
}

void launch_fill_zero(void* ptr, size_t bytes, hipStream_t stream) {
  size_t n = bytes / 4;
  if (n == 0) return;
  int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, stream, (uint32_t*)ptr, n);
}

}  // namespace rtg

// Device-side building blocks of the wavefront path tracer (gfx950, wave64).
// Arithmetic follows the reference's FP32 operation order, without FMA contraction
// (the library is compiled with -ffp-contract=off), so BVH first hits and BSDF
// eval/pdf match a CPU restatement bit for bit; sampling uses the device's
// sin/cos/atan/acos/log, which differ from glibc by a few ulp.
#pragma once
#include "rtg_internal.h"

namespace rtg {
namespace dev {

// Path-slot flags word: bits 0..11 depth of the pending hit, 12..19 job slot, bit 20 LAST (only the
// emission of the pending segment remains), bit 21 EMISSION (depth-0 emission enabled), bits 22-23
// volumetric path state (rtg_kernels.hip), bit 24 VOL (the path belongs to a VolPathTracing job),
// bits 25/26 XEDGE/YEDGE (the sample's pixel jitter is exactly 0 in x/y: all the box filter's splat
// needs of the jitter, recorded when the sample is claimed).
static const uint32_t DEPTH_MASK = 0xFFFu;
static const uint32_t JOB_SHIFT = 12;
static const uint32_t JOB_MASK = 0xFFu;
static const uint32_t FL_LAST = 1u << 20;
static const uint32_t FL_EMIT = 1u << 21;
static const uint32_t FL_VOL = 1u << 24;
static const uint32_t FL_XEDGE = 1u << 25;
static const uint32_t FL_YEDGE = 1u << 26;
__device__ __forceinline__ uint32_t fl_depth(uint32_t f) { return f & DEPTH_MASK; }
__device__ __forceinline__ uint32_t fl_job(uint32_t f) { return (f >> JOB_SHIFT) & JOB_MASK; }

// core/BasicMath.h:16-21 (float branch)
constexpr float PI = 3.14159265358979323846f;
constexpr float INV_PI = 0.3183098861837906715377f;
constexpr float MAX_REAL = 10e7f;
constexpr float EPS = 1e-4f;
constexpr float ONE_MINUS_EPS = 0.99999994f;
// |det| < 1e-8 in Triangle.cpp:66 is a double comparison; for a float |det| it is
// equivalent to |det| < the smallest float >= 1e-8.
constexpr float DET_EPS = 0x1.5798f0p-27f;

struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ f3 cmul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 cdiv(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ f3 cadd(f3 a, float s) { return mk(a.x + s, a.y + s, a.z + s); }
// 3-term reductions in Eigen's unrolled order a0 + (a1 + a2)
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float sqnorm(f3 a) { return dot(a, a); }
__device__ __forceinline__ float norm(f3 a) { return sqrtf(sqnorm(a)); }
__device__ __forceinline__ f3 normalized(f3 a) {
  float z = sqnorm(a);
  if (z > 0.f) return a / sqrtf(z);
  return a;
}
__device__ __forceinline__ f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
__device__ __forceinline__ f3 xyz(float4 v) { return mk(v.x, v.y, v.z); }
__device__ __forceinline__ float luminance(f3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
__device__ __forceinline__ bool is_zero(f3 c) {
  return fabsf(c.x) <= 1e-5f && fabsf(c.y) <= 1e-5f && fabsf(c.z) <= 1e-5f;
}
// every component exactly 0 (a NaN component is not): a contribution that adds nothing
__device__ __forceinline__ bool exactly_zero(f3 c) { return c.x == 0.f && c.y == 0.f && c.z == 0.f; }
__device__ __forceinline__ float average(float a) {  // Color(a).average(), Color.h:151-158
  float s = 0.f;
  s += a; s += a; s += a;
  return s / 3.f;
}
// std::min/std::max argument semantics (NaN handling follows the C++ definitions)
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------- frames (DifferentialGeometry.h:36-66)
struct Frame {
  f3 s, t, n;
};
__device__ __forceinline__ Frame make_frame(f3 n) {
  Frame f;
  f.n = n;
  f3 c;
  if (fabsf(n.x) > fabsf(n.y)) {
    float inv = 1.0f / sqrtf(n.x * n.x + n.z * n.z);
    c = mk(n.z * inv, 0.0f, -n.x * inv);
  } else {
    float inv = 1.0f / sqrtf(n.y * n.y + n.z * n.z);
    c = mk(0.0f, n.z * inv, -n.y * inv);
  }
  f.t = c;
  f.s = cross(c, n);
  return f;
}
__device__ __forceinline__ f3 to_local(const Frame& f, f3 v) { return mk(dot(v, f.s), dot(v, f.t), dot(v, f.n)); }
__device__ __forceinline__ f3 to_world(const Frame& f, f3 v) { return f.s * v.x + f.t * v.y + f.n * v.z; }

// ---------------------------------------------------------------- transforms (Transform.h:87-116,176-184)
__device__ __forceinline__ f3 xform_point(const float4* M, f3 p) {
  float4 r0 = M[0], r1 = M[1], r2 = M[2], r3 = M[3];
  f3 o;
  o.x = r0.x * p.x + r0.y * p.y + r0.z * p.z + r0.w;
  o.y = r1.x * p.x + r1.y * p.y + r1.z * p.z + r1.w;
  o.z = r2.x * p.x + r2.y * p.y + r2.z * p.z + r2.w;
  float w = r3.x * p.x + r3.y * p.y + r3.z * p.z + r3.w;
  if (w != 1.f && w != 0.f) o = o / w;
  return o;
}
__device__ __forceinline__ f3 xform_vector(const float4* M, f3 v) {
  float4 r0 = M[0], r1 = M[1], r2 = M[2];
  return mk(r0.x * v.x + (r0.y * v.y + r0.z * v.z), r1.x * v.x + (r1.y * v.y + r1.z * v.z),
            r2.x * v.x + (r2.y * v.y + r2.z * v.z));
}
__device__ __forceinline__ f3 xform_normal(const float4* Inv, f3 n) {
  float4 r0 = Inv[0], r1 = Inv[1], r2 = Inv[2];
  return mk(r0.x * n.x + (r1.x * n.y + r2.x * n.z), r0.y * n.x + (r1.y * n.y + r2.y * n.z),
            r0.z * n.x + (r1.z * n.y + r2.z * n.z));
}

// ---------------------------------------------------------------- sampler
// Counter-based streams: every (seed, pixel, sample, dimension) has its own value, so a path's
// result does not depend on which slot, iteration or concurrent request runs it.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {  // SplitMix64 finaliser (Steele et al. 2014)
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ uint32_t draw32(uint64_t key, uint32_t counter) {
  return (uint32_t)(mix64(key + (uint64_t)counter * 0xD1B54A32D192ED03ull) >> 32);
}
// tools/Rng.h:132-141 bit trick: uniform float in [0,1) from the top 23 bits
__device__ __forceinline__ float to_unit(uint32_t u) { return __uint_as_float((u >> 9) | 0x3f800000u) - 1.0f; }
// Hash-based permutation of [0, l) (Kensler, "Correlated Multi-Jittered Sampling", 2013)
__device__ __forceinline__ uint32_t permute(uint32_t i, uint32_t l, uint32_t p) {
  uint32_t w = l - 1;
  w |= w >> 1; w |= w >> 2; w |= w >> 4; w |= w >> 8; w |= w >> 16;
  do {
    i ^= p; i *= 0xe170893du;
    i ^= p >> 16;
    i ^= (i & w) >> 4;
    i ^= p >> 8; i *= 0x0929eb3fu;
    i ^= p >> 23;
    i ^= (i & w) >> 1; i *= 1u | p >> 27;
    i *= 0x6935fa69u;
    i ^= (i & w) >> 11; i *= 0x74dcb303u;
    i ^= (i & w) >> 2; i *= 0x9e501cc3u;
    i ^= (i & w) >> 2; i *= 0xc860a3dfu;
    i &= w;
    i ^= i >> 5;
  } while (i >= l);
  return (i + p) % l;
}
struct PathRng {
  uint64_t kpix;   // per (seed, pixel): stratum permutations shared by the pixel's samples
  uint64_t kpath;  // per (seed, pixel, sample)
  uint32_t sample, spp, sqrt_spp;
  int32_t dimension;  // stratified dimensions (0 for the random sampler)
};
__device__ __forceinline__ PathRng make_rng(uint64_t seed, uint32_t pixel, uint32_t sample, const DevJob& J) {
  PathRng r;
  r.kpix = mix64(mix64(seed) ^ ((uint64_t)pixel * 0x9E3779B97F4A7C15ull));
  r.kpath = mix64(r.kpix ^ ((uint64_t)(sample + 1u) * 0xC2B2AE3D27D4EB4Full));
  r.sample = sample;
  r.spp = (uint32_t)J.spp;
  r.sqrt_spp = (uint32_t)J.sqrt_spp;
  r.dimension = J.sampler == 1 ? J.dimension : 0;
  return r;
}
// (cell + jitter) / n for a jitter in [0, 1), kept inside [cell / n, (cell + 1) / n) exactly: the
// float sum cell + jitter rounds up to cell + 1 for a jitter within an ulp of 1, and the product
// with a rounded 1 / n can leave the cell at either end. The fused multiply-add's sign is that of
// the exact x n - bound, so the corrections step one ulp at a time exactly while needed (the two
// roundings put x at most a few ulps outside the cell, which is many ulps wide).
__device__ __forceinline__ float in_stratum(uint32_t cell, float jitter, uint32_t n) {
  float x = ((float)cell + jitter) * (1.f / (float)n);
  for (int i = 0; i < 8 && fmaf(x, (float)n, -(float)(cell + 1u)) >= 0.f; i++)
    x = __uint_as_float(__float_as_uint(x) - 1u);
  for (int i = 0; i < 8 && fmaf(x, (float)n, -(float)cell) < 0.f; i++) x = __uint_as_float(__float_as_uint(x) + 1u);
  return smin(x, ONE_MINUS_EPS);
}
// 2D draw number k: stratified over the pixel's samples for k < dimension
// (sampler/StratifiedSampler.cpp:145-151,174-215), independent uniform beyond.
__device__ __forceinline__ void sample2d(const PathRng& r, uint32_t k, float& x, float& y) {
  if ((int32_t)k < r.dimension) {
    uint32_t j = permute(r.sample, r.spp, draw32(r.kpix, 2u * k));
    uint32_t sx = j / r.sqrt_spp, sy = j % r.sqrt_spp;
    float jx = to_unit(draw32(r.kpath, 0x40000000u + 2u * k));
    float jy = to_unit(draw32(r.kpath, 0x40000001u + 2u * k));
    x = in_stratum(sx, jx, r.sqrt_spp);
    y = in_stratum(sy, jy, r.sqrt_spp);
  } else {
    x = to_unit(draw32(r.kpath, 2u * k));
    y = to_unit(draw32(r.kpath, 2u * k + 1u));
  }
}
__device__ __forceinline__ float sample1d(const PathRng& r, uint32_t k) {
  if ((int32_t)k < r.dimension) {
    uint32_t j = permute(r.sample, r.spp, draw32(r.kpix, 0x20000001u + 2u * k));
    float u = to_unit(draw32(r.kpath, 0x60000000u + k));
    return in_stratum(j, u, r.spp);
  }
  return to_unit(draw32(r.kpath, 0x80000000u + k));
}

// ---------------------------------------------------------------- wave primitives
// Wave-aggregated append: one atomic per wave, lanes get consecutive positions.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
  unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return 0xffffffffu;
  uint32_t lane = __lane_id();
  uint32_t leader = (uint32_t)__builtin_ctzll(mask);
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, (int)leader);
  return pred ? base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull)) : 0xffffffffu;
}
// Block-aggregated append into NQ queues at once: lane class c in [0, NQ) or -1. One atomic per
// (block, queue). Every thread of the block must call it (it synchronises the block).
// s: LDS scratch of NWAVES * NQ + NQ words.
template <int NWAVES, int NQ>
__device__ __forceinline__ uint32_t block_append(uint32_t* const* counters, int c, uint32_t* s) {
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  unsigned long long mine = 0ull;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    unsigned long long m = __ballot(c == q);
    if (c == q) mine = m;
    if (lane == 0) s[wave * NQ + q] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < NWAVES; w++) tot += s[w * NQ + threadIdx.x];
    s[NWAVES * NQ + threadIdx.x] = tot ? atomicAdd(counters[threadIdx.x], tot) : 0u;
  }
  __syncthreads();
  uint32_t pos = 0xffffffffu;
  if (c >= 0) {
    pos = s[NWAVES * NQ + c];
    for (uint32_t w = 0; w < wave; w++) pos += s[w * NQ + c];
    pos += (uint32_t)__popcll(mine & lt);
  }
  __syncthreads();
  return pos;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ void wave_count64(unsigned long long* counter, bool pred) {
  unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return;
  uint32_t lane = __lane_id();
  if (lane == (uint32_t)__builtin_ctzll(mask)) atomicAdd(counter, (unsigned long long)__popcll(mask));
}

// ---------------------------------------------------------------- ray / box / primitives
struct Ray {
  f3 o, d, inv;
  float tmin, tmax;
};
__device__ __forceinline__ bool has_zero_dir(const Ray& r) { return r.d.x == 0.f || r.d.y == 0.f || r.d.z == 0.f; }
// accelStruct/Bvh.cpp:411-413 adaptive epsilon
__device__ __forceinline__ void adapt_eps(Ray& r) {
  if (r.tmin == EPS) r.tmin = smax(r.tmin, r.tmin * smax(smax(fabsf(r.o.x), fabsf(r.o.y)), fabsf(r.o.z)));
}
// core/BoundingBox.h:154-183
__device__ __forceinline__ bool slab(const Ray& r, float bx0, float by0, float bz0, float bx1, float by1,
                                     float bz1, float& nearT) {
  float nT = -MAX_REAL, fT = MAX_REAL;
  const float o[3] = {r.o.x, r.o.y, r.o.z};
  const float d[3] = {r.d.x, r.d.y, r.d.z};
  const float iv[3] = {r.inv.x, r.inv.y, r.inv.z};
  const float mn[3] = {bx0, by0, bz0};
  const float mx[3] = {bx1, by1, bz1};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    if (d[i] == 0.f) {
      if (o[i] < mn[i] || o[i] > mx[i]) return false;
    } else {
      float t1 = (mn[i] - o[i]) * iv[i];
      float t2 = (mx[i] - o[i]) * iv[i];
      if (t1 > t2) {
        float tmp = t1; t1 = t2; t2 = tmp;
      }
      nT = smax(t1, nT);
      fT = smin(t2, fT);
      if (!(nT <= fT)) return false;
    }
  }
  nearT = nT;
  return (r.tmin <= fT && r.tmax >= nT);
}
// geometricShape/Triangle.cpp:46-89
__device__ __forceinline__ bool tri_hit(const Ray& r, const float* v, float& t, float& b1, float& b2) {
  f3 p0 = mk(v[0], v[1], v[2]);
  f3 p1 = mk(v[3], v[4], v[5]);
  f3 p2 = mk(v[6], v[7], v[8]);
  f3 e1 = p1 - p0;
  f3 e2 = p2 - p0;
  f3 s = r.o - p0;
  f3 s1 = cross(r.d, e2);
  f3 s2 = cross(s, e1);
  float det = dot(s1, e1);
  if (fabsf(det) < DET_EPS) return false;
  float inv = 1.f / det;
  float u = inv * dot(s1, s);
  if (u < 0.f || u > 1.f) return false;
  float w = inv * dot(s2, r.d);
  if (w < 0.f || u + w > 1.f) return false;
  float tt = inv * dot(s2, e2);
  if (tt < r.tmin || tt > r.tmax) return false;
  t = tt; b1 = u; b2 = w;
  return true;
}
// geometricShape/Sphere.cpp:38-111 (sph: inverse rows 0..3, then radius)
__device__ __forceinline__ bool sphere_hit(const Ray& r, const float4* sph, float& t) {
  f3 lo = xform_point(sph, r.o);
  f3 ld = xform_vector(sph, r.d);
  float radius = sph[4].x;
  float a = 1.f;
  float b = 2.f * dot(ld, lo);
  float c = dot(lo, lo) - radius * radius;
  float delta = b * b - 4.f * a * c;
  if (delta < 0.f) return false;
  float t1 = (-b + sqrtf(delta)) / (2.f * a);
  float t2 = (-b - sqrtf(delta)) / (2.f * a);
  float tn = t1, tf = t2;
  if (tn > tf) { float x = tn; tn = tf; tf = x; }
  if (tn <= r.tmin) {
    if (tf > r.tmin && tf < r.tmax) { t = tf; return true; }
  } else if (tn > r.tmin && tn < r.tmax) {
    t = tn; return true;
  }
  return false;
}

struct HitRec {
  float t, b1, b2;
  int32_t prim;  // leaf-order primitive index, -1 on miss
};

// Traversal stack: the top kLdsStack entries of each lane live in LDS ([entry][block thread] layout,
// conflict-free), deeper entries spill to a global buffer ([entry][grid thread] layout, coalesced).
// An entry is a subtree ref plus the entry distance of its box, so a popped subtree that starts
// beyond the current closest hit is skipped without fetching its node.
struct TravStack {
  uint2* l;        // LDS entries (ref, entry distance bits), this lane's column
  int lstride;     // threads per block
  uint2* g;        // global spill, this lane's column
  int gstride;     // threads in the grid
  int cap;         // total entries (>= the deepest stack any ray needs, set by the host)
  uint32_t* spill; // counts pushes into the global spill area (null: not counted)
  uint32_t* ovf;   // set to 1 when a push finds the stack full (the host reports it as an error)
};
__device__ __forceinline__ void stk_push(const TravStack& s, int& sp, uint32_t ref, float t) {
  if (sp < kLdsStack) {
    s.l[sp * s.lstride] = make_uint2(ref, __float_as_uint(t));
    sp++;
  } else if (sp < s.cap) {
    s.g[(size_t)(sp - kLdsStack) * s.gstride] = make_uint2(ref, __float_as_uint(t));
    sp++;
    if (s.spill) atomicAdd(s.spill, 1u);
  } else {
    // The builder's bound on the stack depth was exceeded: the subtree would be lost. Never
    // expected; flagged so that the render fails instead of dropping geometry silently.
    *s.ovf = 1u;
  }
}
// Pushes the nh - 1 farther hit children of a node (sorted near to far: 1, 2, 3), farthest at the
// bottom. With room for three entries in LDS it writes three slots unconditionally (slots above the
// new top are free) instead of branching per entry.
__device__ __forceinline__ void stk_push_children(const TravStack& s, int& sp, int nh, uint32_t c1, float k1,
                                                  uint32_t c2, float k2, uint32_t c3, float k3) {
  if (sp + 3 <= kLdsStack) {
    const uint2 e1 = make_uint2(c1, __float_as_uint(k1));
    const uint2 e2 = make_uint2(c2, __float_as_uint(k2));
    const uint2 e3 = make_uint2(c3, __float_as_uint(k3));
    s.l[sp * s.lstride] = nh == 4 ? e3 : (nh == 3 ? e2 : e1);
    s.l[(sp + 1) * s.lstride] = nh == 4 ? e2 : e1;
    s.l[(sp + 2) * s.lstride] = e1;
    sp += nh - 1;
  } else {
    if (nh > 3) stk_push(s, sp, c3, k3);
    if (nh > 2) stk_push(s, sp, c2, k2);
    stk_push(s, sp, c1, k1);
  }
}
__device__ __forceinline__ void stk_top(const TravStack& s, int sp, uint32_t& ref, float& t) {
  const uint2 e = sp < kLdsStack ? s.l[sp * s.lstride] : s.g[(size_t)(sp - kLdsStack) * s.gstride];
  ref = e.x;
  t = __uint_as_float(e.y);
}

// BVH traversal over 4-wide nodes, nearest child first. Closest hit keeps the reference's
// acceptance rules (triangles accept t == maxT, the last accepted primitive wins); ANY stops at
// the first accepted primitive. The traversal is resumable one step at a time (TravState), so a
// persistent kernel can give a lane a new ray as soon as its previous one is done.
struct TravState {
  Ray r;
  uint32_t ref;  // node or leaf to visit next
  int sp;        // stack entries in use
  bool zdir;     // a direction component is zero: the reference's per-axis slab branch
  bool hit;
  HitRec h;
};
// The scene arrays a traversal reads, kept in registers for the duration of a ray.
struct SceneGeo {
  const float4* nodes;  // compressed 4-wide nodes, 4 x 16 bytes each (rtg_bvh.h quantize_nodes)
  const float* tri_v;
  const float4* spheres;
  int32_t ntri;
};
__device__ __forceinline__ SceneGeo scene_geo(const DevScene& S) {
  SceneGeo g;
  g.nodes = S.nodes;
  g.tri_v = S.tri_v;
  g.spheres = S.spheres;
  g.ntri = S.ntri;
  return g;
}

// Starts the traversal of s.r; false when there is nothing to traverse (empty scene, empty ray
// interval or a miss of the scene box), with s.hit false and s.h.prim -1.
__device__ __forceinline__ bool trav_begin(const DevScene& S, TravState& s) {
  s.h.prim = -1;
  s.hit = false;
  s.sp = 0;
  adapt_eps(s.r);
  if (S.empty || s.r.tmax < s.r.tmin) return false;
  float tn;
  if (!slab(s.r, S.root_min[0], S.root_min[1], S.root_min[2], S.root_max[0], S.root_max[1], S.root_max[2], tn))
    return false;
  s.ref = S.root_ref;
  // the reference's per-axis branch also takes directions whose reciprocal overflowed (the fused
  // node test below needs finite reciprocals)
  s.zdir = has_zero_dir(s.r) || !(fabsf(s.r.inv.x) < INFINITY && fabsf(s.r.inv.y) < INFINITY &&
                                   fabsf(s.r.inv.z) < INFINITY);
  return true;
}

// Next subtree from the stack that can still hold a closer hit (its box starts at or before maxT),
// or kTravDone when the stack is empty.
constexpr uint32_t kTravDone = 0xFFFFFFFFu;
__device__ __forceinline__ uint32_t trav_pop(const TravStack& stk, TravState& s) {
  while (s.sp > 0) {
    s.sp--;
    float pt;
    uint32_t pr;
    stk_top(stk, s.sp, pr, pt);
    if (s.r.tmax >= pt && pr != kEmptyChild) return pr;
  }
  return kTravDone;
}

// Visits the interior node s.ref: descends into the nearest hit child and pushes the other hit
// children farthest first (so the next nearest is popped next), or pops when no child is hit.
// A compressed node's child bound: origin + q * step (exact product, one rounded addition; the host
// rounded q outward for exactly this arithmetic, so the decoded box contains the child's box).
__device__ __forceinline__ float4 dq4(uint32_t w, float origin, float step) {
  return make_float4(origin + (float)(w & 0xFFu) * step, origin + (float)((w >> 8) & 0xFFu) * step,
                     origin + (float)((w >> 16) & 0xFFu) * step, origin + (float)(w >> 24) * step);
}
// Fused test of the four children for rays without zero direction components: per axis, with
// A = step / d and B = (origin - o) / d, a child bound q gives the entry or exit distance q A + B
// (one fma after the byte conversion instead of a decode and a slab), and the bytes of the near
// and far bound of every child are chosen once per node by the sign of 1 / d. The distances are
// widened by delta = 2^-20 (|origin / d| + |B| + 256 |A|), which bounds the rounding difference
// to the reference's slab test on the exact child box (which the decoded box contains) for these
// operations, so the test accepts every box the reference's would: it only culls, and hits are
// decided by the triangle test.
__device__ __forceinline__ void fused_axis(float o, float step, float ro, float inv, uint32_t qlo, uint32_t qhi,
                                           float& A, float& Bn, float& Bf, uint32_t& qn, uint32_t& qf) {
  A = step * inv;
  const float B = (o - ro) * inv;
  const float delta = 0x1p-20f * (fabsf(o * inv) + fabsf(B) + 256.f * fabsf(A));
  Bn = B - delta;
  Bf = B + delta;
  qn = inv >= 0.f ? qlo : qhi;
  qf = inv >= 0.f ? qhi : qlo;
}
__device__ __forceinline__ bool fused_child(const Ray& r, float qnx, float qfx, float qny, float qfy, float qnz,
                                            float qfz, float Ax, float Bnx, float Bfx, float Ay, float Bny,
                                            float Bfy, float Az, float Bnz, float Bfz, float& nearT) {
  const float nT = fmaxf(fmaxf(fmaf(qnx, Ax, Bnx), fmaf(qny, Ay, Bny)), fmaf(qnz, Az, Bnz));
  const float fT = fminf(fminf(fmaf(qfx, Ax, Bfx), fmaf(qfy, Ay, Bfy)), fmaf(qfz, Az, Bfz));
  nearT = nT;
  return nT <= fT && r.tmin <= fT && r.tmax >= nT;
}
__device__ __forceinline__ float qbyte(uint32_t w, int j) { return (float)((w >> (8 * j)) & 0xFFu); }

// Box tests of the four children of a compressed node (header hd, bounds q0, q1): hit flags and
// entry distances. zdir: the ray has a zero direction component (or an overflowed reciprocal) and
// takes the reference's per-axis slab test on the decoded boxes.
__device__ __forceinline__ void node_children(const uint4& hd, const uint4& q0, const uint4& q1, const Ray& r,
                                              bool zdir, bool& h0, bool& h1, bool& h2, bool& h3, float& t0,
                                              float& t1, float& t2, float& t3) {
  const float ox = __uint_as_float(hd.x), oy = __uint_as_float(hd.y), oz = __uint_as_float(hd.z);
  const float sx = __uint_as_float((hd.w & 0xFFu) << 23), sy = __uint_as_float(((hd.w >> 8) & 0xFFu) << 23),
              sz = __uint_as_float(((hd.w >> 16) & 0xFFu) << 23);
  if (!zdir) {
    float Ax, Bnx, Bfx, Ay, Bny, Bfy, Az, Bnz, Bfz;
    uint32_t nx, fx, ny, fy, nz, fz;
    fused_axis(ox, sx, r.o.x, r.inv.x, q0.x, q0.w, Ax, Bnx, Bfx, nx, fx);
    fused_axis(oy, sy, r.o.y, r.inv.y, q0.y, q1.x, Ay, Bny, Bfy, ny, fy);
    fused_axis(oz, sz, r.o.z, r.inv.z, q0.z, q1.y, Az, Bnz, Bfz, nz, fz);
#define RTG_CHILD(j, h, t)                                                                               \
  h = fused_child(r, qbyte(nx, j), qbyte(fx, j), qbyte(ny, j), qbyte(fy, j), qbyte(nz, j), qbyte(fz, j), Ax, \
                  Bnx, Bfx, Ay, Bny, Bfy, Az, Bnz, Bfz, t)
    RTG_CHILD(0, h0, t0);
    RTG_CHILD(1, h1, t1);
    RTG_CHILD(2, h2, t2);
    RTG_CHILD(3, h3, t3);
#undef RTG_CHILD
  } else {
    const float4 lx = dq4(q0.x, ox, sx), ly = dq4(q0.y, oy, sy), lz = dq4(q0.z, oz, sz);
    const float4 hx = dq4(q0.w, ox, sx), hy = dq4(q1.x, oy, sy), hz = dq4(q1.y, oz, sz);
    h0 = slab(r, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, t0);
    h1 = slab(r, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, t1);
    h2 = slab(r, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, t2);
    h3 = slab(r, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, t3);
  }
}

template <bool ANY>
__device__ __forceinline__ void trav_node(const SceneGeo& G, TravState& s, const TravStack& stk,
                                          uint32_t& nodes_visited) {
  // 64-byte compressed node: origin + step exponents, 8-bit child bounds, child refs
  const uint4* nd = reinterpret_cast<const uint4*>(G.nodes) + 4u * s.ref;
  const uint4 hd = nd[0], q0 = nd[1], q1 = nd[2], rf = nd[3];
  nodes_visited++;
  float t0, t1, t2, t3;
  bool h0, h1, h2, h3;
  node_children(hd, q0, q1, s.r, s.zdir, h0, h1, h2, h3, t0, t1, t2, t3);
  h0 = h0 && rf.x != kEmptyChild;
  h1 = h1 && rf.y != kEmptyChild;
  h2 = h2 && rf.z != kEmptyChild;
  h3 = h3 && rf.w != kEmptyChild;
  const int nh = (int)h0 + (int)h1 + (int)h2 + (int)h3;
  if (nh == 0) {
    s.ref = trav_pop(stk, s);
    return;
  }
  if (ANY) {
    // Any hit: the ray interval never shrinks, so the visiting order changes neither the result nor
    // the work much; no sort. Descend into the first hit child in slot order, push the others.
    // (entry distances as the sorted visit keeps them: infinite or NaN ones as the largest float)
    const bool f1 = h1 && !h0, f2 = h2 && !h0 && !h1, f3 = h3 && !h0 && !h1 && !h2;
    const float k1 = t1 < INFINITY ? t1 : 3.0e38f, k2 = t2 < INFINITY ? t2 : 3.0e38f,
                k3 = t3 < INFINITY ? t3 : 3.0e38f;
    const uint32_t next = h0 ? rf.x : (h1 ? rf.y : (h2 ? rf.z : rf.w));
    if (s.sp + 3 <= kLdsStack) {
      if (h1 && !f1) { stk.l[s.sp * stk.lstride] = make_uint2(rf.y, __float_as_uint(k1)); s.sp++; }
      if (h2 && !f2) { stk.l[s.sp * stk.lstride] = make_uint2(rf.z, __float_as_uint(k2)); s.sp++; }
      if (h3 && !f3) { stk.l[s.sp * stk.lstride] = make_uint2(rf.w, __float_as_uint(k3)); s.sp++; }
    } else {
      if (h1 && !f1) stk_push(stk, s.sp, rf.y, k1);
      if (h2 && !f2) stk_push(stk, s.sp, rf.z, k2);
      if (h3 && !f3) stk_push(stk, s.sp, rf.w, k3);
    }
    s.ref = next;
    return;
  }
  // Misses sort last (key +inf); a hit's key is always below +inf (a NaN or infinite entry
  // distance sorts as the largest finite float), so the first nh children after the sort are
  // exactly the hit ones.
  auto key = [](bool h, float t) { return h ? (t < INFINITY ? t : 3.0e38f) : INFINITY; };
  float k0 = key(h0, t0), k1 = key(h1, t1), k2 = key(h2, t2), k3 = key(h3, t3);
  uint32_t c0 = rf.x, c1 = rf.y, c2 = rf.z, c3 = rf.w;
#define RTG_CSWAP(ka, ca, kb, cb)          \
  {                                        \
    const bool sw = kb < ka;               \
    const float kt = sw ? kb : ka;         \
    kb = sw ? ka : kb;                     \
    ka = kt;                               \
    const uint32_t ct = sw ? cb : ca;      \
    cb = sw ? ca : cb;                     \
    ca = ct;                               \
  }
  RTG_CSWAP(k0, c0, k1, c1);
  RTG_CSWAP(k2, c2, k3, c3);
  RTG_CSWAP(k0, c0, k2, c2);
  RTG_CSWAP(k1, c1, k3, c3);
  RTG_CSWAP(k1, c1, k2, c2);
#undef RTG_CSWAP
  if (nh > 1) stk_push_children(stk, s.sp, nh, c1, k1, c2, k2, c3, k3);
  s.ref = c0;
}

// Tests the primitives of leaf `ref`; true when ANY found an accepted primitive.
template <bool ANY>
__device__ __forceinline__ bool trav_leaf(const SceneGeo& G, TravState& s, uint32_t ref, uint32_t& prims_tested) {
  const uint32_t first = leaf_first(ref), cnt = leaf_count(ref);
  prims_tested += cnt;
  if ((int32_t)first < G.ntri) {
    for (uint32_t i = 0; i < cnt; i++) {
      float t, b1, b2;
      if (tri_hit(s.r, G.tri_v + 9u * (first + i), t, b1, b2)) {
        s.h.t = t; s.h.b1 = b1; s.h.b2 = b2; s.h.prim = (int32_t)(first + i);
        s.hit = true;
        if (ANY) return true;
        s.r.tmax = t;
      }
    }
  } else {
    for (uint32_t i = 0; i < cnt; i++) {
      float t;
      uint32_t si = first + i - (uint32_t)G.ntri;
      if (sphere_hit(s.r, G.spheres + 5u * si, t)) {
        s.h.t = t; s.h.b1 = 0.f; s.h.b2 = 0.f; s.h.prim = (int32_t)(first + i);
        s.hit = true;
        if (ANY) return true;
        s.r.tmax = t;
      }
    }
  }
  return false;
}

// One step: an interior node, or a leaf followed by the pops up to the next subtree that can
// still hold a closer hit. True when the traversal has finished; the result is in s.hit / s.h.
// (Speculative while-while traversal, where lanes that reached a leaf keep visiting nodes until
// every lane of the wave has one, measured slower here: it visits more nodes, and node fetches
// are what bounds this traversal.)
template <bool ANY>
__device__ __forceinline__ bool trav_step(const SceneGeo& G, TravState& s, const TravStack& stk,
                                          uint32_t& nodes_visited, uint32_t& prims_tested) {
  if (!(s.ref & kLeafBit)) {
    trav_node<ANY>(G, s, stk, nodes_visited);
  } else {
    if (trav_leaf<ANY>(G, s, s.ref, prims_tested)) return true;
    s.ref = trav_pop(stk, s);
  }
  return s.ref == kTravDone;
}

template <bool ANY>
__device__ __forceinline__ bool traverse(const DevScene& S, Ray r, const TravStack& stk, HitRec& h,
                                         uint32_t& nodes_visited, uint32_t& prims_tested) {
  TravState s;
  s.r = r;
  if (trav_begin(S, s)) {
    const SceneGeo G = scene_geo(S);
    while (!trav_step<ANY>(G, s, stk, nodes_visited, prims_tested)) {
    }
  }
  h = s.h;
  return s.hit;
}

// Surface data at a hit (Triangle.cpp:130-215, Sphere.cpp:115-139, Primitive.cpp:28)
struct Surface {
  f3 p;
  Frame sh;
  uint32_t material;
  uint32_t light;
};
__device__ __forceinline__ Surface surface_at(const DevScene& S, const HitRec& h, f3 ro, f3 rd) {
  Surface s;
  uint4 info = S.prim_info[h.prim];
  s.material = info.y;
  s.light = info.z;
  if (h.prim < S.ntri) {
    const float* v = S.tri_v + 9u * (uint32_t)h.prim;
    f3 p0 = mk(v[0], v[1], v[2]), p1 = mk(v[3], v[4], v[5]), p2 = mk(v[6], v[7], v[8]);
    float b0 = 1.f - (h.b1 + h.b2);
    s.p = b0 * p0 + h.b1 * p1 + h.b2 * p2;
    f3 n;
    if (S.tri_has_n && S.tri_has_n[h.prim]) {
      const int32_t* vi = S.tri_vidx + 3u * (uint32_t)h.prim;
      f3 n0 = ld3(S.vnormals + 3 * vi[0]), n1 = ld3(S.vnormals + 3 * vi[1]), n2 = ld3(S.vnormals + 3 * vi[2]);
      n = normalized(b0 * n0 + h.b1 * n1 + h.b2 * n2);
    } else {
      n = normalized(cross(p1 - p0, p2 - p0));
    }
    s.sh = make_frame(n);
  } else {
    const float4* sph = S.spheres + 5u * (uint32_t)(h.prim - S.ntri);
    s.p = ro + h.t * rd;
    f3 local = xform_point(sph, s.p);
    f3 nl = normalized(local);
    s.sh = make_frame(normalized(xform_normal(sph, nl)));
  }
  return s;
}

// ---------------------------------------------------------------- microfacets (MicrofacetDistribution.cpp)
__device__ __forceinline__ float tan2(f3 v) {
  float s2 = 1.f - v.z * v.z;
  if (s2 <= 0.f) return 0.f;
  return s2 / (v.z * v.z);
}
__device__ __forceinline__ f3 mf_sample(int dist, float ux, float uy, float alpha) {
  float theta;
  if (dist == 0)
    theta = atanf(sqrtf(-alpha * alpha * logf(1.f - ux)));
  else
    theta = atanf(alpha * sqrtf(ux) / sqrtf(1.f - ux));
  float phi = 2.f * PI * uy;
  float st = sinf(theta);
  return mk(st * cosf(phi), st * sinf(phi), cosf(theta));
}
__device__ __forceinline__ float mf_D(int dist, f3 wh, float alpha) {
  if (wh.z <= 0.f) return 0.f;
  if (dist == 0) {
    float c2 = wh.z * wh.z;
    float t2 = (1.f - wh.z * wh.z) / c2;
    float a2 = alpha * alpha;
    return expf(-t2 / a2) / (PI * a2 * c2 * c2);
  }
  float a2 = alpha * alpha;
  float c2 = wh.z * wh.z;
  float t2 = tan2(wh);
  float ss = a2 + t2;
  return a2 / (PI * c2 * c2 * ss * ss);
}
__device__ __forceinline__ float mf_G1(int dist, f3 v, f3 m, float alpha) {
  if (dot(v, m) / v.z <= 0.f) return 0.f;
  if (dist == 0) {
    float s2 = 1.f - v.z * v.z;
    float st = s2 < 0.f ? 0.f : sqrtf(s2);
    float tv = st / v.z;
    if (tv == 0.f) return 1.f;
    float b = fabsf(1.f / (alpha * tv));
    if (b < 1.6f) return (3.535f * b + 2.181f * b * b) / (1.f + 2.276f * b + 2.577f * b * b);
    return 1.f;
  }
  float a2 = alpha * alpha;
  float t2 = tan2(v);
  return 2.f / (1.f + sqrtf(1.f + a2 * t2));
}
__device__ __forceinline__ float mf_G(int dist, f3 wi, f3 wo, f3 wh, float alpha) {
  return mf_G1(dist, wi, wh, alpha) * mf_G1(dist, wo, wh, alpha);
}
__device__ __forceinline__ float adj_alpha(float alpha, float cosI) {
  return (1.2f - 0.2f * sqrtf(fabsf(cosI))) * alpha;
}

// ---------------------------------------------------------------- Fresnel (core/Fresnel.h)
__device__ __forceinline__ f3 fresnel_conductor(f3 eta, f3 k, float cosI) {
  f3 epa = cmul(eta, eta) + cmul(k, k);
  float c2 = cosI * cosI;
  f3 epac2 = epa * c2;
  f3 e2c = (2.f * eta) * cosI;
  f3 rpar = cdiv(cadd(epac2 - e2c, 1.f), cadd(epac2 + e2c, 1.f));
  f3 rperp = cdiv(cadd(epa - e2c, c2), cadd(epa + e2c, c2));
  return (cmul(rpar, rpar) + cmul(rperp, rperp)) * 0.5f;
}
__device__ __forceinline__ float fresnel_dielectric(float etaI, float etaT, float cosI, float cosT) {
  float a = etaT * cosI, b = etaI * cosT, c = etaI * cosI, d = etaT * cosT;
  float rpar = (a - b) / (a + b);
  float rperp = (c - d) / (c + d);
  return (rpar * rpar + rperp * rperp) * 0.5f;
}
__device__ __forceinline__ float est_dielectric(float etaExt, float etaInt, float cosI, float& etaI, float& etaT,
                                                float& rel, float& cosT) {
  etaI = etaExt;
  etaT = etaInt;
  bool entering = true;
  float fr;
  if (cosI <= 0.f) {
    entering = false;
    float x = etaI; etaI = etaT; etaT = x;
  }
  rel = etaI / etaT;
  float sinT = rel * sqrtf(smax(0.f, 1.f - cosI * cosI));
  cosT = sqrtf(smax(0.f, 1.f - sinT * sinT));
  if (sinT >= 1.f) {
    cosT = 0.f;
    fr = 1.f;
  } else {
    fr = fresnel_dielectric(etaI, etaT, fabsf(cosI), cosT);
    if (entering) cosT = -cosT;
  }
  return fr;
}
__device__ __forceinline__ float est_dielectric(float etaExt, float etaInt, float cosI) {
  float a, b, c, d;
  return est_dielectric(etaExt, etaInt, cosI, a, b, c, d);
}

// ---------------------------------------------------------------- BSDFs
struct BsdfSample {
  f3 wo;
  f3 f;        // f * |cos| / pdf
  float pdf;
  float rel_eta;
};
__device__ __forceinline__ f3 mat_col(const float* c) { return mk(c[0], c[1], c[2]); }
__device__ __forceinline__ f3 reflect_n(f3 v, f3 n) { return (2.f * dot(v, n)) * n - v; }
__device__ __forceinline__ float sgn(float v) { return v >= 0.f ? 1.f : -1.f; }

// bsdf/Diffuse.cpp:28-106
__device__ __forceinline__ f3 diffuse_eval(const DevMaterial& m, f3 wi, f3 wo) {
  if (wi.z <= 0.f || wo.z <= 0.f) return mk(0.f, 0.f, 0.f);
  return mat_col(m.albedo) * INV_PI;
}
__device__ __forceinline__ float diffuse_pdf(f3 wi, f3 wo) {
  if (wi.z <= 0.f || wo.z <= 0.f) return 0.f;
  return INV_PI * wo.z;
}
__device__ __forceinline__ f3 concentric_disk_cosine(float ux, float uy) {  // Mapping.h:62-67,110-140
  const float r1 = 2.f * ux - 1.f;
  const float r2 = 2.f * uy - 1.f;
  float rx, ry;
  if (r1 > -r2 && r1 > r2) {
    rx = r1; ry = (PI * r2) / (4.f * r1);
  } else if (r1 < r2 && r1 > -r2) {
    rx = r2; ry = (PI / 4.f) * (2.f - r1 / r2);
  } else if (r1 < -r2 && r1 < r2) {
    rx = -r1; ry = (PI / 4.f) * (4.f + r2 / r1);
  } else {
    rx = -r2; ry = (PI / 4.f) * (6.f - r1 / r2);
  }
  float dx = rx * cosf(ry), dy = rx * sinf(ry);
  float z = sqrtf(smax(0.f, 1.f - dx * dx - dy * dy));
  return mk(dx, dy, z);
}
__device__ __forceinline__ bool diffuse_sample(const DevMaterial& m, f3 wi, float ux, float uy, BsdfSample& bs) {
  if (wi.z <= 0.f) return false;
  bs.wo = concentric_disk_cosine(ux, uy);
  bs.pdf = diffuse_pdf(wi, bs.wo);
  bs.rel_eta = 1.f;
  bs.f = mat_col(m.albedo);
  return true;
}

// bsdf/RoughConductor.cpp:77-145
__device__ __forceinline__ f3 conductor_eval(const DevMaterial& m, f3 wi, f3 wo) {
  float ci = wi.z, co = wo.z;
  if (ci <= 0.f || co <= 0.f) return mk(0.f, 0.f, 0.f);
  f3 wh = normalized(wi + wo);
  float D = mf_D(m.dist, wh, m.alpha);
  float G = mf_G(m.dist, wi, wo, wh, m.alpha);
  f3 F = fresnel_conductor(mat_col(m.eta), mat_col(m.k), ci);
  return cmul(mat_col(m.albedo) * D, F) * G / (4.f * ci * co);
}
__device__ __forceinline__ float conductor_pdf(const DevMaterial& m, f3 wi, f3 wo) {
  float ci = wi.z, co = wo.z;
  if (ci <= 0.f || co <= 0.f) return 0.f;
  f3 wh = normalized(wi + wo);
  float Jh = 1.f / (4.f * fabsf(dot(wh, wo)));
  float D = mf_D(m.dist, wh, m.alpha);
  return D * wh.z * Jh;
}
__device__ __forceinline__ bool conductor_sample(const DevMaterial& m, f3 wi, float ux, float uy, BsdfSample& bs) {
  if (wi.z <= 0.f) return false;
  float a = adj_alpha(m.alpha, wi.z);
  f3 n = mf_sample(m.dist, ux, uy, a);
  bs.wo = reflect_n(wi, n);
  bs.rel_eta = 1.f;
  float co = bs.wo.z;
  if (co <= 0.f) return false;
  bs.pdf = conductor_pdf(m, wi, bs.wo);
  if (bs.pdf <= 0.f) return false;
  bs.f = conductor_eval(m, wi, bs.wo) * fabsf(co) / bs.pdf;
  return true;
}

// bsdf/RoughDielectric.cpp:83-329
__device__ __forceinline__ f3 half_t(f3 wi, f3 wo, float etaI, float etaT) {
  return normalized(-1.f * (etaI * wi + etaT * wo));
}
__device__ __forceinline__ f3 diel_eval_refl(const DevMaterial& m, f3 wi, f3 wo) {
  float ci = wi.z, co = wo.z;
  f3 wh = normalized(wi + wo);
  wh = sgn(ci) * wh;
  float F = est_dielectric(m.eta_ext, m.eta_int, dot(wi, wh));
  f3 r = mat_col(m.albedo) * mf_D(m.dist, wh, m.alpha);
  r = r * F;
  return r * mf_G(m.dist, wi, wo, wh, m.alpha) / (4.f * ci * co);
}
__device__ __forceinline__ f3 diel_eval_refr(const DevMaterial& m, f3 wi, f3 wo) {
  float ci = wi.z, co = wo.z;
  float etaI, etaT, rel, cosT;
  float fr = est_dielectric(m.eta_ext, m.eta_int, ci, etaI, etaT, rel, cosT);
  f3 wt = half_t(wi, wo, etaI, etaT);
  float left = dot(wi, wt) * dot(wo, wt) / (ci * co);
  f3 term = mat_col(m.trans) * left * etaT * etaT;
  term = term * ((-1.f * fr) + 1.f);
  term = term * mf_G(m.dist, wi, wo, wt, m.alpha) * mf_D(m.dist, wt, m.alpha) * rel * rel;
  float denom = (etaI * (dot(wi, wt)) + etaT * dot(wo, wt));
  return term / (denom * denom);
}
__device__ __forceinline__ f3 dielectric_eval(const DevMaterial& m, f3 wi, f3 wo) {
  return (wi.z * wo.z > 0.f) ? diel_eval_refl(m, wi, wo) : diel_eval_refr(m, wi, wo);
}
__device__ __forceinline__ float diel_pdf_refr(const DevMaterial& m, f3 wi, f3 wo) {
  float ci = wi.z;
  float etaI, etaT, rel, cosT;
  est_dielectric(m.eta_ext, m.eta_int, ci, etaI, etaT, rel, cosT);
  f3 wt = half_t(wi, wo, etaI, etaT);
  float a = (1.2f - 0.2f * sqrtf(fabsf(ci))) * m.alpha;
  float hd = (etaI * dot(wi, wt) + etaT * dot(wo, wt));
  float Jh = (etaT * etaT * dot(wo, wt)) / (hd * hd);
  return fabsf(mf_D(m.dist, wt, a) * wt.z * Jh);
}
__device__ __forceinline__ float diel_pdf_refl(const DevMaterial& m, f3 wi, f3 wo) {
  float ci = wi.z;
  float a = (1.2f - 0.2f * sqrtf(fabsf(ci))) * m.alpha;
  f3 wh = normalized(wi + wo);
  wh = sgn(ci) * wh;
  float Jh = 1.f / (4.f * dot(wh, wo));
  return fabsf(mf_D(m.dist, wh, a) * wh.z * Jh);
}
__device__ __forceinline__ float dielectric_pdf(const DevMaterial& m, f3 wi, f3 wo) {
  float ci = wi.z, co = wo.z;
  if (ci * co > 0.f) {
    f3 wh = normalized(wi + wo);
    wh = sgn(ci) * wh;
    float fr = average(est_dielectric(m.eta_ext, m.eta_int, dot(wi, wh)));
    return diel_pdf_refl(m, wi, wo) * fr;
  }
  float etaI = m.eta_ext, etaT = m.eta_int;
  if (ci < 0.f) { float x = etaI; etaI = etaT; etaT = x; }
  f3 wt = half_t(wi, wo, etaI, etaT);
  float fr = average(est_dielectric(m.eta_ext, m.eta_int, dot(wi, wt)));
  return diel_pdf_refr(m, wi, wo) * (1.f - fr);
}
// The sampler after its microfacet normal n is drawn (with alpha' = a from ux, uy); ux, which chose
// n's polar angle, also chooses reflection (ux <= Fresnel at wi.n) or refraction.
__device__ __forceinline__ bool dielectric_sample_n(const DevMaterial& m, f3 wi, float a, f3 n, float ux,
                                                    BsdfSample& bs) {
  float ci = wi.z;
  float etaI, etaT, rel, cosT;
  float fr = average(est_dielectric(m.eta_ext, m.eta_int, dot(wi, n), etaI, etaT, rel, cosT));
  bs.rel_eta = 1.f;
  if (ux <= fr) {
    bs.wo = reflect_n(wi, n);
    if (wi.z * bs.wo.z <= 0.f) return false;
    bs.pdf = diel_pdf_refl(m, wi, bs.wo) * fr;
    if (bs.pdf <= 0.f) return false;
    f3 wh = normalized(wi + bs.wo);
    wh = sgn(ci) * wh;
    float w = 1.f * dot(wi, n);
    w = w * mf_G(m.dist, wi, bs.wo, wh, a);
    w = w / (ci * n.z);
    bs.f = mk(w, w, w);
    return true;
  }
  bs.wo = n * (dot(wi, n) * rel + cosT) - wi * rel;
  bs.wo = normalized(bs.wo);
  bs.rel_eta = rel;
  float co = dot(bs.wo, n);
  if (wi.z * bs.wo.z >= 0.f) return false;
  bs.pdf = diel_pdf_refr(m, wi, bs.wo) * (1.f - fr);
  if (bs.pdf <= 0.f) return false;
  bs.f = diel_eval_refr(m, wi, bs.wo) * fabsf(co) / bs.pdf;
  return true;
}
__device__ __forceinline__ bool dielectric_sample(const DevMaterial& m, f3 wi, float ux, float uy, BsdfSample& bs) {
  float a = adj_alpha(m.alpha, wi.z);
  return dielectric_sample_n(m, wi, a, mf_sample(m.dist, ux, uy, a), ux, bs);
}

// ---------------------------------------------------------------- rough-dielectric exit proposals
// (no counterpart in the reference; used by guided renders, rtg_render_desc.guiding 1 and 2)
// The sampler above is a map from the microfacet normal n (density D_a(n) n.z, a = alpha') to a
// weight w(n), so the plain estimator integrates w(n) D_a(n) n.z over normals. Where wi lies on the
// denser side and the macro-surface transmits, w is unbounded in two narrow bands of c = wi.n:
//   * near the critical angle, x = -c - cos(theta_c) -> 0+: refraction through n has 1 - F ~ sqrt(x)
//     in the pdf's denominator while the BSDF value carries the macro-surface Fresnel;
//   * at grazing front-facing normals, c -> 0+: the sampler takes n for an entering facet,
//     1 - F ~ c.
// Refraction there needs ux > F, i.e. steep normals where D_a is small, so these samples are rare
// and carry weights up to 10^3-10^6: the dominant fireflies of rough glass. In both bands the
// integrand w D_a n.z grows like 1 / sqrt(distance to the band's edge), so two proposals draw n
// around wi with densities of that shape (azimuth uniform, solid angle dc dphi):
//   critical: x = B1 s^2, density 1 / (2 sqrt(B1 x)) / 2pi on x in (0, B1];
//   grazing:  c = G1 s^2, density 1 / (2 sqrt(G1 c)) / 2pi on c in (0, G1].
// One-sample MIS (balance heuristic) in the space of normals over {sampler, critical, grazing}
// with probabilities {1 - P1 - P2, P1, P2}: the estimate for a normal n from any of them is
// w(n) D_a(n) n.z / sum_k P_k p_k(n) (at most 1 / (1 - P1 - P2) times the sampler's weight),
// whose expectation is the plain estimator's. A proposed normal goes through the sampler's own map
// (dielectric_sample_n with the ux that mf_sample would have needed for it), so every outcome,
// weight and reported pdf is one the sampler can produce. The bands' densities are evaluated with
// their edge distance floored at 1e-10 and 1e-12 (and extended over the float rounding of the
// edge: est_dielectric decides total internal reflection from sqrt(1 - c^2), not from x), so a
// normal the sampler transmits at the very edge never meets a zero band density.
constexpr float kExitP1 = 0.05f;   // probability of the critical-band proposal
constexpr float kExitP2 = 0.10f;   // probability of the grazing-band proposal
constexpr float kExitB1 = 0.1f;    // width of the critical band in x
constexpr float kExitG1 = 0.1f;    // width of the grazing band in c

// the first number mf_sample needs to produce a normal of this polar angle (its inverse)
__device__ __forceinline__ float mf_u_of(int dist, f3 n, float a) {
  const float t2 = tan2(n);
  const float u = dist == 0 ? -expm1f(-t2 / (a * a)) : t2 / (a * a + t2);
  return smin(u, ONE_MINUS_EPS);
}
// densities (solid angle of n) of the two band proposals around wi, for cos_c = cos(theta_c)
__device__ __forceinline__ float exit_band_density(float c, float cos_c, float& p2) {
  float x = -c - cos_c;
  if (x <= 1e-10f && x > -1e-5f) x = 1e-10f;
  const float p1 = (x > 0.f && x <= kExitB1) ? 1.f / (2.f * sqrtf(kExitB1 * x)) * (0.5f * INV_PI) : 0.f;
  p2 = (c > -1e-7f && c <= kExitG1) ? 1.f / (2.f * sqrtf(kExitG1 * smax(c, 1e-12f))) * (0.5f * INV_PI) : 0.f;
  return p1;
}
// us picks the technique; (ux, uy) drive it. factor (optional): the MIS factor applied to the
// sampler's weight, D_a(n) n.z / sum_k P_k p_k(n).
__device__ __forceinline__ bool dielectric_exit_mis(const DevMaterial& m, f3 wi, float us, float ux, float uy,
                                                    BsdfSample& bs, float* factor = nullptr) {
  const float ci = wi.z;
  const float eI = ci < 0.f ? m.eta_int : m.eta_ext, eT = ci < 0.f ? m.eta_ext : m.eta_int;
  const float a = adj_alpha(m.alpha, ci);
  if (factor) *factor = 1.f;
  // only where wi is on the denser side and the macro-surface transmits (elsewhere plain sampling)
  if (!(eI > eT) || est_dielectric(m.eta_ext, m.eta_int, ci) >= 1.f) return dielectric_sample(m, wi, ux, uy, bs);
  const float rr = eT / eI;
  const float cos_c = sqrtf(smax(0.f, 1.f - rr * rr));
  f3 n;
  float un = ux;
  if (us < kExitP1 + kExitP2) {
    float c;
    if (us < kExitP1) c = -smin(cos_c + kExitB1 * ux * ux, 1.f);
    else c = kExitG1 * ux * ux;
    const float sc = sqrtf(smax(0.f, 1.f - c * c));
    float sp, cp;
    sincosf(2.f * PI * uy, &sp, &cp);
    n = to_world(make_frame(wi), mk(sc * cp, sc * sp, c));
    if (!(n.z > 0.f)) return false;  // D_a(n) = 0: the integrand vanishes
    un = mf_u_of(m.dist, n, a);
  } else {
    n = mf_sample(m.dist, ux, uy, a);
  }
  if (!dielectric_sample_n(m, wi, a, n, un, bs)) return false;
  const float ps = mf_D(m.dist, n, a) * n.z;
  float p2;
  const float p1 = exit_band_density(dot(wi, n), cos_c, p2);
  const float den = (1.f - kExitP1 - kExitP2) * ps + kExitP1 * p1 + kExitP2 * p2;
  if (!(den > 0.f) || !(ps > 0.f)) return false;
  const float k = ps / den;
  bs.f = bs.f * k;
  if (factor) *factor = k;
  return true;
}

// bsdf/Conductor.cpp:50-64: mirror reflection, weight R * Fresnel(cos_i), pdf 1
__device__ __forceinline__ bool smooth_conductor_sample(const DevMaterial& m, f3 wi, BsdfSample& bs) {
  if (wi.z <= 0.f) return false;
  bs.wo = mk(-wi.x, -wi.y, wi.z);
  bs.pdf = 1.f;
  bs.rel_eta = 1.f;
  bs.f = cmul(mat_col(m.albedo), fresnel_conductor(mat_col(m.eta), mat_col(m.k), wi.z));
  return true;
}
// bsdf/Dielectric.cpp:37-95: reflect with probability fr (weight R), else refract without
// renormalising (weight T * eta_rel^2); the relative IOR is reported for both branches.
__device__ __forceinline__ bool smooth_dielectric_sample(const DevMaterial& m, f3 wi, float ux, BsdfSample& bs) {
  float etaI, etaT, rel, cosT;
  const float fr = est_dielectric(m.eta_ext, m.eta_int, wi.z, etaI, etaT, rel, cosT);
  bs.rel_eta = rel;
  if (ux <= fr) {
    bs.wo = mk(-wi.x, -wi.y, wi.z);
    bs.pdf = fr;
    bs.f = mat_col(m.albedo);
  } else {
    bs.wo = mk(rel * -wi.x, rel * -wi.y, cosT);
    bs.pdf = 1.f - fr;
    bs.f = mat_col(m.trans) * (rel * rel);
  }
  return true;
}
// Smooth conductors and dielectrics share the shade kernels (and template MT) of their rough
// kind; in the solid-angle measure their eval and pdf are 0 (Conductor.cpp:26-35, 68-77,
// Dielectric.cpp:28-35, 98-104), so next-event estimation from them contributes nothing.
__device__ __forceinline__ bool is_delta(const DevMaterial& m) { return m.type >= MAT_CONDUCTOR; }
template <int MT>
__device__ __forceinline__ f3 bsdf_eval(const DevMaterial& m, f3 wi, f3 wo) {
  if (MT != 0 && is_delta(m)) return mk(0.f, 0.f, 0.f);
  if (MT == 0) return diffuse_eval(m, wi, wo);
  if (MT == 1) return conductor_eval(m, wi, wo);
  return dielectric_eval(m, wi, wo);
}
template <int MT>
__device__ __forceinline__ float bsdf_pdf(const DevMaterial& m, f3 wi, f3 wo) {
  if (MT != 0 && is_delta(m)) return 0.f;
  if (MT == 0) return diffuse_pdf(wi, wo);
  if (MT == 1) return conductor_pdf(m, wi, wo);
  return dielectric_pdf(m, wi, wo);
}
template <int MT>
__device__ __forceinline__ bool bsdf_sample(const DevMaterial& m, f3 wi, float ux, float uy, BsdfSample& bs) {
  if (MT == 0) return diffuse_sample(m, wi, ux, uy, bs);
  if (MT == 1) return m.type == MAT_CONDUCTOR ? smooth_conductor_sample(m, wi, bs) : conductor_sample(m, wi, ux, uy, bs);
  return m.type == MAT_DIELECTRIC ? smooth_dielectric_sample(m, wi, ux, bs) : dielectric_sample(m, wi, ux, uy, bs);
}

// ---------------------------------------------------------------- guiding: one-sample MIS support
// Density with which bsdf_sample<MT>(m, wi, u) produces the direction wo (solid angle), together
// with the weight the sampler returns for it (f*cos/pdf), the pdf it reports (the value the MIS
// weights use) and the relative IOR. Defined for the diffuse and rough-conductor BSDFs; the
// conductor samples microfacet normals with density D(alpha')(m) cos(theta_m) (mf_sample) and
// reflects, so p = D(alpha')(wh) wh.z / (4 |wo.wh|), while its reported pdf uses alpha.
template <int MT>
__device__ __forceinline__ bool bsdf_density(const DevMaterial& m, f3 wi, f3 wo, f3& weight, float& ps, float& pref,
                                             float& rel) {
  rel = 1.f;
  if (wi.z <= 0.f || wo.z <= 0.f || is_delta(m)) return false;
  if (MT == 0) {
    ps = wo.z * INV_PI;
    pref = diffuse_pdf(wi, wo);
    weight = mat_col(m.albedo);
    return ps > 0.f;
  }
  f3 wh = normalized(wi + wo);
  float a = adj_alpha(m.alpha, wi.z);
  ps = mf_D(m.dist, wh, a) * wh.z / (4.f * fabsf(dot(wh, wo)));
  pref = conductor_pdf(m, wi, wo);
  if (!(ps > 0.f) || !(pref > 0.f)) return false;
  weight = conductor_eval(m, wi, wo) * fabsf(wo.z) / pref;
  return true;
}

// A direction from the vMF lobe (mu, kappa) with (u1, u2), by inverting the CDF of cos(theta)
// (Wenzel Jakob, "Numerically stable sampling of the von Mises Fisher distribution on S2", 2012).
__device__ __forceinline__ f3 vmf_sample(f3 mu, float kappa, float u1, float u2) {
  float w = 1.f + logf(u1 + (1.f - u1) * expf(-2.f * kappa)) / kappa;
  w = fminf(fmaxf(w, -1.f), 1.f);
  const float sr = sqrtf(fmaxf(0.f, 1.f - w * w));
  float sp, cp;
  sincosf(2.f * PI * u2, &sp, &cp);
  Frame F = make_frame(mu);
  return to_world(F, mk(sr * cp, sr * sp, w));
}
// First lobe whose cumulative weight exceeds ul (the last one absorbs rounding); pi(i) gives weight i.
template <class Pi>
__device__ __forceinline__ int mix_select(Pi pi, float ul) {
  int sel = 7;
  float acc = 0.f;
#pragma unroll
  for (int i = 0; i < 7; i++) {
    acc += pi(i);
    if (sel == 7 && ul < acc) sel = i;
  }
  return sel;
}
// Mixtures of 8 von Mises-Fisher lobes, density sum_k pi_k kappa_k / (2 pi (1 - e^{-2 kappa_k}))
// e^{kappa_k (mu_k.w - 1)}. The network gives (pi, mu xyz, kappa) per lobe; the renderer keeps each
// guided vertex's mixture in this layout (kGuideMixR floats, 16-byte aligned): pi[8],
// c[8] = pi * kappa / (2 pi (1 - e^{-2 kappa})), mu[8][3], kappa[8]. The normalisers are formed once
// when the network writes the mixture, so a density costs one exponential per lobe, and the fields
// load as whole float4s. mixr_store_half writes lobes 4h..4h+3 (m: their (pi, mu xyz, kappa)).
__device__ __forceinline__ void mixr_store_half(float* mr, int h, const float m[20]) {
  float4* d4 = reinterpret_cast<float4*>(mr);
  float c[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const float ka = m[5 * j + 4];
    c[j] = m[5 * j] * (ka / (2.f * PI * (-expm1f(-2.f * ka))));
  }
  d4[h] = make_float4(m[0], m[5], m[10], m[15]);
  d4[2 + h] = make_float4(c[0], c[1], c[2], c[3]);
  d4[4 + 3 * h] = make_float4(m[1], m[2], m[3], m[6]);
  d4[5 + 3 * h] = make_float4(m[7], m[8], m[11], m[12]);
  d4[6 + 3 * h] = make_float4(m[13], m[16], m[17], m[18]);
  d4[10 + h] = make_float4(m[4], m[9], m[14], m[19]);
}
__device__ __forceinline__ float mixr_pdf(const float* mx, f3 w) {
  const float4* m4 = reinterpret_cast<const float4*>(mx);
  const float4 c0 = m4[2], c1 = m4[3];
  const float c[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
  float mu[24];
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const float4 t = m4[4 + q];
    mu[4 * q] = t.x; mu[4 * q + 1] = t.y; mu[4 * q + 2] = t.z; mu[4 * q + 3] = t.w;
  }
  const float4 k0 = m4[10], k1 = m4[11];
  const float ka[8] = {k0.x, k0.y, k0.z, k0.w, k1.x, k1.y, k1.z, k1.w};
  float p = 0.f;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const float d = mu[3 * k] * w.x + mu[3 * k + 1] * w.y + mu[3 * k + 2] * w.z;
    p += c[k] * expf(ka[k] * (d - 1.f));
  }
  return p;
}
// Sample a lobe with ul, then its vMF with (u1, u2).
__device__ __forceinline__ f3 mixr_sample(const float* mx, float ul, float u1, float u2) {
  const float4* m4 = reinterpret_cast<const float4*>(mx);
  const float4 p0 = m4[0], p1 = m4[1];
  const float pi[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
  const int sel = mix_select([&](int i) { return pi[i]; }, ul);
  const float* mu = mx + 16 + 3 * sel;
  return vmf_sample(mk(mu[0], mu[1], mu[2]), mx[40 + sel], u1, u2);
}

// ---------------------------------------------------------------- CDF (core/CDF.h:44-81)
__device__ __forceinline__ int cdf_sample(const float* c, int n /* entries incl. leading 0 */, float s) {
  int lo = 0, hi = n;  // lower_bound
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (c[mid] < s) lo = mid + 1; else hi = mid;
  }
  int index = lo - 1;
  if (index < 0) index = 0;
  if (index > n - 2) index = n - 2;
  return index;
}
// cdf_sample with the CDF's cut points (kCdfGuide + 1 entries, rtg_host.cpp cdf_guide): the
// lower bound of s lies between guide[k] and guide[k + 1] for k = floor(s kCdfGuide) (exact: a
// power-of-two scale), so the binary search covers that range only and returns the same index.
__device__ __forceinline__ int cdf_sample_guided(const float* c, int n, float s, const int32_t* guide) {
  const int k = min(max((int)(s * (float)kCdfGuide), 0), kCdfGuide - 1);
  int lo = guide[k], hi = guide[k + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (c[mid] < s) lo = mid + 1; else hi = mid;
  }
  int index = lo - 1;
  if (index < 0) index = 0;
  if (index > n - 2) index = n - 2;
  return index;
}
__device__ __forceinline__ int cdf_sample_reuse(const float* c, int n, float& s, float& pdf) {
  int i = cdf_sample(c, n, s);
  s = (s - c[i]) / (c[i + 1] - c[i]);
  pdf = c[i + 1] - c[i];
  return i;
}

// ---------------------------------------------------------------- lights
struct LightSample {
  f3 dir;
  float dist;
  f3 intensity;
  float pdf;
  f3 point;   // sampled point (environment: at the bounding-sphere radius)
  f3 normal;  // its normal (environment: zero, the default of LightSamplingInfos)
};
// light/MeshLight.cpp:32-64 with geometricShape/Mesh.cpp:38-68, Mapping.h:94-105
__device__ __forceinline__ LightSample mesh_light_sample(const DevScene& S, const DevLight& L, f3 pFrom, float ux,
                                                         float uy) {
  LightSample ls;
  float dummy;
  float sx = ux;
  int idx = cdf_sample_reuse(S.emit_cdf + L.cdf_offset, L.emit_count + 1, sx, dummy);
  const float* v = S.emit_tri + 9u * (uint32_t)(L.emit_offset + idx);
  f3 p0 = mk(v[0], v[1], v[2]), p1 = mk(v[3], v[4], v[5]), p2 = mk(v[6], v[7], v[8]);
  float sq = sqrtf(sx);
  float al = 1.f - sq, be = (1.f - uy) * sq, ga = uy * sq;
  f3 pt = p0 * al + p1 * be + p2 * ga;
  f3 n;
  if (L.has_normals) {
    const float* nn = S.emit_nrm + 9u * (uint32_t)(L.emit_offset + idx);
    n = normalized(al * mk(nn[0], nn[1], nn[2]) + be * mk(nn[3], nn[4], nn[5]) + ga * mk(nn[6], nn[7], nn[8]));
  } else {
    n = normalized(cross(p1 - p0, p2 - p0));
    n = normalized(n);
  }
  ls.point = pt;
  ls.normal = n;
  ls.dir = pt - pFrom;
  ls.dist = norm(ls.dir);
  ls.dir = ls.dir / ls.dist;
  float cosine = dot(n, -ls.dir);
  if (cosine <= 0.f) {
    ls.intensity = mk(0.f, 0.f, 0.f);
    ls.pdf = 0.f;
  } else {
    ls.intensity = mat_col(L.radiance);
    ls.pdf = ls.dist * ls.dist / (L.area * cosine);
  }
  return ls;
}
// geometricShape/Mesh.cpp:70-93
__device__ __forceinline__ float mesh_light_pdf(const DevLight& L, f3 pFrom, f3 pt, f3 n) {
  f3 v = pFrom - pt;
  float d = norm(v);
  v = normalized(v);
  float c = dot(n, v);
  if (c > 0.f) return d * d / (L.area * c);
  return 0.f;
}
// geometricShape/Geometry.h:163-190
__device__ __forceinline__ float sph_phi(f3 v) {
  float r = atan2f(v.x, -v.z);
  return r < 0.f ? r + 2.f * 3.14159265358979323846f : r;
}
__device__ __forceinline__ float sph_theta(f3 v) { return acosf(v.y); }
// light/EnvironmentLight.cpp:176-186
__device__ __forceinline__ void check_xy(const DevScene& S, float& x, float& y) {
  int W = S.env_w, H = S.env_h;
  if (x >= (float)W) x = (float)((int)x % W) + (x - (float)(int)x);
  else if (x < 0.f) x = (float)(W - abs((int)x) % W) + (x - (float)(int)x);
  if (y >= (float)H) y = (float)((int)y % H) + (y - (float)(int)y);
  else if (y < 0.f) y = (float)(H - abs((int)y) % H) + (y - (float)(int)y);
}
// core/BasicMath.h:74-90 on an H x W RGB array
__device__ __forceinline__ f3 interp2_map(const float* m, int W, int H, float x, float y) {
  float ax = x - (float)(int)x;
  float ay = y - (float)(int)y;
  int x0 = min(max((int)x, 0), W - 2);
  int y0 = min(max((int)y, 0), H - 2);
  int x1 = x0 + 1 < W ? x0 + 1 : x0;
  x1 = min(max(x1, 0), W - 1);
  int y1 = y0 + 1 < H ? y0 + 1 : y0;
  y1 = min(max(y1, 0), H - 1);
  f3 A = ld3(m + 3 * (y0 * W + x0)), B = ld3(m + 3 * (y1 * W + x0));
  f3 C = ld3(m + 3 * (y0 * W + x1)), D = ld3(m + 3 * (y1 * W + x1));
  return ((1.f - ax) * (1.f - ay)) * A + ((1.f - ax) * ay) * B + (ax * (1.f - ay)) * C + (ax * ay) * D;
}
// core/BasicMath.h:74-90
__device__ __forceinline__ f3 interp2(const DevScene& S, float x, float y) {
  int W = S.env_w, H = S.env_h;
  float ax = x - (float)(int)x;
  float ay = y - (float)(int)y;
  int x0 = min(max((int)x, 0), W - 2);
  int y0 = min(max((int)y, 0), H - 2);
  int x1 = x0 + 1 < W ? x0 + 1 : x0;
  x1 = min(max(x1, 0), W - 1);
  int y1 = y0 + 1 < H ? y0 + 1 : y0;
  y1 = min(max(y1, 0), H - 1);
  const float* m = S.env_map;
  f3 A = ld3(m + 3 * (y0 * W + x0)), B = ld3(m + 3 * (y1 * W + x0));
  f3 C = ld3(m + 3 * (y0 * W + x1)), D = ld3(m + 3 * (y1 * W + x1));
  return ((1.f - ax) * (1.f - ay)) * A + ((1.f - ax) * ay) * B + (ax * (1.f - ay)) * C + (ax * ay) * D;
}
__device__ __forceinline__ void env_coords(const DevScene& S, f3 d, float& fx, float& fy) {
  f3 dir = normalized(d);
  fx = (sph_phi(dir) / (2.f * PI)) * (float)S.env_w;
  fy = (sph_theta(dir) / PI) * (float)S.env_h;
  check_xy(S, fx, fy);
}
// light/EnvironmentLight.cpp:137-166
__device__ __forceinline__ f3 env_le(const DevScene& S, f3 d) {
  float fx, fy;
  env_coords(S, d, fx, fy);
  return interp2(S, fx, fy) * S.env_factor;
}
// light/EnvironmentLight.cpp:105-135 (sin of the texel coordinate, as in the reference)
__device__ __forceinline__ float env_pdf(const DevScene& S, f3 d) {
  float fx, fy;
  env_coords(S, d, fx, fy);
  float sinT = sinf(fy);
  int row = min(max((int)fy, 0), S.env_h - 1);
  float nf = S.env_weight[row] / (S.env_marg_sum * (2.f * PI / (float)S.env_w) * (PI / (float)S.env_h));
  float lum = luminance(interp2(S, fx, fy) * S.env_factor);
  float pdf = lum * nf;
  pdf /= smax(fabsf(sinT), EPS);
  return pdf;
}
// light/EnvironmentLight.cpp:56-103
__device__ __forceinline__ LightSample env_sample(const DevScene& S, f3 pFrom, float ux, float uy) {
  LightSample ls;
  int W = S.env_w, H = S.env_h;
  int row = cdf_sample_guided(S.env_marg_cdf, H + 1, ux, S.env_marg_guide);
  int col = cdf_sample_guided(S.env_row_cdf + (size_t)row * (W + 1), W + 1, uy,
                              S.env_row_guide + (size_t)row * (kCdfGuide + 1));
  float rcol = (float)col + 0.f * (float)W;
  float rrow = (float)row + 0.f * (float)H;
  check_xy(S, rcol, rrow);
  row = (int)rrow;
  col = (int)rcol;
  float theta = PI * (float)row / (float)H;
  float phi = 2.f * PI * (float)col / (float)W;
  float sinTheta = sinf(theta);
  float cosTheta = cosf(theta);
  float cosPhi = cosf(phi), sinPhi = sinf(phi);
  ls.dir = mk(sinPhi * sinTheta, cosTheta, -cosPhi * sinTheta);
  ls.dist = S.env_radius;
  ls.dir = normalized(ls.dir);
  ls.point = pFrom + S.env_radius * ls.dir;
  ls.normal = mk(0.f, 0.f, 0.f);
  int cc = min(max(col, 0), W - 1), rr = min(max(row, 0), H - 1);
  ls.intensity = ld3(S.env_map + 3 * (rr * W + cc)) * S.env_factor;
  ls.pdf = luminance(ls.intensity) * S.env_weight[row] /
           (S.env_marg_sum * (2.f * PI / (float)W) * (PI / (float)H));
  ls.pdf /= smax(fabsf(sinTheta), EPS);
  return ls;
}

// ---------------------------------------------------------------- image textures
// Texture coordinates of a hit: a triangle's are its barycentrics (b1, b2) unless its mesh has
// uvs, which are then interpolated (geometricShape/Triangle.cpp:141-171); a sphere's are the
// spherical angles of the normalized world-space hit point, phi / 2pi and theta / pi
// (Sphere.cpp:100-104: the world ray's point, not the point relative to the centre).
__device__ __forceinline__ void surface_uv(const DevScene& S, const HitRec& h, f3 ro, f3 rd, float& u, float& v) {
  if (h.prim < S.ntri) {
    u = h.b1;
    v = h.b2;
    if (S.tri_has_uv && S.tri_has_uv[h.prim]) {
      const int32_t* vi = S.tri_vidx + 3u * (uint32_t)h.prim;
      const float b0 = 1.f - (h.b1 + h.b2);
      const float* a = S.vuv + 2 * vi[0];
      const float* b = S.vuv + 2 * vi[1];
      const float* c = S.vuv + 2 * vi[2];
      u = b0 * a[0] + h.b1 * b[0] + h.b2 * c[0];
      v = b0 * a[1] + h.b1 * b[1] + h.b2 * c[1];
    }
  } else {
    const f3 n = normalized(ro + h.t * rd);
    u = sph_phi(n) / (2.f * PI);
    v = sph_theta(n) / PI;
  }
}
// ImageTexture::eval (texture/ImageTexture.cpp:27-64) times the ScaleTexture factor
__device__ __forceinline__ f3 tex_eval(const DevScene& S, int ti, float u, float v) {
  const DevTexture& T = S.textures[ti];
  const float W = (float)T.w, H = (float)T.h;
  float x, y;
  if (!T.invert_y) y = v * H;
  else y = (float)(int)v + (sgn(v) * 1.f - (v - (float)(int)v)) * H;
  if (!T.invert_x) x = u * (W - 1.f);
  else x = (float)(int)u + (sgn(u) * 1.f - (u - (float)(int)u)) * W;
  x *= T.uv_scale;
  y *= T.uv_scale;
  if (x >= W) x = (float)((int)x % T.w) + (x - (float)(int)x);
  else if (x < 0.f) x = (float)(T.w - abs((int)x) % T.w) + (x - (float)(int)x);
  if (y >= H) y = (float)((int)y % T.h) + (y - (float)(int)y);
  else if (y < 0.f) y = (float)(T.h - abs((int)y) % T.h) + (y - (float)(int)y);
  return interp2_map(S.texels + T.offset, T.w, T.h, x, y) * T.value_scale;
}
// The BSDF parameters at a hit: each textured one evaluated at the hit's uv (Diffuse.cpp:34,91,
// RoughConductor.cpp:87-90, RoughDielectric.cpp:108-140, Conductor.cpp:42, Dielectric.cpp); a
// roughness texture contributes its channel average (Color.h:151-158).
__device__ __forceinline__ void resolve_material(const DevScene& S, DevMaterial& M, const HitRec& h, f3 ro, f3 rd) {
  if (M.tex[0] < 0 && M.tex[1] < 0 && M.tex[2] < 0) return;
  float u, v;
  surface_uv(S, h, ro, rd, u, v);
  if (M.tex[0] >= 0) {
    const f3 c = tex_eval(S, M.tex[0], u, v);
    M.albedo[0] = c.x; M.albedo[1] = c.y; M.albedo[2] = c.z;
  }
  if (M.tex[1] >= 0) {
    const f3 c = tex_eval(S, M.tex[1], u, v);
    float a = 0.f;
    a += c.x; a += c.y; a += c.z;
    M.alpha = a / 3.f;
  }
  if (M.tex[2] >= 0) {
    const f3 c = tex_eval(S, M.tex[2], u, v);
    M.trans[0] = c.x; M.trans[1] = c.y; M.trans[2] = c.z;
  }
}

// ---------------------------------------------------------------- participating medium
// The camera medium: medium/Heterogeneous.cpp over volume/GridVolume.cpp with
// medium/HGPhaseFunction.cpp. A camera medium is never placed in an owner box, so its
// world-to-array transform is the identity and the grid spans [0,1]^3 in world space.

// volume/GridVolume.cpp:118-148 (x Heterogeneous.cpp:403-406 scale): trilinear, 0 outside the
// cells [0, res-1) on each axis
__device__ __forceinline__ float medium_sigma(const DevScene& S, f3 p) {
  const float qx = p.x * (float)S.gres[0], qy = p.y * (float)S.gres[1], qz = p.z * (float)S.gres[2];
  const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
  if (!(fx >= 0.f && fy >= 0.f && fz >= 0.f && fx < (float)(S.gres[0] - 1) && fy < (float)(S.gres[1] - 1) &&
        fz < (float)(S.gres[2] - 1)))
    return 0.f;
  const size_t row = (size_t)S.gres[0], slab = row * (size_t)S.gres[1];
  const float* g = S.grid + (size_t)fz * slab + (size_t)fy * row + (size_t)fx;
  const float d000 = g[0], d001 = g[1], d010 = g[row], d011 = g[row + 1];
  const float d100 = g[slab], d101 = g[slab + 1], d110 = g[slab + row], d111 = g[slab + row + 1];
  const float w1x = qx - fx, w1y = qy - fy, w1z = qz - fz;
  const float w0x = 1.f - w1x, w0y = 1.f - w1y, w0z = 1.f - w1z;
  const float v = ((d000 * w0x + d001 * w1x) * w0y + (d010 * w0x + d011 * w1x) * w1y) * w0z +
                  ((d100 * w0x + d101 * w1x) * w0y + (d110 * w0x + d111 * w1x) * w1y) * w1z;
  return v * S.gscale;
}

// core/BoundingBox.h:154-183 on the box [0,1]^3: unclipped entry/exit distances; true when
// [minT, maxT] reaches the box
__device__ __forceinline__ bool unit_box(f3 o, f3 d, float minT, float maxT, float& nearT, float& farT) {
  nearT = -MAX_REAL;
  farT = MAX_REAL;
  const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    if (dd[i] == 0.f) {
      if (oo[i] < 0.f || oo[i] > 1.f) return false;
    } else {
      const float inv = 1.f / dd[i];
      float t1 = (0.f - oo[i]) * inv, t2 = (1.f - oo[i]) * inv;
      if (t1 > t2) {
        const float x = t1;
        t1 = t2;
        t2 = x;
      }
      nearT = fmaxf(t1, nearT);
      farT = fminf(t2, farT);
      if (!(nearT <= farT)) return false;
    }
  }
  return minT <= farT && maxT >= nearT;
}

// Delta tracking (Heterogeneous.cpp:491-546): true on a scattering event at distance t. The
// reference starts its tentative collisions at minT; before the box entry the density is 0 and
// every one of them is rejected, so starting at max(minT, entry) draws the same distribution.
// At most kMediumSteps tentative collisions per call (the reference has no bound; the expected
// count is sigma_max times the length inside the box).
constexpr int kMediumSteps = 16384;
template <class U>
__device__ __forceinline__ bool medium_sample_distance(const DevScene& S, f3 o, f3 d, float minT, float maxT, U& u,
                                                       float& t) {
  float nearT, farT;
  if (!unit_box(o, d, minT, maxT, nearT, farT)) return false;
  const float tMax = fminf(maxT, farT);
  t = fmaxf(minT, nearT);
  for (int it = 0; it < kMediumSteps; it++) {
    t += -logf(1.f - u()) / S.sigma_max;
    if (t > tMax) return false;
    const float sigma = medium_sigma(S, o + t * d);
    if (u() < sigma / S.sigma_max) return sigma != 0.f;
  }
  return false;
}

// Ratio tracking with Russian roulette below 0.1 (Heterogeneous.cpp:419-489), including its
// segment length maxT - minT and its limit of 1025 steps
template <class U>
__device__ __forceinline__ float medium_transmittance(const DevScene& S, f3 o, f3 d, float minT, float maxT, U& u) {
  float tMax = maxT - minT;
  float nearT, farT;
  if (!unit_box(o, d, minT, maxT, nearT, farT)) return 1.f;
  tMax = fminf(tMax, farT);
  float s = fmaxf(nearT, minT);
  float tr = 1.f;
  for (int breaker = 0;; breaker++) {
    s += -logf(1.f - u()) / S.sigma_max;
    if (s >= tMax) return tr;
    const float sigma = medium_sigma(S, o + s * d);
    tr *= 1.f - fmaxf(0.f, sigma / S.sigma_max);
    if (tr < 0.1f) {
      const float q = fmaxf(0.05f, 1.f - tr);
      if (u() < q) return 0.f;
      tr /= 1.f - q;
    }
    if (breaker >= 1024) return 0.f;
  }
}

// The camera medium's two operations for the integrator, for either kind (RGB weights):
// Homogeneous (medium/Homogeneous.cpp:26-60) samples a channel uniformly, an exponential distance
// with that channel's extinction (+1e-4) measured from minT but compared with maxT as an absolute
// distance, and returns the spectral-MIS weight sigma_s T / avg(sigma_t T) on a collision,
// T / avg(T) otherwise (T = exp(-sigma_t t)); its transmittance is exp(-sigma_t (maxT - minT)),
// without bounds. Heterogeneous: delta tracking (weight = albedo on a collision, 1 otherwise) and
// ratio tracking as above.
template <class U>
__device__ __forceinline__ bool medium_distance(const DevScene& S, f3 o, f3 d, float minT, float maxT, U& u,
                                                float& t, f3& weight) {
  if (S.medium_kind == 2) {
    const float uc = u();
    const int ch = min((int)(uc * 3.f), 2);
    const float sig = (ch == 0 ? S.sig_t[0] : ch == 1 ? S.sig_t[1] : S.sig_t[2]) + EPS;
    t = -logf(1.f - u()) / sig;
    const f3 st = mk(S.sig_t[0], S.sig_t[1], S.sig_t[2]);
    if (t > maxT) {
      t = maxT;
      const f3 e = mk(expf(-st.x * t), expf(-st.y * t), expf(-st.z * t));
      const float pdf = (e.x + e.y + e.z) / 3.f;
      const float len = maxT - minT;
      weight = mk(expf(-st.x * len), expf(-st.y * len), expf(-st.z * len)) / pdf;
      return false;
    }
    const f3 e = mk(expf(-st.x * t), expf(-st.y * t), expf(-st.z * t));
    const f3 se = cmul(st, e);
    const float pdf = (se.x + se.y + se.z) / 3.f;
    const float len = t;  // transmittance over [minT, minT + t]
    t += minT;
    weight = cmul(mk(S.sig_s[0], S.sig_s[1], S.sig_s[2]), mk(expf(-st.x * len), expf(-st.y * len), expf(-st.z * len))) /
             pdf;
    return true;
  }
  weight = mk(1.f, 1.f, 1.f);
  const bool sc = medium_sample_distance(S, o, d, minT, maxT, u, t);
  if (sc) weight = mk(S.albedo[0], S.albedo[1], S.albedo[2]);
  return sc;
}
template <class U>
__device__ __forceinline__ f3 medium_tr(const DevScene& S, f3 o, f3 d, float minT, float maxT, U& u) {
  if (S.medium_kind == 2) {
    const float len = maxT - minT;
    return mk(expf(-S.sig_t[0] * len), expf(-S.sig_t[1] * len), expf(-S.sig_t[2] * len));
  }
  const float tr = medium_transmittance(S, o, d, minT, maxT, u);
  return mk(tr, tr, tr);
}

// HGPhaseFunction.cpp:61-79: cos = (-wi).wo
__device__ __forceinline__ float hg_eval(float g, f3 wi, f3 wo) {
  const float c = dot(-wi, wo);
  return (0.0795774715459476678844f * (1.f - g * g)) / powf(1.f + g * g - 2.f * g * c, 1.5f);
}
// HGPhaseFunction.cpp:27-45 as written: the inversion divides (1 - g^2)^2 by (1 - g + 2 g u) once,
// and the direction is built around world +y (core/Geometry.h:181-188), not around wi; weight 1
__device__ __forceinline__ f3 hg_sample(float g, f3 wi, float ux, float uy, float& pdf) {
  float cosTheta;
  if (g < EPS)
    cosTheta = 1.f - 2.f * ux;
  else
    cosTheta = (1.f + g * g - ((1.f - g * g) * (1.f - g * g) / (1.f - g + 2.f * g * ux))) / (2.f * g);
  const float theta = acosf(cosTheta);
  const float phi = 2.f * PI * uy;
  const float sinT = sinf(theta), cosT = cosf(theta), cosP = cosf(phi), sinP = sinf(phi);
  const f3 wo = mk(sinP * sinT, cosT, -cosP * sinT);
  pdf = hg_eval(g, wi, wo);
  return wo;
}

// Mesh::pdf (geometricShape/Mesh.cpp:70-93) as VolPathTracing.cpp:152 calls it after a medium
// vertex: from the point where the sampled ray met the light to the point that light sampling
// drew at that vertex. On a planar emitter the two points are coplanar and the cosine is 0 in
// exact arithmetic (pdf 0, MIS weight 1); rounding makes the reference's sign arbitrary, so a
// cosine within 1e-4 of 0 is taken as 0 here.
__device__ __forceinline__ float medium_hit_light_pdf(const DevLight& L, f3 pFrom, f3 sampled, f3 n) {
  f3 v = pFrom - sampled;
  const float dist = norm(v);
  v = normalized(v);
  const float c = dot(n, v);
  if (c > 1e-4f) return dist * dist / (L.area * c);
  return 0.f;
}

}  // namespace dev
}  // namespace rtg

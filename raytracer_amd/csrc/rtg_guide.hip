// Path-guiding network on gfx950: inference and training with v_mfma_f32_32x32x16_bf16
// (bf16 operands, FP32 accumulation). No counterpart in the reference renderer.
//
// Network (rtg_internal.h): 144 sin/cos features of (position in the scene box, shading normal)
// -> 64 -> 64 -> 64 (ReLU) -> 40 = 8 vMF lobes x (logit, mean xyz, log kappa).
//
// One wave owns a tile of 32 rows (queries or training samples). Activations of a tile live in
// LDS, row-major with padded strides (152 and 72 bf16) so that a lane's 8 consecutive bf16 of a
// row form one 16-byte load. MFMA operand layout (checked on the device): lane l holds
// A[l%32][8*(l/32) + j] and B[8*(l/32) + j][l%32], j < 8; the accumulator holds
// D[8*(v/4) + 4*(l/32) + v%4][l%32], v < 16. Weights are read from global memory (21504 + 12288
// bf16, cache resident) in the forward layout W[out][in] (the B operand of X * W^T) and the
// backward layout W^T[in][out] (the B operand of dY * W).
//
// Training step (65536 samples): a weight-sum kernel, a fused forward + loss + backward-data kernel
// that writes every layer input and output gradient to HBM in a column-blocked layout
// ([row/8][column][8] bf16, so 8 rows of one column are one 16-byte load), a weight-gradient
// kernel (dW = dY^T X as MFMA over row chunks, FP32 atomics into the gradient), and Adam.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "rtg_device.h"
#include "rtg_internal.h"

namespace rtg {
using namespace dev;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kXS = 152;  // LDS row stride of the 144-feature tile (bf16)
constexpr int kHS = 72;   // LDS row stride of 64-column tiles (bf16)
constexpr int kOS = 41;   // LDS row stride of the FP32 output tile

__device__ __forceinline__ f32x16 mfma_bf16(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ uint16_t f2bf(float x) {
  __bf16 b = (__bf16)x;  // round to nearest even
  return __builtin_bit_cast(uint16_t, b);
}
__device__ __forceinline__ float bf2f(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }
__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }

// c0/c1 (output columns 0..31 / 32..63) += A (32 x 16*KS, LDS, row stride sa) * W^T, W = [64][16*KS]
// row-major in global memory (row stride sw).
template <int KS>
__device__ __forceinline__ void gemm_tile(const uint16_t* A, int sa, const uint16_t* W, int sw, f32x16& c0,
                                          f32x16& c1) {
  const int l = __lane_id(), r = l & 31, h = l >> 5;
  const uint16_t* ar = A + r * sa + h * 8;
  const uint16_t* w0 = W + r * sw + h * 8;
  const uint16_t* w1 = W + (32 + r) * sw + h * 8;
#pragma unroll
  for (int ks = 0; ks < KS; ks++) {
    bf16x8 a = ld8(ar + ks * 16);
    bf16x8 b0 = ld8(w0 + ks * 16);
    bf16x8 b1 = ld8(w1 + ks * 16);
    c0 = mfma_bf16(a, b0, c0);
    c1 = mfma_bf16(a, b1, c1);
  }
}

// Accumulator (columns col0..col0+31) + bias -> ReLU -> bf16 into an LDS tile.
__device__ __forceinline__ void store_relu(uint16_t* H, const f32x16& c, int col0, const float* bias) {
  const int l = __lane_id(), r = l & 31, h = l >> 5;
  const int col = col0 + r;
  const float b = bias[col];
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int row = 8 * (v >> 2) + 4 * h + (v & 3);
    float x = c[v] + b;
    H[row * kHS + col] = f2bf(x > 0.f ? x : 0.f);
  }
}
// Accumulator + bias -> FP32 output tile (columns < 40).
__device__ __forceinline__ void store_out(float* O, const f32x16& c, int col0, const float* bias) {
  const int l = __lane_id(), r = l & 31, h = l >> 5;
  const int col = col0 + r;
  if (col >= kGuideOut) return;
  const float b = bias[col];
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int row = 8 * (v >> 2) + 4 * h + (v & 3);
    O[row * kOS + col] = c[v] + b;
  }
}
// Gradient accumulator masked by ReLU'(Hmask) -> bf16 into an LDS tile.
__device__ __forceinline__ void store_grad(uint16_t* Gt, const f32x16& c, int col0, const uint16_t* Hmask) {
  const int l = __lane_id(), r = l & 31, h = l >> 5;
  const int col = col0 + r;
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int row = 8 * (v >> 2) + 4 * h + (v & 3);
    float x = bf2f(Hmask[row * kHS + col]) > 0.f ? c[v] : 0.f;
    Gt[row * kHS + col] = f2bf(x);
  }
}

// The 24 features of one coordinate v: sin and cos of pi * 2^f * v for f = 0..11. sincospi at
// f = 0 in FP64, the higher frequencies by the double-angle identities sin 2x = 2 sin x cos x,
// cos 2x = (cos x - sin x)(cos x + sin x), also in FP64: each doubling at most doubles the error,
// about 2^11 ulp of FP64 at f = 11, so the features are FP32 values within a rounding of the exact
// ones (the bf16 features then differ from correctly rounded ones only at exact rounding ties).
// freq_block gives the next four frequencies (8 features) and advances the state past them.
struct FreqState {
  double s, c;
};
__device__ __forceinline__ FreqState freq_start(float v) {
  FreqState q;
  sincospi((double)v, &q.s, &q.c);
  return q;
}
__device__ __forceinline__ void freq_block(FreqState& q, bool first, float e[8]) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (!first || k > 0) {
      const double s2 = 2.0 * (q.s * q.c);
      q.c = (q.c - q.s) * (q.c + q.s);
      q.s = s2;
    }
    e[2 * k] = (float)q.s;
    e[2 * k + 1] = (float)q.c;
  }
}

// 72 features of this lane's half (h = 0: position, h = 1: normal) for row r into the X tile.
// Feature 72h + 24d + 2f (+1) = sin (cos) of pi * 2^f * v_d.
__device__ __forceinline__ void encode_half(uint16_t* X, int r, int h, const float v[3]) {
  uint16_t* x = X + r * kXS + 72 * h;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    FreqState q = freq_start(v[d]);
#pragma unroll
    for (int b = 0; b < 3; b++) {
      float e[8];
      freq_block(q, b == 0, e);
#pragma unroll
      for (int k = 0; k < 4; k++)  // sin and cos of one frequency: one 32-bit LDS store
        *reinterpret_cast<uint32_t*>(x + 24 * d + 8 * b + 2 * k) =
            (uint32_t)f2bf(e[2 * k]) | ((uint32_t)f2bf(e[2 * k + 1]) << 16);
    }
  }
}

// First layer with the features computed in registers as MFMA A operands: in k-step b, lane
// (r, h) supplies feature block b of its half (columns 72h + 8b .. +7 of row r) and the B operand
// the same 8 columns of W1's output rows r and 32 + r, so each k-step sums 16 of the 144 columns.
__device__ __forceinline__ void layer1_from_features(const float v[3], const uint16_t* W1, int sw, f32x16& c0,
                                                     f32x16& c1) {
  const int l = __lane_id(), r = l & 31, h = l >> 5;
  const uint16_t* w0 = W1 + r * sw + 72 * h;
  const uint16_t* w1 = W1 + (32 + r) * sw + 72 * h;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    FreqState q = freq_start(v[d]);
#pragma unroll
    for (int bb = 0; bb < 3; bb++) {
      const int b = 3 * d + bb;
      float e[8];
      freq_block(q, bb == 0, e);
      bf16x8 a;
#pragma unroll
      for (int j = 0; j < 8; j++) a[j] = (__bf16)e[j];
      c0 = mfma_bf16(a, ld8(w0 + 8 * b), c0);
      c1 = mfma_bf16(a, ld8(w1 + 8 * b), c1);
    }
  }
}

// Every LDS tile here belongs to one wave, so ordering the wave's own LDS accesses is enough: a
// wave's LDS instructions complete in issue order; this keeps the compiler from moving them.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward through the four layers for the wave's tile; returns with the FP32 outputs in O
// (which may alias X). HA/HB are scratch tiles (HB may alias X, which is dead after layer 1);
// when keep != nullptr, H1..H3 go to keep[0..2].
__device__ __forceinline__ void forward_tile(const DevGuide& G, uint16_t* X, uint16_t* HA, uint16_t* HB, float* O,
                                             uint16_t* const* keep) {
  uint16_t* h1 = keep ? keep[0] : HA;
  uint16_t* h2 = keep ? keep[1] : HB;
  uint16_t* h3 = keep ? keep[2] : HA;
  f32x16 c0 = {}, c1 = {};
  gemm_tile<9>(X, kXS, G.wf + kGWf1, kGuideIn, c0, c1);
  store_relu(h1, c0, 0, G.bias + kGBias1);
  store_relu(h1, c1, 32, G.bias + kGBias1);
  wave_sync();
  c0 = f32x16{}; c1 = f32x16{};
  gemm_tile<4>(h1, kHS, G.wf + kGWf2, kGuideHid, c0, c1);
  store_relu(h2, c0, 0, G.bias + kGBias2);
  store_relu(h2, c1, 32, G.bias + kGBias2);
  wave_sync();
  c0 = f32x16{}; c1 = f32x16{};
  gemm_tile<4>(h2, kHS, G.wf + kGWf3, kGuideHid, c0, c1);
  wave_sync();  // h3 may alias h1 (inference)
  store_relu(h3, c0, 0, G.bias + kGBias3);
  store_relu(h3, c1, 32, G.bias + kGBias3);
  wave_sync();
  c0 = f32x16{}; c1 = f32x16{};
  gemm_tile<4>(h3, kHS, G.wf + kGWf4, kGuideHid, c0, c1);
  store_out(O, c0, 0, G.bias + kGBias4);
  store_out(O, c1, 32, G.bias + kGBias4);
  wave_sync();
}

// Mixture parameters (pi, mu xyz, kappa) x 8 from the 40 raw outputs of one row: the softmax's
// maximum and normaliser over the 8 logits, then lobe k's 5 parameters into m[0..4].
__device__ __forceinline__ void mixture_softmax(const float* o, float& amax, float& z) {
  amax = o[0];
#pragma unroll
  for (int k = 1; k < kGuideLobes; k++) amax = fmaxf(amax, o[5 * k]);
  z = 0.f;
#pragma unroll
  for (int k = 0; k < kGuideLobes; k++) z += expf(o[5 * k] - amax);
}
__device__ __forceinline__ void mixture_lobe(const float* o, int k, float amax, float z, float* m) {
  float m0 = o[5 * k + 1], m1 = o[5 * k + 2], m2 = o[5 * k + 3];
  float n = fmaxf(sqrtf(m0 * m0 + (m1 * m1 + m2 * m2)), 1e-6f);
  float s = fminf(fmaxf(o[5 * k + 4], kGuideLogKappaMin), kGuideLogKappaMax);
  m[0] = expf(o[5 * k] - amax) / z;
  m[1] = m0 / n;
  m[2] = m1 / n;
  m[3] = m2 / n;
  m[4] = expf(s);
}
__device__ __forceinline__ void mixture_from_raw(const float* o, float* mix) {
  float amax, z;
  mixture_softmax(o, amax, z);
#pragma unroll
  for (int k = 0; k < kGuideLobes; k++) mixture_lobe(o, k, amax, z, mix + 5 * k);
}

// ---------------------------------------------------------------- inference
// MODE 0: rows are the guide queue's pool slots (hit surface -> inputs, mixture -> G.mix[slot]).
// MODE 1: rows are in6[] inputs (position in [0,1]^3, normal), mixture -> out[row].
// The first layer takes its features from registers (layer1_from_features); one LDS tile per wave
// holds the hidden activations (bf16, stride kHS) and, aliasing them, the FP32 outputs, each layer
// overwriting its input once the wave has read it. The block copies the bf16 weights into LDS once
// (rows padded to the activation tiles' strides, which keep the MFMA operand reads conflict-free):
// read from global memory, every 32-row tile would fetch all 43 KB of them again through the
// vector memory path. Blocks of 16 waves share one copy: 46 KB of weights + 16 x 5.1 KB of tiles
// = 128 KB, one block per CU, 4 waves per SIMD (blocks of 4 waves, two per CU, kept 2 per SIMD).
constexpr int kInferWaves = 16;
template <int MODE>
__global__ void __launch_bounds__(64 * kInferWaves) k_guide_infer(DevGuide G, DevPool P, const DevScene* scenes,
                                                                   const DevJob* jobs, const float* in6, int n_arr,
                                                                   float* out) {
  __shared__ __attribute__((aligned(16))) float s_t[kInferWaves][32 * kOS];
  __shared__ __attribute__((aligned(16))) uint16_t s_w1[64 * kXS];
  __shared__ __attribute__((aligned(16))) uint16_t s_wh[3][64 * kHS];
  for (int i = threadIdx.x; i < 64 * (kGuideIn / 8); i += blockDim.x) {
    const int row = i / (kGuideIn / 8), c = i - row * (kGuideIn / 8);
    *reinterpret_cast<uint4*>(s_w1 + row * kXS + 8 * c) =
        *reinterpret_cast<const uint4*>(G.wf + kGWf1 + row * kGuideIn + 8 * c);
  }
  for (int i = threadIdx.x; i < 3 * 64 * (kGuideHid / 8); i += blockDim.x) {
    const int m = i / (64 * (kGuideHid / 8)), j = i - m * (64 * (kGuideHid / 8));
    const int row = j / (kGuideHid / 8), c = j - row * (kGuideHid / 8);
    const int off = m == 0 ? kGWf2 : (m == 1 ? kGWf3 : kGWf4);
    *reinterpret_cast<uint4*>(s_wh[m] + row * kHS + 8 * c) =
        *reinterpret_cast<const uint4*>(G.wf + off + row * kGuideHid + 8 * c);
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, l = __lane_id(), r = l & 31, h = l >> 5;
  float* O = s_t[wave];
  uint16_t* H = reinterpret_cast<uint16_t*>(O);
  static_assert(32 * kHS * 2 <= 32 * kOS * 4, "hidden tiles alias the output tile");
  const uint32_t n = MODE == 0 ? *G.gq_count : (uint32_t)n_arr;
  if (MODE == 0 && blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(G.infer_count, (unsigned long long)n);
  const uint32_t ntiles = (n + 31) / 32;
  for (uint32_t t0 = blockIdx.x * kInferWaves; t0 < ntiles; t0 += gridDim.x * kInferWaves) {
    const uint32_t e = (t0 + wave) * 32 + r;
    const bool valid = e < n;
    float v[3] = {0.f, 0.f, 0.f};
    uint32_t slot = 0;
    if (valid) {
      if (MODE == 0) {
        slot = G.gq[e];
        const uint32_t f = P.flags[slot];
        const DevScene& S = scenes[jobs[fl_job(f)].scene];
        float4 ro = P.ray_o[slot], rd = P.ray_d[slot], hv = P.hit[slot];
        HitRec hr;
        hr.t = hv.x; hr.b1 = hv.y; hr.b2 = hv.z; hr.prim = __float_as_int(hv.w);
        Surface sf = surface_at(S, hr, mk(ro.x, ro.y, ro.z), mk(rd.x, rd.y, rd.z));
        if (h == 0) {
          v[0] = (sf.p.x - G.bmin[0]) * G.inv_ext[0];
          v[1] = (sf.p.y - G.bmin[1]) * G.inv_ext[1];
          v[2] = (sf.p.z - G.bmin[2]) * G.inv_ext[2];
        } else {
          v[0] = (sf.sh.n.x + 1.f) * 0.5f;
          v[1] = (sf.sh.n.y + 1.f) * 0.5f;
          v[2] = (sf.sh.n.z + 1.f) * 0.5f;
        }
      } else {
        const float* q = in6 + 6 * (size_t)e;
        if (h == 0) {
          v[0] = q[0]; v[1] = q[1]; v[2] = q[2];
        } else {
          v[0] = (q[3] + 1.f) * 0.5f; v[1] = (q[4] + 1.f) * 0.5f; v[2] = (q[5] + 1.f) * 0.5f;
        }
      }
    }
    f32x16 c0 = {}, c1 = {};
    layer1_from_features(v, s_w1, kXS, c0, c1);
    store_relu(H, c0, 0, G.bias + kGBias1);
    store_relu(H, c1, 32, G.bias + kGBias1);
    wave_sync();
    const int boffs[2] = {kGBias2, kGBias3};
#pragma unroll
    for (int k = 0; k < 2; k++) {
      c0 = f32x16{}; c1 = f32x16{};
      gemm_tile<4>(H, kHS, s_wh[k], kHS, c0, c1);
      wave_sync();
      store_relu(H, c0, 0, G.bias + boffs[k]);
      store_relu(H, c1, 32, G.bias + boffs[k]);
      wave_sync();
    }
    c0 = f32x16{}; c1 = f32x16{};
    gemm_tile<4>(H, kHS, s_wh[2], kHS, c0, c1);
    wave_sync();
    store_out(O, c0, 0, G.bias + kGBias4);
    store_out(O, c1, 32, G.bias + kGBias4);
    wave_sync();
    if (valid) {
      // both halves of the wave: half h forms lobes 4h..4h+3
      const float* o = O + r * kOS;
      float amax, z, m[kGuideOut / 2];
      mixture_softmax(o, amax, z);
#pragma unroll
      for (int j = 0; j < kGuideLobes / 2; j++) mixture_lobe(o, (kGuideLobes / 2) * h + j, amax, z, m + 5 * j);
      if (MODE == 0) {
        // the renderer's layout (kGuideMixR: pi, pi x normaliser, mu, kappa), this half's lobes
        mixr_store_half(G.mix + (size_t)slot * kGuideMixR, h, m);
      } else {
        // the parameter layout (pi, mu xyz, kappa) x 8: floats 20h..20h+19 of the row
        float4* d4 = reinterpret_cast<float4*>(out + (size_t)e * kGuideOut + (kGuideOut / 2) * h);
#pragma unroll
        for (int k = 0; k < kGuideOut / 8; k++) d4[k] = make_float4(m[4 * k], m[4 * k + 1], m[4 * k + 2], m[4 * k + 3]);
      }
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------- training
// Column-blocked copy of a 32-row LDS tile (stride st, `valid` columns, the rest zero) into
// dst[(row0/8 + blk) * ncb + c][8].
__device__ __forceinline__ void tile_to_cb(const uint16_t* T, int st, int valid, int ncb, uint16_t* dst,
                                           uint32_t row0) {
  const int l = __lane_id();
  for (int p = l; p < 4 * ncb; p += 64) {
    const int blk = p / ncb, c = p - blk * ncb;
    uint16_t e[8];
#pragma unroll
    for (int j = 0; j < 8; j++) e[j] = c < valid ? T[(8 * blk + j) * st + c] : (uint16_t)0;
    uint4 u;
    u.x = (uint32_t)e[0] | ((uint32_t)e[1] << 16);
    u.y = (uint32_t)e[2] | ((uint32_t)e[3] << 16);
    u.z = (uint32_t)e[4] | ((uint32_t)e[5] << 16);
    u.w = (uint32_t)e[6] | ((uint32_t)e[7] << 16);
    *reinterpret_cast<uint4*>(dst + ((size_t)(row0 / 8 + blk) * ncb + c) * 8) = u;
  }
}

// d(loss)/d(raw outputs) for one sample with scale c; returns c * (-log q).
__device__ __forceinline__ float vmf_loss_grad(const float* o, const float* w, float c, float* g) {
  float mix[kGuideOut];
  mixture_from_raw(o, mix);
  float lg[kGuideLobes], cosv[kGuideLobes];
  float mx = -3.0e38f;
#pragma unroll
  for (int k = 0; k < kGuideLobes; k++) {
    const float kap = mix[5 * k + 4];
    cosv[k] = mix[5 * k + 1] * w[0] + (mix[5 * k + 2] * w[1] + mix[5 * k + 3] * w[2]);
    lg[k] = logf(mix[5 * k]) + (logf(kap) - logf(2.f * PI) - logf(-expm1f(-2.f * kap)) + kap * (cosv[k] - 1.f));
    mx = fmaxf(mx, lg[k]);
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kGuideLobes; k++) s += expf(lg[k] - mx);
  const float lq = mx + logf(s);
#pragma unroll
  for (int k = 0; k < kGuideLobes; k++) {
    const float gam = expf(lg[k] - lq);
    const float kap = mix[5 * k + 4];
    const float m0 = o[5 * k + 1], m1 = o[5 * k + 2], m2 = o[5 * k + 3];
    const float mn = fmaxf(sqrtf(m0 * m0 + (m1 * m1 + m2 * m2)), 1e-6f);
    g[5 * k] = c * (mix[5 * k] - gam);
#pragma unroll
    for (int d = 0; d < 3; d++) g[5 * k + 1 + d] = -c * gam * kap * (w[d] - cosv[k] * mix[5 * k + 1 + d]) / mn;
    const float sr = o[5 * k + 4];
    const bool clamped = sr < kGuideLogKappaMin || sr > kGuideLogKappaMax;
    const float dl = 1.f - 2.f * kap / expm1f(2.f * kap) + kap * (cosv[k] - 1.f);
    g[5 * k + 4] = clamped ? 0.f : -c * gam * dl;
  }
  return -c * lq;
}

// Weight sum of the batch, zeroed loss and gradient (one block).
// Weight normalisation of a training batch. Sample weights are clamped at kGuideWClamp times the
// batch mean: one path that found a rare bright connection must not decide a whole step (the
// clamp shapes the training target only; the renderer's estimator is unbiased for any mixture).
__device__ __forceinline__ float block_sum_1024(float acc, float* s_sum) {
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (__lane_id() == 0) s_sum[threadIdx.x >> 6] = acc;
  __syncthreads();
  float t = 0.f;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += s_sum[w];
  __syncthreads();
  return t;
}
constexpr int kPrepPer = 64;  // weights each thread keeps in registers (1024 x 64 = one batch)
__global__ void __launch_bounds__(1024) k_guide_prep(const float* samples, uint64_t base, uint32_t mask, int n,
                                                     GuideTrainScratch S) {
  __shared__ float s_sum[16];
  // all loads of a thread are issued before they are summed; the clamped pass reuses them
  float wv[kPrepPer];
#pragma unroll
  for (int j = 0; j < kPrepPer; j++) {
    const int i = threadIdx.x + j * (int)blockDim.x;
    wv[j] = i < n ? samples[10 * (size_t)((base + i) & mask) + 9] : 0.f;
  }
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < kPrepPer; j++) acc += wv[j];
  for (int i = threadIdx.x + kPrepPer * (int)blockDim.x; i < n; i += blockDim.x)
    acc += samples[10 * (size_t)((base + i) & mask) + 9];
  for (int i = threadIdx.x; i < kGuideParams; i += blockDim.x) S.grad[i] = 0.f;
  const float total = block_sum_1024(acc, s_sum);
  const float cap = n > 0 ? kGuideWClamp * (total / (float)n) : 0.f;
  acc = 0.f;
#pragma unroll
  for (int j = 0; j < kPrepPer; j++) acc += fminf(wv[j], cap);
  for (int i = threadIdx.x + kPrepPer * (int)blockDim.x; i < n; i += blockDim.x)
    acc += fminf(samples[10 * (size_t)((base + i) & mask) + 9], cap);
  const float clamped = block_sum_1024(acc, s_sum);
  if (threadIdx.x == 0) {
    S.scal[0] = clamped;
    S.scal[1] = 0.f;
    S.scal[2] = cap;
  }
}

constexpr int kTrainWaves = 2;
__global__ void __launch_bounds__(64 * kTrainWaves) k_guide_train(DevGuide G, const float* samples, uint64_t base,
                                                                   uint32_t mask, int n, GuideTrainScratch S) {
  __shared__ __attribute__((aligned(16))) uint16_t s_x[kTrainWaves][32 * kXS];
  __shared__ __attribute__((aligned(16))) uint16_t s_h[kTrainWaves][3][32 * kHS];
  __shared__ __attribute__((aligned(16))) uint16_t s_g[kTrainWaves][2][32 * kHS];
  const int wave = threadIdx.x >> 6, l = __lane_id(), r = l & 31, h = l >> 5;
  uint16_t* X = s_x[wave];
  float* O = reinterpret_cast<float*>(X);
  uint16_t* H[3] = {s_h[wave][0], s_h[wave][1], s_h[wave][2]};
  uint16_t* G0 = s_g[wave][0];
  uint16_t* G1 = s_g[wave][1];
  const float wsum = S.scal[0];
  const float inv_w = wsum > 0.f ? 1.f / wsum : 0.f;
  const float cap = S.scal[2];
  const uint32_t ntiles = ((uint32_t)n + 31) / 32;
  for (uint32_t t0 = blockIdx.x * kTrainWaves; t0 < ntiles; t0 += gridDim.x * kTrainWaves) {
    const uint32_t tile = t0 + wave;
    const bool tile_ok = tile < ntiles;
    const uint32_t row0 = tile * 32;
    const uint32_t e = row0 + r;
    const bool valid = e < (uint32_t)n;
    const float* smp = samples + 10 * (size_t)((base + e) & mask);
    float v[3] = {0.f, 0.f, 0.f};
    if (valid) {
      if (h == 0) {
        v[0] = smp[0]; v[1] = smp[1]; v[2] = smp[2];
      } else {
        v[0] = (smp[3] + 1.f) * 0.5f; v[1] = (smp[4] + 1.f) * 0.5f; v[2] = (smp[5] + 1.f) * 0.5f;
      }
    }
    encode_half(X, r, h, v);
    __syncthreads();
    if (tile_ok) tile_to_cb(X, kXS, kGuideIn, 160, S.x0, row0);
    __syncthreads();
    forward_tile(G, X, nullptr, nullptr, O, H);
    if (tile_ok) {
      tile_to_cb(H[0], kHS, 64, 64, S.h[0], row0);
      tile_to_cb(H[1], kHS, 64, 64, S.h[1], row0);
      tile_to_cb(H[2], kHS, 64, 64, S.h[2], row0);
    }
    // loss and output gradient (one lane per row)
    float lsum = 0.f;
    if (h == 0) {
      float g[kGuideOut];
#pragma unroll
      for (int k = 0; k < kGuideOut; k++) g[k] = 0.f;
      if (valid && smp[9] > 0.f) {
        float raw[kGuideOut];
#pragma unroll
        for (int k = 0; k < kGuideOut; k++) raw[k] = O[r * kOS + k];
        const float w[3] = {smp[6], smp[7], smp[8]};
        lsum = vmf_loss_grad(raw, w, fminf(smp[9], cap) * inv_w, g);
      }
#pragma unroll
      for (int k = 0; k < kGuideOut; k++) G0[r * kHS + k] = f2bf(g[k]);
#pragma unroll
      for (int k = kGuideOut; k < 64; k++) G0[r * kHS + k] = 0;
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (l == 0 && lsum != 0.f) unsafeAtomicAdd(&S.scal[1], lsum);
    __syncthreads();
    if (tile_ok) tile_to_cb(G0, kHS, 64, 64, S.g[3], row0);
    // dH3 = dOut W4 (.) relu'(H3)
    f32x16 c0 = {}, c1 = {};
    gemm_tile<4>(G0, kHS, G.wb + kGWb4, 64, c0, c1);
    store_grad(G1, c0, 0, H[2]);
    store_grad(G1, c1, 32, H[2]);
    __syncthreads();
    if (tile_ok) tile_to_cb(G1, kHS, 64, 64, S.g[2], row0);
    c0 = f32x16{}; c1 = f32x16{};
    gemm_tile<4>(G1, kHS, G.wb + kGWb3, 64, c0, c1);
    __syncthreads();
    store_grad(G0, c0, 0, H[1]);
    store_grad(G0, c1, 32, H[1]);
    __syncthreads();
    if (tile_ok) tile_to_cb(G0, kHS, 64, 64, S.g[1], row0);
    c0 = f32x16{}; c1 = f32x16{};
    gemm_tile<4>(G0, kHS, G.wb + kGWb2, 64, c0, c1);
    store_grad(G1, c0, 0, H[0]);
    store_grad(G1, c1, 32, H[0]);
    __syncthreads();
    if (tile_ok) tile_to_cb(G1, kHS, 64, 64, S.g[0], row0);
    __syncthreads();
  }
}

// Weight gradients dW = dY^T X over row chunks (22 output tiles of 32 x 32), bias gradients from
// the same dY fragments. One wave per (tile, chunk).
constexpr int kWgradTiles = 22;
__global__ void __launch_bounds__(64) k_guide_wgrad(GuideTrainScratch S, uint32_t nrows, uint32_t chunk_rows) {
  const int tile = blockIdx.x % kWgradTiles;
  const uint32_t chunk = blockIdx.x / kWgradTiles;
  int layer, ot, it;
  if (tile < 10) {
    layer = 0; ot = tile / 5; it = tile % 5;
  } else {
    layer = 1 + (tile - 10) / 4; ot = ((tile - 10) % 4) / 2; it = (tile - 10) % 2;
  }
  const uint16_t* dY = S.g[layer];
  const uint16_t* Xp = layer == 0 ? S.x0 : S.h[layer - 1];
  const int xcols = layer == 0 ? 160 : 64;
  const int nin = layer == 0 ? kGuideIn : kGuideHid;
  const int nout = layer == 3 ? kGuideOut : kGuideHid;
  const int offW = layer == 0 ? kGOffW1 : layer == 1 ? kGOffW2 : layer == 2 ? kGOffW3 : kGOffW4;
  const int offB = layer == 0 ? kGOffB1 : layer == 1 ? kGOffB2 : layer == 2 ? kGOffB3 : kGOffB4;
  const int l = __lane_id(), r = l & 31, h = l >> 5;
  const uint32_t rb0 = chunk * chunk_rows;
  const uint32_t rb1 = min(rb0 + chunk_rows, nrows);
  f32x16 acc = {};
  float bsum = 0.f;
  for (uint32_t rb = rb0; rb < rb1; rb += 16) {
    const uint32_t blk = rb / 8 + h;
    bf16x8 a = ld8(dY + ((size_t)blk * 64 + ot * 32 + r) * 8);
    bf16x8 b = ld8(Xp + ((size_t)blk * xcols + it * 32 + r) * 8);
    acc = mfma_bf16(a, b, acc);
    if (it == 0) {
#pragma unroll
      for (int j = 0; j < 8; j++) bsum += (float)a[j];
    }
  }
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int o = ot * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
    const int i = it * 32 + r;
    if (o < nout && i < nin) unsafeAtomicAdd(&S.grad[offW + o * nin + i], acc[v]);
  }
  if (it == 0) {
    bsum += __shfl_xor(bsum, 32);
    const int o = ot * 32 + r;
    if (h == 0 && o < nout) unsafeAtomicAdd(&S.grad[offB + o], bsum);
  }
}

// Master parameter index -> bf16 copies (forward layout always, backward layout for layers 2..4)
// and FP32 biases.
__device__ __forceinline__ void scatter_param(int p, float w, uint16_t* wf, uint16_t* wb, float* bias) {
  if (p < kGOffB1) {
    wf[kGWf1 + p] = f2bf(w);
  } else if (p < kGOffW2) {
    bias[kGBias1 + (p - kGOffB1)] = w;
  } else if (p < kGOffB2) {
    int q = p - kGOffW2, o = q / 64, i = q % 64;
    wf[kGWf2 + q] = f2bf(w);
    wb[kGWb2 + i * 64 + o] = f2bf(w);
  } else if (p < kGOffW3) {
    bias[kGBias2 + (p - kGOffB2)] = w;
  } else if (p < kGOffB3) {
    int q = p - kGOffW3, o = q / 64, i = q % 64;
    wf[kGWf3 + q] = f2bf(w);
    wb[kGWb3 + i * 64 + o] = f2bf(w);
  } else if (p < kGOffW4) {
    bias[kGBias3 + (p - kGOffB3)] = w;
  } else if (p < kGOffB4) {
    int q = p - kGOffW4, o = q / 64, i = q % 64;
    wf[kGWf4 + q] = f2bf(w);
    wb[kGWb4 + i * 64 + o] = f2bf(w);
  } else if (p < kGuideParams) {
    bias[kGBias4 + (p - kGOffB4)] = w;
  }
}

__global__ void k_guide_refresh(const float* master, uint16_t* wf, uint16_t* wb, float* bias) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kGuideParams) scatter_param(p, master[p], wf, wb, bias);
}

// Adam (Kingma & Ba 2015), beta1 0.9, beta2 0.999, eps 1e-8.
__global__ void k_guide_adam(float* master, float* m, float* v, const float* grad, int step, float lr, uint16_t* wf,
                             uint16_t* wb, float* bias) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kGuideParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float g = grad[p];
  if (!(fabsf(g) < 1e30f)) g = 0.f;  // a non-finite gradient does not poison the state
  float mm = b1 * m[p] + (1.f - b1) * g;
  float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  const float mh = mm / (1.f - powf(b1, (float)step));
  const float vh = vv / (1.f - powf(b2, (float)step));
  const float w = master[p] - lr * mh / (sqrtf(vh) + eps);
  master[p] = w;
  scatter_param(p, w, wf, wb, bias);
}

// ---------------------------------------------------------------- launchers
void launch_guide_refresh(const float* master, uint16_t* wf, uint16_t* wb, float* bias, hipStream_t stream) {
  hipLaunchKernelGGL(k_guide_refresh, dim3((kGuideParams + 255) / 256), dim3(256), 0, stream, master, wf, wb, bias);
}

int guide_infer_blocks_per_cu() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(k_guide_infer<0>),
                                                   64 * kInferWaves, 0) != hipSuccess)
    return 0;
  return nb;
}

void launch_guide_infer_pool(const DevGuide& G, const DevPool& P, const DevScene* scenes, const DevJob* jobs,
                             int grid, hipStream_t stream) {
  hipLaunchKernelGGL(k_guide_infer<0>, dim3(grid), dim3(64 * kInferWaves), 0, stream, G, P, scenes, jobs,
                     (const float*)nullptr, 0, (float*)nullptr);
}

void launch_guide_infer_array(const DevGuide& G, const float* in6, int n, float* mix_out, hipStream_t stream) {
  if (n <= 0) return;
  DevPool P;
  memset(&P, 0, sizeof(P));
  int tiles = (n + 31) / 32;
  int blocks = std::min((tiles + kInferWaves - 1) / kInferWaves, 4096);
  hipLaunchKernelGGL(k_guide_infer<1>, dim3(blocks), dim3(64 * kInferWaves), 0, stream, G, P,
                     (const DevScene*)nullptr, (const DevJob*)nullptr, in6, n, mix_out);
}

void launch_guide_grad(const DevGuide& G, const float* samples, uint64_t base, uint32_t mask, int n,
                       const GuideTrainScratch& S, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_guide_prep, dim3(1), dim3(1024), 0, stream, samples, base, mask, n, S);
  const int tiles = (n + 31) / 32;
  const int blocks = std::min((tiles + kTrainWaves - 1) / kTrainWaves, 2048);
  hipLaunchKernelGGL(k_guide_train, dim3(blocks), dim3(64 * kTrainWaves), 0, stream, G, samples, base, mask, n, S);
  const uint32_t nrows = (uint32_t)tiles * 32;
  const uint32_t chunk = 2048;
  const uint32_t nchunks = (nrows + chunk - 1) / chunk;
  hipLaunchKernelGGL(k_guide_wgrad, dim3(kWgradTiles * nchunks), dim3(64), 0, stream, S, nrows, chunk);
}

void launch_guide_adam(float* master, float* m, float* v, const float* grad, int step, float lr, uint16_t* wf,
                       uint16_t* wb, float* bias, hipStream_t stream) {
  hipLaunchKernelGGL(k_guide_adam, dim3((kGuideParams + 255) / 256), dim3(256), 0, stream, master, m, v, grad, step,
                     lr, wf, wb, bias);
}

}  // namespace rtg

// CPU oracle of the path-guiding network (FP64), the checker of the GPU guiding kernels.
//
// TEST INFRASTRUCTURE ONLY, like rt_oracle.cpp: nothing in the product includes, links or calls
// this file. The guiding network has no counterpart in the reference renderer; this file
// restates the network the library defines (encoding, layers, vMF-mixture head, loss) in double
// precision, optionally rounding to bfloat16 at the same points as the GPU kernels:
//   * the 144 input features, the weights, and each hidden activation after its ReLU;
//   * the output gradient and each hidden-layer gradient before they feed a product (the weight
//     gradients are sums of bf16 x bf16 products, the bias gradients sums of the rounded
//     gradients).
// Pinned by: central finite differences of the FP64 loss (no rounding), the vMF density's
// normalisation over the sphere, and its closed-form sampling (tests/test_oracle_guide.py).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace gor {

constexpr int kIn = 144, kHid = 64, kLobes = 8, kOut = 40;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kIn;       // 9216
constexpr int kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid * kHid;
constexpr int kOffW4 = kOffB3 + kHid, kOffB4 = kOffW4 + kOut * kHid;
constexpr int kParams = kOffB4 + kOut;                          // 20200
constexpr double kPi = 3.14159265358979323846;
constexpr double kLogKappaMin = -3.0, kLogKappaMax = 8.0;

// round-to-nearest-even to bfloat16 (through float), returned as double
double bf16(double x) {
  float f = (float)x;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return f;  // inf / nan unchanged
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return f;
}
inline double rnd(double x, bool emu) { return emu ? bf16(x) : x; }

// 12-frequency sin/cos encoding of (pos in [0,1]^3, normal in [-1,1]^3 mapped to [0,1]).
// Feature d*24 + 2f (+1): sin (cos) of pi * 2^f * v_d, with the argument reduced exactly mod 2.
void encode(const float* in6, double* feat) {
  for (int d = 0; d < 6; d++) {
    float v = d < 3 ? in6[d] : (in6[d] + 1.0f) * 0.5f;
    for (int f = 0; f < 12; f++) {
      double t = std::ldexp((double)v, f);
      t = t - 2.0 * std::floor(t * 0.5);
      feat[d * 24 + 2 * f] = std::sin(kPi * t);
      feat[d * 24 + 2 * f + 1] = std::cos(kPi * t);
    }
  }
}

struct Net {
  std::vector<double> w;  // weights (bf16-rounded when emulating), biases unrounded
  bool emu;
  Net(const float* p, bool e) : w(kParams), emu(e) {
    for (int i = 0; i < kParams; i++) w[i] = p[i];
    if (emu) {
      for (int i = kOffW1; i < kOffB1; i++) w[i] = bf16(w[i]);
      for (int i = kOffW2; i < kOffB2; i++) w[i] = bf16(w[i]);
      for (int i = kOffW3; i < kOffB3; i++) w[i] = bf16(w[i]);
      for (int i = kOffW4; i < kOffB4; i++) w[i] = bf16(w[i]);
    }
  }
  // y = W x + b (W: nout x nin row-major)
  void dense(int offW, int offB, int nin, int nout, const double* x, double* y) const {
    for (int o = 0; o < nout; o++) {
      double s = 0;
      for (int i = 0; i < nin; i++) s += w[offW + o * nin + i] * x[i];
      y[o] = s + w[offB + o];
    }
  }
};

struct Acts {
  double x0[kIn], h1[kHid], h2[kHid], h3[kHid], out[kOut];
};

void forward(const Net& N, const float* in6, Acts& A) {
  encode(in6, A.x0);
  for (int i = 0; i < kIn; i++) A.x0[i] = rnd(A.x0[i], N.emu);
  double t[kHid];
  N.dense(kOffW1, kOffB1, kIn, kHid, A.x0, t);
  for (int i = 0; i < kHid; i++) A.h1[i] = rnd(t[i] > 0 ? t[i] : 0.0, N.emu);
  N.dense(kOffW2, kOffB2, kHid, kHid, A.h1, t);
  for (int i = 0; i < kHid; i++) A.h2[i] = rnd(t[i] > 0 ? t[i] : 0.0, N.emu);
  N.dense(kOffW3, kOffB3, kHid, kHid, A.h2, t);
  for (int i = 0; i < kHid; i++) A.h3[i] = rnd(t[i] > 0 ? t[i] : 0.0, N.emu);
  N.dense(kOffW4, kOffB4, kHid, kOut, A.h3, A.out);
}

struct Mix {
  double pi[kLobes], mu[kLobes][3], kappa[kLobes], mnorm[kLobes];
  bool clamped[kLobes];
};

void mixture(const double* out, Mix& M) {
  double amax = -1e300;
  for (int k = 0; k < kLobes; k++) amax = std::max(amax, out[5 * k]);
  double z = 0;
  for (int k = 0; k < kLobes; k++) z += std::exp(out[5 * k] - amax);
  for (int k = 0; k < kLobes; k++) {
    M.pi[k] = std::exp(out[5 * k] - amax) / z;
    const double* m = out + 5 * k + 1;
    double n = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    n = std::max(n, 1e-6);
    M.mnorm[k] = n;
    for (int c = 0; c < 3; c++) M.mu[k][c] = m[c] / n;
    double s = out[5 * k + 4];
    M.clamped[k] = s < kLogKappaMin || s > kLogKappaMax;
    M.kappa[k] = std::exp(std::min(std::max(s, kLogKappaMin), kLogKappaMax));
  }
}

// log of the vMF density on the unit sphere: kappa / (2 pi (1 - e^{-2 kappa})) e^{kappa (mu.w - 1)}
double vmf_log(double kappa, double cosang) {
  return std::log(kappa) - std::log(2.0 * kPi) - std::log(-std::expm1(-2.0 * kappa)) + kappa * (cosang - 1.0);
}

double mix_logpdf(const Mix& M, const double* w, double* lg /* kLobes: log pi_k + log v_k */) {
  double mx = -1e300;
  for (int k = 0; k < kLobes; k++) {
    double c = M.mu[k][0] * w[0] + M.mu[k][1] * w[1] + M.mu[k][2] * w[2];
    lg[k] = std::log(M.pi[k]) + vmf_log(M.kappa[k], c);
    mx = std::max(mx, lg[k]);
  }
  double s = 0;
  for (int k = 0; k < kLobes; k++) s += std::exp(lg[k] - mx);
  return mx + std::log(s);
}

// d(-log q)/d(out) for one sample, times scale c
void loss_grad_out(const double* out, const Mix& M, const double* w, double c, double* g) {
  double lg[kLobes];
  double lq = mix_logpdf(M, w, lg);
  for (int k = 0; k < kLobes; k++) {
    double gam = std::exp(lg[k] - lq);
    double kap = M.kappa[k];
    double cosang = M.mu[k][0] * w[0] + M.mu[k][1] * w[1] + M.mu[k][2] * w[2];
    g[5 * k] = c * (M.pi[k] - gam);
    for (int d = 0; d < 3; d++)
      g[5 * k + 1 + d] = -c * gam * kap * (w[d] - cosang * M.mu[k][d]) / M.mnorm[k];
    double dlogv_dlogk = 1.0 - 2.0 * kap / std::expm1(2.0 * kap) + kap * (cosang - 1.0);
    g[5 * k + 4] = M.clamped[k] ? 0.0 : -c * gam * dlogv_dlogk;
  }
  (void)out;
}

}  // namespace gor

using namespace gor;

extern "C" {

int gor_num_params() { return kParams; }

double gor_bf16(double x) { return bf16(x); }

void gor_encode(const float* in6, int n, double* feat) {
  for (int i = 0; i < n; i++) encode(in6 + 6 * i, feat + (size_t)kIn * i);
}

// Network outputs (n x 40) and the mixture per input (n x 40: per lobe pi, mu xyz, kappa).
void gor_forward(const float* params, int emu, const float* in6, int n, double* out, double* mix) {
  Net N(params, emu != 0);
  for (int i = 0; i < n; i++) {
    Acts A;
    forward(N, in6 + 6 * i, A);
    if (out) std::memcpy(out + (size_t)kOut * i, A.out, sizeof(A.out));
    if (mix) {
      Mix M;
      mixture(A.out, M);
      for (int k = 0; k < kLobes; k++) {
        double* m = mix + (size_t)kOut * i + 5 * k;
        m[0] = M.pi[k];
        m[1] = M.mu[k][0];
        m[2] = M.mu[k][1];
        m[3] = M.mu[k][2];
        m[4] = M.kappa[k];
      }
    }
  }
}

// log density of the mixture given per-lobe (pi, mu xyz, kappa) at directions w (n x 3)
void gor_mix_logpdf(const double* mix, const double* w, int n, double* out) {
  for (int i = 0; i < n; i++) {
    Mix M;
    for (int k = 0; k < kLobes; k++) {
      const double* m = mix + (size_t)kOut * i + 5 * k;
      M.pi[k] = m[0];
      M.mu[k][0] = m[1];
      M.mu[k][1] = m[2];
      M.mu[k][2] = m[3];
      M.kappa[k] = m[4];
    }
    double lg[kLobes];
    out[i] = mix_logpdf(M, w + 3 * i, lg);
  }
}

// Loss L = -(1/sum w') sum_i w'_i log q(omega_i | x_i) and dL/dparams, with the weights clamped,
// w'_i = min(w_i, 16 * mean w) (the GPU training step's clamp).
// samples: n x 10 = pos (3, in [0,1]), normal (3), omega (3, unit), weight.
// The bf16 emulation mirrors the GPU training kernels.
static const double kWClamp = 16.0;
double gor_loss_grad(const float* params, int emu, const float* samples, int n, double* grad) {
  Net N(params, emu != 0);
  std::vector<double> g(kParams, 0.0);
  double wraw = 0;
  for (int i = 0; i < n; i++) wraw += samples[10 * i + 9];
  const double cap = n > 0 ? kWClamp * wraw / n : 0.0;
  double wsum = 0;
  for (int i = 0; i < n; i++) wsum += std::min((double)samples[10 * i + 9], cap);
  if (!(wsum > 0)) {
    if (grad) std::memset(grad, 0, sizeof(double) * kParams);
    return 0.0;
  }
  double loss = 0;
  for (int i = 0; i < n; i++) {
    const float* s = samples + 10 * i;
    double c = std::min((double)s[9], cap) / wsum;
    if (c == 0) continue;
    Acts A;
    forward(N, s, A);
    Mix M;
    mixture(A.out, M);
    double w[3] = {s[6], s[7], s[8]};
    double lg[kLobes];
    loss -= c * mix_logpdf(M, w, lg);
    double dout[kOut];
    loss_grad_out(A.out, M, w, c, dout);
    // layer 4
    double dq[kOut];
    for (int o = 0; o < kOut; o++) {
      dq[o] = rnd(dout[o], N.emu);
      g[kOffB4 + o] += dq[o];
      for (int h = 0; h < kHid; h++) g[kOffW4 + o * kHid + h] += dq[o] * A.h3[h];
    }
    // back through layers 3, 2, 1
    const double* hs[3] = {A.h3, A.h2, A.h1};
    const double* xs[3] = {A.h2, A.h1, A.x0};
    const int offW[3] = {kOffW3, kOffW2, kOffW1}, offB[3] = {kOffB3, kOffB2, kOffB1};
    const int nin[3] = {kHid, kHid, kIn};
    // gradient w.r.t. h3 from the output layer
    double dh[kHid];
    for (int h = 0; h < kHid; h++) {
      double sum = 0;
      for (int o = 0; o < kOut; o++) sum += dq[o] * N.w[kOffW4 + o * kHid + h];
      dh[h] = hs[0][h] > 0 ? sum : 0.0;
    }
    for (int L = 0; L < 3; L++) {
      double dhq[kHid];
      for (int o = 0; o < kHid; o++) {
        dhq[o] = rnd(dh[o], N.emu);
        g[offB[L] + o] += dhq[o];
        for (int j = 0; j < nin[L]; j++) g[offW[L] + o * nin[L] + j] += dhq[o] * xs[L][j];
      }
      if (L == 2) break;
      double nd[kHid];
      for (int j = 0; j < kHid; j++) {
        double sum = 0;
        for (int o = 0; o < kHid; o++) sum += dhq[o] * N.w[offW[L] + o * kHid + j];
        nd[j] = hs[L + 1][j] > 0 ? sum : 0.0;
      }
      std::memcpy(dh, nd, sizeof(dh));
    }
  }
  if (grad) std::memcpy(grad, g.data(), sizeof(double) * kParams);
  return loss;
}

}  // extern "C"

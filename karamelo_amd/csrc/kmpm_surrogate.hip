// kmpm_surrogate.hip -- the learned Johnson-Cook surrogate: tuples from the
// plain strength/damage functions, BF16-MFMA training (fused forward and
// backward, FP32 accumulation, Adam) and evaluation, all on the device.
//
// Two networks per material (kmpm_surrogate.h):
//   net 0 (flow stress)   : (eps_p^(1/4), ln epsdot*)             -> sigma_y / A
//   net 1 (failure strain): (triaxiality, ln(epsdot / epsdot0_d)) -> ln eps_f
// Their targets are read off the outputs of the plain FP64
// StrengthJohnsonCook / DamageJohnsonCook restatements (kmpm_constitutive.h)
// called on sampled states: sigma_y is the von Mises stress the radial return
// lands on, eps_f = dps / (damage_init increment).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "kmpm_ctx.h"
#include "kmpm_surrogate.h"

namespace kmpm {

__device__ inline double u01(uint64_t seed, uint64_t i, uint64_t k) {
  // splitmix64 of (seed, sample, stream) -> [0, 1)
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (2 * i + 1) + 0xD1B54A32D192ED03ull * (k + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

constexpr double EPS_P_MIN = 1.0e-10;   // strength_jc.cpp:118: below it sigma_y = A

// A flow-stress tuple at (eps_p, epsdot): a deviatoric rate large enough
// that the trial state is plastic; the returned stress then sits on the
// yield surface.  Target NaN if it is not plastic (never for a valid d_rate).
__device__ inline void tuple_y(const DevMat &m, double ep, double epsdot, double d_rate, float *x,
                               float *y) {
  const double sig[6] = {0, 0, 0, 0, 0, 0};
  const double D[6] = {d_rate, -d_rate, 0, 0, 0, 0};
  double out[6], dps = 0;
  jc_strength(m, sig, D, 1.0, ep, epsdot, 0.0, out, &dps);
  const double ys = SQRT_3_OVER_2 * norm_sym(out);
  double ratio = epsdot / m.strength_p[4];
  ratio = ratio > 1.0 ? ratio : 1.0;
  x[0] = sqrtf(sqrtf((float)ep));
  x[1] = __builtin_amdgcn_logf((float)ratio) * 0.69314718055994531f;
  *y = dps > 0 ? (float)(ys / m.strength_p[1]) : NAN;
}

// A failure-strain tuple at triaxiality tr in (-3, ...) and rate epsdot: a
// unit von Mises deviator and the pressure that gives tr; eps_f = dps /
// damage_init increment.
__device__ inline void tuple_f(const DevMat &m, double tr, double epsdot, float *x, float *y) {
  const double d4 = m.damage_p[3], epsdot0 = m.damage_p[5];
  const double a = 0.57735026918962576;   // 1/sqrt(3): sqrt(3/2) |s| ~ 1
  const double sdev[6] = {a, -a, 0, 0, 0, 0};
  const double vm = SQRT_3_OVER_2 * norm_sym(sdev);
  const double pH = tr <= 0 ? -tr * vm / (1.0 + 0.001 * tr) : -tr * vm / (1.0 - 0.001 * tr);
  const double tr_used = (pH != 0.0 && vm != 0.0) ? -pH / (vm + 0.001 * fabs(pH)) : 0.0;
  double dinit = 0, dmg = 0;
  const double dps = 1.0e-3;
  jc_damage(m, dinit, dmg, pH, sdev, epsdot, dps);
  const double feat_r = (d4 > 0.0 && epsdot > epsdot0) ? log(epsdot / epsdot0) : 0.0;
  x[0] = (float)tr_used;
  x[1] = (float)feat_r;
  *y = dinit > 0 ? (float)log(dps / dinit) : NAN;
}

// Flow-stress tuples at sampled states: eps_p^(1/4) uniform in [q_lo, q_hi],
// ln epsdot* uniform in [0, lnr_hi].
__global__ void k_tuples_y(DevMat m, double q_lo, double q_hi, double lnr_hi, double d_rate,
                           uint64_t seed, int64_t n, float *X, float *Y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double q = q_lo + (q_hi - q_lo) * u01(seed, i, 0);
  const double ep = fmax((q * q) * (q * q), EPS_P_MIN);
  const double lnr = lnr_hi * u01(seed, i, 1);
  tuple_y(m, ep, m.strength_p[4] * exp(lnr), d_rate, X + 2 * i, Y + i);
}

// Failure-strain tuples at sampled states: triaxiality in (-3, triax_hi]
// (-3 itself is the reset branch, damage_jc.cpp:119-122), ln(epsdot /
// epsdot0_d) uniform in [0, lnr_hi].
__global__ void k_tuples_f(DevMat m, double triax_hi, double lnr_hi, uint64_t seed, int64_t n,
                           float *X, float *Y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double tr = -3.0 + 1e-6 + (triax_hi + 3.0 - 1e-6) * (1.0 - u01(seed, i, 0));
  const double lnr = lnr_hi * u01(seed, i, 1);
  tuple_f(m, tr, m.damage_p[5] * exp(lnr), X + 2 * i, Y + i);
}

// A field of particle slot i of a context's current state (FP64 layout or
// the mixed layout's double and float blocks).
__device__ inline double state_field(const KArgs &a, int f, int i) {
  if (a.prec == KMPM_PRECISION_MIXED)
    return mx_is_double(f) ? a.pa[(size_t)mx_dslot(f) * a.cap + i]
                           : (double)a.fa[(size_t)mx_fslot(f) * a.cap + i];
  return a.pa[(size_t)f * a.cap + i];
}

// Tuples at the states a simulation visits: slot k * stride of the context's
// state gives the features of network `net` (net 0: its eps_p and rate;
// net 1: the triaxiality of its stress, -p / (vm + 0.001 |p|) with p =
// -tr(sigma) / 3 and vm the von Mises stress of dev(sigma), and its rate),
// labelled by the plain functions of material m.  Features and target are
// NaN where the particle has no such state (eps_p below 1e-10; a zero
// pressure or deviator, or a triaxiality at or below -3).
__global__ void k_ctx_tuples(KArgs a, DevMat m, int net, int64_t stride, int64_t n,
                             double d_rate, float *X, float *Y) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = (int)(k * stride);
  const double epdot = state_field(a, EPDOT, i);
  if (net == 0) {
    const double ep = state_field(a, EP, i);
    if (ep >= EPS_P_MIN) {
      tuple_y(m, ep, epdot, d_rate, X + 2 * k, Y + k);
      return;
    }
  } else {
    double s[6];
    for (int e = 0; e < 6; e++) s[e] = state_field(a, SXX + e, i);
    const double p = -(s[0] + s[1] + s[2]) / 3.0;
    const double d[6] = {s[0] + p, s[1] + p, s[2] + p, s[3], s[4], s[5]};
    const double vm = SQRT_3_OVER_2 * norm_sym(d);
    if (p != 0.0 && vm != 0.0) {
      const double tr = -p / (vm + 0.001 * fabs(p));
      if (tr > -3.0) {
        tuple_f(m, tr, epdot, X + 2 * k, Y + k);
        return;
      }
    }
  }
  X[2 * k] = NAN;
  X[2 * k + 1] = NAN;
  Y[k] = NAN;
}

// Network outputs for n feature pairs (inference path of the step kernels).
__global__ __launch_bounds__(256) void k_mlp_eval(const MlpNet *net, const float *X, int64_t n,
                                                  float *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  const float f0 = live ? X[2 * i] : 0.0f, f1 = live ? X[2 * i + 1] : 0.0f;
  const float y = mlp_eval_wave(*net, f0, f1);
  if (live) out[i] = y;
}

// Fused forward + backward of the loss L = (1/n) sum (z - t)^2, z the raw
// network output and t the normalised residual of the target over the
// baseline, over n tuples.  A block (four
// waves, 256 tuples) writes its sum of dL/dparams and of the loss to
// partial[block][0..MLP_NPARAM].  The 32x32 layer and both of its backward
// products run on the matrix cores:
//   forward        D[o][p]   = sum_k W2[o][k] H1[k][p]
//   weight grad    dW2[o][k] = sum_p G2[o][p] H1[k][p]   (K = tuples)
//   input grad     G1[k][p]  = sum_o W2[o][k] G2[o][p]
// G2 and H1 are staged in LDS (BF16) to change operand layouts.
__global__ __launch_bounds__(256) void k_mlp_grad(const MlpNet *netp, const float *X,
                                                  const float *Y, int64_t n, float inv_n,
                                                  float *partial) {
  __shared__ float gacc[MLP_NPARAM + 1];
  __shared__ uint16_t sh1[4][MLP_H][MLP_H + 2];
  __shared__ float sh1f[4][MLP_H][MLP_H + 1];
  __shared__ uint16_t sg2[4][MLP_H][MLP_H + 2];
  const MlpNet &nt = *netp;
  const MlpParams &P = nt.p;
  for (int q = threadIdx.x; q <= MLP_NPARAM; q += blockDim.x) gacc[q] = 0;
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, half = lane >> 5, col = lane & 31;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = s < n;
  const float x0 = valid ? fmaf(X[2 * s], nt.in_scale[0], nt.in_shift[0]) : 0.0f;
  const float x1 = valid ? fmaf(X[2 * s + 1], nt.in_scale[1], nt.in_shift[1]) : 0.0f;
  const float tg = valid ? (Y[s] - mlp_base(nt, x0, x1) - nt.out_shift) / nt.out_scale : 0.0f;
  const float vf = valid ? 1.0f : 0.0f;
  // flat parameter offsets
  constexpr int OW1 = 0, OB1 = OW1 + MLP_H * MLP_IN, OW2 = OB1 + MLP_H, OB2 = OW2 + MLP_H * MLP_H,
                OW3 = OB2 + MLP_H, OB3 = OW3 + MLP_H;
  bf16x8_t a0, a1, at0, at1;   // W2 and W2^T operands
  mlp_load_w2(nt, a0, a1);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    at0[i] = bf16_from_bits(nt.w2_bf16[8 * half + i][col]);
    at1[i] = bf16_from_bits(nt.w2_bf16[16 + 8 * half + i][col]);
  }
  f32x16_t dw2 = {};
  for (int T = 0; T < 2; T++) {
    const int src = 32 * T + col;
    const float g0 = __shfl(x0, src, 64), g1 = __shfl(x1, src, 64);
    const float tt = __shfl(tg, src, 64), vv = __shfl(vf, src, 64);
    float h[16];
    bf16x8_t b0, b1;
    mlp_layer1(P, g0, g1, h, b0, b1);
    f32x16_t d = {};
    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, d, 0, 0, 0);
    float h2[16];
    float part = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int o = mlp_drow(i);
      h2[i] = mlp_act(d[i] + P.b2[o]);
      part = fmaf(P.w3[o], h2[i], part);
    }
    part += __shfl_xor(part, 32, 64);
    const float z = part + P.b3;
    const float r = (z - tt) * vv;
    const float g = 2.0f * r * inv_n;
    if (half == 0) {
      atomicAdd(&gacc[MLP_NPARAM], r * r * inv_n);
      atomicAdd(&gacc[OB3], g);
    }
    // layer 3 and the layer-2 activation
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int o = mlp_drow(i);
      atomicAdd(&gacc[OW3 + o], g * h2[i]);
      const float g2 = g * P.w3[o] * mlp_act_grad(h2[i]);
      atomicAdd(&gacc[OB2 + o], g2);
      sg2[wave][o][col] = __builtin_bit_cast(uint16_t, to_bf16(g2));
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int k = (i < 8 ? 0 : 16) + 8 * half + (i & 7);
      sh1[wave][k][col] = __builtin_bit_cast(uint16_t, to_bf16(h[i]));
      sh1f[wave][k][col] = h[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // dW2[o][k] += sum_p G2[o][p] H1[k][p]
#pragma unroll
    for (int st = 0; st < 2; st++) {
      bf16x8_t ag, bh;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        ag[i] = bf16_from_bits(sg2[wave][col][16 * st + 8 * half + i]);
        bh[i] = bf16_from_bits(sh1[wave][col][16 * st + 8 * half + i]);
      }
      dw2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ag, bh, dw2, 0, 0, 0);
    }
    // G1[k][p] = (sum_o W2[o][k] G2[o][p]) act'(H1[k][p])
    f32x16_t g1a = {};
#pragma unroll
    for (int st = 0; st < 2; st++) {
      bf16x8_t bg;
#pragma unroll
      for (int i = 0; i < 8; i++) bg[i] = bf16_from_bits(sg2[wave][16 * st + 8 * half + i][col]);
      g1a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(st == 0 ? at0 : at1, bg, g1a, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int k = mlp_drow(i);
      const float hv = sh1f[wave][k][col];
      const float gk = g1a[i] * mlp_act_grad(hv);
      atomicAdd(&gacc[OW1 + k * MLP_IN + 0], gk * g0);
      atomicAdd(&gacc[OW1 + k * MLP_IN + 1], gk * g1);
      atomicAdd(&gacc[OB1 + k], gk);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
#pragma unroll
  for (int i = 0; i < 16; i++) atomicAdd(&gacc[OW2 + mlp_drow(i) * MLP_H + col], dw2[i]);
  __syncthreads();
  float *dst = partial + (size_t)blockIdx.x * (MLP_NPARAM + 1);
  for (int q = threadIdx.x; q <= MLP_NPARAM; q += blockDim.x) dst[q] = gacc[q];
}

// Sums the block partials in a fixed order: gradient (and mean loss) into
// out[0..MLP_NPARAM]; with adam != 0 also takes one Adam step on the
// network's FP32 parameters and refreshes their BF16 copy.
__global__ void k_mlp_reduce(MlpNet *net, const float *partial, int nblocks, float *out,
                             float *mom, float *vel, int adam, float lr, float b1, float b2,
                             float eps, float c1, float c2) {
  float *pf = reinterpret_cast<float *>(&net->p);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q <= MLP_NPARAM; q += gridDim.x * blockDim.x) {
    float g = 0;
    for (int b = 0; b < nblocks; b++) g += partial[(size_t)b * (MLP_NPARAM + 1) + q];
    out[q] = g;
    if (!adam || q == MLP_NPARAM) continue;
    const float m = b1 * mom[q] + (1 - b1) * g;
    const float v = b2 * vel[q] + (1 - b2) * g * g;
    mom[q] = m;
    vel[q] = v;
    const float w = pf[q] - lr * (m * c1) / (sqrtf(v * c2) + eps);
    pf[q] = w;
    constexpr int OW2 = MLP_H * MLP_IN + MLP_H;
    if (q >= OW2 && q < OW2 + MLP_H * MLP_H) {
      const int r = (q - OW2) / MLP_H, c = (q - OW2) % MLP_H;
      net->w2_bf16[r][c] = __builtin_bit_cast(uint16_t, to_bf16(w));
    }
  }
}

}  // namespace kmpm

using namespace kmpm;

struct kmpm_surrogate {
  kmpm_surrogate_config cfg;
  kmpm_material mat;
  DevMat dmat;
  MlpNet host[2];
  MlpNet *d_net = nullptr;           // [2]
  float *d_X[2] = {nullptr, nullptr}, *d_Y[2] = {nullptr, nullptr};
  float *d_mom = nullptr, *d_vel = nullptr;   // [2][MLP_NPARAM + 1]
  float *d_partial = nullptr, *d_out = nullptr;
  int partial_blocks = 0;
  double lo[2][2], hi[2][2];         // feature ranges per network
  double q_lo, q_hi, d_rate;         // flow-stress tuple sampling
  double sy_virgin;
  hipStream_t stream = nullptr;
};

namespace {

#define SCHK(x)                                   \
  do {                                            \
    if ((x) != hipSuccess) return KMPM_ERR_HIP;   \
  } while (0)

int ensure_partial(kmpm_surrogate *s, int nblocks) {
  if (nblocks <= s->partial_blocks) return KMPM_OK;
  if (s->d_partial) (void)hipFree(s->d_partial);
  s->d_partial = nullptr;
  SCHK(hipMalloc(&s->d_partial, (size_t)nblocks * (MLP_NPARAM + 1) * sizeof(float)));
  s->partial_blocks = nblocks;
  return KMPM_OK;
}

int gen_tuples(kmpm_surrogate *s, int net, int64_t n, uint64_t seed, float *X, float *Y) {
  const int g = (int)((n + 255) / 256);
  if (net == 0)
    hipLaunchKernelGGL(k_tuples_y, dim3(g), dim3(256), 0, s->stream, s->dmat, s->q_lo, s->q_hi,
                       s->hi[0][1], s->d_rate, seed, n, X, Y);
  else
    hipLaunchKernelGGL(k_tuples_f, dim3(g), dim3(256), 0, s->stream, s->dmat, s->hi[1][0],
                       s->hi[1][1], seed, n, X, Y);
  return hipGetLastError() == hipSuccess ? KMPM_OK : KMPM_ERR_HIP;
}

int upload_net(kmpm_surrogate *s, int net) {
  MlpNet &h = s->host[net];
  for (int r = 0; r < MLP_H; r++)
    for (int c = 0; c < MLP_H; c++) {
      // round to nearest even BF16 (host copy of the device rounding)
      uint32_t b;
      std::memcpy(&b, &h.p.w2[r][c], 4);
      b += 0x7FFFu + ((b >> 16) & 1u);
      h.w2_bf16[r][c] = (uint16_t)(b >> 16);
    }
  SCHK(hipMemcpyAsync(s->d_net + net, &h, sizeof(MlpNet), hipMemcpyHostToDevice, s->stream));
  SCHK(hipStreamSynchronize(s->stream));
  return KMPM_OK;
}

int download_net(kmpm_surrogate *s, int net) {
  SCHK(hipMemcpyAsync(&s->host[net], s->d_net + net, sizeof(MlpNet), hipMemcpyDeviceToHost,
                      s->stream));
  SCHK(hipStreamSynchronize(s->stream));
  return KMPM_OK;
}

// One fused forward/backward pass over tuples [X, Y) of length n; sums into
// s->d_out (gradient, then loss).  adam != 0 also steps the parameters.
int grad_pass(kmpm_surrogate *s, int net, const float *X, const float *Y, int64_t n, int adam,
              float lr, int64_t t) {
  const int nb = (int)((n + 255) / 256);
  int st = ensure_partial(s, nb);
  if (st) return st;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(256), 0, s->stream, s->d_net + net, X, Y, n,
                     1.0f / (float)n, s->d_partial);
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f / (1.0f - powf(b1, (float)t)), c2 = 1.0f / (1.0f - powf(b2, (float)t));
  hipLaunchKernelGGL(k_mlp_reduce, dim3(5), dim3(256), 0, s->stream, s->d_net + net,
                     s->d_partial, nb, s->d_out, s->d_mom + (size_t)net * (MLP_NPARAM + 1),
                     s->d_vel + (size_t)net * (MLP_NPARAM + 1), adam, lr, b1, b2, 1e-8f, c1, c2);
  return hipGetLastError() == hipSuccess ? KMPM_OK : KMPM_ERR_HIP;
}

// Least-squares quadratic baseline of the targets in the network inputs
// (normal equations of the 6 monomials, solved by Gaussian elimination with
// partial pivoting in double).
bool fit_baseline(MlpNet &h, const std::vector<float> &x, const std::vector<float> &y) {
  double M[6][7] = {};
  for (size_t i = 0; i < y.size(); i++) {
    const double x0 = std::fma(x[2 * i], h.in_scale[0], h.in_shift[0]);
    const double x1 = std::fma(x[2 * i + 1], h.in_scale[1], h.in_shift[1]);
    const double m[6] = {1, x0, x1, x0 * x0, x0 * x1, x1 * x1};
    for (int a = 0; a < 6; a++) {
      for (int b = 0; b < 6; b++) M[a][b] += m[a] * m[b];
      M[a][6] += m[a] * y[i];
    }
  }
  for (int c = 0; c < 6; c++) {
    int piv = c;
    for (int r = c + 1; r < 6; r++)
      if (std::fabs(M[r][c]) > std::fabs(M[piv][c])) piv = r;
    if (std::fabs(M[piv][c]) < 1e-300) {   // a constant feature: drop its monomials
      for (int r = 0; r < 6; r++) M[r][c] = 0;
      M[c][c] = 1;
      M[c][6] = 0;
      continue;
    }
    std::swap(M[c], M[piv]);
    for (int r = 0; r < 6; r++) {
      if (r == c) continue;
      const double f = M[r][c] / M[c][c];
      for (int k = c; k < 7; k++) M[r][k] -= f * M[c][k];
    }
  }
  for (int c = 0; c < 6; c++) {
    const double v = M[c][6] / M[c][c];
    if (!std::isfinite(v)) return false;
    h.base[c] = (float)v;
  }
  return true;
}

void init_params(MlpParams &p, uint64_t seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  const double a1 = std::sqrt(6.0 / (MLP_IN + MLP_H)), a2 = std::sqrt(6.0 / (2 * MLP_H)),
               a3 = std::sqrt(6.0 / (MLP_H + 1));
  for (int k = 0; k < MLP_H; k++) {
    for (int j = 0; j < MLP_IN; j++) p.w1[k][j] = (float)(a1 * U(rng));
    p.b1[k] = 0;
    for (int j = 0; j < MLP_H; j++) p.w2[k][j] = (float)(a2 * U(rng));
    p.b2[k] = 0;
    p.w3[k] = (float)(a3 * U(rng));
  }
  p.b3 = 0;
}

}  // namespace

extern "C" {

int kmpm_surrogate_create(const kmpm_material *mat, const kmpm_surrogate_config *cfg,
                          kmpm_surrogate **out) {
  if (!mat || !cfg || !out) return KMPM_ERR_ARG;
  *out = nullptr;
  if (mat->type != KMPM_MAT_EOS_STRENGTH || mat->strength != KMPM_STRENGTH_JOHNSON_COOK)
    return KMPM_ERR_ARG;
  if (!(cfg->eps_p_max > EPS_P_MIN) || !(cfg->rate_max >= 1.0) || !(cfg->triax_max > -3.0) ||
      cfg->samples < 256 || cfg->batch < 256 || cfg->batch % 256 || cfg->iterations < 0 ||
      !(cfg->learning_rate > 0))
    return KMPM_ERR_ARG;
  const double A = mat->strength_p[1], B = mat->strength_p[2], nexp = mat->strength_p[3],
               C = mat->strength_p[5];
  if (!(A > 0) || !(mat->strength_p[4] > 0)) return KMPM_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return KMPM_ERR_NO_DEVICE;
  kmpm_surrogate *s = new kmpm_surrogate();
  s->cfg = *cfg;
  s->mat = *mat;
  s->dmat = make_devmat(*mat);
  const double lnr = std::log(cfg->rate_max);
  s->lo[0][0] = std::sqrt(std::sqrt(EPS_P_MIN));
  s->hi[0][0] = std::sqrt(std::sqrt(cfg->eps_p_max));
  s->lo[0][1] = 0;
  s->hi[0][1] = lnr;
  s->lo[1][0] = -3.0;
  s->hi[1][0] = cfg->triax_max;
  s->lo[1][1] = 0;
  s->hi[1][1] = mat->damage == KMPM_DAMAGE_JOHNSON_COOK && mat->damage_p[5] > 0
                    ? std::log(cfg->rate_max * mat->strength_p[4] / mat->damage_p[5]) : 0;
  if (s->hi[1][1] < 0) s->hi[1][1] = 0;
  s->q_lo = s->lo[0][0];
  s->q_hi = s->hi[0][0];
  // trial von Mises of the tuple rate, 4x the largest flow stress of the envelope
  const double ys_max = (A + B * std::pow(cfg->eps_p_max, nexp)) * std::pow(1.0 + cfg->rate_max, C);
  const double G = mat->strength_p[0];
  s->d_rate = 4.0 * ys_max / (SQRT_3_OVER_2 * std::sqrt(2.0) * 2.0 * G);
  s->sy_virgin = A * (C != 0 ? std::pow(2.0, C) : 1.0) * s->dmat.jc_thermal;
  bool ok = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&s->d_net, 2 * sizeof(MlpNet)) == hipSuccess &&
            hipMalloc(&s->d_mom, 2 * (MLP_NPARAM + 1) * sizeof(float)) == hipSuccess &&
            hipMalloc(&s->d_vel, 2 * (MLP_NPARAM + 1) * sizeof(float)) == hipSuccess &&
            hipMalloc(&s->d_out, (MLP_NPARAM + 1) * sizeof(float)) == hipSuccess;
  for (int k = 0; k < 2 && ok; k++)
    ok = hipMalloc(&s->d_X[k], 2 * cfg->samples * sizeof(float)) == hipSuccess &&
         hipMalloc(&s->d_Y[k], cfg->samples * sizeof(float)) == hipSuccess;
  if (!ok) {
    kmpm_surrogate_destroy(s);
    return KMPM_ERR_HIP;
  }
  // tuples, normalisation and initial parameters
  for (int k = 0; k < 2; k++) {
    int st = gen_tuples(s, k, cfg->samples, cfg->seed + 17 * k, s->d_X[k], s->d_Y[k]);
    if (st) { kmpm_surrogate_destroy(s); return st; }
    std::vector<float> y(cfg->samples);
    if (hipMemcpyAsync(y.data(), s->d_Y[k], y.size() * sizeof(float), hipMemcpyDeviceToHost,
                       s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess) {
      kmpm_surrogate_destroy(s);
      return KMPM_ERR_HIP;
    }
    std::vector<float> x(2 * cfg->samples);
    if (hipMemcpyAsync(x.data(), s->d_X[k], x.size() * sizeof(float), hipMemcpyDeviceToHost,
                       s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess) {
      kmpm_surrogate_destroy(s);
      return KMPM_ERR_HIP;
    }
    MlpNet &h = s->host[k];
    std::memset(&h, 0, sizeof(h));
    init_params(h.p, cfg->seed * 31 + k);
    for (int j = 0; j < 2; j++) {
      const double w = s->hi[k][j] - s->lo[k][j];
      h.in_scale[j] = (float)(w > 0 ? 2.0 / w : 0.0);
      h.in_shift[j] = (float)(w > 0 ? -(s->hi[k][j] + s->lo[k][j]) / w : 0.0);
    }
    for (float v : y)
      if (!std::isfinite(v)) { kmpm_surrogate_destroy(s); return KMPM_ERR_ARG; }
    if (!fit_baseline(h, x, y)) { kmpm_surrogate_destroy(s); return KMPM_ERR_ARG; }
    // residual statistics
    double sum = 0, sq = 0;
    for (size_t i = 0; i < y.size(); i++) {
      const float x0 = std::fma(x[2 * i], h.in_scale[0], h.in_shift[0]);
      const float x1 = std::fma(x[2 * i + 1], h.in_scale[1], h.in_shift[1]);
      const double r = (double)y[i] - mlp_base(h, x0, x1);
      sum += r;
      sq += r * r;
    }
    const double mean = sum / y.size();
    const double sd = std::sqrt(std::max(sq / y.size() - mean * mean, 1e-30));
    h.out_scale = (float)sd;
    h.out_shift = (float)mean;
    if (upload_net(s, k)) { kmpm_surrogate_destroy(s); return KMPM_ERR_HIP; }
  }
  if (hipMemsetAsync(s->d_mom, 0, 2 * (MLP_NPARAM + 1) * sizeof(float), s->stream) != hipSuccess ||
      hipMemsetAsync(s->d_vel, 0, 2 * (MLP_NPARAM + 1) * sizeof(float), s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess) {
    kmpm_surrogate_destroy(s);
    return KMPM_ERR_HIP;
  }
  *out = s;
  return KMPM_OK;
}

int kmpm_surrogate_destroy(kmpm_surrogate *s) {
  if (!s) return KMPM_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  void *ptrs[] = {s->d_net, s->d_X[0], s->d_X[1], s->d_Y[0], s->d_Y[1], s->d_mom, s->d_vel,
                  s->d_partial, s->d_out};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return KMPM_OK;
}

int kmpm_surrogate_train(kmpm_surrogate *s, double final_loss[2]) {
  if (!s) return KMPM_ERR_ARG;
  const int64_t n = s->cfg.samples, B = std::min<int64_t>(s->cfg.batch, n - n % 256);
  const int64_t nbatch = n / B;
  for (int k = 0; k < 2; k++) {
    for (int64_t t = 1; t <= s->cfg.iterations; t++) {
      const int64_t b = (t - 1) % nbatch;
      // cosine decay from the initial step to 1% of it
      const double frac = (double)(t - 1) / std::max<int64_t>(1, s->cfg.iterations - 1);
      const float lr =
          (float)(s->cfg.learning_rate * (0.01 + 0.99 * 0.5 * (1 + std::cos(3.141592653589793 * frac))));
      int st = grad_pass(s, k, s->d_X[k] + 2 * b * B, s->d_Y[k] + b * B, B, 1, lr, t);
      if (st) return st;
    }
    // loss over the whole tuple set with the final parameters
    int st = grad_pass(s, k, s->d_X[k], s->d_Y[k], n, 0, 0.0f, 1);
    if (st) return st;
    float loss = 0;
    SCHK(hipMemcpyAsync(&loss, s->d_out + MLP_NPARAM, sizeof(float), hipMemcpyDeviceToHost,
                        s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    if (final_loss) final_loss[k] = loss;
    if ((st = download_net(s, k))) return st;
  }
  return KMPM_OK;
}

int kmpm_surrogate_tuples(kmpm_surrogate *s, int net, int64_t n, uint64_t seed, float *features,
                          float *targets) {
  if (!s || net < 0 || net > 1 || n < 1 || !features || !targets) return KMPM_ERR_ARG;
  float *dX = nullptr, *dY = nullptr;
  SCHK(hipMalloc(&dX, 2 * n * sizeof(float)));
  if (hipMalloc(&dY, n * sizeof(float)) != hipSuccess) {
    (void)hipFree(dX);
    return KMPM_ERR_HIP;
  }
  int st = gen_tuples(s, net, n, seed, dX, dY);
  if (!st && (hipMemcpyAsync(features, dX, 2 * n * sizeof(float), hipMemcpyDeviceToHost, s->stream) !=
                  hipSuccess ||
              hipMemcpyAsync(targets, dY, n * sizeof(float), hipMemcpyDeviceToHost, s->stream) !=
                  hipSuccess ||
              hipStreamSynchronize(s->stream) != hipSuccess))
    st = KMPM_ERR_HIP;
  (void)hipFree(dX);
  (void)hipFree(dY);
  return st;
}

int kmpm_surrogate_tuples_from_ctx(kmpm_surrogate *s, kmpm_ctx *ctx, int net, int64_t stride,
                                   float *features, float *targets, int64_t n) {
  if (!s || !ctx || net < 0 || net > 1 || stride < 1 || n < 0) return KMPM_ERR_ARG;
  const int64_t np = ctx->a.np;
  if (n != (np + stride - 1) / stride || (n > 0 && (!features || !targets))) return KMPM_ERR_ARG;
  if (n == 0) return KMPM_OK;
  SCHK(hipStreamSynchronize(ctx->stream));   // the state the context's steps left
  float *dX = nullptr, *dY = nullptr;
  SCHK(hipMalloc(&dX, 2 * n * sizeof(float)));
  if (hipMalloc(&dY, n * sizeof(float)) != hipSuccess) {
    (void)hipFree(dX);
    return KMPM_ERR_HIP;
  }
  hipLaunchKernelGGL(k_ctx_tuples, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->stream,
                     ctx->a, s->dmat, net, stride, n, s->d_rate, dX, dY);
  int st = hipGetLastError() == hipSuccess ? KMPM_OK : KMPM_ERR_HIP;
  if (!st && (hipMemcpyAsync(features, dX, 2 * n * sizeof(float), hipMemcpyDeviceToHost, s->stream) !=
                  hipSuccess ||
              hipMemcpyAsync(targets, dY, n * sizeof(float), hipMemcpyDeviceToHost, s->stream) !=
                  hipSuccess ||
              hipStreamSynchronize(s->stream) != hipSuccess))
    st = KMPM_ERR_HIP;
  (void)hipFree(dX);
  (void)hipFree(dY);
  return st;
}

int kmpm_surrogate_set_tuples(kmpm_surrogate *s, int net, int64_t offset, const float *features,
                              const float *targets, int64_t n) {
  if (!s || net < 0 || net > 1 || offset < 0 || n < 0 || offset + n > s->cfg.samples)
    return KMPM_ERR_ARG;
  if (n == 0) return KMPM_OK;
  if (!features || !targets) return KMPM_ERR_ARG;
  for (int64_t i = 0; i < n; i++)
    if (!std::isfinite(features[2 * i]) || !std::isfinite(features[2 * i + 1]) ||
        !std::isfinite(targets[i]))
      return KMPM_ERR_ARG;
  SCHK(hipMemcpyAsync(s->d_X[net] + 2 * offset, features, 2 * n * sizeof(float),
                      hipMemcpyHostToDevice, s->stream));
  SCHK(hipMemcpyAsync(s->d_Y[net] + offset, targets, n * sizeof(float), hipMemcpyHostToDevice,
                      s->stream));
  SCHK(hipStreamSynchronize(s->stream));
  return KMPM_OK;
}

int kmpm_surrogate_eval(kmpm_surrogate *s, int net, const float *features, int64_t n, float *out) {
  if (!s || net < 0 || net > 1 || n < 1 || !features || !out) return KMPM_ERR_ARG;
  float *dX = nullptr, *dO = nullptr;
  SCHK(hipMalloc(&dX, 2 * n * sizeof(float)));
  if (hipMalloc(&dO, n * sizeof(float)) != hipSuccess) {
    (void)hipFree(dX);
    return KMPM_ERR_HIP;
  }
  int st = KMPM_OK;
  if (hipMemcpyAsync(dX, features, 2 * n * sizeof(float), hipMemcpyHostToDevice, s->stream) !=
      hipSuccess)
    st = KMPM_ERR_HIP;
  if (!st) {
    hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->stream,
                       s->d_net + net, dX, n, dO);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out, dO, n * sizeof(float), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
      st = KMPM_ERR_HIP;
  }
  (void)hipFree(dX);
  (void)hipFree(dO);
  return st;
}

int kmpm_surrogate_gradient(kmpm_surrogate *s, int net, const float *features,
                            const float *targets, int64_t n, float *grad, double *loss) {
  if (!s || net < 0 || net > 1 || n < 1 || !features || !targets || !grad) return KMPM_ERR_ARG;
  float *dX = nullptr, *dY = nullptr;
  SCHK(hipMalloc(&dX, 2 * n * sizeof(float)));
  if (hipMalloc(&dY, n * sizeof(float)) != hipSuccess) {
    (void)hipFree(dX);
    return KMPM_ERR_HIP;
  }
  int st = KMPM_OK;
  if (hipMemcpyAsync(dX, features, 2 * n * sizeof(float), hipMemcpyHostToDevice, s->stream) !=
          hipSuccess ||
      hipMemcpyAsync(dY, targets, n * sizeof(float), hipMemcpyHostToDevice, s->stream) != hipSuccess)
    st = KMPM_ERR_HIP;
  if (!st) st = grad_pass(s, net, dX, dY, n, 0, 0.0f, 1);
  std::vector<float> h(MLP_NPARAM + 1);
  if (!st && (hipMemcpyAsync(h.data(), s->d_out, h.size() * sizeof(float), hipMemcpyDeviceToHost,
                             s->stream) != hipSuccess ||
              hipStreamSynchronize(s->stream) != hipSuccess))
    st = KMPM_ERR_HIP;
  (void)hipFree(dX);
  (void)hipFree(dY);
  if (st) return st;
  std::memcpy(grad, h.data(), MLP_NPARAM * sizeof(float));
  if (loss) *loss = h[MLP_NPARAM];
  return KMPM_OK;
}

int kmpm_surrogate_params(kmpm_surrogate *s, int net, float *params, int nparam, int set) {
  if (!s || net < 0 || net > 1 || !params || nparam != MLP_NPARAM) return KMPM_ERR_ARG;
  if (set) {
    std::memcpy(&s->host[net].p, params, sizeof(MlpParams));
    return upload_net(s, net);
  }
  int st = download_net(s, net);
  if (st) return st;
  std::memcpy(params, &s->host[net].p, sizeof(MlpParams));
  return KMPM_OK;
}

int kmpm_surrogate_normalization(kmpm_surrogate *s, int net, double map[12]) {
  if (!s || net < 0 || net > 1 || !map) return KMPM_ERR_ARG;
  const MlpNet &h = s->host[net];
  map[0] = h.in_scale[0];
  map[1] = h.in_scale[1];
  map[2] = h.in_shift[0];
  map[3] = h.in_shift[1];
  map[4] = h.out_scale;
  map[5] = h.out_shift;
  for (int c = 0; c < 6; c++) map[6 + c] = h.base[c];
  return KMPM_OK;
}

int kmpm_surrogate_envelope(kmpm_surrogate *s, double lo[4], double hi[4]) {
  if (!s || !lo || !hi) return KMPM_ERR_ARG;
  for (int k = 0; k < 2; k++)
    for (int j = 0; j < 2; j++) {
      lo[2 * k + j] = s->lo[k][j];
      hi[2 * k + j] = s->hi[k][j];
    }
  return KMPM_OK;
}

}  // extern "C"

// Device image of a trained surrogate for the step kernels.
namespace kmpm {
int surrogate_device_image(const kmpm_surrogate *s, SurrogateDev *out) {
  if (!s) return KMPM_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  out->y = s->host[0];
  out->f = s->host[1];
  out->enabled = 1;
  out->ep_hi = s->cfg.eps_p_max;
  out->lnr_y_hi = s->hi[0][1];
  out->triax_hi = s->hi[1][0];
  out->lnr_f_hi = s->hi[1][1];
  out->sy_virgin = s->sy_virgin;
  return KMPM_OK;
}
bool surrogate_matches(const kmpm_surrogate *s, const kmpm_material &m) {
  return std::memcmp(s->mat.strength_p, m.strength_p, sizeof(m.strength_p)) == 0 &&
         std::memcmp(s->mat.damage_p, m.damage_p, sizeof(m.damage_p)) == 0 &&
         s->mat.damage == m.damage && s->mat.strength == m.strength;
}
}  // namespace kmpm

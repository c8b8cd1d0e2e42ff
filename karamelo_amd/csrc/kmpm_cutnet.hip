// kmpm_cutnet.hip -- the cut network (kmpm_cutnet.h): training tuples from
// the plain strength_jc / damage_jc functions, BF16-MFMA training (forward
// and backward, FP32 accumulation, Adam), evaluation, and its use in the
// USL step as the constitutive update of the JC + shock + JC damage class
// on the mixed path, with the plain FP64 functions as fallback outside the
// training envelope and as the judge of its error.
#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "kmpm_cutnet.h"
#include "kmpm_kernels.h"

namespace kmpm {

// ---- training tuples ---------------------------------------------------------

__device__ inline double cn_u01(uint64_t seed, uint64_t i, uint64_t k) {
  // splitmix64 of (seed, sample, stream) -> [0, 1)
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (2 * i + 1) + 0xD1B54A32D192ED03ull * (k + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__device__ inline double cn_normal(uint64_t seed, uint64_t i, uint64_t k) {
  // Box-Muller from streams 2k, 2k + 1
  const double u1 = 1.0 - cn_u01(seed, i, 2 * k), u2 = cn_u01(seed, i, 2 * k + 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// A random deviatoric direction of unit von Mises norm (streams 2k .. 2k + 11).
__device__ inline void cn_unit_dev(uint64_t seed, uint64_t i, uint64_t k, double *d) {
  for (int e = 0; e < 6; e++) d[e] = cn_normal(seed, i, k + e);
  const double m = (d[0] + d[1] + d[2]) / 3.0;
  for (int e = 0; e < 3; e++) d[e] -= m;
  const double vm = SQRT_3_OVER_2 * norm_sym(d);
  for (int e = 0; e < 6; e++) d[e] /= vm;
}

// How the states of the training tuples are sampled.
struct CnSampling {
  double ep_max;                 // eps_p^(1/4) uniform in [0, ep_max^(1/4)] (5%: 0)
  double ratio_lo, ratio_max;    // epsdot_p / epsdot0 log-uniform
  double dmg_max;                // damage: 80% 0, else uniform in [0, dmg_max)
  double dinit_max;              // damage_init uniform in [0, dinit_max]
  double p_max;                  // pH / A uniform in [-p_max, p_max]
  double inc_lo, inc_max;        // 2 G dt vm(dev D) / A log-uniform
  double dt, tav;                // time step, and the eps_p filter's 1000 h / c_signal
};

// The trial deviator t = dev(sig + 2 Gd dt D) and its von Mises stress, in
// the order of strength_jc.cpp:146-157.
template <class T>
__host__ __device__ inline T cn_trial(T G, T dt, T dmg, const T *sig, const T *D, T *t) {
  const T Gd = dmg > T(0) ? G * (T(1) - dmg) : G;
  const T c = dt * T(2) * Gd;
  for (int e = 0; e < 6; e++) t[e] = sig[e] + c * D[e];
  const T t3 = (t[0] + t[1] + t[2]) / T(3);
  t[0] -= t3;
  t[1] -= t3;
  t[2] -= t3;
  T n2 = t[0] * t[0];
  n2 += t[3] * t[3];
  n2 += t[4] * t[4];
  n2 += t[3] * t[3];
  n2 += t[1] * t[1];
  n2 += t[5] * t[5];
  n2 += t[4] * t[4];
  n2 += t[5] * t[5];
  n2 += t[2] * t[2];
  return T(SQRT_3_OVER_2) * sqrt(n2);
}

// The features of a state (kmpm_cutnet.h) from its trial deviator t and
// von Mises stress q; D the rate of deformation (xx yy zz xy xz yz).
template <class T>
__host__ __device__ inline void cn_features(T A, T G, T epsdot0, T dt, const T *t, T q,
                                            const T *D, T ep, T epdot, T dmg, T dinit, T pH,
                                            float *f) {
  const T d3 = (D[0] + D[1] + D[2]) / T(3);
  const T k = T(2) * G * dt / A;
  for (int e = 0; e < 6; e++) {
    f[e] = (float)(t[e] / A);
    f[6 + e] = (float)(k * (D[e] - (e < 3 ? d3 : T(0))));
  }
  T r = epdot / epsdot0;
  r = r > T(1) ? r : T(1);
  f[12] = (float)sqrt(sqrt(ep));
  f[13] = (float)log(r);
  f[14] = (float)dmg;
  f[15] = (float)dinit;
  f[16] = (float)(pH / A);
  f[17] = (float)(q / A);
}

// One tuple: a sampled state, the plain functions of the cut at it
// (strength_jc.cpp:100-184, the eps_p filter of solid.cpp:1319-1329,
// damage_jc.cpp:91-143) and its targets.
__global__ void k_cn_tuples(DevMat m, CnSampling S, uint64_t seed, int64_t n, float *F,
                            float *Y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double G = m.strength_p[0], A = m.strength_p[1], B = m.strength_p[2],
               nn = m.strength_p[3], e0 = m.strength_p[4], C = m.strength_p[5];
  double ep = 0.0;
  if (cn_u01(seed, i, 0) >= 0.05) {
    const double q = cn_u01(seed, i, 1) * sqrt(sqrt(S.ep_max));
    ep = (q * q) * (q * q);
  }
  const double llo = log(S.ratio_lo), lhi = log(S.ratio_max);
  const double epdot = e0 * exp(llo + (lhi - llo) * cn_u01(seed, i, 2));
  const double dmg = cn_u01(seed, i, 3) < 0.8 ? 0.0 : S.dmg_max * cn_u01(seed, i, 4);
  const double dinit = S.dinit_max * cn_u01(seed, i, 5);
  // the state's yield stress sets the scale of its stress
  double ratio = epdot / e0;
  ratio = ratio > 1.0 ? ratio : 1.0;
  double ys = ep < 1.0e-10 ? A : A + B * pow_pos(ep, nn);     // as strength_jc.cpp:115-133
  if (C != 0) ys *= pow_pos(1.0 + ratio, C);
  ys *= m.jc_thermal;
  if (dmg > 0) ys *= 1.0 - dmg;
  double sig[6];
  cn_unit_dev(seed, i, 8, sig);                  // streams 16..27
  const double rho = cn_u01(seed, i, 6) < 0.6 ? 1.0 - 0.03 * fabs(cn_normal(seed, i, 20))
                                              : 1.05 * cn_u01(seed, i, 7);
  for (int e = 0; e < 6; e++) sig[e] *= rho * ys;
  double dd[6];
  cn_unit_dev(seed, i, 24, dd);                  // streams 48..59
  const double ilo = log(S.inc_lo), ihi = log(S.inc_max);
  const double inc = exp(ilo + (ihi - ilo) * cn_u01(seed, i, 60));
  const double kd = inc * A / (2.0 * G);         // vm of dev(dt D)
  const double vol = (2.0 * cn_u01(seed, i, 61) - 1.0) * kd;
  double D[6];
  for (int e = 0; e < 6; e++) D[e] = (kd * dd[e] + (e < 3 ? vol / 3.0 : 0.0)) / S.dt;
  const double pH = (2.0 * cn_u01(seed, i, 62) - 1.0) * S.p_max * A;
  // the plain functions
  double out[6], dps = 0.0;
  jc_strength(m, sig, D, S.dt, ep, epdot, dmg, out, &dps);
  double epd = epdot - epdot * S.dt / S.tav + dps / S.tav;
  epd = epd > 0.0 ? epd : 0.0;
  double di = dinit, dm = dmg;
  bool reset = false;
  if (m.damage == KMPM_DAMAGE_JOHNSON_COOK) {
    const double vm = SQRT_3_OVER_2 * norm_sym(out);
    const double triax = (pH != 0.0 && vm != 0.0) ? -pH / (vm + 0.001 * fabs(pH)) : 0.0;
    reset = !(dps == 0 && dmg >= 1.0) && triax <= -3;
    jc_damage(m, di, dm, pH, out, epd, dps);
  }
  double t[6];
  const double q = cn_trial(G, S.dt, dmg, sig, D, t);
  float *f = F + i * CN_NI;
  cn_features<double>(A, G, e0, S.dt, t, q, D, ep, epdot, dmg, dinit, pH, f);
  float *y = Y + i * CN_NO;
  for (int e = 0; e < 6; e++) y[e] = (float)(out[e] / A);
  // the signed plastic increment: 3 G dps / A where plastic, the yield
  // function G (q - sigma_y) / (Gd A) on the elastic side
  const double Gd = dmg > 0 ? G * (1.0 - dmg) : G;
  y[6] = (float)(dps > 0 ? 3.0 * G * dps / A : G * (q - ys) / (Gd * A));
  // the reset of damage_init is applied exactly by the decoder
  y[7] = reset ? 0.0f : (float)(3.0 * G * (di - dinit) / A);
  y[8] = (float)dm;
}

// ---- weights: MFMA operand images from the flat parameters -----------------

__device__ inline uint16_t cn_pbits(const float *p, int idx) { return cn_bf16_bits(p[idx]); }
// output k of the last layer whose physical row is rho, or -1
__device__ inline int cn_out_of_row(int rho) {
  const int k = (rho / 8) * 4 + rho % 8;
  return (rho % 8 < 4 && k < CN_NO) ? k : -1;
}

__global__ void k_cn_pack(const float *p, CutNetDev *net, CutNetT *nt) {
  constexpr int N_W1 = 2 * 3 * 64, N_WH = 4 * 2 * 4 * 64, N_W6 = 4 * 64, N_TW6 = 2 * 2 * 64;
  constexpr int N_SLOTS = N_W1 + N_WH + N_W6 + N_WH + N_TW6;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N_SLOTS + 5 * CN_NH + 16) {
    return;
  } else if (q >= N_SLOTS) {   // biases
    const int b = q - N_SLOTS;
    if (b < 5 * CN_NH) {
      const int l = b / CN_NH, u = b % CN_NH;
      net->b[l][u] = p[l == 0 ? CN_OB1 + u : cn_ob(l + 1) + u];
    } else {
      const int k = b - 5 * CN_NH;
      net->b6[k] = k < CN_NO ? p[CN_OB6 + k] : 0.0f;
    }
    return;
  }
  int r = q;
  Bf8 v;
  if (r < N_W1) {
    const int lane = r % 64, s = (r / 64) % 3, R = r / 192;
    const int h = lane / 32, c = lane % 32, o = 32 * R + c;
    for (int i = 0; i < 8; i++) {
      const int kk = 16 * s + 8 * h + i;
      v.v[i] = kk < CN_NI ? cn_pbits(p, CN_OW1 + o * CN_NI + kk)
               : kk < 2 * CN_NI ? cn_pbits(p, CN_OW1 + o * CN_NI + kk - CN_NI) : (uint16_t)0;
    }
    net->w.w1[R][s][lane] = v;
    return;
  }
  r -= N_W1;
  if (r < N_WH) {
    const int lane = r % 64, s = (r / 64) % 4, R = (r / 256) % 2, l = r / 512;
    const int h = lane / 32, c = lane % 32, o = 32 * R + c;
    for (int i = 0; i < 8; i++) v.v[i] = cn_pbits(p, cn_ow(l + 2) + o * CN_NH + cn_u(s, h, i));
    net->w.wh[l][R][s][lane] = v;
    return;
  }
  r -= N_WH;
  if (r < N_W6) {
    const int lane = r % 64, s = r / 64;
    const int h = lane / 32, c = lane % 32, k = cn_out_of_row(c);
    for (int i = 0; i < 8; i++)
      v.v[i] = k < 0 ? (uint16_t)0 : cn_pbits(p, CN_OW6 + k * CN_NH + cn_u(s, h, i));
    net->w.w6[s][lane] = v;
    return;
  }
  r -= N_W6;
  if (r < N_WH) {
    const int lane = r % 64, s = (r / 64) % 4, Rn = (r / 256) % 2, l = r / 512;
    const int h = lane / 32, c = lane % 32, in = 32 * Rn + c;
    for (int i = 0; i < 8; i++) v.v[i] = cn_pbits(p, cn_ow(l + 2) + cn_u(s, h, i) * CN_NH + in);
    nt->wh[l][Rn][s][lane] = v;
    return;
  }
  r -= N_WH;
  {
    const int lane = r % 64, s = (r / 64) % 2, Rn = r / 128;
    const int h = lane / 32, c = lane % 32, in = 32 * Rn + c;
    for (int i = 0; i < 8; i++) {
      const int k = cn_out_of_row(cn_r(8 * s + i, h));
      v.v[i] = k < 0 ? (uint16_t)0 : cn_pbits(p, CN_OW6 + k * CN_NH + in);
    }
    nt->w6[Rn][s][lane] = v;
  }
}
constexpr int CN_PACK_THREADS = 2 * 3 * 64 + 2 * (4 * 2 * 4 * 64) + 4 * 64 + 2 * 2 * 64 +
                                5 * CN_NH + 16;

// ---- evaluation ---------------------------------------------------------------

__device__ inline void cn_copy_to_lds(const CutNetDev *src, CutNetDev *dst) {
  const uint4 *s = reinterpret_cast<const uint4 *>(src);
  uint4 *d = reinterpret_cast<uint4 *>(dst);
  for (int q = threadIdx.x; q < (int)(sizeof(CutNetDev) / 16); q += blockDim.x) d[q] = s[q];
}
static_assert(sizeof(CutNetDev) % 16 == 0, "16-byte copies");

__global__ __launch_bounds__(256, 2) void k_cn_eval(const CutNetDev *net, const float *F, int64_t n,
                                                 float *Y) {
  __shared__ CutNetDev s_net;
  __shared__ __attribute__((aligned(16))) uint16_t stage[4][64 * CN_KE];
  cn_copy_to_lds(net, &s_net);
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + threadIdx.x;
    const bool live = i < n;
    float f[CN_NI], y[CN_NO];
    for (int k = 0; k < CN_NI; k++) f[k] = live ? F[i * CN_NI + k] : 0.0f;
    cn_eval_wave(s_net, f, y, stage[wave]);
    if (live)
      for (int k = 0; k < CN_NO; k++) Y[i * CN_NO + k] = y[k];
  }
}

// ---- training: fused forward + backward ---------------------------------------
//
// A wave takes 32 tuples (one column tile) at a time; the lower half of the
// wave loads them.  The hidden pre-activations stay in registers for the
// backward pass.  Per layer l (from the last):
//   weight grad  dW_l[o][u] = sum_p d_l[o][p] h_{l-1}[u][p]   (K = tuples)
//   input grad   d_{l-1}     = (W_l^T d_l) * gelu'(z_{l-1})
// d_l and h_{l-1} are staged in LDS as [unit][tuple] BF16 for the weight
// product; the input-gradient product takes d_l as its B operand straight
// from the accumulators, against the transposed weight images (CutNetT).
// The block's gradient sums stay in LDS and go to partial[block] at the end.
constexpr int CN_GRAD_WAVES = 4;

__device__ inline void cn_wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over the 32 tuples of a half-wave (lanes with the same h).
__device__ inline float cn_half_sum(float v) {
#pragma unroll
  for (int o = 1; o < 32; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Stage the 64 x 32 values of a pair of accumulator tiles as BF16 [unit][tuple].
__device__ inline void cn_stage_t(uint16_t *dst, const f32x16_t *t, int ntiles) {
  const int lane = threadIdx.x & 63, h = lane >> 5, c = lane & 31;
  for (int R = 0; R < ntiles; R++)
#pragma unroll
    for (int e = 0; e < 16; e++) dst[(32 * R + cn_r(e, h)) * 32 + c] = cn_bf16_bits(t[R][e]);
}

// dW[o][u] (+)= sum_p A[o][p] B[u][p] for the staged [unit][tuple] tiles:
// out tile Rm of A (rows >= arows read as 0), in tile Rn of B.
__device__ inline f32x16_t cn_wgrad_tile(const uint16_t *sa, const uint16_t *sb, int Rm, int Rn,
                                         int arows) {
  const int lane = threadIdx.x & 63, h = lane >> 5, c = lane & 31;
  f32x16_t acc = {};
#pragma unroll
  for (int s = 0; s < 2; s++) {
    const int ra = 32 * Rm + c;
    bf16x8_t av = {}, bv;
    if (ra < arows) av = cn_ld(*reinterpret_cast<const Bf8 *>(sa + ra * 32 + 16 * s + 8 * h));
    bv = cn_ld(*reinterpret_cast<const Bf8 *>(sb + (32 * Rn + c) * 32 + 16 * s + 8 * h));
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
  }
  return acc;
}

__global__ __launch_bounds__(256, 1) void k_cn_grad(const CutNetDev *net, const CutNetT *nt,
                                                    const float *F, const float *Y, int64_t n,
                                                    float inv_n, float *partial) {
  __shared__ float gacc[CN_NPARAM + 1];
  __shared__ __attribute__((aligned(16))) uint16_t sx[CN_GRAD_WAVES][32 * CN_KE];   // [tuple][kk]
  __shared__ __attribute__((aligned(16))) uint16_t sxt[CN_GRAD_WAVES][64 * 32];     // [kk][tuple]
  __shared__ __attribute__((aligned(16))) uint16_t sh[CN_GRAD_WAVES][64 * 32];      // [unit][tuple]
  __shared__ __attribute__((aligned(16))) uint16_t sd[CN_GRAD_WAVES][64 * 32];      // [unit][tuple]
  const CutNetDev &N = *net;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, h = lane >> 5, c = lane & 31;
  for (int q = threadIdx.x; q <= CN_NPARAM; q += blockDim.x) gacc[q] = 0.0f;
  for (int q = lane; q < 64 * 32; q += 64) sxt[wave][q] = 0;   // rows >= 34 stay zero
  __syncthreads();
  uint16_t *mx = sx[wave], *mxt = sxt[wave], *mh = sh[wave], *md = sd[wave];
  float loss = 0.0f;
  for (int64_t b0 = ((int64_t)blockIdx.x * CN_GRAD_WAVES + wave) * 32; b0 < n;
       b0 += (int64_t)gridDim.x * CN_GRAD_WAVES * 32) {
    const int64_t smp = b0 + c;
    const bool valid = smp < n;
    float t[CN_NO];
    if (h == 0) {
      float f[CN_NI], x[CN_NI];
      for (int k = 0; k < CN_NI; k++) f[k] = valid ? F[smp * CN_NI + k] : 0.0f;
      cn_inputs(N, f, x);
      cn_encode_row(x, mx + c * CN_KE);
      for (int k = 0; k < 2 * CN_NI; k++) mxt[k * 32 + c] = mx[c * CN_KE + k];
      for (int k = 0; k < CN_NO; k++) {
        float acc = N.skip[k][CN_NI];
        for (int q = 0; q < CN_NI; q++) acc = fmaf(N.skip[k][q], x[q], acc);
        const float yv = valid ? Y[smp * CN_NO + k] : 0.0f;
        t[k] = (yv - acc) / N.out_scale[k];
      }
    }
    cn_wave_fence();
    bf16x8_t xb[3];
#pragma unroll
    for (int s = 0; s < 3; s++) xb[s] = cn_ld(*reinterpret_cast<const Bf8 *>(mx + c * CN_KE + 16 * s + 8 * h));
    f32x16_t zs[5][2];
    const f32x16_t out = cn_forward_tile(N.w, N.b, xb, zs);
    // output error: lane c of the lower half holds outputs 0..8 of tuple c
    f32x16_t d[2];
    d[0] = f32x16_t{};
    d[1] = f32x16_t{};
    if (h == 0 && valid) {
#pragma unroll
      for (int e = 0; e < CN_NO; e++) {
        const float rr = out[e] + N.b6[e] - t[e];
        loss = fmaf(rr, rr, loss);
        d[0][e] = 2.0f * rr * inv_n;
      }
    }
    // ---- layer 6 ----
#pragma unroll
    for (int e = 0; e < CN_NO; e++) {
      const float sb = cn_half_sum(h == 0 ? d[0][e] : 0.0f);
      if (lane == 0) atomicAdd(&gacc[CN_OB6 + e], sb);
    }
    if (h == 0)
#pragma unroll
      for (int e = 0; e < CN_NO; e++) md[e * 32 + c] = cn_bf16_bits(d[0][e]);
    {
      f32x16_t hh[2];
#pragma unroll
      for (int R = 0; R < 2; R++)
#pragma unroll
        for (int e = 0; e < 16; e++) hh[R][e] = cn_gelu(zs[4][R][e] + N.b[4][32 * R + cn_r(e, h)]);
      cn_stage_t(mh, hh, 2);
    }
    cn_wave_fence();
#pragma unroll
    for (int Rn = 0; Rn < 2; Rn++) {
      const f32x16_t g = cn_wgrad_tile(md, mh, 0, Rn, CN_NO);
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const int o = cn_r(e, h);
        if (o < CN_NO) atomicAdd(&gacc[CN_OW6 + o * CN_NH + 32 * Rn + c], g[e]);
      }
    }
    {
      bf16x8_t b6[2];
#pragma unroll
      for (int s = 0; s < 2; s++)
#pragma unroll
        for (int i = 0; i < 8; i++) b6[s][i] = to_bf16(d[0][8 * s + i]);
#pragma unroll
      for (int Rn = 0; Rn < 2; Rn++) {
        f32x16_t acc = {};
#pragma unroll
        for (int s = 0; s < 2; s++)
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cn_ld(nt->w6[Rn][s][lane]), b6[s], acc, 0,
                                                        0, 0);
#pragma unroll
        for (int e = 0; e < 16; e++)
          acc[e] *= cn_gelu_grad(zs[4][Rn][e] + N.b[4][32 * Rn + cn_r(e, h)]);
        d[Rn] = acc;
      }
    }
    cn_wave_fence();
    // ---- layers 5 .. 2 ----
#pragma unroll
    for (int l = 5; l >= 2; l--) {
      // bias gradient, staging of d_l and h_{l-1}
#pragma unroll
      for (int R = 0; R < 2; R++)
#pragma unroll
        for (int e = 0; e < 16; e++) {
          const float sb = cn_half_sum(d[R][e]);
          if (c == 0) atomicAdd(&gacc[cn_ob(l) + 32 * R + cn_r(e, h)], sb);
        }
      cn_stage_t(md, d, 2);
      {
        f32x16_t hh[2];
#pragma unroll
        for (int R = 0; R < 2; R++)
#pragma unroll
          for (int e = 0; e < 16; e++)
            hh[R][e] = cn_gelu(zs[l - 2][R][e] + N.b[l - 2][32 * R + cn_r(e, h)]);
        cn_stage_t(mh, hh, 2);
      }
      cn_wave_fence();
#pragma unroll
      for (int Rm = 0; Rm < 2; Rm++)
#pragma unroll
        for (int Rn = 0; Rn < 2; Rn++) {
          const f32x16_t g = cn_wgrad_tile(md, mh, Rm, Rn, 64);
#pragma unroll
          for (int e = 0; e < 16; e++)
            atomicAdd(&gacc[cn_ow(l) + (32 * Rm + cn_r(e, h)) * CN_NH + 32 * Rn + c], g[e]);
        }
      bf16x8_t bd[4];
#pragma unroll
      for (int R = 0; R < 2; R++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
          for (int i = 0; i < 8; i++) bd[2 * R + j][i] = to_bf16(d[R][8 * j + i]);
#pragma unroll
      for (int Rn = 0; Rn < 2; Rn++) {
        f32x16_t acc = {};
#pragma unroll
        for (int s = 0; s < 4; s++)
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cn_ld(nt->wh[l - 2][Rn][s][lane]), bd[s],
                                                        acc, 0, 0, 0);
#pragma unroll
        for (int e = 0; e < 16; e++)
          acc[e] *= cn_gelu_grad(zs[l - 2][Rn][e] + N.b[l - 2][32 * Rn + cn_r(e, h)]);
        d[Rn] = acc;
      }
      cn_wave_fence();
    }
    // ---- layer 1 ----
#pragma unroll
    for (int R = 0; R < 2; R++)
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const float sb = cn_half_sum(d[R][e]);
        if (c == 0) atomicAdd(&gacc[CN_OB1 + 32 * R + cn_r(e, h)], sb);
      }
    cn_stage_t(md, d, 2);
    cn_wave_fence();
#pragma unroll
    for (int Rm = 0; Rm < 2; Rm++)
#pragma unroll
      for (int Rn = 0; Rn < 2; Rn++) {
        const f32x16_t g = cn_wgrad_tile(md, mxt, Rm, Rn, 64);
        const int kk = 32 * Rn + c;
        const int fcol = kk < CN_NI ? kk : kk - CN_NI;
        if (kk < 2 * CN_NI)
#pragma unroll
          for (int e = 0; e < 16; e++)
            atomicAdd(&gacc[CN_OW1 + (32 * Rm + cn_r(e, h)) * CN_NI + fcol], g[e]);
      }
    cn_wave_fence();
  }
  loss = cn_half_sum(loss);
  if (lane == 0 && loss != 0.0f) atomicAdd(&gacc[CN_NPARAM], loss * inv_n);
  __syncthreads();
  float *dst = partial + (size_t)blockIdx.x * (CN_NPARAM + 1);
  for (int q = threadIdx.x; q <= CN_NPARAM; q += blockDim.x) dst[q] = gacc[q];
}

// Sums the block partials in a fixed order into out[0..NPARAM] (gradient,
// then the loss); with adam != 0 takes one Adam step on the parameters.
// The loss goes to out_loss[0].
__global__ void k_cn_reduce(float *p, const float *partial, int nblocks, float *out,
                            float *out_loss, float *mom, float *vel, int adam, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > CN_NPARAM) return;
  float g = 0.0f;
  for (int b = 0; b < nblocks; b++) g += partial[(size_t)b * (CN_NPARAM + 1) + q];
  if (q == CN_NPARAM) {
    out_loss[0] = g;
    return;
  }
  out[q] = g;
  if (!adam) return;
  const float m = b1 * mom[q] + (1.0f - b1) * g;
  const float v = b2 * vel[q] + (1.0f - b2) * g * g;
  mom[q] = m;
  vel[q] = v;
  p[q] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
}

// ---- the step's learned constitutive update ------------------------------------
//
// Particles in sorted order j, 64 per wave; the state is read from the
// current slot order[j] and written at j (the G2P kernel's permutation),
// with the rate of deformation, pressure and wave speed the G2P kernel
// staged.  MODE 1: the JC material functions from the two small networks
// (kmpm_surrogate.h); MODE 2: the cut network.  Lanes inside the envelope
// take the networks, the others the plain FP64 functions.  MODE 2: one wave
// in judge_stride (rotating with the step) also runs the plain strength for
// its network lanes and sums the squared deviatoric error.
// One particle's inputs to the learned update: its stress state (from the
// current slot order[j]) and what the G2P kernel staged at j.
struct LearnedIn {
  float sig[6], eel[6], Dr[6];
  float ep, epdot, dinit, dmg, pH, cwave;
};
__device__ inline void learned_load(const KArgs &a, int j, LearnedIn &in) {
  constexpr int PREC = KMPM_PRECISION_MIXED;
  const int slot = a.order[j];
  const float *st = a.cut_stage;
  const size_t cs = (size_t)a.cap;
  for (int e = 0; e < 6; e++) {
    in.sig[e] = (float)LDA(SXX + e, slot);
    in.eel[e] = (float)LDA(EXX + e, slot);
    in.Dr[e] = st[e * cs + j];
  }
  in.ep = (float)LDA(EP, slot);
  in.epdot = (float)LDA(EPDOT, slot);
  in.dinit = (float)LDA(DINIT, slot);
  in.dmg = (float)LDA(DMG, slot);
  in.pH = st[6 * cs + j];
  in.cwave = st[7 * cs + j];
}

template <int MODE>
__global__ __launch_bounds__(256, 2) void k_stress_learned(KArgs a) {
  constexpr int PREC = KMPM_PRECISION_MIXED;
  if (frozen(a)) return;   // the G2P kernel copied the whole state
  constexpr size_t NETB = MODE == 2 ? sizeof(CutNetDev) : sizeof(SurrogateDev);
  __shared__ __attribute__((aligned(16))) unsigned char s_net[NETB];
  __shared__ __attribute__((aligned(16))) uint16_t stage[MODE == 2 ? 4 : 1][64 * CN_KE];
  {
    static_assert(NETB % 4 == 0, "4-byte copies");
    const uint32_t *src = reinterpret_cast<const uint32_t *>(
        MODE == 2 ? (const void *)a.cut : (const void *)a.surr);
    uint32_t *dst = reinterpret_cast<uint32_t *>(s_net);
    for (int q = threadIdx.x; q < (int)(NETB / 4); q += blockDim.x) dst[q] = src[q];
  }
  __syncthreads();
  const CutNetDev &N = *reinterpret_cast<const CutNetDev *>(s_net);
  const SurrogateDev &SD = *reinterpret_cast<const SurrogateDev *>(s_net);
  const int wave = threadIdx.x >> 6;
  const DevMat &m = a.mats[0];
  const double dt = a.sc->dt;
  const float dtf = (float)dt;
  const float tav = (float)(1000 * a.h / m.signal_velocity);          // solid.cpp:1324-1329
  const int64_t step = a.sc->step;
  double cmax = 0.0;
  SurrCount cnt = {0, 0};
  unsigned long long jn = 0;
  double je2 = 0.0, jr2 = 0.0;
  // The next round's inputs are loaded before this round's updates, so
  // their memory latency overlaps the networks (the state is read from the
  // current buffers and written to the next ones: no load waits on a store).
  const int stride = gridDim.x * blockDim.x;
  LearnedIn nxt;
  if ((int)(blockIdx.x * blockDim.x) < a.np)
    learned_load(a, min((int)(blockIdx.x * blockDim.x + threadIdx.x), a.np - 1), nxt);
  for (int base = blockIdx.x * blockDim.x; base < a.np; base += stride) {
    const int j0 = base + (int)threadIdx.x;
    const bool live = j0 < a.np;
    const int j = live ? j0 : a.np - 1;
    LearnedIn in = nxt;
    if (base + stride < a.np) learned_load(a, min(j0 + stride, a.np - 1), nxt);
    float *sig = in.sig, *eel = in.eel, *Dr = in.Dr;
    float ep = in.ep, epdot = in.epdot, dinit = in.dinit, dmg = in.dmg;
    const float pH = in.pH, cwave = in.cwave;
    if constexpr (MODE == 1) {
      update_stress_jc_surrogate_p(m, SD, live, dt, a.h, pH, Dr, sig, eel, ep, epdot, dinit, dmg,
                                   cnt);
    } else {
      const float A = N.A, G = N.G, epsdot0 = N.epsdot0;
      // the trial deviator (strength_jc.cpp:146-157) and the envelope
      float t[6];
      const float q = cn_trial<float>(G, dtf, dmg, sig, Dr, t);
      float f[CN_NI];
      cn_features<float>(A, G, epsdot0, dtf, t, q, Dr, ep, epdot, dmg, dinit, pH, f);
      float dD[6];
      for (int e = 0; e < 6; e++) dD[e] = f[6 + e];
      const float inc = (float)SQRT_3_OVER_2 * norm_sym_t(dD);
      const bool inside = ep <= N.ep_max && epdot <= N.ratio_max * epsdot0 && dmg < N.dmg_max &&
                          dinit <= N.dinit_max && fabsf(pH) <= N.p_max * A &&
                          q <= N.q_max * A && inc <= N.inc_max;
      const bool use_net = live && inside;
      const bool fb = live && !inside;
      float y[CN_NO];
      if (__ballot(use_net)) cn_eval_wave(N, f, y, stage[wave]);
      float sdev[6], dps = 0.0f;
      bool plastic = false;
      if (use_net) {
        plastic = y[6] > 0.0f;
        for (int e = 0; e < 6; e++) sdev[e] = plastic ? A * y[e] : t[e];
        dps = plastic ? A * y[6] / (3.0f * G) : 0.0f;
      } else {
        double sig_c[6], D_c[6], out_c[6], dps_c = 0.0;
        for (int e = 0; e < 6; e++) {
          sig_c[e] = sig[e];
          D_c[e] = Dr[e];
        }
        if (fb) jc_strength_fallback(m, sig_c, D_c, dt, ep, epdot, dmg, out_c, &dps_c);
        for (int e = 0; e < 6; e++) sdev[e] = (float)out_c[e];
        dps = (float)dps_c;
      }
      // the judge: the plain FP64 strength on one wave in judge_stride
      const int64_t wid = base / 64 + wave;
      if (N.judge_stride > 0 && (wid + step) % N.judge_stride == 0 && use_net) {
        double sig_c[6], D_c[6], out_c[6], dps_c = 0.0;
        for (int e = 0; e < 6; e++) {
          sig_c[e] = sig[e];
          D_c[e] = Dr[e];
        }
        jc_strength_fallback(m, sig_c, D_c, dt, ep, epdot, dmg, out_c, &dps_c);
        double diff[6];
        for (int e = 0; e < 6; e++) diff[e] = (double)sdev[e] - out_c[e];
        const double nd = norm_sym(diff), nr = norm_sym(out_c);
        je2 += nd * nd;
        jr2 += nr * nr;
        jn += 1;
      }
      ep += dps;                                                        // solid.cpp:1319
      epdot -= __fdividef(epdot * dtf, tav);                            // solid.cpp:1324-1329
      epdot += __fdividef(dps, tav);
      epdot = epdot > 0.0f ? epdot : 0.0f;
      // damage (damage_jc.cpp:91-143): the network's increment, the reset and
      // the damage closed form applied exactly
      if (m.damage == KMPM_DAMAGE_JOHNSON_COOK) {
        if (use_net) {
          if (plastic) dinit += fmaxf(A * y[7] / (3.0f * G), 0.0f);
          const float vm = (float)SQRT_3_OVER_2 * norm_sym_t(sdev);
          float triax = 0.0f;
          if (pH != 0.0f && vm != 0.0f) triax = __fdividef(-pH, vm + 0.001f * fabsf(pH));
          if (triax <= -3.0f) dinit = 0.0f;
          if (dinit >= 1.0f) dmg = fminf(fmaxf(y[8], dmg), 1.0f);
        } else if (fb) {
          double di_c = dinit, dm_c = dmg, sdev_c[6];
          for (int e = 0; e < 6; e++) sdev_c[e] = sdev[e];
          jc_damage_fallback(m, di_c, dm_c, pH, sdev_c, epdot, dps);
          dinit = (float)di_c;
          dmg = (float)dm_c;
        }
      }
      surrogate_count(cnt, use_net, fb);
      // assembly (solid.cpp:1353-1366)
      const float trD = Dr[0] + Dr[1] + Dr[2];
      const float trE = eel[0] + eel[1] + eel[2];
      if (dmg == 0.0f || pH >= 0.0f) {
        for (int e = 0; e < 3; e++) sig[e] = -pH * 1.0f + sdev[e];
      } else {
        const float pd = -pH * (1.0f - dmg);
        for (int e = 0; e < 3; e++) sig[e] = pd * 1.0f + sdev[e];
      }
      for (int e = 3; e < 6; e++) sig[e] = sdev[e];
      const float e3 = (dtf * trD + trE) / 3.0f;
      const float Ge = dmg > 1e-10f ? (float)m.G * (1.0f - dmg) : (float)m.G;
      for (int e = 0; e < 3; e++) eel[e] = e3 * 1.0f + __fdividef(sdev[e], Ge);
      for (int e = 3; e < 6; e++) eel[e] = __fdividef(sdev[e], Ge);
    }
    if (!live) continue;
    // the wave speed of particles that are not fully damaged (solid.cpp:1376-1385)
    if (dmg < 1.0f) {
      if (isnan(cwave)) set_error(a.sc, KMPM_ERR_WAVE_SPEED, IFLD(a.ib, PID, j));
      if (cwave > cmax) cmax = cwave;
    }
    for (int e = 0; e < 6; e++) {
      STB(SXX + e, j, sig[e]);
      STB(EXX + e, j, eel[e]);
    }
    STB(EP, j, ep);
    STB(EPDOT, j, epdot);
    STB(DINIT, j, dinit);
    STB(DMG, j, dmg);
  }
  finish_cfl<false>(a, cmax, 1.0);
  surrogate_count_flush(cnt, a.sc);
  if (MODE == 2 && jn) {
    atomicAdd(&a.sc->judge_n, jn);
    atomicAdd(&a.sc->judge_err2, je2);
    atomicAdd(&a.sc->judge_ref2, jr2);
  }
}

hipError_t launch_learned_stress(kmpm_ctx *c) {
  if (c->a.np == 0) return hipSuccess;
  // One round of resident workgroups: a grid larger than what fits at once
  // would run its last blocks at a fraction of the occupancy.
  static int resident[2] = {0, 0};
  const int mode = c->a.cut_on ? 1 : 0;
  if (resident[mode] == 0) {
    int n = 0;
    const hipError_t e =
        c->a.cut_on
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stress_learned<2>, 256, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stress_learned<1>, 256, 0);
    if (e != hipSuccess) return e;
    resident[mode] = n > 0 ? n : 1;
  }
  int g = (c->a.np + 255) / 256;
  if (g > c->cu_count * resident[mode]) g = c->cu_count * resident[mode];
  if (c->a.cut_on)
    hipLaunchKernelGGL(k_stress_learned<2>, dim3(g), dim3(256), 0, c->stream, c->a);
  else
    hipLaunchKernelGGL(k_stress_learned<1>, dim3(g), dim3(256), 0, c->stream, c->a);
  return hipGetLastError();
}

}  // namespace kmpm

// ---- host side ----------------------------------------------------------------------

using namespace kmpm;

struct kmpm_cutnet {
  kmpm_cutnet_config cfg;
  kmpm_material mat;
  DevMat dmat;
  CnSampling samp;
  CutNetDev *h_net = nullptr;        // host image (maps, envelope; weights refreshed from the device)
  CutNetDev *d_net = nullptr;
  CutNetT *d_nt = nullptr;
  float *d_p = nullptr, *d_mom = nullptr, *d_vel = nullptr;   // [NPARAM + 1]
  float *d_F = nullptr, *d_Y = nullptr;                        // training tuples
  int64_t ntup = 0;
  float *d_partial = nullptr, *d_out = nullptr;
  int partial_blocks = 0;
  int64_t iter_done = 0;
  int device = 0;
  hipStream_t stream = nullptr;
};

namespace {

#define CCHK(x)                                   \
  do {                                            \
    if ((x) != hipSuccess) return KMPM_ERR_HIP;   \
  } while (0)

int cn_pack(kmpm_cutnet *s) {
  hipLaunchKernelGGL(k_cn_pack, dim3((CN_PACK_THREADS + 255) / 256), dim3(256), 0, s->stream,
                     s->d_p, s->d_net, s->d_nt);
  return hipGetLastError() == hipSuccess ? KMPM_OK : KMPM_ERR_HIP;
}

// the host image's maps and envelope into the device image (weights stay)
int cn_upload_maps(kmpm_cutnet *s) {
  const size_t off = offsetof(CutNetDev, in_scale);
  CCHK(hipMemcpyAsync(reinterpret_cast<char *>(s->d_net) + off,
                      reinterpret_cast<const char *>(s->h_net) + off, sizeof(CutNetDev) - off,
                      hipMemcpyHostToDevice, s->stream));
  CCHK(hipStreamSynchronize(s->stream));
  return KMPM_OK;
}

int cn_gen(kmpm_cutnet *s, int64_t n, uint64_t seed, float *F, float *Y) {
  if (n <= 0) return KMPM_OK;
  hipLaunchKernelGGL(k_cn_tuples, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->stream,
                     s->dmat, s->samp, seed, n, F, Y);
  return hipGetLastError() == hipSuccess ? KMPM_OK : KMPM_ERR_HIP;
}

int cn_ensure_partial(kmpm_cutnet *s, int nb) {
  if (nb <= s->partial_blocks) return KMPM_OK;
  if (s->d_partial) (void)hipFree(s->d_partial);
  s->d_partial = nullptr;
  CCHK(hipMalloc(&s->d_partial, (size_t)nb * (CN_NPARAM + 1) * sizeof(float)));
  s->partial_blocks = nb;
  return KMPM_OK;
}

// blocks of the gradient kernel for n tuples (at most 256 partial sums)
int cn_grad_blocks(int64_t n) {
  const int64_t per = (int64_t)CN_GRAD_WAVES * 32;
  int64_t nb = (n + per - 1) / per;
  return (int)std::max<int64_t>(1, std::min<int64_t>(nb, 256));
}

int cn_grad_pass(kmpm_cutnet *s, const float *F, const float *Y, int64_t n, int adam, float lr,
                 int64_t t) {
  const int nb = cn_grad_blocks(n);
  int st = cn_ensure_partial(s, nb);
  if (st) return st;
  hipLaunchKernelGGL(k_cn_grad, dim3(nb), dim3(256), 0, s->stream, s->d_net, s->d_nt, F, Y, n,
                     1.0f / (float)n, s->d_partial);
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = adam ? 1.0f / (1.0f - powf(b1, (float)t)) : 1.0f;
  const float c2 = adam ? 1.0f / (1.0f - powf(b2, (float)t)) : 1.0f;
  hipLaunchKernelGGL(k_cn_reduce, dim3((CN_NPARAM + 1 + 255) / 256), dim3(256), 0, s->stream,
                     s->d_p, s->d_partial, nb, s->d_out, s->d_out + CN_NPARAM, s->d_mom, s->d_vel,
                     adam, lr, b1, b2, 1e-8f, c1, c2);
  if (hipGetLastError() != hipSuccess) return KMPM_ERR_HIP;
  return adam ? cn_pack(s) : KMPM_OK;
}

// Solves the normal equations M c = r (n x n, symmetric positive
// semi-definite) by Gaussian elimination with partial pivoting; a
// vanishing pivot (a constant feature) drops that unknown.
void cn_solve(std::vector<double> &M, std::vector<double> &r, int n) {
  for (int col = 0; col < n; col++) {
    int piv = col;
    for (int row = col + 1; row < n; row++)
      if (std::fabs(M[row * n + col]) > std::fabs(M[piv * n + col])) piv = row;
    if (std::fabs(M[piv * n + col]) < 1e-12 * (1.0 + std::fabs(M[col * n + col]))) {
      for (int k = 0; k < n; k++) M[col * n + k] = M[k * n + col] = 0.0;
      M[col * n + col] = 1.0;
      r[col] = 0.0;
      continue;
    }
    if (piv != col) {
      for (int k = 0; k < n; k++) std::swap(M[col * n + k], M[piv * n + k]);
      std::swap(r[col], r[piv]);
    }
    for (int row = 0; row < n; row++) {
      if (row == col) continue;
      const double fct = M[row * n + col] / M[col * n + col];
      if (fct == 0.0) continue;
      for (int k = col; k < n; k++) M[row * n + k] -= fct * M[col * n + k];
      r[row] -= fct * r[col];
    }
  }
  for (int k = 0; k < n; k++) r[k] /= M[k * n + k];
}

// Input maps (standardisation) and the linear skip fitted by least squares
// on a sample of the tuples; output scales = RMS of the residuals.
int cn_fit_maps(kmpm_cutnet *s) {
  const int64_t m = std::min<int64_t>(s->ntup, 1 << 20);
  std::vector<float> F((size_t)m * CN_NI), Y((size_t)m * CN_NO);
  CCHK(hipMemcpyAsync(F.data(), s->d_F, F.size() * 4, hipMemcpyDeviceToHost, s->stream));
  CCHK(hipMemcpyAsync(Y.data(), s->d_Y, Y.size() * 4, hipMemcpyDeviceToHost, s->stream));
  CCHK(hipStreamSynchronize(s->stream));
  CutNetDev &H = *s->h_net;
  for (int k = 0; k < CN_NI; k++) {
    double mu = 0, var = 0;
    for (int64_t i = 0; i < m; i++) mu += F[i * CN_NI + k];
    mu /= (double)m;
    for (int64_t i = 0; i < m; i++) {
      const double d = F[i * CN_NI + k] - mu;
      var += d * d;
    }
    const double sd = std::sqrt(var / (double)m);
    const double sc = sd > 1e-12 ? 1.0 / sd : 1.0;
    H.in_scale[k] = (float)sc;
    H.in_shift[k] = (float)(-mu * sc);
  }
  constexpr int NB = CN_NI + 1;
  std::vector<double> M(NB * NB, 0.0), R((size_t)NB * CN_NO, 0.0);
  std::vector<double> xr(NB);
  for (int64_t i = 0; i < m; i++) {
    for (int k = 0; k < CN_NI; k++)
      xr[k] = (double)std::fma(F[i * CN_NI + k], H.in_scale[k], H.in_shift[k]);
    xr[CN_NI] = 1.0;
    for (int a = 0; a < NB; a++) {
      for (int b = a; b < NB; b++) M[a * NB + b] += xr[a] * xr[b];
      for (int o = 0; o < CN_NO; o++) R[o * NB + a] += xr[a] * Y[i * CN_NO + o];
    }
  }
  for (int a = 0; a < NB; a++)
    for (int b = 0; b < a; b++) M[a * NB + b] = M[b * NB + a];
  for (int o = 0; o < CN_NO; o++) {
    std::vector<double> Mo = M, ro(R.begin() + o * NB, R.begin() + (o + 1) * NB);
    cn_solve(Mo, ro, NB);
    for (int k = 0; k < NB; k++) H.skip[o][k] = (float)ro[k];
  }
  std::vector<double> ss(CN_NO, 0.0);
  for (int64_t i = 0; i < m; i++) {
    for (int k = 0; k < CN_NI; k++)
      xr[k] = (double)std::fma(F[i * CN_NI + k], H.in_scale[k], H.in_shift[k]);
    for (int o = 0; o < CN_NO; o++) {
      double v = H.skip[o][CN_NI];
      for (int k = 0; k < CN_NI; k++) v += (double)H.skip[o][k] * xr[k];
      const double r = Y[i * CN_NO + o] - v;
      ss[o] += r * r;
    }
  }
  for (int o = 0; o < CN_NO; o++) {
    const double rms = std::sqrt(ss[o] / (double)m);
    H.out_scale[o] = (float)(rms > 1e-7 ? rms : 1e-7);
  }
  return cn_upload_maps(s);
}

// Initial parameters: weights normal with variance 1 / fan_in (the last
// layer's scaled by 0.1), biases zero.
void cn_init_params(std::vector<float> &p, uint64_t seed) {
  std::mt19937_64 rng(seed ^ 0x5DEECE66Dull);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  std::fill(p.begin(), p.end(), 0.0f);
  for (int i = 0; i < CN_NH * CN_NI; i++) p[CN_OW1 + i] = nd(rng) / std::sqrt((float)CN_NI);
  for (int l = 2; l <= 5; l++)
    for (int i = 0; i < CN_NH * CN_NH; i++) p[cn_ow(l) + i] = nd(rng) / std::sqrt((float)CN_NH);
  for (int i = 0; i < CN_NO * CN_NH; i++) p[CN_OW6 + i] = 0.1f * nd(rng) / std::sqrt((float)CN_NH);
}

bool cn_valid_material(const kmpm_material &m) {
  return m.type == KMPM_MAT_EOS_STRENGTH && m.strength == KMPM_STRENGTH_JOHNSON_COOK &&
         m.eos == KMPM_EOS_SHOCK;
}

}  // namespace

extern "C" {

int kmpm_cutnet_create(const kmpm_material *mat, const kmpm_cutnet_config *cfg,
                       kmpm_cutnet **out) {
  if (!mat || !cfg || !out) return KMPM_ERR_ARG;
  *out = nullptr;
  if (!cn_valid_material(*mat)) return KMPM_ERR_ARG;
  const kmpm_cutnet_config &c = *cfg;
  if (!(c.eps_p_max > 0) || !(c.rate_min > 0) || !(c.rate_max > c.rate_min) ||
      !(c.damage_max >= 0 && c.damage_max < 1) || !(c.damage_init_max >= 0) ||
      !(c.pressure_max > 0) || !(c.increment_min > 0) ||
      !(c.increment_max > c.increment_min) || !(c.dt > 0) || !(c.cellsize > 0) ||
      c.samples < 1024 || c.iterations < 0 || c.batch < 128 || c.batch % 128 != 0 ||
      c.judge_stride < 0 || !(c.learning_rate > 0))
    return KMPM_ERR_ARG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return KMPM_ERR_NO_DEVICE;
  kmpm_cutnet *s = new kmpm_cutnet();
  s->device = dev;
  s->cfg = c;
  s->mat = *mat;
  s->dmat = make_devmat(*mat);
  CnSampling &S = s->samp;
  S.ep_max = c.eps_p_max;
  S.ratio_lo = c.rate_min;
  S.ratio_max = c.rate_max;
  S.dmg_max = c.damage_max;
  S.dinit_max = c.damage_init_max;
  S.p_max = c.pressure_max;
  S.inc_lo = c.increment_min;
  S.inc_max = c.increment_max;
  S.dt = c.dt;
  S.tav = 1000.0 * c.cellsize / s->dmat.signal_velocity;
  s->h_net = new CutNetDev();
  std::memset(s->h_net, 0, sizeof(CutNetDev));
  CutNetDev &H = *s->h_net;
  const double A = mat->strength_p[1], G = mat->strength_p[0], e0 = mat->strength_p[4];
  H.A = (float)A;
  H.G = (float)G;
  H.epsdot0 = (float)e0;
  H.ep_max = (float)c.eps_p_max;
  H.ratio_max = (float)c.rate_max;
  H.dmg_max = (float)c.damage_max;
  H.dinit_max = (float)c.damage_init_max;
  H.p_max = (float)c.pressure_max;
  // the largest trial von Mises stress the sampling produces: 1.05 of the
  // yield stress at the envelope's strain and rate, plus the increment
  double ratio = std::max(c.rate_max, 1.0);
  double ys = A + mat->strength_p[2] * std::pow(c.eps_p_max, mat->strength_p[3]);
  if (mat->strength_p[5] != 0) ys *= std::pow(1.0 + ratio, mat->strength_p[5]);
  ys *= s->dmat.jc_thermal;
  H.q_max = (float)(1.05 * ys / A + c.increment_max);
  H.inc_max = (float)c.increment_max;
  H.judge_stride = c.judge_stride;
  int st = KMPM_OK;
  auto fail = [&](int code) {
    kmpm_cutnet_destroy(s);
    return code;
  };
  if (hipStreamCreate(&s->stream) != hipSuccess) return fail(KMPM_ERR_HIP);
  s->ntup = c.samples;
  if (hipMalloc(&s->d_net, sizeof(CutNetDev)) != hipSuccess ||
      hipMalloc(&s->d_nt, sizeof(CutNetT)) != hipSuccess ||
      hipMalloc(&s->d_p, (CN_NPARAM + 1) * sizeof(float)) != hipSuccess ||
      hipMalloc(&s->d_mom, (CN_NPARAM + 1) * sizeof(float)) != hipSuccess ||
      hipMalloc(&s->d_vel, (CN_NPARAM + 1) * sizeof(float)) != hipSuccess ||
      hipMalloc(&s->d_out, (CN_NPARAM + 1) * sizeof(float)) != hipSuccess ||
      hipMalloc(&s->d_F, (size_t)s->ntup * CN_NI * sizeof(float)) != hipSuccess ||
      hipMalloc(&s->d_Y, (size_t)s->ntup * CN_NO * sizeof(float)) != hipSuccess)
    return fail(KMPM_ERR_HIP);
  if (hipMemsetAsync(s->d_net, 0, sizeof(CutNetDev), s->stream) != hipSuccess ||
      hipMemsetAsync(s->d_mom, 0, (CN_NPARAM + 1) * sizeof(float), s->stream) != hipSuccess ||
      hipMemsetAsync(s->d_vel, 0, (CN_NPARAM + 1) * sizeof(float), s->stream) != hipSuccess)
    return fail(KMPM_ERR_HIP);
  std::vector<float> p(CN_NPARAM + 1);
  cn_init_params(p, c.seed);
  if (hipMemcpyAsync(s->d_p, p.data(), p.size() * 4, hipMemcpyHostToDevice, s->stream) !=
      hipSuccess)
    return fail(KMPM_ERR_HIP);
  if ((st = cn_gen(s, s->ntup, c.seed, s->d_F, s->d_Y))) return fail(st);
  if ((st = cn_pack(s))) return fail(st);
  if ((st = cn_fit_maps(s))) return fail(st);
  *out = s;
  return KMPM_OK;
}

int kmpm_cutnet_destroy(kmpm_cutnet *s) {
  if (!s) return KMPM_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (void *ptr : {(void *)s->d_net, (void *)s->d_nt, (void *)s->d_p, (void *)s->d_mom,
                    (void *)s->d_vel, (void *)s->d_out, (void *)s->d_F, (void *)s->d_Y,
                    (void *)s->d_partial})
    if (ptr) (void)hipFree(ptr);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s->h_net;
  delete s;
  return KMPM_OK;
}

int kmpm_cutnet_train(kmpm_cutnet *s, int64_t iterations, double *final_loss, double *seconds) {
  if (!s || iterations < 0) return KMPM_ERR_ARG;
  const int64_t B = std::min<int64_t>(s->cfg.batch, s->ntup);
  const int64_t nbatches = s->ntup / B;
  const double lr0 = s->cfg.learning_rate;
  CCHK(hipStreamSynchronize(s->stream));
  const auto t0 = std::chrono::steady_clock::now();
  // the losses of the last iterations (at most 16), read back at the end
  const int64_t tail = std::min<int64_t>(iterations, 16);
  std::vector<float> losses;
  float *d_tail = nullptr;
  if (tail > 0) CCHK(hipMalloc(&d_tail, (size_t)tail * sizeof(float)));
  for (int64_t it = 0; it < iterations; it++) {
    // batches in a pseudo-random order over the tuple set
    uint64_t z = (uint64_t)(s->iter_done + it) * 0x9E3779B97F4A7C15ull + s->cfg.seed;
    z = (z ^ (z >> 31)) * 0xBF58476D1CE4E5B9ull;
    z ^= z >> 29;
    const int64_t off = (int64_t)(z % (uint64_t)nbatches) * B;
    // cosine decay to 1% of the initial step
    const double fr = iterations > 1 ? (double)it / (double)(iterations - 1) : 1.0;
    const float lr = (float)(lr0 * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(M_PI * fr))));
    int st = cn_grad_pass(s, s->d_F + off * CN_NI, s->d_Y + off * CN_NO, B, 1, lr,
                          s->iter_done + it + 1);
    if (st) {
      (void)hipFree(d_tail);
      return st;
    }
    if (it >= iterations - tail)
      (void)hipMemcpyAsync(d_tail + (it - (iterations - tail)), s->d_out + CN_NPARAM,
                           sizeof(float), hipMemcpyDeviceToDevice, s->stream);
  }
  s->iter_done += iterations;
  if (hipStreamSynchronize(s->stream) != hipSuccess) {
    (void)hipFree(d_tail);
    return KMPM_ERR_HIP;
  }
  const auto t1 = std::chrono::steady_clock::now();
  double mean = 0.0;
  if (tail > 0) {
    losses.resize((size_t)tail);
    CCHK(hipMemcpy(losses.data(), d_tail, (size_t)tail * sizeof(float), hipMemcpyDeviceToHost));
    for (float v : losses) mean += v;
    mean /= (double)tail;
    (void)hipFree(d_tail);
  }
  if (final_loss) *final_loss = mean;
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return KMPM_OK;
}

int kmpm_cutnet_tuples(kmpm_cutnet *s, int64_t n, uint64_t seed, float *features,
                       float *targets) {
  if (!s || n < 0 || (n > 0 && (!features || !targets))) return KMPM_ERR_ARG;
  if (n == 0) return KMPM_OK;
  float *dF = nullptr, *dY = nullptr;
  CCHK(hipMalloc(&dF, (size_t)n * CN_NI * 4));
  if (hipMalloc(&dY, (size_t)n * CN_NO * 4) != hipSuccess) {
    (void)hipFree(dF);
    return KMPM_ERR_HIP;
  }
  int st = cn_gen(s, n, seed, dF, dY);
  if (!st && (hipMemcpyAsync(features, dF, (size_t)n * CN_NI * 4, hipMemcpyDeviceToHost,
                             s->stream) != hipSuccess ||
              hipMemcpyAsync(targets, dY, (size_t)n * CN_NO * 4, hipMemcpyDeviceToHost,
                             s->stream) != hipSuccess ||
              hipStreamSynchronize(s->stream) != hipSuccess))
    st = KMPM_ERR_HIP;
  (void)hipFree(dF);
  (void)hipFree(dY);
  return st;
}

int kmpm_cutnet_eval(kmpm_cutnet *s, const float *features, int64_t n, float *outputs) {
  if (!s || n < 0 || (n > 0 && (!features || !outputs))) return KMPM_ERR_ARG;
  if (n == 0) return KMPM_OK;
  float *dF = nullptr, *dY = nullptr;
  CCHK(hipMalloc(&dF, (size_t)n * CN_NI * 4));
  if (hipMalloc(&dY, (size_t)n * CN_NO * 4) != hipSuccess) {
    (void)hipFree(dF);
    return KMPM_ERR_HIP;
  }
  int st = KMPM_OK;
  if (hipMemcpyAsync(dF, features, (size_t)n * CN_NI * 4, hipMemcpyHostToDevice, s->stream) !=
      hipSuccess)
    st = KMPM_ERR_HIP;
  if (!st) {
    const int g = (int)std::min<int64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_cn_eval, dim3(g), dim3(256), 0, s->stream, s->d_net, dF, n, dY);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(outputs, dY, (size_t)n * CN_NO * 4, hipMemcpyDeviceToHost, s->stream) !=
            hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
      st = KMPM_ERR_HIP;
  }
  (void)hipFree(dF);
  (void)hipFree(dY);
  return st;
}

int kmpm_cutnet_gradient(kmpm_cutnet *s, const float *features, const float *targets, int64_t n,
                         float *grad, double *loss) {
  if (!s || n <= 0 || !features || !targets || !grad || !loss) return KMPM_ERR_ARG;
  float *dF = nullptr, *dY = nullptr;
  CCHK(hipMalloc(&dF, (size_t)n * CN_NI * 4));
  if (hipMalloc(&dY, (size_t)n * CN_NO * 4) != hipSuccess) {
    (void)hipFree(dF);
    return KMPM_ERR_HIP;
  }
  int st = KMPM_OK;
  std::vector<float> out(CN_NPARAM + 1);
  if (hipMemcpyAsync(dF, features, (size_t)n * CN_NI * 4, hipMemcpyHostToDevice, s->stream) !=
          hipSuccess ||
      hipMemcpyAsync(dY, targets, (size_t)n * CN_NO * 4, hipMemcpyHostToDevice, s->stream) !=
          hipSuccess)
    st = KMPM_ERR_HIP;
  if (!st) st = cn_grad_pass(s, dF, dY, n, 0, 0.0f, 1);
  if (!st && (hipMemcpyAsync(out.data(), s->d_out, out.size() * 4, hipMemcpyDeviceToHost,
                             s->stream) != hipSuccess ||
              hipStreamSynchronize(s->stream) != hipSuccess))
    st = KMPM_ERR_HIP;
  (void)hipFree(dF);
  (void)hipFree(dY);
  if (st) return st;
  std::memcpy(grad, out.data(), CN_NPARAM * sizeof(float));
  *loss = out[CN_NPARAM];
  return KMPM_OK;
}

int kmpm_cutnet_params(kmpm_cutnet *s, float *params, int nparam, int set) {
  if (!s || !params || nparam != CN_NPARAM) return KMPM_ERR_ARG;
  if (set) {
    CCHK(hipMemcpyAsync(s->d_p, params, CN_NPARAM * 4, hipMemcpyHostToDevice, s->stream));
    int st = cn_pack(s);
    if (st) return st;
  } else {
    CCHK(hipMemcpyAsync(params, s->d_p, CN_NPARAM * 4, hipMemcpyDeviceToHost, s->stream));
  }
  CCHK(hipStreamSynchronize(s->stream));
  return KMPM_OK;
}

int kmpm_cutnet_maps(kmpm_cutnet *s, float *in_scale, float *in_shift, float *out_scale,
                     float *skip, int set) {
  if (!s || !in_scale || !in_shift || !out_scale || !skip) return KMPM_ERR_ARG;
  CutNetDev &H = *s->h_net;
  if (set) {
    std::memcpy(H.in_scale, in_scale, sizeof(H.in_scale));
    std::memcpy(H.in_shift, in_shift, sizeof(H.in_shift));
    std::memcpy(H.out_scale, out_scale, sizeof(H.out_scale));
    std::memcpy(H.skip, skip, sizeof(H.skip));
    return cn_upload_maps(s);
  }
  std::memcpy(in_scale, H.in_scale, sizeof(H.in_scale));
  std::memcpy(in_shift, H.in_shift, sizeof(H.in_shift));
  std::memcpy(out_scale, H.out_scale, sizeof(H.out_scale));
  std::memcpy(skip, H.skip, sizeof(H.skip));
  return KMPM_OK;
}

// ---- the gradient without allocation ------------------------------------------
//
// kmpm_cutnet_gradient_noalloc keeps the block partial sums of the gradient
// in `scratch`, a file-scope device array of CN_NOALLOC_BLOCKS partial
// vectors (4.7 MB) that every GPU the library runs on holds until the
// process ends.  Its address on each device is resolved once, with that
// device current, on the entry point's first call there.  Calls that
// overlap on one GPU share the array and race.
constexpr int CN_NOALLOC_BLOCKS = 64;
static __device__ __align__(256) float scratch[CN_NOALLOC_BLOCKS * (CN_NPARAM + 1)];

namespace {
constexpr int CN_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[CN_MAX_DEVICES];   // by device; nullptr until resolved

int cn_scratch_ptr(int dev, float **out) {
  if (dev < 0 || dev >= CN_MAX_DEVICES) return KMPM_ERR_NO_DEVICE;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return KMPM_ERR_HIP;
    g_scratch_ptr[dev] = static_cast<float *>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return KMPM_OK;
}
}  // namespace

int kmpm_cutnet_gradient_noalloc(kmpm_cutnet *s, const float *features, const float *targets,
                                 int64_t n, float *grad, float *loss) {
  if (!s || !features || !targets || !grad || !loss) return KMPM_ERR_ARG;
  if (n <= 0 || n > KMPM_CUTNET_NOALLOC_MAX_TUPLES) return KMPM_ERR_CAPACITY;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return KMPM_ERR_NO_DEVICE;
  if (dev != s->device) return KMPM_ERR_ARG;
  float *part = nullptr;
  const int st = cn_scratch_ptr(dev, &part);
  if (st) return st;
  int64_t nb = (n + (int64_t)CN_GRAD_WAVES * 32 - 1) / ((int64_t)CN_GRAD_WAVES * 32);
  if (nb > CN_NOALLOC_BLOCKS) nb = CN_NOALLOC_BLOCKS;
  hipLaunchKernelGGL(k_cn_grad, dim3((unsigned)nb), dim3(256), 0, s->stream, s->d_net, s->d_nt,
                     features, targets, n, 1.0f / (float)n, part);
  hipLaunchKernelGGL(k_cn_reduce, dim3((CN_NPARAM + 1 + 255) / 256), dim3(256), 0, s->stream,
                     s->d_p, part, (int)nb, grad, loss, s->d_mom, s->d_vel, 0, 0.0f, 0.9f, 0.999f,
                     1e-8f, 1.0f, 1.0f);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess)
    return KMPM_ERR_HIP;
  return KMPM_OK;
}

int kmpm_cutnet_envelope(kmpm_cutnet *s, double env[7]) {
  if (!s || !env) return KMPM_ERR_ARG;
  const CutNetDev &H = *s->h_net;
  const double v[7] = {H.ep_max, H.ratio_max, H.dmg_max, H.dinit_max, H.p_max, H.q_max, H.inc_max};
  for (int k = 0; k < 7; k++) env[k] = v[k];
  return KMPM_OK;
}

}  // extern "C"

namespace kmpm {

bool cutnet_matches(const kmpm_cutnet *s, const kmpm_material &m) {
  return std::memcmp(&s->mat, &m, sizeof(kmpm_material)) == 0;
}

// The network image a context copies (weights from the device, maps and
// envelope from the host image).
int cutnet_device_image(const kmpm_cutnet *s, CutNetDev *dst_dev, hipStream_t stream) {
  const size_t off = offsetof(CutNetDev, in_scale);
  if (hipStreamSynchronize(s->stream) != hipSuccess) return KMPM_ERR_HIP;
  if (hipMemcpyAsync(dst_dev, s->d_net, off, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
      hipMemcpyAsync(reinterpret_cast<char *>(dst_dev) + off,
                     reinterpret_cast<const char *>(s->h_net) + off, sizeof(CutNetDev) - off,
                     hipMemcpyHostToDevice, stream) != hipSuccess)
    return KMPM_ERR_HIP;
  if (hipStreamSynchronize(stream) != hipSuccess) return KMPM_ERR_HIP;
  return KMPM_OK;
}

}  // namespace kmpm

// kmpm_cutnet.h -- the cut network: one MLP that stands in for both
// constitutive calls of Solid::update_stress for a johnson_cook +
// damage_johnson_cook material (solid.cpp:1314-1316 strength and
// :1337-1339 damage), 18 features of the state at the cut to 9 outputs:
//   18 -> 64 -> 64 -> 64 -> 64 -> 64 -> 9, tanh-form GELU activations,
//   every product on the gfx950 BF16 matrix cores
//   (v_mfma_f32_32x32x16_bf16, FP32 accumulation), over an FP32 linear skip
//   of the inputs.
//
// Features f (FP32, A = JC A, G the shear modulus, Gd = G (1 - damage),
// epsdot0 the JC rate):
//   f0..5   t / A, t = dev(sigma_n + 2 Gd dt D) the trial deviator
//           (strength_jc.cpp:151-152; xx yy zz xy xz yz)
//   f6..11  2 G dt dev(D) / A                      (the elastic increment)
//   f12     eps_p^(1/4)          f13  ln max(epsdot_p / epsdot0, 1)
//   f14     damage               f15  damage_init      f16  pH / A
//   f17     q / A, q = sqrt(3/2) |t| the trial von Mises stress
// Outputs y (target units):
//   y0..5   dev(sigma_{n+1}) / A
//   y6      the signed plastic increment G (q - sigma_y) / (Gd A): 3 G dps / A
//           where positive (strength_jc.cpp:169-181), and the yield
//           function's distance on the elastic side
//   y7      3 G (damage_init_{n+1} - damage_init_n) / A
//   y8      damage_{n+1}
// The inputs enter the first product as two BF16 halves (x = hi + lo), so
// they keep ~16 significant bits.
//
// Operand layouts of v_mfma_f32_32x32x16_bf16 (lane l, h = l / 32,
// c = l % 32, i = 0..7 within a 16-deep k-step, e = 0..15):
//   A : lane l holds A[c][8 h + i]      B : lane l holds B[8 h + i][c]
//   D : lane l holds D[r(e, h)][c],  r(e, h) = 8 (e / 4) + 4 h + e % 4
// A tile is 32 samples (the columns).  The accumulators of one layer feed
// the next layer's B operand unchanged: k-step s = 2 R + j of a 64-wide
// layer takes elements e = 8 j .. 8 j + 7 of output tile R, i.e. the input
// unit u(s, h, i) = 32 R + r(8 j + i, h); the weights are stored permuted to
// match (CutNetW).  Output k of the last layer is row r(k, 0), so lane c of
// the lower half holds all 9 outputs of sample c.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kmpm_mlp.h"

namespace kmpm {

constexpr int CN_NI = 18, CN_NH = 64, CN_NO = 9;
constexpr int CN_KE = 48;      // encoded input columns: hi[18], lo[18], zeros
// flat FP32 parameters: w1[64][18] b1[64] (w_l[64][64] b_l[64], l = 2..5)
// w6[9][64] b6[9]; weights [out][in]
constexpr int CN_OW1 = 0;
constexpr int CN_OB1 = CN_OW1 + CN_NH * CN_NI;
__host__ __device__ constexpr int cn_ow(int l) {   // l = 2..6
  return CN_OB1 + CN_NH + (l - 2) * (CN_NH * CN_NH + CN_NH);
}
__host__ __device__ constexpr int cn_ob(int l) { return cn_ow(l) + CN_NH * CN_NH; }   // l = 2..5
constexpr int CN_OW6 = cn_ow(6);
constexpr int CN_OB6 = CN_OW6 + CN_NO * CN_NH;
constexpr int CN_NPARAM = CN_OB6 + CN_NO;
static_assert(CN_NPARAM == 18441, "flat parameter count");

// Eight BF16 values: one lane's operand of one matrix instruction.
struct alignas(16) Bf8 {
  uint16_t v[8];
};

__host__ __device__ constexpr int cn_r(int e, int h) { return 8 * (e / 4) + 4 * h + (e % 4); }
// input unit of k-step s, half h, element i of a 64-wide chained layer
__host__ __device__ constexpr int cn_u(int s, int h, int i) {
  return 32 * (s / 2) + cn_r(8 * (s % 2) + i, h);
}

// MFMA-ready BF16 images of the weights (forward products).
struct CutNetW {
  Bf8 w1[2][3][64];        // [R][s][lane] = W1enc[32 R + c][16 s + 8 h + i]
  Bf8 wh[4][2][4][64];     // layers 2..5: [R][s][lane] = W[32 R + c][cn_u(s, h, i)]
  Bf8 w6[4][64];           // [s][lane] = W6phys[c][cn_u(s, h, i)], W6phys[r(k, 0)] = W6[k]
};
// Transposed images for the input-gradient products of training.
struct CutNetT {
  Bf8 wh[4][2][4][64];     // layers 2..5: [Rn][s][lane] = W[cn_u(s, h, i)][32 Rn + c]
  Bf8 w6[2][2][64];        // [Rn][s][lane] = W6phys[cn_r(8 s + i, h)][32 Rn + c]
};

// What the step and evaluation kernels read (copied to LDS per block).
struct CutNetDev {
  CutNetW w;
  float b[5][CN_NH];       // b1..b5
  float b6[16];            // b6, zero padded
  float in_scale[CN_NI], in_shift[CN_NI];
  float out_scale[CN_NO];
  float skip[CN_NO][CN_NI + 1];   // output = skip x + skip[.][17] + out_scale z6
  // material (FP32) and the training envelope
  float A, G, epsdot0, pad0;
  float ep_max, ratio_max, dmg_max, dinit_max;   // eps_p, epsdot / epsdot0, damage, damage_init
  float p_max, q_max, inc_max, pad1;             // |pH| / A, q_trial / A, relative increment
  int judge_stride;                              // one wave in judge_stride checks the plain functions
  int pad2[3];
};

__device__ inline bf16x8_t cn_ld(const Bf8 &b) { return *reinterpret_cast<const bf16x8_t *>(&b); }

__device__ inline uint16_t cn_bf16_bits(float x) { return __builtin_bit_cast(uint16_t, to_bf16(x)); }

// gelu(z) = z sigmoid(2 u), u = sqrt(2/pi) (z + 0.044715 z^3); also returns
// the sigmoid s for the derivative.
constexpr float CN_K2L = 2.3022081983f;    // 2 sqrt(2/pi) log2(e)
constexpr float CN_C = 0.044715f;
__device__ inline float cn_gelu(float z, float &s) {
  const float e = __builtin_amdgcn_exp2f(-CN_K2L * z * fmaf(CN_C * z, z, 1.0f));
  s = __builtin_amdgcn_rcpf(1.0f + e);
  return z * s;
}
__device__ inline float cn_gelu(float z) {
  float s;
  return cn_gelu(z, s);
}
// d gelu / dz = s + z s (1 - s) 2 u'(z), 2 u' = 2 sqrt(2/pi) (1 + 3 C z^2)
__device__ inline float cn_gelu_grad(float z) {
  float s;
  (void)cn_gelu(z, s);
  const float du2 = fmaf(0.2140644f * z, z, 1.5957691216f);
  return fmaf(z * s * (1.0f - s), du2, s);
}

// Bias, activation and BF16 packing of one 32-wide output tile's
// accumulators into the two B operands (k-steps 2 R and 2 R + 1) of the next
// layer.  bias: the layer's 64 biases; R: the tile.
__device__ inline void cn_act_pack(const f32x16_t &acc, const float *bias, int R, bf16x8_t &b0,
                                   bf16x8_t &b1) {
  const int h = (threadIdx.x & 63) >> 5;
#pragma unroll
  for (int e = 0; e < 16; e++) {
    const float v = cn_gelu(acc[e] + bias[32 * R + cn_r(e, h)]);
    if (e < 8) b0[e] = to_bf16(v);
    else b1[e - 8] = to_bf16(v);
  }
}

// Forward of one 32-sample tile.  xb[3]: the lane's layer-1 B operands
// (encoded inputs of sample c, columns 16 s + 8 h .. + 7).  Returns the
// output accumulators (lane c of the lower half: outputs 0..8 of sample c,
// without b6).  ZS != nullptr keeps every hidden pre-activation
// (zs[layer][R] for layers 1..5) for the backward pass.
template <class W>
__device__ inline f32x16_t cn_forward_tile(const W &w, const float (*bias)[CN_NH],
                                           const bf16x8_t xb[3], f32x16_t (*zs)[2] = nullptr) {
  // The opaque barriers keep the compiler from hoisting every weight load
  // of the network into registers up front (each operand is read from LDS
  // or the cache just before its matrix instruction).
  const int lane = threadIdx.x & 63;
  bf16x8_t in[4], nx[4];
#pragma unroll
  for (int R = 0; R < 2; R++) {
    asm volatile("" ::: "memory");
    f32x16_t acc = {};
#pragma unroll
    for (int s = 0; s < 3; s++)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cn_ld(w.w1[R][s][lane]), xb[s], acc, 0, 0, 0);
    if (zs) zs[0][R] = acc;
    cn_act_pack(acc, bias[0], R, nx[2 * R], nx[2 * R + 1]);
  }
#pragma unroll
  for (int l = 0; l < 4; l++) {
#pragma unroll
    for (int s = 0; s < 4; s++) in[s] = nx[s];
#pragma unroll
    for (int R = 0; R < 2; R++) {
      asm volatile("" ::: "memory");
      f32x16_t acc = {};
#pragma unroll
      for (int s = 0; s < 4; s++)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cn_ld(w.wh[l][R][s][lane]), in[s], acc, 0,
                                                      0, 0);
      if (zs) zs[l + 1][R] = acc;
      cn_act_pack(acc, bias[l + 1], R, nx[2 * R], nx[2 * R + 1]);
    }
  }
  asm volatile("" ::: "memory");
  f32x16_t out = {};
#pragma unroll
  for (int s = 0; s < 4; s++)
    out = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cn_ld(w.w6[s][lane]), nx[s], out, 0, 0, 0);
  return out;
}

// FP32 inputs of features f: x = f in_scale + in_shift.
__device__ inline void cn_inputs(const CutNetDev &n, const float *f, float *x) {
#pragma unroll
  for (int k = 0; k < CN_NI; k++) x[k] = fmaf(f[k], n.in_scale[k], n.in_shift[k]);
}

// The encoded input row of one sample: hi[18], lo[18], zeros (48 BF16).
__device__ inline void cn_encode_row(const float *x, uint16_t *row) {
#pragma unroll
  for (int k = 0; k < CN_NI; k++) {
    const uint16_t hb = cn_bf16_bits(x[k]);
    const float hv = __builtin_bit_cast(float, (uint32_t)hb << 16);
    row[k] = hb;
    row[CN_NI + k] = cn_bf16_bits(x[k] - hv);
  }
#pragma unroll
  for (int k = 2 * CN_NI; k < CN_KE; k++) row[k] = 0;
}

// Network outputs (target units) of a whole wave: every lane passes its
// sample's 18 features and receives its 9 outputs.  `n` and the staging
// area (64 rows of CN_KE BF16 for this wave, 16-byte aligned) are in LDS.
// Every lane of the wave must call it.
__device__ inline void cn_eval_wave(const CutNetDev &n, const float *f, float *y,
                                    uint16_t *stage) {
  const int lane = threadIdx.x & 63, h = lane >> 5, c = lane & 31;
  float x[CN_NI];
  cn_inputs(n, f, x);
  cn_encode_row(x, stage + lane * CN_KE);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  float o1[CN_NO];
  f32x16_t out0;
#pragma unroll
  for (int T = 0; T < 2; T++) {
    bf16x8_t xb[3];
    const Bf8 *row = reinterpret_cast<const Bf8 *>(stage + (32 * T + c) * CN_KE);
#pragma unroll
    for (int s = 0; s < 3; s++) xb[s] = cn_ld(row[2 * s + h]);
    const f32x16_t out = cn_forward_tile(n.w, n.b, xb);
    if (T == 0) out0 = out;
    else
#pragma unroll
      for (int k = 0; k < CN_NO; k++) o1[k] = out[k];
  }
  // the lower half holds the outputs of samples c (T = 0) and 32 + c
  // (T = 1); the upper half takes the latter from lane c
#pragma unroll
  for (int k = 0; k < CN_NO; k++) {
    const float t1 = __shfl(o1[k], c, 64);
    const float z = (h == 0 ? out0[k] : t1) + n.b6[k];
    float acc = n.skip[k][CN_NI];
#pragma unroll
    for (int q = 0; q < CN_NI; q++) acc = fmaf(n.skip[k][q], x[q], acc);
    y[k] = fmaf(n.out_scale[k], z, acc);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace kmpm

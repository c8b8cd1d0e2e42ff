// kmpm_mlp.h -- the learned constitutive surrogate: a 2 -> 32 -> 32 -> 1
// ISRU-activated MLP whose hidden 32x32 layer runs on the gfx950 BF16 matrix cores
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation).  Layers 1 and 3 are FP32
// vector arithmetic.
//
// A wave evaluates 64 samples (one per lane) as two tiles of 32: for tile T
// the MFMA computes D[o][p] = sum_k W2[o][k] H1[k][p] over the tile's 32
// samples p.  Operand layouts (lane l, half = l / 32, col = l % 32):
//   A = W2 : lane l holds W2[col][16 s + 8 half + i], i = 0..7, k-step s
//   B = H1 : lane l holds H1[16 s + 8 half + i][col]
//   D      : lane l holds D[8 (i / 4) + 4 half + i % 4][col], i = 0..15
// so a lane computes the 16 layer-1 units it feeds for sample 32 T + col and
// receives 16 layer-2 units of the same sample; the two halves of the wave
// combine their layer-3 partial sums with one cross-half shuffle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kmpm {

constexpr int MLP_IN = 2;
constexpr int MLP_H = 32;

// Parameters in the order of the flat FP32 parameter vector.
struct MlpParams {
  float w1[MLP_H][MLP_IN];
  float b1[MLP_H];
  float w2[MLP_H][MLP_H];   // [out][in]
  float b2[MLP_H];
  float w3[MLP_H];
  float b3;
};
constexpr int MLP_NPARAM = MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1;
static_assert(sizeof(MlpParams) == MLP_NPARAM * sizeof(float), "flat parameter layout");

// A network as the kernels read it: FP32 parameters, the BF16 copy of W2
// the matrix cores consume, the affine maps of inputs and output, and a
// quadratic baseline in the inputs x (least-squares fit of the targets,
// coefficients of 1, x0, x1, x0^2, x0 x1, x1^2).  The network learns the
// residual over the baseline:
//   target = base(x) + out_shift + out_scale * z.
// W2's BF16 rows are padded to MLP_W2_LD values (80 bytes): the A-operand
// reads of 16 lanes (16 bytes each, one row per lane) then fall on 16
// distinct groups of 4 LDS banks.
constexpr int MLP_W2_LD = MLP_H + 8;
struct MlpNet {
  MlpParams p;
  uint16_t w2_bf16[MLP_H][MLP_W2_LD];
  float in_scale[MLP_IN], in_shift[MLP_IN];   // feature -> network input
  float out_scale, out_shift;                 // network output -> residual
  float base[6];
};

__host__ __device__ inline float mlp_base(const MlpNet &n, float x0, float x1) {
  return n.base[0] + x0 * (n.base[1] + n.base[3] * x0 + n.base[4] * x1) +
         x1 * (n.base[2] + n.base[5] * x1);
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

__device__ inline __bf16 to_bf16(float x) { return (__bf16)x; }   // round to nearest even

__device__ inline __bf16 bf16_from_bits(uint16_t b) {
  return __builtin_bit_cast(__bf16, b);
}

// The activation: the inverse square root unit x / sqrt(1 + x^2) (Carlile
// et al. 2017, "ISRU"), odd, bounded by +-1 like tanh, one hardware
// reciprocal square root where tanh needs an exponential and a reciprocal.
__device__ inline float mlp_act(float x) {
  return x * __builtin_amdgcn_rsqf(fmaf(x, x, 1.0f));
}
// Its derivative from the activation h: (1 + x^2)^(-3/2) = (1 - h^2)^(3/2).
__device__ inline float mlp_act_grad(float h) {
  const float t = fmaf(-h, h, 1.0f);
  return t * __builtin_amdgcn_sqrtf(t);
}

// The A operand (W2) of this lane for k-steps 0 and 1.
__device__ inline void mlp_load_w2(const MlpNet &n, bf16x8_t &a0, bf16x8_t &a1) {
  const int lane = threadIdx.x & 63, half = lane >> 5, col = lane & 31;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    a0[i] = bf16_from_bits(n.w2_bf16[col][8 * half + i]);
    a1[i] = bf16_from_bits(n.w2_bf16[col][16 + 8 * half + i]);
  }
}

// Layer 1 for sample (x0, x1) of this lane's tile column: the 16 units the
// lane feeds into the MFMA, as FP32 (h) and as the B operands.
__device__ inline void mlp_layer1(const MlpParams &p, float x0, float x1, float h[16],
                                  bf16x8_t &b0, bf16x8_t &b1) {
  const int half = (threadIdx.x & 63) >> 5;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int k = (i < 8 ? 0 : 16) + 8 * half + (i & 7);
    h[i] = mlp_act(fmaf(p.w1[k][1], x1, fmaf(p.w1[k][0], x0, p.b1[k])));
    if (i < 8) b0[i] = to_bf16(h[i]);
    else b1[i - 8] = to_bf16(h[i]);
  }
}

// Row (output unit) of accumulator element i of this lane.
__device__ inline int mlp_drow(int i) {
  const int half = (threadIdx.x & 63) >> 5;
  return 8 * (i >> 2) + 4 * half + (i & 3);
}

// Network output (in target units) for this lane's own features (f0, f1);
// the network's affine maps are applied here.  Every lane of the wave must
// call it (the matrix instructions use the whole wave).  Layer 1 is formed
// one k-step (8 units) at a time, just before the matrix instruction that
// consumes it, to keep few values live.
__device__ inline float mlp_eval_wave(const MlpNet &n, float f0, float f1) {
  const int lane = threadIdx.x & 63, half = lane >> 5, col = lane & 31;
  const float x0 = fmaf(f0, n.in_scale[0], n.in_shift[0]);
  const float x1 = fmaf(f1, n.in_scale[1], n.in_shift[1]);
  float own = 0.0f;
#pragma unroll
  for (int T = 0; T < 2; T++) {
    const int src = 32 * T + col;
    const float g0 = __shfl(x0, src, 64), g1 = __shfl(x1, src, 64);
    f32x16_t d = {};
#pragma unroll
    for (int st = 0; st < 2; st++) {
      // Weights are re-read at each use: the opaque barriers keep the
      // compiler from hoisting all of them into registers up front, which
      // would cost ~90 registers in the kernels that call this.
      asm volatile("" ::: "memory");
      bf16x8_t a, b;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int k = 16 * st + 8 * half + i;
        a[i] = bf16_from_bits(n.w2_bf16[col][k]);
        b[i] = to_bf16(mlp_act(fmaf(n.p.w1[k][1], g1, fmaf(n.p.w1[k][0], g0, n.p.b1[k]))));
      }
      d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, d, 0, 0, 0);
    }
    float part = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      if ((i & 7) == 0) asm volatile("" ::: "memory");
      const int o = mlp_drow(i);
      part = fmaf(n.p.w3[o], mlp_act(d[i] + n.p.b2[o]), part);
    }
    part += __shfl_xor(part, 32, 64);
    if (T == half) own = part + n.p.b3;
  }
  return fmaf(own, n.out_scale, n.out_shift) + mlp_base(n, x0, x1);
}

}  // namespace kmpm

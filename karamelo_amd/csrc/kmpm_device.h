// kmpm_device.h -- device-side data layout and point functions of the
// MI355X explicit MPM step.  Built with -ffp-contract=off: every rounding in
// the shape functions and cell indices matches the reference expression
// order; accumulation loops use explicit fma().
#pragma once

#include <cstddef>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/kmpm.h"

namespace kmpm {

// ---- particle state (FP64 plain path), structure of arrays --------------
// Double fields, each an array of `cap` doubles at base + field * cap.
// Symmetric tensors are stored as xx, yy, zz, xy, xz, yz.
enum DField {
  PX = 0, PY, PZ,            // position x
  VX, VY, VZ,                // velocity v
  F00, F01, F02, F10, F11, F12, F20, F21, F22,  // deformation gradient, row-major
  SXX, SYY, SZZ, SXY, SXZ, SYZ,                  // Cauchy stress sigma
  EXX, EYY, EZZ, EXY, EXZ, EYZ,                  // elastic strain strain_el
  EP, EPDOT, DINIT, DMG,     // eff_plastic_strain, its rate, damage_init, damage
  MASS, VOL0, VOL,           // mass, vol0, vol
  NDF                        // number of double fields
};
// TLMPM fields after the common ones: reference position x0 and vol0 PK1
// (first Piola-Kirchhoff stress times vol0, row-major), solid.h:50-97.
enum TLField {
  X0X = NDF, X0Y, X0Z,
  P00, P01, P02, P10, P11, P12, P20, P21, P22,
  NDF_TL
};
// APIC family: the affine matrix C of the transfers (L for ULMPM, Fdot for
// TLMPM, row-major), kept from one step to the next (solid.cpp:392-426).
// The APIC layout holds every TLMPM field too, so one layout serves both
// methods.
enum ApicField {
  C00 = NDF_TL, C01, C02, C10, C11, C12, C20, C21, C22,
  NDF_AP
};
// Thermo-mechanical contexts (FP64): temperature, heat source and heat flux
// (solid.h: T, gamma, q), after every other field.
enum ThermField { TEMP = NDF_AP, GAMMA, QX, QY, QZ, NDF_TH };
// CPDI contexts (FP64, 2D): the particle domain after the TLMPM fields.
// DOM0..7 the current one (R4: the vectors r1, r2 as x, y pairs; Q4: the
// four corners' x, y), REF0..7 the reference one (rp0 / xpc0), solid.h.
enum CpdiField { DOM0 = NDF_TL, REF0 = DOM0 + 8, NDF_CP = REF0 + 8 };
// Int fields: PID = index into the static (upload-order) arrays, MAT = material.
enum IField { PID = 0, MATI, NIF };

// Mixed-precision layout (KMPM_PRECISION_MIXED): position and deformation
// gradient stay FP64 (a double block of 12 fields: x, y, z, F00..F22), every
// other field is FP32 (a float block of NDF - 12 fields in DField order).
// TLMPM adds the reference position to the double block (x0, slots 12-14;
// the weights and cell keys are computed from it) and vol0 PK1 to the float
// block (after the common fields); the APIC layout adds C to the float block
// after those.
constexpr int MX_ND = 12;
constexpr int MX_NF = NDF - MX_ND;
constexpr int MX_ND_TL = MX_ND + 3;
constexpr int MX_NF_TL = MX_NF + 9;
constexpr int MX_NF_AP = MX_NF_TL + 9;
__host__ __device__ constexpr bool mx_is_double(int f) {
  return f <= PZ || (f >= F00 && f <= F22) || (f >= X0X && f <= X0Z);
}
__host__ __device__ constexpr int mx_dslot(int f) {
  return f <= PZ ? f : f <= F22 ? f - F00 + 3 : f - X0X + MX_ND;
}
__host__ __device__ constexpr int mx_fslot(int f) {
  return f < F00 ? f - VX : f < NDF ? f - SXX + 3 : f - P00 + MX_NF;
}
static_assert(mx_fslot(VOL) == MX_NF - 1 && mx_dslot(F22) == MX_ND - 1, "mixed layout");
static_assert(mx_fslot(C00) == MX_NF_TL && mx_fslot(C22) == MX_NF_AP - 1, "mixed APIC layout");
// State layouts: ULMPM, TLMPM, the APIC layout (TLMPM's fields and C), the
// thermo-mechanical layout (all of those and the thermal fields; FP64
// only) and the CPDI layout (TLMPM's fields and the domains; FP64 only).
enum Layout { LAY_UL = 0, LAY_TL = 1, LAY_AP = 2, LAY_TH = 3, LAY_CP = 4 };
// Field counts of the double and float blocks per layout and precision.
__host__ __device__ constexpr int state_nd(bool mixed, int lay) {
  return mixed ? (lay != LAY_UL ? MX_ND_TL : MX_ND)
               : (lay == LAY_UL   ? NDF
                  : lay == LAY_TL ? NDF_TL
                  : lay == LAY_AP ? NDF_AP
                  : lay == LAY_TH ? NDF_TH
                                  : NDF_CP);
}
__host__ __device__ constexpr int state_nf(bool mixed, int lay) {
  return mixed ? (lay == LAY_UL ? MX_NF : lay == LAY_TL ? MX_NF_TL : MX_NF_AP) : 0;
}

// Real type of the bulk state and of the transfer arithmetic per precision.
template <int PREC> struct Real { using T = double; using T2 = double2; };
template <> struct Real<KMPM_PRECISION_MIXED> { using T = float; using T2 = float2; };

// Fixed-point grid accumulation (mixed path): scaled node sums are 64-bit
// integers, so their value does not depend on the order of the adds.
enum FxField { FX_MASS = 0, FX_MOM, FX_FORCE, FX_N };
// A single flush (a run of particles of one cell) may reach 2^54 in scaled
// units before the step reports KMPM_ERR_FIXED_POINT; the scales put one
// particle's bound at 2^46, leaving room for 2^9 flush sums per node.
constexpr double FX_FLUSH_LIMIT = 18014398509481984.0;   // 2^54
constexpr int FX_CONTRIB_BITS = 46;

// ---- material table entry ------------------------------------------------
struct DevMat {
  int type, eos, strength, damage;
  double rho0, K, G, lambda, signal_velocity;
  double cfl_k43g;            // K + FOUR_THIRD * G  (solid.cpp:1382)
  double jc_thermal;          // JC thermal factor at T = 0 (strength_jc.cpp:127-133), host-computed
  double eos_p[9];
  double strength_p[9];
  double damage_p[8];
  // temperature law (material.cpp:726-742): temp_on when cp != 0, then
  // chi, 1/cp, kappa, the thermal-pressure alpha, T0 and Tm of
  // temperature_plastic_work.cpp:50-55
  int temp_on, pad_t;
  double chi, invcp, kappa, talpha, T0, Tmelt;
};

// Material classes the stress kernels are specialised for (chosen on the
// host from the material table): only the plugins of the class are compiled
// into the kernel, which keeps its register footprint small.
enum MatClass { MC_GENERIC = 0, MC_LINEAR = 1, MC_JC_SHOCK = 2, MC_FLUID = 3 };

// A velocity fix: constant values (val, prev) or, with expr != 0, the
// programs KArgs::fexpr[6 * index + 0..2] (value) and [+ 3..5] (previous).
struct DevFixVel { uint32_t bit; int set[3]; int expr; double val[3], prev[3]; };
struct DevFixBF { uint32_t bit; int set[3]; double val[3]; };
constexpr int MAX_FIX = 8;
// Blocks of the particle energy reduction (k_energy), summed in a fixed order.
constexpr int ENERGY_BLOCKS = 64;
// Per-tile cell sort: cells per tile covered by its LDS offset table, and the
// particles per tile it holds in registers (TILE_SORT_KPL per lane of one
// wavefront) and orders in LDS; larger tiles order through memory.
constexpr int TILE_SORT_MAX_CPT = 1024;
constexpr int TILE_SORT_KPL = 16;
constexpr int TILE_SORT_CAP = 64 * TILE_SORT_KPL;

// ---- device scalars (time, dt, CFL accumulator, sticky error) -------------
struct DevScalars {
  double dt, time;
  unsigned long long max_c_bits;  // max wave speed of the current step (bits of a
                                  // non-negative double; ordered like the value)
  int64_t step;
  int error;                      // first error code (sticky)
  int frozen;                     // set by the step that raised `error`: later steps
                                  // launch but change nothing
  int64_t err_pid;                // pid of the first offending particle
  int tile_count;                 // active tiles this step
  int dil_count;                  // tiles whose owned nodes receive contributions
  unsigned long long min_ratio_bits;  // TLMPM: smallest |Re eig(F)| of the step, at most 1
                                      // (bits of a non-negative double)
  unsigned long long surr_evals;      // particle updates that used a surrogate network
  unsigned long long surr_fallbacks;  // particle updates outside the surrogate's envelope
  // reaction force of each velocity fix, double-buffered by step parity: the
  // step that starts with `step` completed steps sums into [step & 1], and
  // k_adjust_dt clears the other buffer for the next step
  double fv_force[2][MAX_FIX][3];
  double fv_last[MAX_FIX][6];            // the last grid update's fix values (val, prev)
  // cut network: network updates the plain functions also judged, and the
  // judged sums of squared deviatoric-stress error and of the plain stress
  unsigned long long judge_n;
  double judge_err2, judge_ref2;
};
// k_tiles appends to both tile lists with one 64-bit atomic on the pair
static_assert(offsetof(DevScalars, tile_count) % 8 == 0 &&
                  offsetof(DevScalars, dil_count) == offsetof(DevScalars, tile_count) + 4,
              "tile_count and dil_count form one aligned 64-bit word");

// Per-instance scalars of a batched context (independent instances packed
// side by side on one grid, each with its own dt, time and CFL).
struct InstScalars {
  double dt, time;
  unsigned long long max_c_bits;
  unsigned long long min_ratio_bits;
  int tlo[3], thi[3];   // the instance's tiles: [tlo, thi) per axis
};

constexpr double SQRT_3_OVER_2 = 1.224744871;                            // strength_jc.cpp:29
constexpr double FOUR_THIRD = 1.333333333333333333333333333333333333333; // solid.cpp:39

// ---- shape functions: basis_functions.h:21-104 ------------------------------
__device__ __host__ inline double bf_linear(double r_) {
  double r = fabs(r_);
  if (r >= 1.0) return 0.0;
  return 1.0 - r;
}
__device__ __host__ inline double bf_dlinear(double r, double inv_h) {
  if (r >= 1.0 || r <= -1.0 || r == 0) return 0.0;
  else if (r > 0.0) return -inv_h;
  else return inv_h;
}
__device__ __host__ inline double bf_cubic(double r, int ntype) {
  if (r >= 1 && r < 2) {
    if (ntype == 1) return 0;
    return ((-1.0 / 6.0 * r + 1) * r - 2) * r + 4.0 / 3.0;
  } else if (r >= 0 && r < 1) {
    if (ntype == -2) return (1.0 / 6.0 * r * r - 1) * r + 1;
    else if (ntype == 2) return 1;
    else if (ntype == 1) return (1.0 / 3.0 * r - 1) * r * r + 2.0 / 3.0;
    else return (0.5 * r - 1) * r * r + 2.0 / 3.0;
  } else if (r >= -1 && r < 0) {
    if (ntype == 2) return (-1.0 / 6.0 * r * r + 1) * r + 1;
    else if (ntype == -1) return (-1.0 / 3.0 * r - 1) * r * r + 2.0 / 3.0;
    else return (-0.5 * r - 1) * r * r + 2.0 / 3.0;
  } else if (r >= -2 && r < -1) {
    return ((1.0 / 6.0 * r + 1) * r + 2) * r + 4.0 / 3.0;
  }
  return 0;
}
__device__ __host__ inline double bf_dcubic(double r, int ntype, double ih) {
  if (r >= 1 && r < 2) {
    if (ntype == 1) return -ih;
    return ih * ((-0.5 * r + 2) * r - 2);
  } else if (r >= 0 && r < 1) {
    if (ntype == -2) return ih * (0.5 * r * r - 1);
    else if (ntype == 2) return ih;
    else if (ntype == 1) return ih * r * (r - 2);
    else return ih * (3.0 / 2.0 * r - 2) * r;
  } else if (r >= -1 && r < 0) {
    if (ntype == 2) return ih * (-0.5 * r * r + 1);
    else if (ntype == -1) return ih * (-r - 2) * r;
    else return ih * (-3.0 / 2.0 * r - 2) * r;
  } else if (r >= -2 && r < -1) {
    return ih * ((0.5 * r + 2) * r + 2);
  }
  return 0;
}

// Node type of node index i on an axis with n nodes (grid.cpp:236-239).
__device__ __host__ inline int node_type(int i, int n) {
  int a = i < 2 ? i : 2;
  int b = (n - 1 - i) < 2 ? (n - 1 - i) : 2;
  return a - b;
}

// Cell of a coordinate for the sort: trunc((x - lo) * inv_h) clamped to
// [0, n-1]; flags NaN and out-of-range values.
__device__ inline int cell_of(double x, double lo, double inv_h, int n, bool &bad) {
  double u = (x - lo) * inv_h;
  if (!(u >= 0.0)) { bad = true; return 0; }          // negative or NaN
  if (u >= (double)(n - 1)) { if (u > (double)(n - 1)) bad = true; return n - 1; }
  return (int)u;
}

// 1D weights and derivatives of the S1 stencil nodes base..base+S1-1 of a
// particle at coordinate x on an axis (ulmpm.cpp:244-311 with the pair
// filter of :252-265 applied per factor: a zero value zeroes the derivative,
// so a dropped pair contributes nothing).  Nodes outside [0, n) get 0.
template <int S1>
__device__ inline void weights_1d(double x, double lo, double h, double inv_h, int n, int base,
                                  double *w, double *dw) {
  if (S1 == 4 && base >= 2 && base + 3 <= n - 3) {
    // Interior stencil (every node type 0): stencil slot a lies on the
    // spline piece r in [1 - a, 2 - a), so each slot evaluates its piece's
    // expression of bf_cubic / bf_dcubic directly (the same operations, so
    // the same roundings).  A position whose r falls outside the expected
    // piece by rounding takes the general branches below.
    double r[4];
#pragma unroll
    for (int a = 0; a < 4; a++) r[a] = (x - (lo + (0 + base + a) * h)) * inv_h;
    if (r[0] >= 1 && r[0] < 2 && r[1] >= 0 && r[1] < 1 && r[2] >= -1 && r[2] < 0 &&
        r[3] >= -2 && r[3] < -1) {
      w[0] = ((-1.0 / 6.0 * r[0] + 1) * r[0] - 2) * r[0] + 4.0 / 3.0;
      dw[0] = inv_h * ((-0.5 * r[0] + 2) * r[0] - 2);
      w[1] = (0.5 * r[1] - 1) * r[1] * r[1] + 2.0 / 3.0;
      dw[1] = inv_h * (3.0 / 2.0 * r[1] - 2) * r[1];
      w[2] = (-0.5 * r[2] - 1) * r[2] * r[2] + 2.0 / 3.0;
      dw[2] = inv_h * (-3.0 / 2.0 * r[2] - 2) * r[2];
      w[3] = ((1.0 / 6.0 * r[3] + 1) * r[3] + 2) * r[3] + 4.0 / 3.0;
      dw[3] = inv_h * ((0.5 * r[3] + 2) * r[3] + 2);
#pragma unroll
      for (int a = 0; a < 4; a++)
        if (w[a] == 0) dw[a] = 0;
      return;
    }
  }
#pragma unroll
  for (int a = 0; a < S1; a++) {
    int i = base + a;
    double s = 0, d = 0;
    if (i >= 0 && i < n) {
      double xn = lo + (0 + i) * h;          // grid.cpp:227
      double r = (x - xn) * inv_h;           // ulmpm.cpp:248
      if (S1 == 2) {
        s = bf_linear(r);
        d = bf_dlinear(r, inv_h);
      } else {
        int nt = node_type(i, n);
        s = bf_cubic(r, nt);
        d = bf_dcubic(r, nt, inv_h);
      }
      if (s == 0) d = 0;
    }
    w[a] = s;
    dw[a] = d;
  }
}

// ---- 3x3 helpers ------------------------------------------------------------
__device__ inline double det3(const double *m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// Frobenius norm of a symmetric tensor stored xx,yy,zz,xy,xz,yz, summed in the
// row-major order of the full matrix.
__device__ inline double norm_sym(const double *s) {
  double t = s[0] * s[0];
  t += s[3] * s[3];
  t += s[4] * s[4];
  t += s[3] * s[3];
  t += s[1] * s[1];
  t += s[5] * s[5];
  t += s[4] * s[4];
  t += s[5] * s[5];
  t += s[2] * s[2];
  return sqrt(t);
}

}  // namespace kmpm

// kmpm_surrogate.h -- Solid::update_stress for a johnson_cook + shock
// (+ damage_johnson_cook) material with learned material functions.
//
// The surrogate replaces the Johnson-Cook flow stress and failure strain,
// the transcendental parts of StrengthJohnsonCook::update_deviatoric_stress
// (strength_jc.cpp:115-133) and DamageJohnsonCook::compute_damage
// (damage_jc.cpp:124-138), with two small MLPs (kmpm_mlp.h):
//   net y : (eps_p^(1/4), ln epsdot*)              -> sigma_y / A
//   net f : (triaxiality, ln(epsdot / epsdot0_d))  -> ln eps_f
// The radial return, the damage accumulation and every branch of the
// reference (damage >= 1, eps_p < 1e-10, triax <= -3, the epsdot filter,
// the tension rule) stay exact.  Inputs outside the training envelope use
// the plain FP64 functions (the fallback), counted in DevScalars.
#pragma once

#include "kmpm_constitutive.h"
#include "kmpm_mlp.h"

namespace kmpm {

struct SurrogateDev {
  MlpNet y, f;
  int enabled;
  int pad;
  // training envelope
  double ep_hi;          // eps_p in [1e-10, ep_hi]
  double lnr_y_hi;       // ln epsdot* in [0, lnr_y_hi]
  double triax_hi;       // triaxiality in (-3, triax_hi]
  double lnr_f_hi;       // ln(epsdot / epsdot0_d) in [0, lnr_f_hi]
  // yield stress of a particle that has not yielded (eps_p < 1e-10) at
  // epsdot* = 1: A (1 + 1)^C times the thermal factor (strength_jc.cpp:115-133)
  double sy_virgin;
};

// The plain functions as out-of-line calls: the fallback is rare, and kept
// out of line its registers do not add to the step kernel's footprint.
__device__ __noinline__ void jc_strength_fallback(const DevMat &m, const double *sig,
                                                  const double *D, double dt, double eps_p,
                                                  double epsdot, double damage, double *out,
                                                  double *dps) {
  jc_strength(m, sig, D, dt, eps_p, epsdot, damage, out, dps);
}
__device__ __noinline__ void jc_damage_fallback(const DevMat &m, double &damage_init,
                                                double &damage, double pH, const double *sdev,
                                                double epsdot, double dps) {
  jc_damage(m, damage_init, damage, pH, sdev, epsdot, dps);
}

// A wave's count of network evaluations and fallbacks, kept in registers
// (every lane holds the wave's totals) and added to the context's counters
// once per wave at the end of the kernel.
struct SurrCount {
  unsigned long long ev, fb;
};
__device__ inline void surrogate_count(SurrCount &c, bool eval, bool fallback) {
  c.ev += (unsigned long long)__popcll(__ballot(eval));
  c.fb += (unsigned long long)__popcll(__ballot(fallback));
}
__device__ inline void surrogate_count_flush(const SurrCount &c, DevScalars *sc) {
  if ((threadIdx.x & 63) == 0) {
    if (c.ev) atomicAdd(&sc->surr_evals, c.ev);
    if (c.fb) atomicAdd(&sc->surr_fallbacks, c.fb);
  }
}

// Solid::update_stress eos-strength branch (solid.cpp:1292-1374) for the
// JC + shock class on the mixed path, from the EOS pressure pH (formed in
// FP64 from J by the G2P kernel, since a small compression cancels in
// rho / rho0 - 1, then rounded to FP32) and the rate of deformation Df
// (FP32); `dt` is the step's.  Every lane of the wave calls it; `live` is
// false for lanes without a particle (their results are discarded).  The
// arithmetic is FP32, that of the mixed state, except the plain-function
// fallbacks, which stay FP64.
__device__ inline void update_stress_jc_surrogate_p(const DevMat &m, const SurrogateDev &sd,
                                                    bool live, double dt, double h, float pH,
                                                    const float *Df, float *sig, float *eel,
                                                    float &ep, float &epdot, float &dinit,
                                                    float &dmg, SurrCount &cnt) {
  const float dtf = (float)dt;
  double D[6];
  for (int e = 0; e < 6; e++) D[e] = Df[e];
  float epf = ep, epdf = epdot, dif = dinit, dmf = dmg;
  // ---- strength (strength_jc.cpp:100-184) -------------------------------
  const float G_ = (float)m.strength_p[0], A = (float)m.strength_p[1],
              epsdot0 = (float)m.strength_p[4];
  const bool dead = dmf >= 1.0f;
  float ratio = __fdividef(epdf, epsdot0);
  ratio = ratio > 1.0f ? ratio : 1.0f;
  const float fy0 = sqrtf(sqrtf(epf));
  const float fy1 = __builtin_amdgcn_logf(ratio) * 0.69314718055994531f;
  const bool virgin = epf < 1.0e-10f && ratio == 1.0f;
  const bool use_y = live && !dead && !virgin && epf >= 1.0e-10f && epf <= (float)sd.ep_hi &&
                     fy1 <= (float)sd.lnr_y_hi;
  const bool fb_y = live && !dead && !virgin && !use_y;
  float ny = 1.0f;
  if (__ballot(use_y)) ny = mlp_eval_wave(sd.y, fy0, fy1);
  float sdev[6], dps = 0.0f;
  if (fb_y) {
    // the call takes addresses: copies keep the stress, D and sdev in
    // registers on the other paths (only these copies live in the stack
    // frame)
    double sig_c[6], D_c[6], out_c[6], dps_c = 0.0;
    for (int e = 0; e < 6; e++) {
      sig_c[e] = sig[e];
      D_c[e] = D[e];
    }
    jc_strength_fallback(m, sig_c, D_c, dt, epf, epdf, dmf, out_c, &dps_c);
    for (int e = 0; e < 6; e++) sdev[e] = (float)out_c[e];
    dps = (float)dps_c;
  } else if (dead) {
    for (int e = 0; e < 6; e++) sdev[e] = 0;
  } else {
    float ys = virgin ? (float)sd.sy_virgin : A * ny;
    float Gd = G_;
    if (dmf > 0) {
      Gd *= (1 - dmf);
      ys *= (1 - dmf);
    }
    const float c = dtf * 2.0f * Gd;
    float t[6];
    for (int e = 0; e < 6; e++) t[e] = sig[e] + c * Df[e];
    const float t3 = (t[0] + t[1] + t[2]) / 3.0f;
    t[0] -= t3;
    t[1] -= t3;
    t[2] -= t3;
    const float J2 = (float)SQRT_3_OVER_2 * norm_sym_t(t);
    if (J2 < ys) {
      for (int e = 0; e < 6; e++) sdev[e] = t[e];
    } else {
      dps = __fdividef(J2 - ys, 3.0f * Gd);
      const float f = __fdividef(ys, J2);
      for (int e = 0; e < 6; e++) sdev[e] = t[e] * f;
    }
  }
  epf += dps;                                                       // solid.cpp:1319
  const float tav = (float)(1000 * h / m.signal_velocity);          // solid.cpp:1324-1329
  epdf -= __fdividef(epdf * dtf, tav);
  epdf += __fdividef(dps, tav);
  epdf = epdf > 0.0f ? epdf : 0.0f;
  // ---- damage (damage_jc.cpp:91-143) -------------------------------------
  bool use_f = false, fb_f = false;
  if (m.damage == KMPM_DAMAGE_JOHNSON_COOK) {
    const float d4 = (float)m.damage_p[3], epsdot0_d = (float)m.damage_p[5];
    const bool skip = dps == 0 && dmf >= 1.0f;
    const float vm = (float)SQRT_3_OVER_2 * norm_sym_t(sdev);
    float triax = 0.0f;
    if (pH != 0.0f && vm != 0.0f) triax = __fdividef(-pH, vm + 0.001f * fabsf(pH));
    const bool reset = !skip && triax <= -3;
    const float lnr = (d4 > 0.0f && epdf > epsdot0_d) ? __logf(__fdividef(epdf, epsdot0_d)) : 0.0f;
    // eps_f only matters when plastic strain accumulates (dps / eps_f)
    use_f = live && !skip && !reset && dps > 0 && triax <= (float)sd.triax_hi &&
            lnr <= (float)sd.lnr_f_hi;
    fb_f = live && !skip && !reset && dps > 0 && !use_f;
    float nf = 0.0f;
    if (__ballot(use_f)) nf = mlp_eval_wave(sd.f, triax, lnr);
    if (fb_f) {
      double di_c = dif, dm_c = dmf, sdev_c[6];
      for (int e = 0; e < 6; e++) sdev_c[e] = sdev[e];
      jc_damage_fallback(m, di_c, dm_c, pH, sdev_c, epdf, dps);
      dif = (float)di_c;
      dmf = (float)dm_c;
    } else if (reset) {
      dif = 0;
    } else if (!skip) {
      if (use_f) dif += __fdividef(dps, __builtin_amdgcn_exp2f(nf * 1.4426950408889634f));
      if (dif >= 1.0f) {
        const float dd = (dif - 1.0f) * 10;
        dmf = dd < 1.0f ? dd : 1.0f;
      }
    }
  }
  surrogate_count(cnt, use_y || use_f, fb_y || fb_f);
  // ---- assembly (solid.cpp:1353-1366) -----------------------------------
  const float trD = Df[0] + Df[1] + Df[2];
  const float trE = eel[0] + eel[1] + eel[2];
  if (dmf == 0 || pH >= 0) {
    for (int e = 0; e < 3; e++) sig[e] = -pH * 1.0f + sdev[e];
  } else {
    const float pd = -pH * (1.0f - dmf);
    for (int e = 0; e < 3; e++) sig[e] = pd * 1.0f + sdev[e];
  }
  for (int e = 3; e < 6; e++) sig[e] = sdev[e];
  const float t3 = (dtf * trD + trE) / 3.0f;
  const float Gd = dmf > 1e-10f ? (float)m.G * (1 - dmf) : (float)m.G;
  for (int e = 0; e < 3; e++) eel[e] = t3 * 1.0f + __fdividef(sdev[e], Gd);
  for (int e = 3; e < 6; e++) eel[e] = __fdividef(sdev[e], Gd);
  ep = epf;
  epdot = epdf;
  dinit = dif;
  dmg = dmf;
}

}  // namespace kmpm

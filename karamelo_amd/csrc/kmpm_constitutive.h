// kmpm_constitutive.h -- device constitutive point functions.
// Expression order follows the reference plugins so the FP64 path differs
// from the oracle only through summation order upstream and the device
// math library (pow/exp/log/sqrt are the ROCm device implementations).
#pragma once

#include "kmpm_device.h"

namespace kmpm {

// Symmetric tensor index helpers: xx,yy,zz,xy,xz,yz
__device__ inline double tr_sym(const double *s) { return s[0] + s[1] + s[2]; }

// x^y for x > 0 as exp(y log x): a few ulp, and far fewer registers than
// the general-case pow (used where the base is positive by construction).
__device__ inline double pow_pos(double x, double y) { return exp(y * log(x)); }

// StrengthJohnsonCook::update_deviatoric_stress, strength_jc.cpp:100-184 (T = 0).
// out = final deviatoric stress, *dps = plastic strain increment (untouched
// when damage >= 1, the caller holds 0).
__device__ inline void jc_strength(const DevMat &m, const double *sig, const double *D, double dt,
                                   double eps_p, double epsdot, double damage, double *out,
                                   double *dps) {
  const double G_ = m.strength_p[0], A = m.strength_p[1], B = m.strength_p[2],
               n = m.strength_p[3], epsdot0 = m.strength_p[4], C = m.strength_p[5],
               mm = m.strength_p[6], Tr = m.strength_p[7], Tm = m.strength_p[8];
  const double Tmr = Tm - Tr, T = 0;
  if (damage >= 1.0) {
    for (int e = 0; e < 6; e++) out[e] = 0;
    return;
  }
  double epsdot_ratio = epsdot / epsdot0;
  epsdot_ratio = epsdot_ratio > 1.0 ? epsdot_ratio : 1.0;
  double ys;
  if (eps_p < 1.0e-10) ys = A;
  else ys = A + B * pow_pos(eps_p, n);                 // eps_p >= 1e-10
  if (C != 0) ys *= pow_pos(1.0 + epsdot_ratio, C);   // base >= 2
  // Thermal softening: with no temperature law T = 0 and the factor
  // (1 - ((T-Tr)/Tmr)^m, 1, or 0 above Tm) depends on the material only.
  ys *= m.jc_thermal;
  (void)T; (void)Tmr; (void)Tr; (void)Tm; (void)mm;
  double Gd = G_;
  if (damage > 0) {
    Gd *= (1 - damage);
    ys *= (1 - damage);
  }
  const double c = dt * 2.0 * Gd;
  double t[6];
  for (int e = 0; e < 6; e++) t[e] = sig[e] + c * D[e];
  const double t3 = (t[0] + t[1] + t[2]) / 3.0;   // MPM_Math::Deviator, mpm_math.h:28-33
  t[0] = t[0] - t3;
  t[1] = t[1] - t3;
  t[2] = t[2] - t3;
  const double J2 = SQRT_3_OVER_2 * norm_sym(t);
  if (J2 < ys) {
    *dps = 0.0;
    for (int e = 0; e < 6; e++) out[e] = t[e];
  } else {
    *dps = (J2 - ys) / (3.0 * Gd);
    const double f = ys / J2;
    for (int e = 0; e < 6; e++) out[e] = t[e] * f;
  }
}

// EOSShock::compute_pressure, eos_shock.cpp:99-133 (T = 0).
__device__ inline double shock_pressure(const DevMat &m, double J, double rho, double damage,
                                        const double *D, double cellsize) {
  const double rho0_ = m.eos_p[0], c0 = m.eos_p[2], S = m.eos_p[3], Gamma = m.eos_p[4],
               cv = m.eos_p[5], Tr = m.eos_p[6], Q1 = m.eos_p[7], Q2 = m.eos_p[8];
  const double alpha = cv * rho0_, e0 = 0, T = 0;
  const double mu = rho / rho0_ - 1.0;
  const double q1 = 1.0 - (S - 1.0) * mu;
  const double pH = rho0_ * (c0 * c0) * mu * (1.0 + mu) / (q1 * q1);
  const double e = T > Tr ? alpha * (T - Tr) : 0;
  double p = pH + Gamma * (e - e0);
  if (damage > 0.0) {
    if (p < 0.0) {
      if (damage >= 1.0) p = 0;
      else p *= 1.0 - damage;
    }
  }
  if (Q1 != 0 || Q2 != 0) {
    const double tr_eps = tr_sym(D);
    if (tr_eps < 0) {
      const double q =
          rho * cellsize * (Q1 * cellsize * tr_eps * tr_eps - Q2 * c0 * sqrt(J) * tr_eps);
      p += q;
    }
  }
  return p;
}

// DamageJohnsonCook::compute_damage, damage_jc.cpp:91-143 (T = 0).
__device__ inline void jc_damage(const DevMat &m, double &damage_init, double &damage, double pH,
                                 const double *sdev, double epsdot, double dps) {
  const double d1 = m.damage_p[0], d2 = m.damage_p[1], d3 = m.damage_p[2], d4 = m.damage_p[3],
               d5 = m.damage_p[4], epsdot0 = m.damage_p[5], Tr = m.damage_p[6],
               Tm = m.damage_p[7];
  const double Tmr = Tm - Tr, T = 0;
  if (dps == 0 && damage >= 1.0) return;
  const double vm = SQRT_3_OVER_2 * norm_sym(sdev);
  double triax = 0.0;
  if (pH != 0.0 && vm != 0.0) triax = -pH / (vm + 0.001 * fabs(pH));
  if (triax <= -3) {
    damage_init = 0;
    return;
  }
  double jc = d1 + d2 * exp(d3 * triax);
  if (d4 > 0.0) {
    if (epsdot > epsdot0) {
      const double r = epsdot / epsdot0;
      jc *= (1.0 + d4 * log(r));
    }
  }
  if (d5 > 0.0 && T >= Tr) jc *= 1 + d5 * (T - Tr) / Tmr;
  damage_init += dps / jc;
  if (damage_init >= 1.0) {
    const double dd = (damage_init - 1.0) * 10;
    damage = dd < 1.0 ? dd : 1.0;
  }
}

// Solid::update_stress per particle, solid.cpp:1262-1374, for one particle.
// sig, eel: symmetric (in/out); D symmetric rate of deformation.
template <int MC>
__device__ inline void update_stress_point(const DevMat &m, double dt, double h, double J,
                                           double rho, const double *D, double *sig, double *eel,
                                           double &ep, double &epdot, double &dinit,
                                           double &dmg) {
  const bool linear = MC == MC_LINEAR || (MC == MC_GENERIC && m.type == KMPM_MAT_LINEAR);
  if (linear) {   // :1268-1280
    double si[6];
    for (int e = 0; e < 6; e++) si[e] = dt * D[e];
    for (int e = 0; e < 6; e++) eel[e] += si[e];
    const double tr = si[0] + si[1] + si[2];
    const double lt = m.lambda * tr;
    const double g2 = 2 * m.G;
    for (int e = 0; e < 3; e++) sig[e] += g2 * si[e] + lt * 1.0;
    for (int e = 3; e < 6; e++) sig[e] += g2 * si[e] + lt * 0.0;
    return;
  }
  if (MC == MC_LINEAR) return;
  // eos-strength branch, :1292-1374
  const bool shock = MC == MC_JC_SHOCK || (MC == MC_GENERIC && m.eos == KMPM_EOS_SHOCK);
  const bool efluid = MC == MC_FLUID || (MC == MC_GENERIC && m.eos == KMPM_EOS_FLUID);
  const bool jc = MC == MC_JC_SHOCK || (MC == MC_GENERIC && m.strength == KMPM_STRENGTH_JOHNSON_COOK);
  const bool sfluid = MC == MC_FLUID || (MC == MC_GENERIC && m.strength == KMPM_STRENGTH_FLUID);
  double pH;
  if (shock) {
    pH = shock_pressure(m, J, rho, dmg, D, h);
  } else if (efluid) {                               // eos_fluid.cpp:68-73
    const double mu = rho / m.eos_p[0];
    pH = m.eos_p[1] * (pow(mu, m.eos_p[2]) - 1.0);
  } else {                                           // eos_linear.cpp:71-74
    pH = m.eos_p[1] * (1 - J) * (1 - dmg);
  }
  double sdev[6], dps = 0;
  if (jc) {
    jc_strength(m, sig, D, dt, ep, epdot, dmg, sdev, &dps);
  } else if (sfluid) {                               // strength_fluid.cpp:46-58
    const double t3 = (D[0] + D[1] + D[2]) / 3.0;
    const double c2 = 2.0 * m.strength_p[0];
    sdev[0] = c2 * (D[0] - t3);
    sdev[1] = c2 * (D[1] - t3);
    sdev[2] = c2 * (D[2] - t3);
    for (int e = 3; e < 6; e++) sdev[e] = c2 * D[e];
  } else {                                           // strength_linear.cpp:59-71
    const double t3 = (D[0] + D[1] + D[2]) / 3.0;
    const double s3 = (sig[0] + sig[1] + sig[2]) / 3.0;
    const double c2 = 2.0 * m.strength_p[0] * (1 - dmg);
    for (int e = 0; e < 6; e++) {
      const double dv = e < 3 ? D[e] - t3 : D[e];
      const double ds = e < 3 ? sig[e] - s3 : sig[e];
      sdev[e] = ds + dt * (c2 * dv);
    }
  }
  ep += dps;                                         // :1319
  const double tav = 1000 * h / m.signal_velocity;  // :1324-1329
  epdot -= epdot * dt / tav;
  epdot += dps / tav;
  epdot = epdot > 0.0 ? epdot : 0.0;
  if (MC != MC_FLUID && m.damage == KMPM_DAMAGE_JOHNSON_COOK)
    jc_damage(m, dinit, dmg, pH, sdev, epdot, dps);
  const double trD = D[0] + D[1] + D[2];
  const double trE = eel[0] + eel[1] + eel[2];
  if (dmg == 0 || pH >= 0) {                        // :1353-1356
    for (int e = 0; e < 3; e++) sig[e] = -pH * 1.0 + sdev[e];
  } else {
    const double pd = -pH * (1.0 - dmg);
    for (int e = 0; e < 3; e++) sig[e] = pd * 1.0 + sdev[e];
  }
  for (int e = 3; e < 6; e++) sig[e] = sdev[e];
  const double t3 = (dt * trD + trE) / 3.0;          // :1358-1366
  const double Gd = dmg > 1e-10 ? m.G * (1 - dmg) : m.G;
  for (int e = 0; e < 3; e++) eel[e] = t3 * 1.0 + sdev[e] / Gd;
  for (int e = 3; e < 6; e++) eel[e] = sdev[e] / Gd;
}

}  // namespace kmpm

// kmpm_constitutive.h -- device constitutive point functions.
// Expression order follows the reference plugins so the FP64 path differs
// from the oracle only through summation order upstream and the device
// math library (pow/exp/log/sqrt are the ROCm device implementations).
//
// The functions are templates on the arithmetic type T.  T = double is the
// FP64 path.  T = float is the mixed path's constitutive update: its state
// is stored in FP32, and the increments are formed in FP32 with the
// hardware exponential, logarithm and reciprocal (relative errors near
// 1e-6, the order of the FP32 storage rounding of the updated fields).
#pragma once

#include "kmpm_device.h"

namespace kmpm {

// Math of the arithmetic type: the double versions are the device library's;
// the float versions use the hardware approximations.
__device__ inline double m_exp(double x) { return exp(x); }
__device__ inline float m_exp(float x) { return __expf(x); }
__device__ inline double m_log(double x) { return log(x); }
__device__ inline float m_log(float x) { return __logf(x); }
__device__ inline double m_sqrt(double x) { return sqrt(x); }
__device__ inline float m_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline double m_div(double a, double b) { return a / b; }
__device__ inline float m_div(float a, float b) { return __fdividef(a, b); }

// Symmetric tensor index helpers: xx,yy,zz,xy,xz,yz
template <class T>
__device__ inline T tr_sym(const T *s) { return s[0] + s[1] + s[2]; }

// Frobenius norm of a symmetric tensor (the order of norm_sym).
template <class T>
__device__ inline T norm_sym_t(const T *s) {
  T t = s[0] * s[0];
  t += s[3] * s[3];
  t += s[4] * s[4];
  t += s[3] * s[3];
  t += s[1] * s[1];
  t += s[5] * s[5];
  t += s[4] * s[4];
  t += s[5] * s[5];
  t += s[2] * s[2];
  return m_sqrt(t);
}

// x^y for x > 0 as exp(y log x): a few ulp, and far fewer registers than
// the general-case pow (used where the base is positive by construction).
template <class T>
__device__ inline T pow_pos(T x, T y) { return m_exp(y * m_log(x)); }

// Thermal inputs and output of one particle's stress update on a
// thermo-mechanical context (FP64): the particle temperature T, whether the
// method is thermo-mechanical (the damage then sees T, solid.cpp:1331-1341),
// the volume that scales the heat source (vol, or vol0 on TLMPM) and the
// heat source gamma (solid.cpp:1343-1351).
struct ThermalPt {
  double T;
  int method_temp;
  double volg;
  double gamma;
};

// StrengthJohnsonCook::update_deviatoric_stress, strength_jc.cpp:100-184 (T = 0).
// out = final deviatoric stress, *dps = plastic strain increment (untouched
// when damage >= 1, the caller holds 0).
// hot: the material's temperature law is on and Tt is the particle's
// temperature (strength_jc.cpp:127-133 evaluated per particle).
template <class T>
__device__ inline void jc_strength(const DevMat &m, const T *sig, const T *D, T dt, T eps_p,
                                   T epsdot, T damage, T *out, T *dps, bool hot = false,
                                   double Tt = 0.0) {
  const T G_ = (T)m.strength_p[0], A = (T)m.strength_p[1], B = (T)m.strength_p[2],
          n = (T)m.strength_p[3], epsdot0 = (T)m.strength_p[4], C = (T)m.strength_p[5];
  if (damage >= T(1.0)) {
    for (int e = 0; e < 6; e++) out[e] = 0;
    return;
  }
  T epsdot_ratio = m_div(epsdot, epsdot0);
  epsdot_ratio = epsdot_ratio > T(1.0) ? epsdot_ratio : T(1.0);
  T ys;
  if (eps_p < T(1.0e-10)) ys = A;
  else ys = A + B * pow_pos(eps_p, n);                 // eps_p >= 1e-10
  if (C != T(0)) ys *= pow_pos(T(1.0) + epsdot_ratio, C);   // base >= 2
  // Thermal softening: with no temperature law T = 0 and the factor
  // (1 - ((T-Tr)/Tmr)^m, 1, or 0 above Tm) depends on the material only.
  if (hot) {
    const double mm = m.strength_p[6], Tr = m.strength_p[7], Tm = m.strength_p[8];
    if (Tt < Tm) {
      if (mm != 0 && Tt >= Tr) ys *= (T)(1.0 - pow((Tt - Tr) / (Tm - Tr), mm));
    } else {
      ys = 0;
    }
  } else {
    ys *= (T)m.jc_thermal;
  }
  T Gd = G_;
  if (damage > T(0)) {
    Gd *= (T(1) - damage);
    ys *= (T(1) - damage);
  }
  const T c = dt * T(2.0) * Gd;
  T t[6];
  for (int e = 0; e < 6; e++) t[e] = sig[e] + c * D[e];
  const T t3 = (t[0] + t[1] + t[2]) / T(3.0);   // MPM_Math::Deviator, mpm_math.h:28-33
  t[0] = t[0] - t3;
  t[1] = t[1] - t3;
  t[2] = t[2] - t3;
  const T J2 = T(SQRT_3_OVER_2) * norm_sym_t(t);
  if (J2 < ys) {
    *dps = T(0.0);
    for (int e = 0; e < 6; e++) out[e] = t[e];
  } else {
    *dps = m_div(J2 - ys, T(3.0) * Gd);
    const T f = m_div(ys, J2);
    for (int e = 0; e < 6; e++) out[e] = t[e] * f;
  }
}

// EOSShock::compute_pressure, eos_shock.cpp:99-133 at temperature Tt (0
// without a temperature law).
template <class T>
__device__ inline T shock_pressure(const DevMat &m, T J, T rho, T damage, const T *D, T cellsize,
                                   T Tt = 0) {
  const T rho0_ = (T)m.eos_p[0], c0 = (T)m.eos_p[2], S = (T)m.eos_p[3], Gamma = (T)m.eos_p[4],
          cv = (T)m.eos_p[5], Tr = (T)m.eos_p[6], Q1 = (T)m.eos_p[7], Q2 = (T)m.eos_p[8];
  const T alpha = cv * rho0_, e0 = 0;
  const T mu = m_div(rho, rho0_) - T(1.0);
  const T q1 = T(1.0) - (S - T(1.0)) * mu;
  const T pH = m_div(rho0_ * (c0 * c0) * mu * (T(1.0) + mu), q1 * q1);
  const T e = Tt > Tr ? alpha * (Tt - Tr) : T(0);
  T p = pH + Gamma * (e - e0);
  if (damage > T(0.0)) {
    if (p < T(0.0)) {
      if (damage >= T(1.0)) p = 0;
      else p *= T(1.0) - damage;
    }
  }
  if (Q1 != T(0) || Q2 != T(0)) {
    const T tr_eps = tr_sym(D);
    if (tr_eps < T(0)) {
      const T q =
          rho * cellsize * (Q1 * cellsize * tr_eps * tr_eps - Q2 * c0 * m_sqrt(J) * tr_eps);
      p += q;
    }
  }
  return p;
}

// DamageJohnsonCook::compute_damage, damage_jc.cpp:91-143 at temperature Tt
// (0 unless the method is thermo-mechanical).
template <class T>
__device__ inline void jc_damage(const DevMat &m, T &damage_init, T &damage, T pH, const T *sdev,
                                 T epsdot, T dps, T Tt = 0) {
  const T d1 = (T)m.damage_p[0], d2 = (T)m.damage_p[1], d3 = (T)m.damage_p[2],
          d4 = (T)m.damage_p[3], d5 = (T)m.damage_p[4], epsdot0 = (T)m.damage_p[5],
          Tr = (T)m.damage_p[6], Tm = (T)m.damage_p[7];
  const T Tmr = Tm - Tr;
  if (dps == T(0) && damage >= T(1.0)) return;
  const T vm = T(SQRT_3_OVER_2) * norm_sym_t(sdev);
  T triax = T(0.0);
  if (pH != T(0.0) && vm != T(0.0)) triax = m_div(-pH, vm + T(0.001) * fabs(pH));
  if (triax <= T(-3)) {
    damage_init = 0;
    return;
  }
  T jc = d1 + d2 * m_exp(d3 * triax);
  if (d4 > T(0.0)) {
    if (epsdot > epsdot0) {
      const T r = m_div(epsdot, epsdot0);
      jc *= (T(1.0) + d4 * m_log(r));
    }
  }
  if (d5 > T(0.0) && Tt >= Tr) jc *= T(1) + m_div(d5 * (Tt - Tr), Tmr);
  damage_init += m_div(dps, jc);
  if (damage_init >= T(1.0)) {
    const T dd = (damage_init - T(1.0)) * T(10);
    damage = dd < T(1.0) ? dd : T(1.0);
  }
}

// Solid::update_stress per particle, solid.cpp:1262-1374, for one particle.
// sig, eel: symmetric (in/out); D symmetric rate of deformation.  J and rho
// stay FP64 and so does the pressure: mu = rho / rho0 - 1 and (rho /
// rho0)^Gamma - 1 cancel, and FP32 would keep only a few digits of a small
// compression.
// th (thermo-mechanical contexts): the particle's temperature state; a
// material with a temperature law evaluates its EOS and strength at T, adds
// the thermal pressure alpha (T0 - T) and sets the heat source
// (solid.cpp:1301-1351).
template <int MC, class T = double>
__device__ inline void update_stress_point(const DevMat &m, T dt, T h, double J, double rho,
                                           const T *D, T *sig, T *eel, T &ep, T &epdot, T &dinit,
                                           T &dmg, ThermalPt *th = nullptr) {
  const bool linear = MC == MC_LINEAR || (MC == MC_GENERIC && m.type == KMPM_MAT_LINEAR);
  if (linear) {   // :1268-1280
    T si[6];
    for (int e = 0; e < 6; e++) si[e] = dt * D[e];
    for (int e = 0; e < 6; e++) eel[e] += si[e];
    const T tr = si[0] + si[1] + si[2];
    const T lt = (T)m.lambda * tr;
    const T g2 = T(2) * (T)m.G;
    for (int e = 0; e < 3; e++) sig[e] += g2 * si[e] + lt * T(1.0);
    for (int e = 3; e < 6; e++) sig[e] += g2 * si[e] + lt * T(0.0);
    return;
  }
  if (MC == MC_LINEAR) return;
  // eos-strength branch, :1292-1374
  const bool shock = MC == MC_JC_SHOCK || (MC == MC_GENERIC && m.eos == KMPM_EOS_SHOCK);
  const bool efluid = MC == MC_FLUID || (MC == MC_GENERIC && m.eos == KMPM_EOS_FLUID);
  const bool jc = MC == MC_JC_SHOCK || (MC == MC_GENERIC && m.strength == KMPM_STRENGTH_JOHNSON_COOK);
  const bool sfluid = MC == MC_FLUID || (MC == MC_GENERIC && m.strength == KMPM_STRENGTH_FLUID);
  const bool hot = th != nullptr && m.temp_on;
  const double Tc = hot ? th->T : 0.0;
  double pHd;
  if (shock) {
    double Dd[6];
    for (int e = 0; e < 6; e++) Dd[e] = D[e];
    pHd = shock_pressure<double>(m, J, rho, dmg, Dd, h, Tc);
  } else if (efluid) {                               // eos_fluid.cpp:68-73
    const double mu = rho / m.eos_p[0];
    pHd = m.eos_p[1] * (pow(mu, m.eos_p[2]) - 1.0);
  } else {                                           // eos_linear.cpp:71-74
    pHd = m.eos_p[1] * (1 - J) * (1 - (double)dmg);
  }
  if (hot) pHd += m.talpha * (m.T0 - th->T);        // temperature_plastic_work.cpp:74-76
  const T pH = (T)pHd;
  T sdev[6], dps = 0;
  if (jc) {
    jc_strength(m, sig, D, dt, ep, epdot, dmg, sdev, &dps, hot, Tc);
  } else if (sfluid) {                               // strength_fluid.cpp:46-58
    const T t3 = (D[0] + D[1] + D[2]) / T(3.0);
    const T c2 = T(2.0) * (T)m.strength_p[0];
    sdev[0] = c2 * (D[0] - t3);
    sdev[1] = c2 * (D[1] - t3);
    sdev[2] = c2 * (D[2] - t3);
    for (int e = 3; e < 6; e++) sdev[e] = c2 * D[e];
  } else {                                           // strength_linear.cpp:59-71
    const T t3 = (D[0] + D[1] + D[2]) / T(3.0);
    const T s3 = (sig[0] + sig[1] + sig[2]) / T(3.0);
    const T c2 = T(2.0) * (T)m.strength_p[0] * (T(1) - dmg);
    for (int e = 0; e < 6; e++) {
      const T dv = e < 3 ? D[e] - t3 : D[e];
      const T ds = e < 3 ? sig[e] - s3 : sig[e];
      sdev[e] = ds + dt * (c2 * dv);
    }
  }
  ep += dps;                                         // :1319
  const T tav = m_div(T(1000) * h, (T)m.signal_velocity);  // :1324-1329
  epdot -= m_div(epdot * dt, tav);
  epdot += m_div(dps, tav);
  epdot = epdot > T(0.0) ? epdot : T(0.0);
  if (MC != MC_FLUID && m.damage == KMPM_DAMAGE_JOHNSON_COOK)
    jc_damage(m, dinit, dmg, pH, sdev, epdot, dps,
              (T)(th != nullptr && th->method_temp ? th->T : 0.0));
  if (hot) {                                          // :1343-1351
    const double flow = SQRT_3_OVER_2 * (double)norm_sym_t(sdev);
    double g = th->T < m.Tmelt ? m.chi * flow * (double)epdot : 0.0;   // :65-71
    g *= th->volg * m.invcp;
    th->gamma = g;
  }
  const T trD = D[0] + D[1] + D[2];
  const T trE = eel[0] + eel[1] + eel[2];
  if (dmg == T(0) || pH >= T(0)) {                  // :1353-1356
    for (int e = 0; e < 3; e++) sig[e] = -pH * T(1.0) + sdev[e];
  } else {
    const T pd = -pH * (T(1.0) - dmg);
    for (int e = 0; e < 3; e++) sig[e] = pd * T(1.0) + sdev[e];
  }
  for (int e = 3; e < 6; e++) sig[e] = sdev[e];
  const T t3 = (dt * trD + trE) / T(3.0);            // :1358-1366
  const T Gd = dmg > T(1e-10) ? (T)m.G * (T(1) - dmg) : (T)m.G;
  for (int e = 0; e < 3; e++) eel[e] = t3 * T(1.0) + m_div(sdev[e], Gd);
  for (int e = 3; e < 6; e++) eel[e] = m_div(sdev[e], Gd);
}

}  // namespace kmpm

// kmpm_reforder.hip -- the reference-order FP64 step (KMPM_ORDER_REFERENCE).
//
// The fast kernels sum the grid and the stencil gathers in orders of their
// own (cell-sorted LDS flushes, per-axis contractions), so FP64 results
// agree with the reference's to rounding, and on chaotic problems (bodies in
// contact over thousands of steps) those roundings grow.  This mode sums in
// the reference's order instead, for small problems:
//   P2G   one thread per node; its particles in ascending particle index
//         (the node-major lists of ulmpm.cpp:269-275): mass
//         (solid.cpp:317-335), velocity per solid divided by the node's
//         total mass (:337-390), internal force (:482-522);
//   grid  grid.cpp:448-466 with the body-force sum of
//         fix_body_force.cpp:117-146 and the velocity fixes;
//   G2P   one thread per particle over its stencil in the reference's
//         candidate order (ulmpm.cpp:167-242) with the pair filter
//         (:250-265): v_update, acceleration (solid.cpp:576-635), L
//         (:860-936), then F, J, D and the stress update as the fast path.
// Every weight of a (particle, node) pair is evaluated by the reference's
// expressions (basis_functions.h:21-104, ulmpm.cpp:244-311) where it is
// used.  The state stays in upload order (no sort).  P2G costs
// O(nodes x particles): the mode is for parity runs, not for speed.
// ULMPM, FP64, one instance.
#include "kmpm_kernels.h"

namespace kmpm {

namespace {

// Storage index of node (i, j, k) in the tile-major grid arrays.
__device__ inline int64_t ref_sidx(const KArgs &a, int i, int j, int k) {
  const TileGeom &g = a.tg;
  const int t = ((i / g.T[0]) * g.nt[1] + j / g.T[1]) * g.nt[2] + k / g.T[2];
  return (int64_t)t * g.cpt + ((i % g.T[0]) * g.T[1] + j % g.T[1]) * g.T[2] + k % g.T[2];
}

// Base cell of a particle (ulmpm.cpp:161-165 linear, :204-208 spline).
template <int S1>
__device__ inline int ref_base(double x, double lo, double inv_h) {
  return S1 == 2 ? (int)((x - lo) * inv_h) : (int)((x - lo) * inv_h - 1);
}

__device__ inline double ref_basis(int S1, double r, int nt, double inv_h, bool der) {
  if (S1 == 2) return der ? bf_dlinear(r, inv_h) : bf_linear(r);
  return der ? bf_dcubic(r, nt, inv_h) : bf_cubic(r, nt);
}

// Weight and gradient of the pair (particle at xp, node (i, j, k)) in the
// reference's expression order (ulmpm.cpp:244-311); false for a pair the
// filter drops (zero weight).
template <int DIM, int S1>
__device__ inline bool ref_pair(const KArgs &a, const double *xp, int i, int j, int k, double &wf,
                                double wd[3]) {
  const int idx[3] = {i, j, k};
  double r[3], sv[3], sd[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++) {
    const double xn = d < DIM ? a.lo[d] + (0 + idx[d]) * a.h : 0.0;   // grid.cpp:227
    r[d] = (xp[d] - xn) * a.inv_h;
  }
  const int nt0 = node_type(i, a.n[0]), nt1 = node_type(j, a.n[1]),
            nt2 = DIM == 3 ? node_type(k, a.n[2]) : 0;
  sv[0] = ref_basis(S1, r[0], nt0, a.inv_h, false);
  wf = sv[0];
  if (wf != 0) {
    sv[1] = ref_basis(S1, r[1], nt1, a.inv_h, false);
    wf *= sv[1];
    if (DIM == 3 && wf != 0) {
      sv[2] = ref_basis(S1, r[2], nt2, a.inv_h, false);
      wf *= sv[2];
    } else {
      sv[2] = 1;
    }
  }
  if (wf == 0) return false;
  sd[0] = ref_basis(S1, r[0], nt0, a.inv_h, true);
  sd[1] = ref_basis(S1, r[1], nt1, a.inv_h, true);
  if (DIM == 3) sd[2] = ref_basis(S1, r[2], nt2, a.inv_h, true);
  if (DIM == 3) {
    wd[0] = sd[0] * sv[1] * sv[2];
    wd[1] = sv[0] * sd[1] * sv[2];
    wd[2] = sv[0] * sv[1] * sd[2];
  } else {
    wd[0] = sd[0] * sv[1];
    wd[1] = sv[0] * sd[1];
    wd[2] = 0;
  }
  return true;
}

// Weight positions of particle slot p: the step's start positions (MUSL
// keeps them in xo before the advance) or the current ones.
__device__ inline void ref_xw(const KArgs &a, int p, bool from_xo, double xp[3]) {
  for (int d = 0; d < 3; d++)
    xp[d] = from_xo ? a.xo[(size_t)d * a.cap + p] : PF(a.pa, PX + d, p);
}

__device__ inline bool ref_touches(const KArgs &a, int S1, int dim, const double *xp, int i, int j,
                                   int k) {
  const int bi = S1 == 2 ? (int)((xp[0] - a.lo[0]) * a.inv_h) : (int)((xp[0] - a.lo[0]) * a.inv_h - 1);
  if (i < bi || i >= bi + S1) return false;
  const int bj = S1 == 2 ? (int)((xp[1] - a.lo[1]) * a.inv_h) : (int)((xp[1] - a.lo[1]) * a.inv_h - 1);
  if (j < bj || j >= bj + S1) return false;
  if (dim == 3) {
    const int bk =
        S1 == 2 ? (int)((xp[2] - a.lo[2]) * a.inv_h) : (int)((xp[2] - a.lo[2]) * a.inv_h - 1);
    if (k < bk || k >= bk + S1) return false;
  }
  return true;
}

}  // namespace

// P2G (REMAP = false) or the MUSL momentum re-map (REMAP = true, weights of
// the step's start positions in xo, ulmpm.cpp:476-496).  Writes mass and v,
// and the internal force into v_update (the grid kernel turns it into the
// updated velocity).
template <int DIM, int S1, bool REMAP>
__global__ __launch_bounds__(256) void k_ref_p2g(KArgs a) {
  if (frozen(a)) return;
  const int64_t tag = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tag >= a.nn) return;
  const int n1 = a.n[1], n2 = a.n[2];
  const int k = (int)(tag % n2), j = (int)((tag / n2) % n1), i = (int)(tag / ((int64_t)n2 * n1));
  const int64_t s = ref_sidx(a, i, j, k);
  const int64_t ns = a.ns;
  double wf, wd[3], xp[3];
  double m = 0;
  if (REMAP) {
    m = a.gmass[s];
  } else {
    for (int p = 0; p < a.np; p++) {                        // solid.cpp:328-332
      ref_xw(a, p, false, xp);
      if (!ref_touches(a, S1, DIM, xp, i, j, k) || !ref_pair<DIM, S1>(a, xp, i, j, k, wf, wd))
        continue;
      m += wf * PF(a.pa, MASS, p);
    }
    a.gmass[s] = m;
  }
  // velocity, per solid: v += (sum_solid (w m) v) / m  (solid.cpp:357-378)
  double gv[3] = {0, 0, 0}, vt[3] = {0, 0, 0};
  int sol = -1;
  bool any = false;
  for (int p = 0; p < a.np; p++) {
    ref_xw(a, p, REMAP, xp);
    if (!ref_touches(a, S1, DIM, xp, i, j, k) || !ref_pair<DIM, S1>(a, xp, i, j, k, wf, wd))
      continue;
    const int sp = a.solid_of[IFLD(a.ia, PID, p)];
    if (any && sp != sol) {
      if (m > 0)
        for (int d = 0; d < 3; d++) gv[d] += vt[d] / m;
      vt[0] = vt[1] = vt[2] = 0;
    }
    sol = sp;
    any = true;
    const double wm = wf * PF(a.pa, MASS, p);
    for (int d = 0; d < 3; d++) vt[d] += wm * PF(a.pa, VX + d, p);
  }
  if (any && m > 0)
    for (int d = 0; d < 3; d++) gv[d] += vt[d] / m;
  for (int d = 0; d < 3; d++) a.gv[d * ns + s] = gv[d];
  if (REMAP) return;
  // internal force f -= vol (sigma . wd)  (solid.cpp:510-515)
  double f[3] = {0, 0, 0};
  for (int p = 0; p < a.np; p++) {
    ref_xw(a, p, false, xp);
    if (!ref_touches(a, S1, DIM, xp, i, j, k) || !ref_pair<DIM, S1>(a, xp, i, j, k, wf, wd))
      continue;
    double sg[9];
    const int full[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    for (int e = 0; e < 9; e++) sg[e] = PF(a.pa, SXX + full[e], p);
    const double vol = PF(a.pa, VOL, p);
    for (int d = 0; d < 3; d++) {
      const double t = sg[3 * d] * wd[0] + sg[3 * d + 1] * wd[1] + sg[3 * d + 2] * wd[2];
      f[d] -= vol * t;
    }
  }
  for (int d = 0; d < 3; d++) a.gvu[d * ns + s] = f[d];
}

// Grid update (grid.cpp:448-466) with the body force mb = sum_fixes g m on
// nodes with mass (fix_body_force.cpp:117-146); the velocity fixes follow in
// k_ref_fixes.
__global__ __launch_bounds__(256) void k_ref_grid(KArgs a) {
  if (frozen(a)) return;
  const int64_t tag = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tag >= a.nn) return;
  const int n1 = a.n[1], n2 = a.n[2];
  const int k = (int)(tag % n2), j = (int)((tag / n2) % n1), i = (int)(tag / ((int64_t)n2 * n1));
  const int64_t s = ref_sidx(a, i, j, k);
  const int64_t ns = a.ns;
  const uint32_t mask = a.gmask[s];
  const double dt = a.sc->dt;
  {
    const double m = a.gmass[s];
    double mb[3] = {0, 0, 0};
    for (int q = 0; q < a.nbf; q++) {
      const DevFixBF &f = a.bf[q];
      if (m > 0 && (mask & f.bit)) {
        double fv[3] = {0, 0, 0};
        for (int d = 0; d < 3; d++)
          if (f.set[d]) fv[d] = f.val[d];
        for (int d = 0; d < 3; d++) fv[d] *= m;
        for (int d = 0; d < 3; d++) mb[d] += fv[d];
      }
    }
    for (int d = 0; d < 3; d++) {
      const double v = a.gv[d * ns + s], f = a.gvu[d * ns + s];
      a.gvu[d * ns + s] = m != 0 ? v + dt * (f + mb[d]) / m : v;
    }
  }
}

// The velocity fixes in the reference's order: fix after fix, each over the
// nodes in ascending tag, with the reaction force summed in that order
// (FixVelocityNodes::post_update_grid_state, fix_velocity_nodes.cpp:130-225;
// after the re-map, post_velocities_to_grid, :227-271).  One thread does the
// sums; the mode is for small problems.
__global__ __launch_bounds__(256) void k_ref_fixes(KArgs a, int remap) {
  if (frozen(a)) return;
  __shared__ double fxv[MAX_FIX][6];
  __shared__ double fsum[MAX_FIX][3];
  fix_values_to_lds(a, fxv, fsum);
  if (threadIdx.x != 0 || a.nfv == 0) return;
  const int64_t ns = a.ns;
  const int n1 = a.n[1], n2 = a.n[2];
  const double inv_dt = 1.0 / a.sc->dt;
  double *ftot = a.sc->fv_force[a.sc->step & 1][0];
  for (int q = 0; q < a.nfv; q++) {
    const DevFixVel &f = a.fv[q];
    double fx[3] = {0, 0, 0};
    for (int64_t tag = 0; tag < a.nn; tag++) {
      const int k = (int)(tag % n2), j = (int)((tag / n2) % n1),
                i = (int)(tag / ((int64_t)n2 * n1));
      const int64_t s = ref_sidx(a, i, j, k);
      if (!(a.gmask[s] & f.bit)) continue;
      double dv[3] = {0, 0, 0};
      for (int d = 0; d < 3; d++) {
        if (!f.set[d]) continue;
        if (remap) {
          a.gv[d * ns + s] = fxv[q][d];
          continue;
        }
        dv[d] = fxv[q][d] - a.gvu[d * ns + s];
        a.gvu[d * ns + s] = fxv[q][d];
        a.gv[d * ns + s] = fxv[q][3 + d];
      }
      const double c = inv_dt * a.gmass[s];
      for (int d = 0; d < 3; d++) fx[d] += c * dv[d];
    }
    if (!remap)
      for (int d = 0; d < 3; d++) ftot[3 * q + d] = fx[d];
  }
}

namespace {

// The stencil gather of particle p in the reference's candidate order:
// vu = sum w v_update, acc = sum w (v_update - v) (solid.cpp:576-600) and
// L = sum v_L (x) wd with v_L = v_update (USL) or v (MUSL re-mapped)
// (solid.cpp:860-936).
template <int DIM, int S1, bool WANT_VA, bool L_FROM_V>
__device__ inline void ref_gather(const KArgs &a, const double *xp, double vu[3], double acc[3],
                                  double L[9]) {
  const int64_t ns = a.ns;
  const int b0 = ref_base<S1>(xp[0], a.lo[0], a.inv_h), b1 = ref_base<S1>(xp[1], a.lo[1], a.inv_h);
  const int b2 = DIM == 3 ? ref_base<S1>(xp[2], a.lo[2], a.inv_h) : 0;
  for (int d = 0; d < 3; d++) vu[d] = acc[d] = 0;
  for (int e = 0; e < 9; e++) L[e] = 0;
  for (int i = b0; i < b0 + S1; i++)
    for (int j = b1; j < b1 + S1; j++)
      for (int k = b2; k < b2 + (DIM == 3 ? S1 : 1); k++) {
        if (i < 0 || i >= a.n[0] || j < 0 || j >= a.n[1] || k < 0 || k >= a.n[2]) continue;
        double w, wd[3];
        if (!ref_pair<DIM, S1>(a, xp, i, j, k, w, wd)) continue;
        const int64_t s = ref_sidx(a, i, j, k);
        double gu[3], gv[3];
        for (int d = 0; d < 3; d++) {
          gu[d] = a.gvu[d * ns + s];
          gv[d] = a.gv[d * ns + s];
        }
        if (WANT_VA) {
          for (int d = 0; d < 3; d++) vu[d] += w * gu[d];
          for (int d = 0; d < 3; d++) acc[d] += w * (gu[d] - gv[d]);
        }
        const double *vl = L_FROM_V ? gv : gu;
        if (DIM == 2) {
          L[0] += vl[0] * wd[0];
          L[1] += vl[0] * wd[1];
          L[3] += vl[1] * wd[0];
          L[4] += vl[1] * wd[1];
        } else {
          for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) L[3 * r + c] += vl[r] * wd[c];
        }
      }
}

// Advance (solid.cpp:603-627, 786-790): x += dt vu, the box check, the
// PIC/FLIP velocity; x and v in place.
__device__ inline void ref_advance(const KArgs &a, double dt, const double *vu, double *acc,
                                   double *x, double *v, int pid) {
  const double inv_dt = 1.0 / dt;
  for (int d = 0; d < 3; d++) acc[d] *= inv_dt;
  for (int d = 0; d < 3; d++) x[d] += dt * vu[d];
  if (outside_box(a, x)) set_error(a.sc, KMPM_ERR_PARTICLE_LEFT_BOX, pid);
  const double alpha = a.flip;
  for (int d = 0; d < 3; d++) v[d] = (1 - alpha) * vu[d] + alpha * (v[d] + dt * acc[d]);
}

// F, J, vol, the stress update and the wave speed of slot p from L; writes
// the particle's whole next state to slot p of pb.
template <int MC>
__device__ inline double ref_finish(const KArgs &a, double dt, int p, const double *x,
                                    const double *v, const double *L) {
  const int pid = IFLD(a.ia, PID, p), mat = IFLD(a.ia, MATI, p);
  double F[9];
  for (int e = 0; e < 9; e++) F[e] = PF(a.pa, F00 + e, p);
  const double J = update_F(dt, L, F);
  const double vol0 = PF(a.pa, VOL0, p);
  double dmg = PF(a.pa, DMG, p);
  if (J <= 0.0 && dmg < 1.0) set_error(a.sc, KMPM_ERR_JACOBIAN, pid);
  double sig[6], eel[6];
  for (int e = 0; e < 6; e++) {
    sig[e] = PF(a.pa, SXX + e, p);
    eel[e] = PF(a.pa, EXX + e, p);
  }
  double ep = PF(a.pa, EP, p), epdot = PF(a.pa, EPDOT, p), dinit = PF(a.pa, DINIT, p);
  const double c = stress_and_speed<MC, KMPM_PRECISION_FP64>(a, dt, L, J, sig, eel, ep, epdot,
                                                             dinit, dmg, mat, v, pid);
  for (int d = 0; d < 3; d++) {
    PF(a.pb, PX + d, p) = x[d];
    PF(a.pb, VX + d, p) = v[d];
  }
  for (int e = 0; e < 9; e++) PF(a.pb, F00 + e, p) = F[e];
  for (int e = 0; e < 6; e++) {
    PF(a.pb, SXX + e, p) = sig[e];
    PF(a.pb, EXX + e, p) = eel[e];
  }
  PF(a.pb, EP, p) = ep;
  PF(a.pb, EPDOT, p) = epdot;
  PF(a.pb, DINIT, p) = dinit;
  PF(a.pb, DMG, p) = dmg;
  PF(a.pb, MASS, p) = PF(a.pa, MASS, p);
  PF(a.pb, VOL0, p) = vol0;
  PF(a.pb, VOL, p) = J * vol0;
  IFLD(a.ib, PID, p) = pid;
  IFLD(a.ib, MATI, p) = mat;
  return c;
}

}  // namespace

// USL: L of v_update, G2P, advance, F, stress, CFL (usl.cpp:59-74).
template <int DIM, int S1, int MC>
__global__ __launch_bounds__(256) void k_ref_g2p_usl(KArgs a) {
  if (frozen(a)) {
    frozen_copy(a);
    return;
  }
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double cmax = 0;
  if (p < a.np) {
    const double dt = a.sc->dt;
    double x[3], v[3], vu[3], acc[3], L[9];
    for (int d = 0; d < 3; d++) {
      x[d] = PF(a.pa, PX + d, p);
      v[d] = PF(a.pa, VX + d, p);
    }
    ref_gather<DIM, S1, true, false>(a, x, vu, acc, L);
    ref_advance(a, dt, vu, acc, x, v, IFLD(a.ia, PID, p));
    const double c = ref_finish<MC>(a, dt, p, x, v, L);
    if (c > cmax) cmax = c;
  }
  block_max_to(cmax, &a.sc->max_c_bits);
}

// MUSL part A: G2P and advance in place; the start positions go to xo
// (musl.cpp:61-67).
template <int DIM, int S1>
__global__ __launch_bounds__(256) void k_ref_g2p_musl_a(KArgs a) {
  if (frozen(a)) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.np) return;
  const double dt = a.sc->dt;
  double x[3], v[3], vu[3], acc[3], L[9];
  for (int d = 0; d < 3; d++) {
    x[d] = PF(a.pa, PX + d, p);
    v[d] = PF(a.pa, VX + d, p);
    a.xo[(size_t)d * a.cap + p] = x[d];
  }
  ref_gather<DIM, S1, true, false>(a, x, vu, acc, L);
  ref_advance(a, dt, vu, acc, x, v, IFLD(a.ia, PID, p));
  for (int d = 0; d < 3; d++) {
    PF(a.pa, PX + d, p) = x[d];
    PF(a.pa, VX + d, p) = v[d];
  }
}

// MUSL part B: L of the re-mapped v with the start-of-step weights, F,
// stress, CFL (musl.cpp:75-77).
template <int DIM, int S1, int MC>
__global__ __launch_bounds__(256) void k_ref_stress_musl_b(KArgs a) {
  if (frozen(a)) {
    frozen_copy(a);
    return;
  }
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double cmax = 0;
  if (p < a.np) {
    const double dt = a.sc->dt;
    double xw[3], x[3], v[3], vu[3], acc[3], L[9];
    for (int d = 0; d < 3; d++) {
      xw[d] = a.xo[(size_t)d * a.cap + p];
      x[d] = PF(a.pa, PX + d, p);
      v[d] = PF(a.pa, VX + d, p);
    }
    ref_gather<DIM, S1, false, true>(a, xw, vu, acc, L);
    const double c = ref_finish<MC>(a, dt, p, x, v, L);
    if (c > cmax) cmax = c;
  }
  block_max_to(cmax, &a.sc->max_c_bits);
}

// ---- launchers --------------------------------------------------------------
namespace {
inline int ref_blocks(int64_t n) { return (int)((n + 255) / 256); }

#define REF_DS(c, CALL)                                              \
  do {                                                               \
    const int ds_ = (c)->a.dim * 10 + (c)->a.S1;                     \
    if (ds_ == 32) { constexpr int DIM = 3, S1 = 2; CALL; }          \
    else if (ds_ == 34) { constexpr int DIM = 3, S1 = 4; CALL; }     \
    else if (ds_ == 22) { constexpr int DIM = 2, S1 = 2; CALL; }     \
    else { constexpr int DIM = 2, S1 = 4; CALL; }                    \
  } while (0)

#define REF_MC(c, CALL)                                              \
  do {                                                               \
    switch ((c)->a.mat_class) {                                      \
      case MC_LINEAR: { constexpr int MC = MC_LINEAR; CALL; break; }       \
      case MC_JC_SHOCK: { constexpr int MC = MC_JC_SHOCK; CALL; break; }   \
      case MC_FLUID: { constexpr int MC = MC_FLUID; CALL; break; }         \
      default: { constexpr int MC = MC_GENERIC; CALL; break; }             \
    }                                                                \
  } while (0)
}  // namespace

hipError_t launch_ref_p2g(kmpm_ctx *c, int remap) {
  const KArgs &a = c->a;
  if (remap)
    REF_DS(c, hipLaunchKernelGGL((k_ref_p2g<DIM, S1, true>), dim3(ref_blocks(a.nn)), dim3(256), 0,
                                 c->stream, a));
  else
    REF_DS(c, hipLaunchKernelGGL((k_ref_p2g<DIM, S1, false>), dim3(ref_blocks(a.nn)), dim3(256),
                                 0, c->stream, a));
  return hipGetLastError();
}

hipError_t launch_ref_grid(kmpm_ctx *c, int remap) {
  const hipError_t e = launch_fix_values(c);
  if (e != hipSuccess) return e;
  if (!remap)
    hipLaunchKernelGGL(k_ref_grid, dim3(ref_blocks(c->a.nn)), dim3(256), 0, c->stream, c->a);
  hipLaunchKernelGGL(k_ref_fixes, dim3(1), dim3(256), 0, c->stream, c->a, remap);
  return hipGetLastError();
}

hipError_t launch_ref_g2p_usl(kmpm_ctx *c) {
  const KArgs &a = c->a;
  const int nb = ref_blocks(a.np > 0 ? a.np : 1);
  REF_DS(c, REF_MC(c, hipLaunchKernelGGL((k_ref_g2p_usl<DIM, S1, MC>), dim3(nb), dim3(256), 0,
                                         c->stream, a)));
  return hipGetLastError();
}

hipError_t launch_ref_g2p_musl_a(kmpm_ctx *c) {
  const KArgs &a = c->a;
  if (a.np == 0) return hipSuccess;
  REF_DS(c, hipLaunchKernelGGL((k_ref_g2p_musl_a<DIM, S1>), dim3(ref_blocks(a.np)), dim3(256), 0,
                               c->stream, a));
  return hipGetLastError();
}

hipError_t launch_ref_stress_musl_b(kmpm_ctx *c) {
  const KArgs &a = c->a;
  const int nb = ref_blocks(a.np > 0 ? a.np : 1);
  REF_DS(c, REF_MC(c, hipLaunchKernelGGL((k_ref_stress_musl_b<DIM, S1, MC>), dim3(nb), dim3(256),
                                         0, c->stream, a)));
  return hipGetLastError();
}

}  // namespace kmpm

// kmpm_cpdi.hip -- the CPDI methods on the device (ULCPDI, TLCPDI; 2D,
// FP64): particle-major kernels over the sorted particles.
//
// A CPDI particle spreads over the union of its domain corners' stencils
// (up to KMPM_CPDI_MAX_SPAN nodes per axis), with weights that are not a
// tensor product, so the tile kernels' column factorisation does not apply.
// Each kernel recomputes a particle's weights from the state they were
// taken from (the reference keeps neighbour lists): P2G adds them into
// node sums with FP64 atomics ([7][ns]: mass, momentum, force), a grid pass
// turns the sums into the node state with the tile path's node update, and
// the G2P and stress kernels gather node by node in ascending node tag, the
// order of the reference's lists.
//
// Weights of a particle (ulcpdi.cpp:111-393, tlcpdi.cpp:96-329): corners
// x -+ r1 -+ r2 (R4) or the tracked corners (Q4); TLCPDI takes them once,
// from x0 and the reference vectors or corners, and its candidate nodes from
// the box's lower corner although its nodes are the solid's (the
// reference's arithmetic, kept).
#include "kmpm_kernels.h"

namespace kmpm {

namespace {

constexpr int PREC = KMPM_PRECISION_FP64;   // CPDI contexts are FP64 (the state macros)

// Tile-major storage index of 2D node (i, j) for the context's tile shape.
__device__ inline int64_t cpdi_sidx(const KArgs &a, int i, int j) {
  const TileGeom &g = a.tg;
  const int t = (i / g.T[0]) * g.nt[1] + j / g.T[1];
  return (int64_t)t * g.cpt + (i % g.T[0]) * g.T[1] + j % g.T[1];
}

// One particle's domain and candidate nodes.
struct CpdiPt {
  double xc[4][2];       // corners
  double r[4];           // R4: r1x, r1y, r2x, r2y
  double inv_V, qa, qb, aov, six;   // Q4: 1/V, a, b, 0.0417/V, 6V
  int c0[4][2];          // candidate stencil origin per corner
  int lo[2], hi[2];      // the candidates' extent on the grid (inclusive)
};

// xp: the weights' position (x, TL x0); dom: the domain (8 doubles: R4
// r1, r2; Q4 corners); V the volume the weights divide by.  False when the
// candidates span more than KMPM_CPDI_MAX_SPAN nodes on an axis.
__device__ inline bool cpdi_setup(const KArgs &a, const double *xp, const double *dom, double V,
                                  CpdiPt &c) {
  const bool q4 = a.cpdi == 1 + KMPM_CPDI_Q4;
  if (!q4) {
    for (int e = 0; e < 4; e++) c.r[e] = dom[e];
    for (int d = 0; d < 2; d++) {
      c.xc[0][d] = xp[d] - c.r[d] - c.r[2 + d];
      c.xc[1][d] = xp[d] + c.r[d] - c.r[2 + d];
      c.xc[2][d] = xp[d] + c.r[d] + c.r[2 + d];
      c.xc[3][d] = xp[d] - c.r[d] + c.r[2 + d];
    }
  } else {
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 2; d++) c.xc[k][d] = dom[2 * k + d];
  }
  c.inv_V = 1.0 / V;                                            // ulcpdi.cpp:329
  if (q4) {                                                     // :331-340
    c.qa = (c.xc[3][0] - c.xc[0][0]) * (c.xc[1][1] - c.xc[2][1]) -
           (c.xc[1][0] - c.xc[2][0]) * (c.xc[3][1] - c.xc[0][1]);
    c.qb = (c.xc[2][0] - c.xc[3][0]) * (c.xc[0][1] - c.xc[1][1]) -
           (c.xc[0][0] - c.xc[1][0]) * (c.xc[2][1] - c.xc[3][1]);
    c.aov = 0.0417 * c.inv_V;
    c.six = 6 * V;
  }
  const bool lin = a.S1 == 2;
  const int m = lin ? 2 : 4;
  for (int d = 0; d < 2; d++) {
    c.lo[d] = 1 << 30;
    c.hi[d] = -(1 << 30);
  }
  for (int k = 0; k < 4; k++)                                   // :224-316
    for (int d = 0; d < 2; d++) {
      const double u = (c.xc[k][d] - a.blo[d]) * a.inv_h;
      if (!(fabs(u) < 1e9)) return false;
      const int o = lin ? (int)u : (int)(u - 1);
      c.c0[k][d] = o;
      const int lo = o < 0 ? 0 : o, hi = o + m - 1 > a.n[d] - 1 ? a.n[d] - 1 : o + m - 1;
      c.lo[d] = lo < c.lo[d] ? lo : c.lo[d];
      c.hi[d] = hi > c.hi[d] ? hi : c.hi[d];
    }
  return c.hi[0] - c.lo[0] < KMPM_CPDI_MAX_SPAN && c.hi[1] - c.lo[1] < KMPM_CPDI_MAX_SPAN;
}

// fn(i, j, wf, wd[2], wfc[4]) for every kept node (wf > 1e-12) in
// ascending tag (x-major): R4 wf = 1/4 sum of the corners' weights above
// 1e-12, Q4 wf = 0.0417/V sum (6V -+ a -+ b) wfc; the gradients of
// ulcpdi.cpp:369-394.
template <class Fn>
__device__ inline void cpdi_nodes(const KArgs &a, const CpdiPt &c, Fn &&fn) {
  const bool q4 = a.cpdi == 1 + KMPM_CPDI_Q4, lin = a.S1 == 2;
  const int m = lin ? 2 : 4;
  for (int i = c.lo[0]; i <= c.hi[0]; i++) {
    const double xn = a.lo[0] + (0 + i) * a.h;                  // grid.cpp:222-227
    const int tx = lin ? 0 : node_type(i, a.n[0]);
    for (int j = c.lo[1]; j <= c.hi[1]; j++) {
      bool cand = false;
      for (int k = 0; k < 4; k++)
        cand |= i >= c.c0[k][0] && i < c.c0[k][0] + m && j >= c.c0[k][1] && j < c.c0[k][1] + m;
      if (!cand) continue;
      const double yn = a.lo[1] + (0 + j) * a.h;
      const int ty = lin ? 0 : node_type(j, a.n[1]);
      double wfc[4], wf = 0;
      for (int k = 0; k < 4; k++) {                             // :343-357
        const double rx = (c.xc[k][0] - xn) * a.inv_h, ry = (c.xc[k][1] - yn) * a.inv_h;
        const double p0 = lin ? bf_linear(rx) : bf_cubic(rx, tx);
        const double p1 = lin ? bf_linear(ry) : bf_cubic(ry, ty);
        wfc[k] = p0 * p1 * 1;
        if (!q4 && wfc[k] > 1.0e-12) wf += wfc[k];
      }
      if (!q4)
        wf *= 0.25;
      else
        wf = c.aov * ((c.six - c.qa - c.qb) * wfc[0] + (c.six - c.qa + c.qb) * wfc[1] +
                      (c.six + c.qa + c.qb) * wfc[2] + (c.six + c.qa - c.qb) * wfc[3]);
      if (!(wf > 1.0e-12)) continue;
      double wd[2];
      if (!q4) {
        wd[0] = (wfc[0] - wfc[2]) * (c.r[1] - c.r[3]) + (wfc[1] - wfc[3]) * (c.r[1] + c.r[3]);
        wd[1] = (wfc[0] - wfc[2]) * (c.r[2] - c.r[0]) - (wfc[1] - wfc[3]) * (c.r[0] + c.r[2]);
        wd[0] *= c.inv_V;
        wd[1] *= c.inv_V;
      } else {
        const double (*x)[2] = c.xc;
        wd[0] = wfc[0] * (x[1][1] - x[3][1]) + wfc[1] * (x[2][1] - x[0][1]) +
                wfc[2] * (x[3][1] - x[1][1]) + wfc[3] * (x[0][1] - x[2][1]);
        wd[1] = wfc[0] * (x[3][0] - x[1][0]) + wfc[1] * (x[0][0] - x[2][0]) +
                wfc[2] * (x[1][0] - x[3][0]) + wfc[3] * (x[2][0] - x[0][0]);
        wd[0] *= 0.5 * c.inv_V;
        wd[1] *= 0.5 * c.inv_V;
      }
      fn(i, j, wf, wd, wfc);
    }
  }
}

// The domain of slot `slot` in the current state as the weights see it:
// ULCPDI the current domain, TLCPDI the reference one (its one weights
// pass); wv: the volume the weights divide by (ULCPDI: vol at the step's
// start, TLCPDI: vol0).
__device__ inline bool cpdi_load(const KArgs &a, int slot, double wv, CpdiPt &c) {
  double xp[2], dom[8];
  for (int d = 0; d < 2; d++) xp[d] = LDA((a.tl ? X0X : PX) + d, slot);
  for (int e = 0; e < 8; e++) dom[e] = LDA((a.tl ? REF0 : DOM0) + e, slot);
  return cpdi_setup(a, xp, dom, wv, c);
}

// Corner polygon area (solid.cpp:1190-1197).
__device__ inline double corner_area(const double *c) {
  return 0.5 * (c[0] * c[3] - c[2] * c[1] + c[2] * c[5] - c[4] * c[3] + c[4] * c[7] -
                c[6] * c[5] + c[6] * c[1] - c[0] * c[7]);
}

// rp = F rp0 (update_particle_domain solid.cpp:2338-2352), 2D.
__device__ inline void r4_follow_F(const double *F, const double *r0, double *r) {
  for (int k = 0; k < 2; k++)
    for (int i = 0; i < 2; i++)
      r[2 * k + i] = F[3 * i] * r0[2 * k] + F[3 * i + 1] * r0[2 * k + 1] + F[3 * i + 2] * 0.0;
}

// The weights' volume of sorted particle j: USF's second P2G and its G2P
// divide by the volume of the step's start, kept in xo by its stress pass.
__device__ inline double weights_volume(const KArgs &a, int slot, int j, bool saved) {
  if (a.tl) return LDA(VOL0, slot);
  return saved ? a.xo[j] : LDA(VOL, slot);
}

__global__ __launch_bounds__(256) void k_cpdi_zero(KArgs a, int f0, int f1) {
  const int64_t n = (int64_t)(f1 - f0) * a.ns;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a.cacc[(int64_t)f0 * a.ns + i] = 0.0;
}

// P2G (ulcpdi.cpp:395-450, tlcpdi.cpp:331-373): m, m v and the internal
// force (UL -vol sigma grad w, TL -vol0 PK1 grad w); remap: the MUSL
// momentum of the advanced velocities with the step's first weights
// (ulcpdi.cpp:472-486).  saved: the weights' volume from xo (USF's second
// P2G).
__global__ __launch_bounds__(256) void k_cpdi_p2g(KArgs a, int remap, int saved) {
  if (frozen(a)) return;
  double *S = a.cacc;
  const int64_t ns = a.ns;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < a.np; j += gridDim.x * blockDim.x) {
    const int slot = a.order[j];
    CpdiPt c;
    if (!cpdi_load(a, slot, weights_volume(a, slot, j, saved != 0), c)) {
      set_error(a.sc, KMPM_ERR_DOMAIN, IFLD(a.ia, PID, slot));
      continue;
    }
    const double m = LDA(MASS, slot);
    // the force's tensor (row-major 2x2) and its factor: TL vol0 PK1 and 1,
    // UL sigma and vol
    double v[2], P[4] = {0, 0, 0, 0}, pv = 1.0;
    for (int d = 0; d < 2; d++) v[d] = remap ? LDB(VX + d, j) : LDA(VX + d, slot);
    if (!remap) {
      if (a.tl) {
        P[0] = LDA(P00, slot);
        P[1] = LDA(P01, slot);
        P[2] = LDA(P10, slot);
        P[3] = LDA(P11, slot);
      } else {
        pv = LDA(VOL, slot);
        P[0] = LDA(SXX, slot);
        P[1] = P[2] = LDA(SXY, slot);
        P[3] = LDA(SYY, slot);
      }
    }
    cpdi_nodes(a, c, [&](int i, int jj, double wf, const double *wd, const double *) {
      const int64_t sn = cpdi_sidx(a, i, jj);
      const double wm = wf * m;
      if (!remap) atomicAdd(&S[sn], wm);
      atomicAdd(&S[ns + sn], wm * v[0]);
      atomicAdd(&S[2 * ns + sn], wm * v[1]);
      if (!remap) {   // solid.cpp:452-522
        atomicAdd(&S[4 * ns + sn], -(pv * (P[0] * wd[0] + P[1] * wd[1])));
        atomicAdd(&S[5 * ns + sn], -(pv * (P[2] * wd[0] + P[3] * wd[1])));
      }
    });
  }
}

// Grid pass over every node (GridMode), from the node sums.
__global__ __launch_bounds__(256) void k_cpdi_grid(KArgs a, int mode) {
  if (frozen(a)) return;
  __shared__ double fxv[MAX_FIX][6];
  __shared__ double fsum[MAX_FIX][3];
  fix_values_to_lds(a, fxv, fsum);
  const TileGeom &g = a.tg;
  const double dt = a.sc->dt;
  for (int64_t sn = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; sn < a.ns;
       sn += (int64_t)gridDim.x * blockDim.x) {
    const int t = (int)(sn / g.cpt), o = (int)(sn % g.cpt);
    const int i = (t / g.nt[1]) * g.T[0] + o / g.T[1], j = (t % g.nt[1]) * g.T[1] + o % g.T[1];
    if (i >= a.n[0] || j >= a.n[1]) continue;   // padding of edge tiles
    double sum[SCR_FIELDS];
    for (int f = 0; f < SCR_FIELDS; f++) sum[f] = a.cacc[f * a.ns + sn];
    grid_node<double>(a, mode, sn, sum, a.gmask[sn], dt, a.gmass, a.gv, a.gvu, fxv, fsum);
  }
  if (mode != GM_USF1) flush_fix_forces(a, fsum);
}

// The block's CFL bounds into the step's (TL also the smallest stretch).
__device__ inline void finish_cfl_any(const KArgs &a, double cmax, double rmin) {
  if (a.tl)
    finish_cfl<true>(a, cmax, rmin);
  else
    finish_cfl<false>(a, cmax, rmin);
}

// The CFL bound of one UL particle after its stress update (solid.cpp:1376-1385).
__device__ inline void cfl_of(double c, double &cmax) {
  if (c > cmax) cmax = c;   // c < 0: fully damaged, skipped
}

// UL F, J, stress, CFL for the rate L into slot o (solid.cpp:1155-1438);
// jv: the Q4 corners' J and vol, else null.
template <int MC>
__device__ inline double ul_stress(const KArgs &a, double dt, const double *L, int slot,
                                   const OutSlot &o, int pid, int mat, const double *v,
                                   const double *jv, double *Fout) {
  double F[9];
  for (int e = 0; e < 9; e++) F[e] = LDA(F00 + e, slot);
  double J = update_F(dt, L, F);
  const double vol0 = LDA(VOL0, slot);
  double vol = J * vol0;
  if (jv) {
    J = jv[0];
    vol = jv[1];
  }
  for (int e = 0; e < 9; e++) {
    STO(F00 + e, F[e]);
    Fout[e] = F[e];
  }
  STO(VOL0, vol0);
  STO(VOL, vol);
  double dmg = LDA(DMG, slot);
  if (J <= 0.0 && dmg < 1.0) set_error(a.sc, KMPM_ERR_JACOBIAN, pid);
  double sig[6], eel[6];
  for (int e = 0; e < 6; e++) {
    sig[e] = LDA(SXX + e, slot);
    eel[e] = LDA(EXX + e, slot);
  }
  double ep = LDA(EP, slot), epdot = LDA(EPDOT, slot), dinit = LDA(DINIT, slot);
  const double c = stress_and_speed<MC, PREC>(a, dt, L, J, sig, eel, ep, epdot, dinit, dmg,
                                              mat, v, pid);
  for (int e = 0; e < 6; e++) {
    STO(SXX + e, sig[e]);
    STO(EXX + e, eel[e]);
  }
  STO(EP, ep);
  STO(EPDOT, epdot);
  STO(DINIT, dinit);
  STO(DMG, dmg);
  return c;
}

// G2P (solid.cpp:696-784, 786-790): v_update, the acceleration and, node by
// node, x += dt w v_update; Q4 corners by their own weights; v blended by
// PIC/FLIP.  CG_USL adds L (from v_update), F, the domain and the stress;
// CG_MUSL keeps the step's first position for the re-map; CG_USF carries
// the state its stress pass updated (with F final: the R4 domain follows it
// here, after the weights of the step have been used).
template <int MC>
__global__ __launch_bounds__(256) void k_cpdi_g2p(KArgs a, int kind) {
  if (frozen(a)) {
    frozen_copy(a);
    return;
  }
  const double dt = a.sc->dt, inv_dt = 1.0 / dt, alpha = a.flip;
  const bool q4 = a.cpdi == 1 + KMPM_CPDI_Q4;
  double cmax = 0, rmin = 1.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < a.np; j += gridDim.x * blockDim.x) {
    const int slot = a.order[j];
    const int pid = IFLD(a.ia, PID, slot), mat = IFLD(a.ia, MATI, slot);
    CpdiPt c;
    if (!cpdi_load(a, slot, weights_volume(a, slot, j, kind == CG_USF), c)) {
      set_error(a.sc, KMPM_ERR_DOMAIN, pid);
      continue;
    }
    double x[3], v[3], vu[3] = {0, 0, 0}, acc[3] = {0, 0, 0}, L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double vc[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int d = 0; d < 3; d++) {
      x[d] = LDA(PX + d, slot);
      v[d] = LDA(VX + d, slot);
    }
    cpdi_nodes(a, c, [&](int i, int jj, double wf, const double *wd, const double *wfc) {
      const int64_t sn = cpdi_sidx(a, i, jj);
      double gvu[2], gv[2];
      for (int d = 0; d < 2; d++) {
        gvu[d] = a.gvu[d * a.ns + sn];
        gv[d] = a.gv[d * a.ns + sn];
      }
      for (int d = 0; d < 2; d++) {
        vu[d] += wf * gvu[d];
        acc[d] += wf * (gvu[d] - gv[d]);
        x[d] += dt * wf * gvu[d];
      }
      if (kind == CG_USL) {   // compute_rate_deformation_gradient_UL/TL (false)
        L[0] += gvu[0] * wd[0];
        L[1] += gvu[0] * wd[1];
        L[3] += gvu[1] * wd[0];
        L[4] += gvu[1] * wd[1];
      }
      if (q4)
        for (int k = 0; k < 4; k++)
          for (int d = 0; d < 2; d++) vc[k][d] += wfc[k] * gvu[d];
    });
    for (int d = 0; d < 3; d++) acc[d] *= inv_dt;
    if (!a.tl && outside_box(a, x)) set_error(a.sc, KMPM_ERR_PARTICLE_LEFT_BOX, pid);
    for (int d = 0; d < 3; d++) v[d] = (1 - alpha) * vu[d] + alpha * (v[d] + dt * acc[d]);
    double dom[8];
    for (int e = 0; e < 8; e++) dom[e] = LDA(DOM0 + e, slot);
    if (q4)
      for (int k = 0; k < 4; k++)
        for (int d = 0; d < 2; d++) dom[2 * k + d] += dt * vc[k][d];
    if (kind == CG_USF) {   // the rest of the state, updated in place by the stress pass
      const int nfld = state_nd(false, a.lay);
      for (int f = F00; f < nfld; f++)
        if (f != MASS && !(f >= X0X && f <= X0Z) && !(f >= DOM0 && f < DOM0 + 8))
          STB(f, j, LDA(f, slot));
      if (!q4 && !a.tl) {
        double F[9], r0[4];
        for (int e = 0; e < 9; e++) F[e] = LDA(F00 + e, slot);
        for (int e = 0; e < 4; e++) r0[e] = LDA(REF0 + e, slot);
        r4_follow_F(F, r0, dom);
      }
    }
    for (int d = 0; d < 3; d++) {
      STB(PX + d, j, x[d]);
      STB(VX + d, j, v[d]);
      STB(X0X + d, j, LDA(X0X + d, slot));
    }
    STB(MASS, j, LDA(MASS, slot));
    for (int e = 0; e < 8; e++) STB(REF0 + e, j, LDA(REF0 + e, slot));
    IFLD(a.ib, PID, j) = pid;
    IFLD(a.ib, MATI, j) = mat;
    if (kind == CG_USL) {
      const OutSlot o{a.pb, a.fb, j};
      double jv[2];
      if (q4) {                          // solid.cpp:1188-1200
        jv[1] = corner_area(dom);
        jv[0] = jv[1] / LDA(VOL0, slot);
      }
      if (a.tl) {
        tl_update<MC, PREC>(a, dt, L, slot, o, pid, mat, v, cmax, rmin, nullptr,
                            q4 ? jv : nullptr);
      } else {
        double F[9];
        const double cw = ul_stress<MC>(a, dt, L, slot, o, pid, mat, v, q4 ? jv : nullptr, F);
        cfl_of(cw, cmax);
        if (!q4) {
          double r0[4];
          for (int e = 0; e < 4; e++) r0[e] = LDA(REF0 + e, slot);
          r4_follow_F(F, r0, dom);
        }
      }
    }
    for (int e = 0; e < 8; e++) STB(DOM0 + e, j, dom[e]);
    // next step's cell: of the advanced x (UL) or of x0 (TL)
    double xk[3];
    for (int d = 0; d < 3; d++) xk[d] = a.tl ? LDA(X0X + d, slot) : x[d];
    a.key[j] = particle_key(a, xk[0], xk[1], xk[2]);
  }
  if (kind == CG_USL) finish_cfl_any(a, cmax, rmin);
}

// MUSL part B and USF's stress pass (musl.cpp:75-77, usf.cpp:58-60): L (or
// Fdot) from the grid velocity with the step's first weights, F, the
// domain (MUSL: R4 follows F; Q4's volume from the advanced corners), the
// stress; MUSL writes the next state, USF the current one in place (and
// keeps the volume its weights divide by in xo).
template <int MC>
__global__ __launch_bounds__(256) void k_cpdi_stress(KArgs a) {
  if (frozen(a)) {
    frozen_copy(a);
    return;
  }
  const double dt = a.sc->dt;
  const bool q4 = a.cpdi == 1 + KMPM_CPDI_Q4, usf = a.scheme == KMPM_SCHEME_USF;
  double cmax = 0, rmin = 1.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < a.np; j += gridDim.x * blockDim.x) {
    const int slot = a.order[j];
    const int pid = IFLD(a.ia, PID, slot), mat = IFLD(a.ia, MATI, slot);
    CpdiPt c;
    const double wv = weights_volume(a, slot, j, false);
    if (!cpdi_load(a, slot, wv, c)) {
      set_error(a.sc, KMPM_ERR_DOMAIN, pid);
      continue;
    }
    double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    cpdi_nodes(a, c, [&](int i, int jj, double, const double *wd, const double *) {
      const int64_t sn = cpdi_sidx(a, i, jj);
      const double g0 = a.gv[sn], g1 = a.gv[a.ns + sn];
      L[0] += g0 * wd[0];
      L[1] += g0 * wd[1];
      L[3] += g1 * wd[0];
      L[4] += g1 * wd[1];
    });
    const OutSlot o = out_slot(a, usf, slot, j);
    if (usf) {
      a.xo[j] = wv;
    } else {
      IFLD(a.ib, PID, j) = pid;
      IFLD(a.ib, MATI, j) = mat;
    }
    double v[3];
    for (int d = 0; d < 3; d++) v[d] = usf ? LDA(VX + d, slot) : LDB(VX + d, j);
    double dom[8];
    for (int e = 0; e < 8; e++) dom[e] = usf ? LDA(DOM0 + e, slot) : LDB(DOM0 + e, j);
    double jv[2];
    if (q4) {
      jv[1] = corner_area(dom);
      jv[0] = jv[1] / LDA(VOL0, slot);
    }
    if (a.tl) {
      tl_update<MC, PREC>(a, dt, L, slot, o, pid, mat, v, cmax, rmin, nullptr, q4 ? jv : nullptr);
    } else {
      double F[9];
      const double cw = ul_stress<MC>(a, dt, L, slot, o, pid, mat, v, q4 ? jv : nullptr, F);
      cfl_of(cw, cmax);
      if (!q4 && !usf) {   // USF: after its G2P, which still needs the step's vectors
        double r0[4];
        for (int e = 0; e < 4; e++) r0[e] = LDA(REF0 + e, slot);
        r4_follow_F(F, r0, dom);
        for (int e = 0; e < 4; e++) STO(DOM0 + e, dom[e]);
      }
    }
  }
  finish_cfl_any(a, cmax, rmin);
}

}  // namespace

#define CPDI_MC_DISPATCH(c, CALL)                                    \
  do {                                                               \
    switch ((c)->a.mat_class) {                                      \
      case MC_LINEAR: { constexpr int MC = MC_LINEAR; CALL; break; } \
      case MC_JC_SHOCK: { constexpr int MC = MC_JC_SHOCK; CALL; break; } \
      case MC_FLUID: { constexpr int MC = MC_FLUID; CALL; break; }   \
      default: { constexpr int MC = MC_GENERIC; CALL; break; }       \
    }                                                                \
  } while (0)

static int cpdi_blocks(kmpm_ctx *c, int64_t n) {
  int64_t g = (n + 255) / 256;
  const int64_t cap = (int64_t)c->cu_count * 8;
  if (g > cap) g = cap;
  return g < 1 ? 1 : (int)g;
}

hipError_t launch_cpdi_p2g(kmpm_ctx *c, int remap, int saved) {
  const int f0 = remap ? 1 : 0, f1 = remap ? 4 : SCR_FIELDS;
  hipLaunchKernelGGL(k_cpdi_zero, dim3(cpdi_blocks(c, (int64_t)(f1 - f0) * c->a.ns)), dim3(256),
                     0, c->stream, c->a, f0, f1);
  if (c->a.np == 0) return hipGetLastError();
  hipLaunchKernelGGL(k_cpdi_p2g, dim3(cpdi_blocks(c, c->a.np)), dim3(256), 0, c->stream, c->a,
                     remap, saved);
  return hipGetLastError();
}

hipError_t launch_cpdi_grid(kmpm_ctx *c, int mode) {
  const hipError_t e = launch_fix_values(c);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cpdi_grid, dim3(cpdi_blocks(c, c->a.ns)), dim3(256), 0, c->stream, c->a,
                     mode);
  return hipGetLastError();
}

hipError_t launch_cpdi_g2p(kmpm_ctx *c, int kind) {
  if (c->a.np == 0) return hipSuccess;
  CPDI_MC_DISPATCH(c, hipLaunchKernelGGL((k_cpdi_g2p<MC>), dim3(cpdi_blocks(c, c->a.np)),
                                         dim3(256), 0, c->stream, c->a, kind));
  return hipGetLastError();
}

hipError_t launch_cpdi_stress(kmpm_ctx *c) {
  if (c->a.np == 0) return hipSuccess;
  CPDI_MC_DISPATCH(c, hipLaunchKernelGGL((k_cpdi_stress<MC>), dim3(cpdi_blocks(c, c->a.np)),
                                         dim3(256), 0, c->stream, c->a));
  return hipGetLastError();
}

}  // namespace kmpm

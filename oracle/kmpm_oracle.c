/*
 * kmpm_oracle.c -- CPU FP64 restatement of Karamelo's explicit MPM step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library; the product path
 * (libkmpm.so) never does.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference repository root, src/).  Expression order follows the
 * reference's C++/Eigen expressions left to right; the library is compiled
 * with -ffp-contract=off so no fused multiply-add changes a rounding.
 *
 * Scope: ULMPM and TLMPM (one solid, its own grid over the solid's limits)
 * with the USL and MUSL schemes, linear and cubic-spline shape
 * functions, 2D and 3D, materials linear / eos-strength with eos
 * linear|shock|fluid, strength linear|johnson_cook|fluid, damage
 * damage_johnson_cook; fixes velocity_nodes (constant values) and
 * body_force (constant values); computes kinetic/strain energy.  One MPI rank.
 */
#include "kmpm_oracle.h"
#include "oracle_var.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* The checker (liboracle.so) is serial.  The CPU-baseline build
 * (liboracle_perf.so, -fopenmp) runs the particle and node loops on every
 * thread; there the node-major neighbour lists fill in thread order, so its
 * sums may differ from the checker's in rounding. */
#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#define OMP_ATOMIC _Pragma("omp atomic")
#define OMP_ATOMIC_CAPTURE _Pragma("omp atomic capture")
#else
#define OMP_FOR
#define OMP_ATOMIC
#define OMP_ATOMIC_CAPTURE
#endif

#define SQRT_3_OVER_2 1.224744871              /* strength_jc.cpp:29, solid.cpp:39 */
#define FOUR_THIRD 1.333333333333333333333333333333333333333   /* solid.cpp:39 */
#define MAXFIX 16

static double dmax(double a, double b) { return a > b ? a : b; } /* MAX macro */
static double dmin(double a, double b) { return a < b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }
static double square(double x) { return x * x; }                /* math_special.h */

/* ------------------------------------------------------------------------ */
/* 3x3 matrix helpers (row-major), in the order Eigen's fixed-size           */
/* expressions evaluate: a product coefficient sums k = 0, 1, 2 in turn.     */
/* ------------------------------------------------------------------------ */
static void m3_mul(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
static void m3_transpose(const double *A, double *T) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T[3 * j + i] = A[3 * i + j];
}
/* cofactor matrix C (C[i][j] = cofactor of A[i][j]); A^-1 = C^T / det A */
static void m3_cofactors(const double *A, double *C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      C[3 * i + j] = A[3 * i1 + j1] * A[3 * i2 + j2] - A[3 * i1 + j2] * A[3 * i2 + j1];
    }
}
/* Matrix3d::inverse (Eigen's 3x3 path: cofactors of column 0 give the
 * determinant, every entry is a cofactor times 1/det) */
static void m3_inverse(const double *A, double *Ai) {
  double C[9];
  m3_cofactors(A, C);
  double det = C[0] * A[0] + C[3] * A[3] + C[6] * A[6];
  double invdet = 1.0 / det;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Ai[3 * i + j] = C[3 * j + i] * invdet;
}
static double m3_det(const double *m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}
/* Rotation R of the polar decomposition F = R U (PolDec, mpm_math.h:104-133,
 * which takes R = U V^T from a JacobiSVD; R is unique for det F > 0).
 * Newton's iteration X <- (X + X^-T) / 2 from X = F, until the largest
 * change is at most 1e-9, then one more iteration (quadratic convergence
 * takes it to rounding level).  Returns 0 when det X <= 0. */
static int polar_rotation(const double *F, double *R) {
  double X[9], C[9];
  memcpy(X, F, 9 * sizeof(double));
  int last = 0;
  for (int it = 0; it < 100; it++) {
    m3_cofactors(X, C);
    double det = X[0] * C[0] + X[1] * C[1] + X[2] * C[2];
    if (!(det > 0)) return 0;
    double inv = 1.0 / det, diff = 0;
    for (int e = 0; e < 9; e++) {
      double xn = 0.5 * (X[e] + C[e] * inv);   /* X^-T = C / det */
      diff = dmax(diff, fabs(xn - X[e]));
      X[e] = xn;
    }
    if (last) break;
    if (diff <= 1e-9) last = 1;
  }
  memcpy(R, X, 9 * sizeof(double));
  return m3_det(R) > 0.0;
}
/* Real parts of the eigenvalues of F (EigenSolver<Matrix3d>, solid.cpp:1401-1408):
 * the characteristic polynomial of G = F - s I, s = tr F / 3, is the depressed
 * cubic t^3 + p t + q (p = sum of the principal 2x2 minors of G, q = -det G);
 * three real roots in trigonometric form, otherwise one real root (Cardano)
 * and a complex pair of real part -t/2. */
static void eig_real_parts(const double *F, double *re) {
  double s = (F[0] + F[4] + F[8]) / 3.0;
  double G[9];
  memcpy(G, F, 9 * sizeof(double));
  G[0] -= s;
  G[4] -= s;
  G[8] -= s;
  double p = (G[0] * G[4] - G[1] * G[3]) + (G[0] * G[8] - G[2] * G[6]) + (G[4] * G[8] - G[5] * G[7]);
  double q = -m3_det(G);
  if (p == 0 && q == 0) {
    re[0] = re[1] = re[2] = s;
    return;
  }
  double disc = q * q / 4.0 + p * p * p / 27.0;
  if (disc > 0) {
    double sq = sqrt(disc);
    double t = cbrt(-q / 2.0 + sq) + cbrt(-q / 2.0 - sq);
    re[0] = s + t;
    re[1] = re[2] = s - t / 2.0;
  } else {
    double m = 2.0 * sqrt(-p / 3.0);
    double arg = (3.0 * q / (2.0 * p)) * sqrt(-3.0 / p);
    arg = dmin(1.0, dmax(-1.0, arg));
    double th = acos(arg) / 3.0;
    for (int k = 0; k < 3; k++) re[k] = s + m * cos(th - 2.0943951023931957 * k);
  }
}
/* vol0PK1 = vol0 J (R sigma R^T) F^-T (solid.cpp:1277-1280,1368-1372) */
static void pk1(double vol0, double J, const double *R, const double *sig, const double *Finv,
                double *out) {
  double Rt[9], RS[9], RSRt[9], FiT[9], a = vol0 * J;
  m3_transpose(R, Rt);
  m3_mul(R, sig, RS);
  m3_mul(RS, Rt, RSRt);
  for (int e = 0; e < 9; e++) RSRt[e] = a * RSRt[e];
  m3_transpose(Finv, FiT);
  m3_mul(RSRt, FiT, out);
}

/* ------------------------------------------------------------------------ */
/* shape functions: basis_functions.h:21-104                                 */
/* ------------------------------------------------------------------------ */
static double bf_linear(double r_, int ntype) {                 /* :21-28 */
  (void)ntype;
  double r = fabs(r_);
  if (r >= 1.0) return 0.0;
  return 1.0 - r;
}
static double bf_dlinear(double r, int ntype, double inv_cellsize) { /* :30-38 */
  (void)ntype;
  if (r >= 1.0 || r <= -1.0 || r == 0) return 0.0;
  else if (r > 0.0) return -inv_cellsize;
  else return inv_cellsize;
}
static double bf_cubic(double r, int ntype) {                   /* :40-71 */
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
static double bf_dcubic(double r, int ntype, double icellsize) { /* :73-104 */
  if (r >= 1 && r < 2) {
    if (ntype == 1) return -icellsize;
    return icellsize * ((-0.5 * r + 2) * r - 2);
  } else if (r >= 0 && r < 1) {
    if (ntype == -2) return icellsize * (0.5 * r * r - 1);
    else if (ntype == 2) return icellsize;
    else if (ntype == 1) return icellsize * r * (r - 2);
    else return icellsize * (3.0 / 2.0 * r - 2) * r;
  } else if (r >= -1 && r < 0) {
    if (ntype == 2) return icellsize * (-0.5 * r * r + 1);
    else if (ntype == -1) return icellsize * (-r - 2) * r;
    else return icellsize * (-3.0 / 2.0 * r - 2) * r;
  } else if (r >= -2 && r < -1) {
    return icellsize * ((0.5 * r + 2) * r + 2);
  }
  return 0;
}
double orc_basis(int shape, int derivative, double r, int ntype, double inv_h) {
  if (shape == KMPM_SHAPE_LINEAR)
    return derivative ? bf_dlinear(r, ntype, inv_h) : bf_linear(r, ntype);
  return derivative ? bf_dcubic(r, ntype, inv_h) : bf_cubic(r, ntype);
}

/* ------------------------------------------------------------------------ */
/* materials: Mat ctors, material.cpp:720-769                                */
/* ------------------------------------------------------------------------ */
int orc_material_init(kmpm_material *m) {
  if (m->type == KMPM_MAT_LINEAR) {                              /* :745-764 */
    m->rho0 = m->lin_p[0];
    m->E = m->lin_p[1];
    m->nu = m->lin_p[2];
    m->G = m->E / (2 * (1 + m->nu));
    m->lambda = m->E * m->nu / ((1 + m->nu) * (1 - 2 * m->nu));
    m->K = m->E / (3 * (1 - 2 * m->nu));
    m->signal_velocity = sqrt(m->K / m->rho0);
    return KMPM_OK;
  }
  if (m->type != KMPM_MAT_EOS_STRENGTH) return KMPM_ERR_ARG;
  /* EOS::rho0(), EOS::K() (eos_*.cpp), Strength::G() (strength_*.cpp) */
  if (m->eos < KMPM_EOS_LINEAR || m->eos > KMPM_EOS_FLUID) return KMPM_ERR_ARG;
  if (m->strength < KMPM_STRENGTH_LINEAR || m->strength > KMPM_STRENGTH_FLUID) return KMPM_ERR_ARG;
  m->rho0 = m->eos_p[0];
  m->K = m->eos_p[1];
  m->G = m->strength_p[0];
  m->E = 9 * m->K * m->G / (3 * m->K + m->G);                  /* :732-736 */
  m->nu = (3 * m->K - 2 * m->G) / (2 * (3 * m->K + m->G));
  m->lambda = m->K - 2 * m->G / 3;
  m->signal_velocity = sqrt((m->lambda + 2 * m->G) / m->rho0);
  if (m->strength == KMPM_STRENGTH_JOHNSON_COOK && m->strength_p[7] == m->strength_p[8])
    return KMPM_ERR_ARG;                                        /* strength_jc.cpp:89-93 */
  if (m->damage == KMPM_DAMAGE_JOHNSON_COOK && m->damage_p[6] == m->damage_p[7])
    return KMPM_ERR_ARG;                                        /* damage_jc.cpp:80-84 */
  return KMPM_OK;
}

/* ------------------------------------------------------------------------ */
/* grid: Grid::init for one rank, grid.cpp:68-275                            */
/* ------------------------------------------------------------------------ */
/* total Lagrangian methods: TLMPM and TLCPDI (is_TL, tlcpdi.cpp:26) */
static int method_tl(int m) { return m == KMPM_METHOD_TLMPM || m == KMPM_METHOD_TLCPDI; }
static int method_cpdi(int m) { return m == KMPM_METHOD_ULCPDI || m == KMPM_METHOD_TLCPDI; }
/* bounds of the grid: the box (UL) or the solid (TL), grid.cpp:91-97 */
static const double *grid_lo(const kmpm_config *c) {
  return method_tl(c->method) ? c->solidlo : c->boxlo;
}
static const double *grid_hi(const kmpm_config *c) {
  return method_tl(c->method) ? c->solidhi : c->boxhi;
}
static void grid_global_counts(const kmpm_config *c, int n[3]) {
  double h = c->cellsize;
  const double *blo = grid_lo(c), *bhi = grid_hi(c);
  double Lx = bhi[0] - blo[0];                                  /* :151-155 */
  n[0] = ((int)(Lx / h)) + 1;
  while (n[0] * h <= Lx + 0.5 * h) n[0]++;
  if (c->dimension >= 2) {                                      /* :157-163 */
    double Ly = bhi[1] - blo[1];
    n[1] = (int)(((int)Ly) / h) + 1;
    while (n[1] * h <= Ly + 0.5 * h) n[1]++;
  } else n[1] = 1;
  if (c->dimension == 3) {                                      /* :165-171 */
    double Lz = bhi[2] - blo[2];
    n[2] = (int)(((int)Lz) / h) + 1;
    while (n[2] * h <= Lz + 0.5 * h) n[2]++;
  } else n[2] = 1;
}
/* local node counts for a rank that owns the whole box (:177-206) */
static void grid_local_counts(const kmpm_config *c, int n[3]) {
  double h = c->cellsize;
  const double *blo = grid_lo(c), *bhi = grid_hi(c);
  for (int d = 0; d < 3; d++) {
    if (d >= c->dimension) { n[d] = 1; continue; }
    /* sub-domain == box: Loffsetlo = max(0, boxlo - lo) = 0 for a solid in the box */
    double L = dmin(c->boxhi[d], bhi[d]) - blo[d];
    n[d] = (int)ceil(L / h);
    if (n[d] < 0) n[d] = 0;
    while (blo[d] + h * (0 + n[d] - 0.5) < dmin(c->boxhi[d], bhi[d])) n[d]++;
  }
}
int orc_grid_size(const kmpm_config *c, int64_t out[3]) {
  int g[3], l[3];
  if (!c || c->cellsize <= 0) return KMPM_ERR_ARG;
  if (method_tl(c->method))
    for (int d = 0; d < c->dimension; d++)
      if (c->solidlo[d] < c->boxlo[d] || c->solidhi[d] > c->boxhi[d] ||
          !(c->solidhi[d] > c->solidlo[d]))
        return KMPM_ERR_ARG;
  grid_global_counts(c, g);
  grid_local_counts(c, l);
  for (int d = 0; d < 3; d++) {
    if (g[d] != l[d]) return KMPM_ERR_GRID;  /* tags would not be a bijection */
    out[d] = g[d];
  }
  return KMPM_OK;
}

/* ------------------------------------------------------------------------ */
/* regions: region_block.cpp:148-154, region_cylinder.cpp:147-168,           */
/* Region::match region.cpp:65-72                                            */
/* ------------------------------------------------------------------------ */
static int region_inside(const kmpm_region *r, double x, double y, double z) {
  if (r->kind == KMPM_REGION_BLOCK) {
    if (x >= r->p[0] && x <= r->p[1] && y >= r->p[2] && y <= r->p[3] &&
        z >= r->p[4] && z <= r->p[5])
      return 1;
    return 0;
  }
  double c1 = r->p[0], c2 = r->p[1], R = r->p[2], lo = r->p[3], hi = r->p[4];
  double RSq = R * R, dSq;
  if (r->axis == 0) {
    dSq = square(y - c1) + square(z - c2);
    if (x >= lo && x <= hi && dSq <= RSq) return 1;
  } else if (r->axis == 1) {
    dSq = square(x - c1) + square(z - c2);
    if (y >= lo && y <= hi && dSq <= RSq) return 1;
  } else {
    dSq = square(x - c1) + square(y - c2);
    if (z >= lo && z <= hi && dSq <= RSq) return 1;
  }
  return 0;
}
/* Region::limits: region_block.cpp:159-168; region_cylinder.cpp:113-137,173-182 */
int orc_region_limits(const kmpm_region *r, double lo[3], double hi[3]) {
  if (r->kind == KMPM_REGION_BLOCK) {
    lo[0] = r->p[0]; hi[0] = r->p[1];
    lo[1] = r->p[2]; hi[1] = r->p[3];
    lo[2] = r->p[4]; hi[2] = r->p[5];
    return KMPM_OK;
  }
  double c1 = r->p[0], c2 = r->p[1], R = r->p[2], alo = r->p[3], ahi = r->p[4];
  if (r->axis == 0) {                 /* x: xlo=lo, ylo=c1-R, zlo=c2-R */
    lo[0] = alo; hi[0] = ahi; lo[1] = c1 - R; hi[1] = c1 + R; lo[2] = c2 - R; hi[2] = c2 + R;
  } else if (r->axis == 1) {          /* y: xlo=c1-R, ylo=lo, zlo=c2-R */
    lo[0] = c1 - R; hi[0] = c1 + R; lo[1] = alo; hi[1] = ahi; lo[2] = c2 - R; hi[2] = c2 + R;
  } else {                            /* z: xlo=c1-R, ylo=c2-R, zlo=lo */
    lo[0] = c1 - R; hi[0] = c1 + R; lo[1] = c2 - R; hi[1] = c2 + R; lo[2] = alo; hi[2] = ahi;
  }
  return KMPM_OK;
}

int orc_region_match(const kmpm_region *r, int dimension, double x, double y, double z) {
  (void)dimension;
  int in = region_inside(r, x, y, z);
  return r->interior ? in : !in;
}

/* ------------------------------------------------------------------------ */
/* populate: Solid::populate, solid.cpp:1810-2336 (ULMPM, one rank)          */
/* ------------------------------------------------------------------------ */
int orc_populate(const kmpm_config *c, const kmpm_region *region, int ppc1d,
                 const kmpm_material *mat, int mat_index, int solid_index,
                 int64_t ptag_offset, kmpm_particles *out, int64_t capacity,
                 int64_t *np_out) {
  double rlo[3], rhi[3];
  const double *boundlo = c->boxlo, *boundhi = c->boxhi;       /* :1927-1936 */
  if (method_tl(c->method)) {
    orc_region_limits(region, rlo, rhi);                        /* :1831-1838 */
    boundlo = rlo;
    boundhi = rhi;
  } else if (c->method != KMPM_METHOD_ULMPM && c->method != KMPM_METHOD_ULCPDI) {
    return KMPM_ERR_ARG;
  }
  const double *sublo = c->boxlo, *subhi = c->boxhi;
  double delta = c->cellsize;                                   /* :1863 */
  int dim = c->dimension;
  int noffsetlo[3], noffsethi[3], nsub[3];
  for (int d = 0; d < 3; d++) {                                 /* :1938-1986 */
    double Llo = dmax(0.0, sublo[d] - boundlo[d]);
    double Lhi = dmax(0.0, dmin(subhi[d], boundhi[d]) - boundlo[d]);
    noffsetlo[d] = (int)floor(Llo / delta);
    noffsethi[d] = (int)ceil(Lhi / delta);
  }
  nsub[0] = noffsethi[0] - noffsetlo[0] > 0 ? noffsethi[0] - noffsetlo[0] : 0;
  nsub[1] = dim >= 2 ? (noffsethi[1] - noffsetlo[1] > 0 ? noffsethi[1] - noffsetlo[1] : 0) : 1;
  nsub[2] = dim >= 3 ? (noffsethi[2] - noffsetlo[2] > 0 ? noffsethi[2] - noffsetlo[2] : 0) : 1;
  for (int d = 0; d < 3; d++) {                                 /* :1988-2002 */
    double lo = d < dim ? boundlo[d] : 0.0, hi = d < dim ? boundhi[d] : 0.0;
    double shi = d < dim ? subhi[d] : 0.0;
    while (lo + delta * (noffsetlo[d] + nsub[d] - 0.5) < dmin(shi, hi)) nsub[d]++;
  }
  double vol_ = dim == 1 ? delta : (dim == 2 ? delta * delta : delta * delta * delta);
  double mass_ = mat->rho0 * vol_;                              /* :2019 */
  /* CPDI half-width of a particle's domain (:2029-2106) */
  double lp = delta;
  if (ppc1d == 1) lp *= 0.5;
  else if (ppc1d == 2) lp *= 0.25;
  else lp *= 1.0 / (2 * ppc1d);
  const int cpdi = method_cpdi(c->method), q4 = cpdi && c->cpdi_style == KMPM_CPDI_Q4;
  int nip = 1;
  double ip3[3 * 64];
  memset(ip3, 0, sizeof(ip3));
  if (ppc1d == 1) {                                             /* :2026-2035 */
    nip = 1;
  } else if (ppc1d == 2) {                                      /* :2036-2051 */
    nip = dim == 1 ? 2 : (dim == 2 ? 4 : 8);
    double xi = 0.25;
    const double t[24] = {-xi, -xi, -xi, -xi, xi, -xi, xi, -xi, -xi, xi, xi, -xi,
                          -xi, -xi, xi,  -xi, xi, xi,  xi, -xi, xi,  xi, xi, xi};
    memcpy(ip3, t, sizeof(t));
  } else if (ppc1d >= 4) {                                      /* :2098-2121 */
    nip = dim == 1 ? ppc1d : (dim == 2 ? ppc1d * ppc1d : ppc1d * ppc1d * ppc1d);
    if (nip > 64) return KMPM_ERR_ARG;
    double dd = 1.0 / ppc1d;
    int l = 0;
    for (int k = 0; k < ppc1d; k++)
      for (int i = 0; i < ppc1d; i++)
        for (int j = 0; j < ppc1d; j++) {
          ip3[l++] = (i + 0.5) * dd - 0.5;
          ip3[l++] = (j + 0.5) * dd - 0.5;
          ip3[l++] = (k + 0.5) * dd - 0.5;
        }
  } else {
    return KMPM_ERR_ARG;  /* ppc1d==3 (Bernstein layout) is not on this path */
  }
  mass_ /= (double)nip;                                         /* :2126-2127 */
  vol_ /= (double)nip;
  int64_t l = 0;
  for (int i = 0; i < nsub[0]; i++)
    for (int j = 0; j < nsub[1]; j++)
      for (int k = 0; k < nsub[2]; k++)
        for (int ip = 0; ip < nip; ip++) {                      /* :2146-2183 */
          double x = boundlo[0] + delta * (noffsetlo[0] + i + 0.5 + ip3[3 * ip + 0]);
          double y = boundlo[1] + delta * (noffsetlo[1] + j + 0.5 + ip3[3 * ip + 1]);
          double z = dim == 3 ? boundlo[2] + delta * (noffsetlo[2] + k + 0.5 + ip3[3 * ip + 2]) : 0;
          /* inside_subdomain (domain.cpp:187-195) && region->inside (== 1) */
          double sl2 = dim == 3 ? sublo[2] : 0, sh2 = dim == 3 ? subhi[2] : 0;
          if (x < sublo[0] || x > subhi[0] || y < sublo[1] || y > subhi[1] || z < sl2 || z > sh2)
            continue;
          if (region_inside(region, x, y, z) != 1) continue;
          if (l < capacity) {
            int64_t q = l;
            out->x0[3 * q] = out->x[3 * q] = x;
            out->x0[3 * q + 1] = out->x[3 * q + 1] = y;
            out->x0[3 * q + 2] = out->x[3 * q + 2] = z;
            /* :2280-2320 */
            for (int d = 0; d < 3; d++) out->v[3 * q + d] = 0;
            for (int e = 0; e < 9; e++) {
              out->F[9 * q + e] = (e % 4 == 0) ? 1.0 : 0.0;
              out->sigma[9 * q + e] = 0;
              out->strain_el[9 * q + e] = 0;
            }
            out->mass[q] = mass_;
            out->vol0[q] = vol_;
            if (out->vol) out->vol[q] = vol_;
            if (out->rho0) out->rho0[q] = mat->rho0;
            if (out->rho) out->rho[q] = mat->rho0;
            if (out->J) out->J[q] = 1;
            out->eff_plastic_strain[q] = 0;
            out->eff_plastic_strain_rate[q] = 0;
            out->damage[q] = 0;
            out->damage_init[q] = 0;
            out->ptag[q] = ptag_offset + q + 1;
            out->solid[q] = solid_index;
            out->mat[q] = mat_index;
            out->mask[q] = 1;
            if (cpdi && !q4) {                                  /* R4, :2188-2205 */
              for (int e = 0; e < 6; e++) {
                const double r = e == 0 || e == 4 ? lp : 0.0;
                if (out->rp0) out->rp0[6 * q + e] = r;
                if (out->rp) out->rp[6 * q + e] = r;
              }
            } else if (q4) {                                    /* Q4, :2208-2222 */
              const double cx[4] = {x - lp, x + lp, x + lp, x - lp};
              const double cy[4] = {y - lp, y - lp, y + lp, y + lp};
              for (int k = 0; k < 4; k++)
                for (int d = 0; d < 3; d++) {
                  const double v = d == 0 ? cx[k] : d == 1 ? cy[k] : 0.0;
                  if (out->xpc0) out->xpc0[12 * q + 3 * k + d] = v;
                  if (out->xpc) out->xpc[12 * q + 3 * k + d] = v;
                }
            }
          }
          l++;
        }
  *np_out = l;
  return l > capacity ? KMPM_ERR_CAPACITY : KMPM_OK;
}

/* ------------------------------------------------------------------------ */
/* simulation state                                                         */
/* ------------------------------------------------------------------------ */
/* velocity_nodes: constant values, or (expr) the Vars of the script's text
 * and of its "time - dt" form; cur/force: this step's values and reaction */
typedef struct {
  uint32_t bit; int set[3]; double val[3], prev[3];
  int expr; ovar val_v[3], prev_v[3];
  double cur[3], cur_prev[3], force[3];
} fixvel_t;
typedef struct { uint32_t bit; int set[3]; double val[3]; } fixbf_t;

struct orc_sim {
  kmpm_config c;
  int dim, S1, S;               /* S1 nodes per dim in the stencil, S per particle */
  double h, inv_h;
  int n[3];
  int64_t nn;
  /* grid (grid.h:81-96) */
  double *gx0; int *gntype; double *gmass, *gv, *gvu, *gf, *gmb; uint32_t *gmask;
  /* materials */
  int nmat; kmpm_material *mats;
  /* particles (solid.h:50-97), ordered by solid */
  int64_t np;
  int64_t *ptag; int32_t *solid, *mat; uint32_t *pmask;
  double *x0, *x, *v, *vu, *a, *F, *sig, *eel, *L, *D;
  double *mass, *vol0, *vol, *rho0, *rho, *J;
  double *ep, *epdot, *dmg, *dinit;
  /* TLMPM: Fdot, F^-1, R (polar), vol0 PK1 (solid.h:50-97) */
  double *Fdot, *Finv, *R, *PK1;
  int tl, wf_done, mass_done;         /* update_wf, update_mass_nodes (tlmpm.cpp:89,345) */
  /* velocity-update sub-method: apic (ULMPM: APIC/AFLIP/ASFLIP/MLS,
   * ulmpm.cpp:79-85; TLMPM: APIC only, tlmpm.cpp:82-84), the MLS force and
   * the ASFLIP position update; alpha = PIC_FLIP after the sub-method's
   * override; Di the inertia tensor of compute_inertia_tensor (diagonal) */
  int apic, mls, asflip; double alpha; double Di[9];
  /* thermo-mechanical (update.cpp:168-180): particle T, heat source gamma
   * and heat flux q; grid T, T_update and the driving forces Qint, Qext
   * (solid.h, grid.h) */
  int temp;
  double *T, *gamma, *q;
  double *gT, *gTu, *gQint, *gQext;
  /* CPDI (ulcpdi.cpp, tlcpdi.cpp; 2D): cpdi = 1 + KMPM_CPDI_*, or 0; R4
   * domain vectors rp0, rp [np][2][3]; Q4 corners xpc0, xpc [np][4][3]; the
   * corners' weights of each kept pair wfc_pn [np][S][4] (wf_pn_corners) */
  int cpdi;
  double *rp0, *rp, *xpc0, *xpc, *wfc_pn;
  /* neighbour lists: particle-major (neigh_pn) and node-major CSR (neigh_np) */
  int32_t *nn_pn; int64_t *neigh_pn; double *wf_pn, *wfd_pn;
  int64_t *np_start; int64_t *neigh_np; double *wf_np, *wfd_np; int64_t npairs_cap;
  int32_t *cell_ijk;
  /* time (update.h) */
  double dt, time, dtCFL; int64_t step;
  int nfv, nbf; fixvel_t fv[MAXFIX]; fixbf_t bf[MAXFIX];
  int status; int64_t err_ptag;
};

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/* first error of the run (the reference aborts on it) */
static void set_status(orc_sim *s, int code, int64_t tag) {
#ifdef _OPENMP
#pragma omp critical(orc_status)
#endif
  {
    if (s->status == KMPM_OK) {
      s->status = code;
      s->err_ptag = tag;
    }
  }
}

orc_sim *orc_create(const kmpm_config *c, const kmpm_material *mats, int nmat) {
  int64_t n3[3];
  if (orc_grid_size(c, n3) != KMPM_OK) return NULL;
  if (c->method < KMPM_METHOD_ULMPM || c->method > KMPM_METHOD_TLCPDI ||
      (c->dimension != 2 && c->dimension != 3))
    return NULL;
  const int cpdi = c->method == KMPM_METHOD_ULCPDI || c->method == KMPM_METHOD_TLCPDI;
  /* the CPDI methods are 2D (ulcpdi.cpp:115-118) and take the PIC/FLIP
   * transfers here */
  if (cpdi && (c->dimension != 2 || c->sub_method >= KMPM_SUB_APIC ||
               (c->cpdi_style != KMPM_CPDI_R4 && c->cpdi_style != KMPM_CPDI_Q4)))
    return NULL;
  if (c->scheme < KMPM_SCHEME_USL || c->scheme > KMPM_SCHEME_USF ||
      c->sub_method < KMPM_SUB_PIC || c->sub_method > KMPM_SUB_MLS)
    return NULL;
  orc_sim *s = (orc_sim *)xcalloc(1, sizeof(orc_sim));
  s->c = *c;
  s->tl = c->method == KMPM_METHOD_TLMPM || c->method == KMPM_METHOD_TLCPDI;
  s->cpdi = cpdi ? 1 + c->cpdi_style : 0;
  /* Update::modify_method (update.cpp:129-146), ULMPM::setup
   * (ulmpm.cpp:79-85), TLMPM::setup (tlmpm.cpp:82-84) */
  {
    int sub = c->sub_method;
    s->alpha = c->pic_flip;
    if (s->tl) {
      s->apic = sub == KMPM_SUB_APIC;
    } else {
      s->apic = sub == KMPM_SUB_APIC || sub == KMPM_SUB_AFLIP || sub == KMPM_SUB_ASFLIP ||
                sub == KMPM_SUB_MLS;
      s->mls = sub == KMPM_SUB_MLS;
      s->asflip = sub == KMPM_SUB_ASFLIP;
    }
    if (sub == KMPM_SUB_APIC || (!s->tl && sub == KMPM_SUB_MLS)) s->alpha = 0;
    if (s->apic) {
      /* Solid::compute_inertia_tensor solid.cpp:1440-1470 */
      double inv2 = 1.0 / (c->cellsize * c->cellsize), di;
      if (c->shape == KMPM_SHAPE_LINEAR) {
        if (!s->tl || (c->np_per_cell != 1 && c->np_per_cell != 2)) {
          free(s);
          return NULL;
        }
        di = (c->np_per_cell == 1 ? 16.0 / 4.0 : 16.0 / 3.0) * inv2;
      } else {
        di = 3.0 * inv2;
      }
      for (int e = 0; e < 9; e++) s->Di[e] = e % 4 == 0 ? di : 0.0;
      if (c->dimension == 2) s->Di[8] = 1;
    }
  }
  const double *glo = grid_lo(c);
  s->dim = c->dimension;
  s->S1 = c->shape == KMPM_SHAPE_LINEAR ? 2 : 4;
  s->S = s->dim == 3 ? s->S1 * s->S1 * s->S1 : s->S1 * s->S1;
  if (s->cpdi) s->S = KMPM_CPDI_MAX_SPAN * KMPM_CPDI_MAX_SPAN;   /* the corners' union */
  s->h = c->cellsize;
  s->inv_h = 1.0 / s->h;                                        /* ulmpm.cpp:122 */
  for (int d = 0; d < 3; d++) s->n[d] = (int)n3[d];
  s->nn = n3[0] * n3[1] * n3[2];
  s->gx0 = (double *)xcalloc(3 * s->nn, sizeof(double));
  s->gntype = (int *)xcalloc(3 * s->nn, sizeof(int));
  s->gmass = (double *)xcalloc(s->nn, sizeof(double));
  s->gv = (double *)xcalloc(3 * s->nn, sizeof(double));
  s->gvu = (double *)xcalloc(3 * s->nn, sizeof(double));
  s->gf = (double *)xcalloc(3 * s->nn, sizeof(double));
  s->gmb = (double *)xcalloc(3 * s->nn, sizeof(double));
  s->gmask = (uint32_t *)xcalloc(s->nn, sizeof(uint32_t));
  s->temp = c->thermo != 0;
  s->gT = (double *)xcalloc(s->nn, sizeof(double));
  s->gTu = (double *)xcalloc(s->nn, sizeof(double));
  s->gQint = (double *)xcalloc(s->nn, sizeof(double));
  s->gQext = (double *)xcalloc(s->nn, sizeof(double));
  s->np_start = (int64_t *)xcalloc(s->nn + 1, sizeof(int64_t));
  int nx = s->n[0], ny = s->n[1], nz = s->n[2];
  int64_t l = 0;
  for (int i = 0; i < nx; i++)                                  /* grid.cpp:218-262 */
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        s->gx0[3 * l] = glo[0] + (0 + i) * s->h;
        s->gx0[3 * l + 1] = s->dim >= 2 ? glo[1] + (0 + j) * s->h : 0;
        s->gx0[3 * l + 2] = s->dim == 3 ? glo[2] + (0 + k) * s->h : 0;
        if (c->shape == KMPM_SHAPE_LINEAR) {
          s->gntype[3 * l] = s->gntype[3 * l + 1] = s->gntype[3 * l + 2] = 0;
        } else {
          s->gntype[3 * l] = imin(2, i) - imin(nx - 1 - i, 2);
          s->gntype[3 * l + 1] = imin(2, j) - imin(ny - 1 - j, 2);
          s->gntype[3 * l + 2] = imin(2, k) - imin(nz - 1 - k, 2);
        }
        s->gmask[l] = 1;                                        /* grid.cpp:429 */
        l++;
      }
  s->nmat = nmat;
  s->mats = (kmpm_material *)xcalloc(nmat, sizeof(kmpm_material));
  memcpy(s->mats, mats, nmat * sizeof(kmpm_material));
  s->dt = c->dt;
  s->time = 0;
  s->step = 0;
  return s;
}

static void free_particles(orc_sim *s) {
  void *ptrs[] = {s->ptag, s->solid, s->mat, s->pmask, s->x0, s->x, s->v, s->vu, s->a,
                  s->F, s->sig, s->eel, s->L, s->D, s->mass, s->vol0, s->vol, s->rho0,
                  s->rho, s->J, s->ep, s->epdot, s->dmg, s->dinit, s->nn_pn, s->neigh_pn,
                  s->wf_pn, s->wfd_pn, s->neigh_np, s->wf_np, s->wfd_np, s->cell_ijk,
                  s->Fdot, s->Finv, s->R, s->PK1, s->T, s->gamma, s->q,
                  s->rp0, s->rp, s->xpc0, s->xpc, s->wfc_pn};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
}

void orc_destroy(orc_sim *s) {
  if (!s) return;
  free_particles(s);
  free(s->gx0); free(s->gntype); free(s->gmass); free(s->gv); free(s->gvu); free(s->gf);
  free(s->gmb); free(s->gmask); free(s->np_start); free(s->mats);
  free(s->gT); free(s->gTu); free(s->gQint); free(s->gQext);
  free(s);
}

/* TLMPM state derived from F and sigma at upload: F^-1, R, J, vol0 PK1 */
static void tl_particle_pk1(orc_sim *s, int64_t i) {
  double *F = &s->F[9 * i];
  m3_inverse(F, &s->Finv[9 * i]);
  if (!polar_rotation(F, &s->R[9 * i]))
    for (int e = 0; e < 9; e++) s->R[9 * i + e] = (e % 4 == 0) ? 1.0 : 0.0;
  double J = m3_det(F);
  pk1(s->vol0[i], J, &s->R[9 * i], &s->sig[9 * i], &s->Finv[9 * i], &s->PK1[9 * i]);
}

int orc_set_particles(orc_sim *s, const kmpm_particles *p) {
  free_particles(s);
  int64_t np = p->np, S = s->S;
  for (int64_t i = 1; i < np; i++)
    if (p->solid[i] < p->solid[i - 1]) return KMPM_ERR_ARG;  /* ordered by solid */
  s->np = np;
#define A(f, t, k) s->f = (t *)xcalloc((size_t)(k) * np, sizeof(t))
  A(ptag, int64_t, 1); A(solid, int32_t, 1); A(mat, int32_t, 1); A(pmask, uint32_t, 1);
  A(x0, double, 3); A(x, double, 3); A(v, double, 3); A(vu, double, 3); A(a, double, 3);
  A(F, double, 9); A(sig, double, 9); A(eel, double, 9); A(L, double, 9); A(D, double, 9);
  A(mass, double, 1); A(vol0, double, 1); A(vol, double, 1); A(rho0, double, 1);
  A(rho, double, 1); A(J, double, 1); A(ep, double, 1); A(epdot, double, 1);
  A(dmg, double, 1); A(dinit, double, 1);
  A(Fdot, double, 9); A(Finv, double, 9); A(R, double, 9); A(PK1, double, 9);
  A(T, double, 1); A(gamma, double, 1); A(q, double, 3);
  A(rp0, double, 6); A(rp, double, 6); A(xpc0, double, 12); A(xpc, double, 12);
  A(wfc_pn, double, s->cpdi ? 4 * S : 1);
  A(nn_pn, int32_t, 1); A(neigh_pn, int64_t, S); A(wf_pn, double, S); A(wfd_pn, double, 3 * S);
  A(cell_ijk, int32_t, 3);
#undef A
  s->npairs_cap = np * S;
  s->wf_done = s->mass_done = 0;
  if (s->tl)
    for (int64_t i = 1; i < np; i++)
      if (p->solid[i] != p->solid[0]) return KMPM_ERR_ARG;   /* one solid, one grid */
  s->neigh_np = (int64_t *)xcalloc(s->npairs_cap, sizeof(int64_t));
  s->wf_np = (double *)xcalloc(s->npairs_cap, sizeof(double));
  s->wfd_np = (double *)xcalloc(3 * s->npairs_cap, sizeof(double));
  for (int64_t i = 0; i < np; i++) {
    int m = p->mat[i];
    if (m < 0 || m >= s->nmat) return KMPM_ERR_ARG;
    s->ptag[i] = p->ptag[i];
    s->solid[i] = p->solid[i];
    s->mat[i] = m;
    s->pmask[i] = p->mask ? p->mask[i] : 1;
    for (int d = 0; d < 3; d++) {
      s->x0[3 * i + d] = p->x0[3 * i + d];
      s->x[3 * i + d] = p->x[3 * i + d];
      s->v[3 * i + d] = p->v[3 * i + d];
    }
    for (int e = 0; e < 9; e++) {
      s->F[9 * i + e] = p->F[9 * i + e];
      s->sig[9 * i + e] = p->sigma[9 * i + e];
      s->eel[9 * i + e] = p->strain_el[9 * i + e];
    }
    s->mass[i] = p->mass[i];
    s->vol0[i] = p->vol0[i];
    s->vol[i] = p->vol ? p->vol[i] : p->vol0[i];
    s->rho0[i] = s->mats[m].rho0;
    s->rho[i] = p->rho ? p->rho[i] : s->mats[m].rho0;
    s->J[i] = p->J ? p->J[i] : 1.0;
    s->ep[i] = p->eff_plastic_strain[i];
    s->epdot[i] = p->eff_plastic_strain_rate[i];
    s->dmg[i] = p->damage[i];
    s->dinit[i] = p->damage_init[i];
    s->T[i] = p->T ? p->T[i] : 0.0;       /* gamma = q = 0 (solid.cpp:2305-2309) */
    if (s->cpdi == 1 + KMPM_CPDI_R4) {
      if (!p->rp0) return KMPM_ERR_ARG;
      for (int e = 0; e < 6; e++) {
        s->rp0[6 * i + e] = p->rp0[6 * i + e];
        s->rp[6 * i + e] = p->rp ? p->rp[6 * i + e] : p->rp0[6 * i + e];
      }
    } else if (s->cpdi == 1 + KMPM_CPDI_Q4) {
      if (!p->xpc0) return KMPM_ERR_ARG;
      for (int e = 0; e < 12; e++) {
        s->xpc0[12 * i + e] = p->xpc0[12 * i + e];
        s->xpc[12 * i + e] = p->xpc ? p->xpc[12 * i + e] : p->xpc0[12 * i + e];
      }
    }
    if (s->tl) tl_particle_pk1(s, i);
  }
  return KMPM_OK;
}

int orc_get_particles(orc_sim *s, kmpm_particles *p) {
  if (p->np < s->np) return KMPM_ERR_CAPACITY;
  for (int64_t i = 0; i < s->np; i++) {
    if (p->ptag) p->ptag[i] = s->ptag[i];
    if (p->solid) p->solid[i] = s->solid[i];
    if (p->mat) p->mat[i] = s->mat[i];
    if (p->mask) p->mask[i] = s->pmask[i];
    for (int d = 0; d < 3; d++) {
      if (p->x0) p->x0[3 * i + d] = s->x0[3 * i + d];
      if (p->x) p->x[3 * i + d] = s->x[3 * i + d];
      if (p->v) p->v[3 * i + d] = s->v[3 * i + d];
    }
    for (int e = 0; e < 9; e++) {
      if (p->F) p->F[9 * i + e] = s->F[9 * i + e];
      if (p->sigma) p->sigma[9 * i + e] = s->sig[9 * i + e];
      if (p->strain_el) p->strain_el[9 * i + e] = s->eel[9 * i + e];
    }
    if (p->mass) p->mass[i] = s->mass[i];
    if (p->vol0) p->vol0[i] = s->vol0[i];
    if (p->vol) p->vol[i] = s->vol[i];
    if (p->rho0) p->rho0[i] = s->rho0[i];
    if (p->rho) p->rho[i] = s->rho[i];
    if (p->J) p->J[i] = s->J[i];
    if (p->eff_plastic_strain) p->eff_plastic_strain[i] = s->ep[i];
    if (p->eff_plastic_strain_rate) p->eff_plastic_strain_rate[i] = s->epdot[i];
    if (p->damage) p->damage[i] = s->dmg[i];
    if (p->damage_init) p->damage_init[i] = s->dinit[i];
    if (p->T) p->T[i] = s->temp ? s->T[i] : 0.0;
    if (p->gamma) p->gamma[i] = s->temp ? s->gamma[i] : 0.0;
    for (int e = 0; e < 6; e++) {
      if (p->rp0) p->rp0[6 * i + e] = s->rp0[6 * i + e];
      if (p->rp) p->rp[6 * i + e] = s->rp[6 * i + e];
    }
    for (int e = 0; e < 12; e++) {
      if (p->xpc0) p->xpc0[12 * i + e] = s->xpc0[12 * i + e];
      if (p->xpc) p->xpc[12 * i + e] = s->xpc[12 * i + e];
    }
  }
  return KMPM_OK;
}

int orc_set_node_mask(orc_sim *s, const uint32_t *mask, int64_t nn) {
  if (nn != s->nn) return KMPM_ERR_ARG;
  memcpy(s->gmask, mask, nn * sizeof(uint32_t));
  return KMPM_OK;
}

int orc_fix_velocity_nodes(orc_sim *s, uint32_t bit, const int set[3], const double val[3],
                           const double prev[3]) {
  if (s->nfv >= MAXFIX) return KMPM_ERR_ARG;
  fixvel_t *f = &s->fv[s->nfv++];
  memset(f, 0, sizeof(*f));
  f->bit = bit;
  for (int d = 0; d < 3; d++) {
    f->set[d] = set[d] && d < s->dim;   /* fix_velocity_nodes.cpp:81,98 */
    f->val[d] = val[d];
    f->prev[d] = prev[d];
  }
  return KMPM_OK;
}

/* The script's dt variable in the current step: 0 in the first step unless
 * set_dt (input.cpp:55, update.cpp:69-76), the step's dt after that
 * (ulmpm.cpp:550). */
static double dt_variable(const orc_sim *s) {
  return (s->step <= 1 && !s->c.dt_constant) ? 0.0 : s->dt;
}

/* fix(id, velocity_nodes, group, vx, vy, vz) with expression texts (NULL or
 * "NULL": free component), parsed at creation as the constructor does
 * (fix_velocity_nodes.cpp:64-116) */
int orc_fix_velocity_nodes_expr(orc_sim *s, uint32_t bit, const char *const texts[3], int nvar,
                                const char *const *names, const double *values) {
  if (s->nfv >= MAXFIX) return KMPM_ERR_ARG;
  fixvel_t f;
  memset(&f, 0, sizeof(f));
  f.bit = bit;
  f.expr = 1;
  for (int d = 0; d < 3; d++) {
    if (!texts[d] || !strcmp(texts[d], "NULL") || d >= s->dim) continue;
    char prev[OV_EQ];
    ovenv env = {s->time, dt_variable(s), nvar, names, values, 0};
    f.val_v[d] = ov_parse(texts[d], &env);
    ov_previous_text(texts[d], prev, sizeof(prev));
    f.prev_v[d] = ov_parse(prev, &env);
    if (env.error) return KMPM_ERR_ARG;
    f.set[d] = 1;
  }
  s->fv[s->nfv++] = f;
  return KMPM_OK;
}

/* One Var of the script at (time, dt): the text (or its "time - dt" form)
 * parsed once, then its result (var.cpp:52-59). */
int orc_var_eval(const char *text, int previous, double time, double dt, int nvar,
                 const char *const *names, const double *values, double *out) {
  char prev[OV_EQ];
  ovenv env = {0.0, 0.0, nvar, names, values, 0};
  if (previous) ov_previous_text(text, prev, sizeof(prev));
  ovar v = ov_parse(previous ? prev : text, &env);
  if (env.error) return KMPM_ERR_ARG;
  int err = 0;
  *out = ov_result(&v, time, dt, &err);
  return err ? KMPM_ERR_ARG : KMPM_OK;
}

/* the Vars <id>_x, _y, _z of the last step (fix_velocity_nodes.cpp:218-224) */
int orc_fix_velocity_force(orc_sim *s, int fix, double out[3]) {
  if (fix < 0 || fix >= s->nfv) return KMPM_ERR_ARG;
  for (int d = 0; d < 3; d++) out[d] = s->fv[fix].force[d];
  return KMPM_OK;
}

int orc_fix_body_force(orc_sim *s, uint32_t bit, const int set[3], const double val[3]) {
  if (s->nbf >= MAXFIX) return KMPM_ERR_ARG;
  fixbf_t *f = &s->bf[s->nbf++];
  f->bit = bit;
  for (int d = 0; d < 3; d++) { f->set[d] = set[d]; f->val[d] = val[d]; }
  return KMPM_OK;
}

/* ------------------------------------------------------------------------ */
/* weights: ULMPM::compute_grid_weight_functions_and_gradients,              */
/* ulmpm.cpp:88-337 (one rank: map_ntag is the identity on [0,nnodes))       */
/* ------------------------------------------------------------------------ */
/* The corners of particle ip's domain, 2D (ulcpdi.cpp:183-222,
 * tlcpdi.cpp:170-196): R4 x -+ r1 -+ r2 from the current vectors (ULCPDI)
 * or, TLCPDI, those of its one weights pass (the reference ones of a fresh
 * run) about x0; Q4 the tracked corners (TLCPDI: their reference
 * positions). */
static void cpdi_corners(const orc_sim *s, int64_t ip, double xc[4][3]) {
  if (s->cpdi == 1 + KMPM_CPDI_R4) {
    const double *xp = s->tl ? &s->x0[3 * ip] : &s->x[3 * ip];
    const double *r = s->tl ? &s->rp0[6 * ip] : &s->rp[6 * ip];
    for (int d = 0; d < 3; d++) {
      xc[0][d] = xp[d] - r[d] - r[3 + d];
      xc[1][d] = xp[d] + r[d] - r[3 + d];
      xc[2][d] = xp[d] + r[d] + r[3 + d];
      xc[3][d] = xp[d] - r[d] + r[3 + d];
    }
  } else {
    const double *c = s->tl ? &s->xpc0[12 * ip] : &s->xpc[12 * ip];
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 3; d++) xc[k][d] = c[3 * k + d];
  }
}

static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return x < y ? -1 : x > y;
}

/* CPDI weights and gradients (ULCPDI::compute_grid_weight_functions_and_
 * gradients ulcpdi.cpp:111-393, TLCPDI tlcpdi.cpp:96-329; 2D).  A
 * particle's candidate nodes are the union of its corners' stencils, taken
 * from (corner - boxlo) / h even on TLCPDI's solid grid, in ascending tag;
 * per node the corners' tensor-product weights wfc give
 *   R4: wf = 1/4 sum of the wfc above 1e-12, and
 *       wfd = ((wfc0 - wfc2)(r1y - r2y) + (wfc1 - wfc3)(r1y + r2y),
 *              (wfc0 - wfc2)(r2x - r1x) - (wfc1 - wfc3)(r1x + r2x)) / V;
 *   Q4: wf = 0.0417/V sum (6V -+ a -+ b) wfc (the reference's rounded 1/24;
 *       a, b from the corners) and wfd = 1/(2V) sum_c wfc_c (y_{c+1} -
 *       y_{c-1}, x_{c-1} - x_{c+1});
 * and the node is kept where wf > 1e-12. */
static void cpdi_weights(orc_sim *s) {
  const double *blo = s->c.boxlo;
  const double inv_h = s->inv_h;
  const int ny = s->n[1], lin = s->c.shape == KMPM_SHAPE_LINEAR, m = lin ? 2 : 4;
  const int q4 = s->cpdi == 1 + KMPM_CPDI_Q4;
  OMP_FOR
  for (int64_t in = 0; in <= s->nn; in++) s->np_start[in] = 0;
  OMP_FOR
  for (int64_t ip = 0; ip < s->np; ip++) {
    double xc[4][3];
    cpdi_corners(s, ip, xc);
    int64_t cand[64];
    int nc = 0, span[2][2] = {{1 << 30, -(1 << 30)}, {1 << 30, -(1 << 30)}};
    for (int c = 0; c < 4; c++) {                               /* :224-316 */
      const int i0 = lin ? (int)((xc[c][0] - blo[0]) * inv_h) : (int)((xc[c][0] - blo[0]) * inv_h - 1);
      const int j0 = lin ? (int)((xc[c][1] - blo[1]) * inv_h) : (int)((xc[c][1] - blo[1]) * inv_h - 1);
      const int lo2[2] = {i0, j0};
      for (int d = 0; d < 2; d++) {
        const int a = lo2[d] < 0 ? 0 : lo2[d], b = lo2[d] + m - 1 > s->n[d] - 1 ? s->n[d] - 1 : lo2[d] + m - 1;
        if (a < span[d][0]) span[d][0] = a;
        if (b > span[d][1]) span[d][1] = b;
      }
      for (int i = i0; i < i0 + m; i++)
        for (int j = j0; j < j0 + m; j++) {
          const int64_t n = (int64_t)ny * i + j;
          /* a candidate off the grid's axes names another node through the
           * tag arithmetic, one far from the corners, whose weights are 0:
           * it is dropped here */
          if (n < s->nn && i >= 0 && i < s->n[0] && j >= 0 && j < ny) cand[nc++] = n;
        }
    }
    /* the device kernels hold KMPM_CPDI_MAX_SPAN nodes per axis of the
     * candidates' extent on the grid */
    if (span[0][1] - span[0][0] >= KMPM_CPDI_MAX_SPAN ||
        span[1][1] - span[1][0] >= KMPM_CPDI_MAX_SPAN) {
      set_status(s, KMPM_ERR_DOMAIN, s->ptag[ip]);
      s->nn_pn[ip] = 0;
      continue;
    }
    qsort(cand, nc, sizeof(int64_t), cmp_i64);                 /* :318-321 */
    int nu = 0;
    for (int k = 0; k < nc; k++)
      if (nu == 0 || cand[k] != cand[nu - 1]) cand[nu++] = cand[k];
    const double Vp = s->vol[ip], inv_Vp = 1.0 / Vp;           /* :329 */
    double qa = 0, qb = 0, alpha_over_Vp = 0, sixVp = 0;
    if (q4) {                                                   /* :331-340 */
      qa = (xc[3][0] - xc[0][0]) * (xc[1][1] - xc[2][1]) -
           (xc[1][0] - xc[2][0]) * (xc[3][1] - xc[0][1]);
      qb = (xc[2][0] - xc[3][0]) * (xc[0][1] - xc[1][1]) -
           (xc[0][0] - xc[1][0]) * (xc[2][1] - xc[3][1]);
      alpha_over_Vp = 0.0417 * inv_Vp;
      sixVp = 6 * Vp;
    }
    const double *r = &s->rp[6 * ip];
    if (s->tl) r = &s->rp0[6 * ip];
    int cnt = 0;
    for (int k = 0; k < nu; k++) {
      const int64_t in = cand[k];
      double wfc[4], wf = 0;
      for (int c = 0; c < 4; c++) {                             /* :345-357 */
        double rr[2];
        for (int d = 0; d < 2; d++) rr[d] = (xc[c][d] - s->gx0[3 * in + d]) * inv_h;
        const int *nt = &s->gntype[3 * in];
        double p0 = orc_basis(s->c.shape, 0, rr[0], nt[0], inv_h);
        double p1 = orc_basis(s->c.shape, 0, rr[1], nt[1], inv_h);
        wfc[c] = p0 * p1 * 1;
        if (!q4 && wfc[c] > 1.0e-12) wf += wfc[c];
      }
      if (!q4) wf *= 0.25;
      else
        wf = alpha_over_Vp * ((sixVp - qa - qb) * wfc[0] + (sixVp - qa + qb) * wfc[1] +
                              (sixVp + qa + qb) * wfc[2] + (sixVp + qa - qb) * wfc[3]);
      if (!(wf > 1.0e-12)) continue;
      double wd[3];
      if (!q4) {                                                /* :369-378 */
        wd[0] = (wfc[0] - wfc[2]) * (r[1] - r[3 + 1]) + (wfc[1] - wfc[3]) * (r[1] + r[3 + 1]);
        wd[1] = (wfc[0] - wfc[2]) * (r[3 + 0] - r[0]) - (wfc[1] - wfc[3]) * (r[0] + r[3 + 0]);
        wd[2] = 0;
        for (int d = 0; d < 3; d++) wd[d] *= inv_Vp;
      } else {                                                  /* :380-394 */
        wd[0] = wfc[0] * (xc[1][1] - xc[3][1]) + wfc[1] * (xc[2][1] - xc[0][1]) +
                wfc[2] * (xc[3][1] - xc[1][1]) + wfc[3] * (xc[0][1] - xc[2][1]);
        wd[1] = wfc[0] * (xc[3][0] - xc[1][0]) + wfc[1] * (xc[0][0] - xc[2][0]) +
                wfc[2] * (xc[1][0] - xc[3][0]) + wfc[3] * (xc[2][0] - xc[0][0]);
        wd[2] = 0;
        for (int d = 0; d < 3; d++) wd[d] *= 0.5 * inv_Vp;
      }
      const int64_t q = ip * s->S + cnt;
      s->neigh_pn[q] = in;
      s->wf_pn[q] = wf;
      for (int d = 0; d < 3; d++) s->wfd_pn[3 * q + d] = wd[d];
      for (int c = 0; c < 4; c++) s->wfc_pn[4 * q + c] = wfc[c];
      OMP_ATOMIC
      s->np_start[in + 1]++;
      cnt++;
    }
    s->nn_pn[ip] = cnt;
    s->cell_ijk[3 * ip] = s->cell_ijk[3 * ip + 1] = s->cell_ijk[3 * ip + 2] = 0;
  }
}

/* node-major lists (neigh_np) in ascending particle order, as the
 * push_back at ulmpm.cpp:269-275 builds them (np_start holds the counts) */
static int node_lists(orc_sim *s) {
  for (int64_t in = 0; in < s->nn; in++) s->np_start[in + 1] += s->np_start[in];
  int64_t *fill = (int64_t *)xcalloc(s->nn, sizeof(int64_t));
  OMP_FOR
  for (int64_t ip = 0; ip < s->np; ip++)
    for (int c = 0; c < s->nn_pn[ip]; c++) {
      int64_t q = ip * s->S + c, in = s->neigh_pn[q], k;
      OMP_ATOMIC_CAPTURE
      k = fill[in]++;
      int64_t o = s->np_start[in] + k;
      s->neigh_np[o] = ip;
      s->wf_np[o] = s->wf_pn[q];
      for (int d = 0; d < 3; d++) s->wfd_np[3 * o + d] = s->wfd_pn[3 * q + d];
    }
  free(fill);
  s->wf_done = 1;
  return s->status;
}

int orc_phase_weights(orc_sim *s) {
  if (s->tl && s->wf_done) return KMPM_OK;                      /* tlmpm.cpp:89,339 */
  if (s->cpdi) {
    cpdi_weights(s);
    return node_lists(s);
  }
  const double *lo = grid_lo(&s->c);
  double inv_h = s->inv_h;
  int nx = s->n[0], ny = s->n[1], nz = s->n[2];
  int64_t nnodes = s->nn;
  int lin = s->c.shape == KMPM_SHAPE_LINEAR;
  int S1 = s->S1;
  (void)nx;
  OMP_FOR
  for (int64_t in = 0; in <= s->nn; in++) s->np_start[in] = 0;
  OMP_FOR
  for (int64_t ip = 0; ip < s->np; ip++) {
    int64_t cand[64];
    /* TLMPM: reference positions against the solid's grid (tlmpm.cpp:152-282) */
    const double *xp = s->tl ? &s->x0[3 * ip] : &s->x[3 * ip];
    int i0, j0, k0;
    if (lin) {                                                  /* :161-165 */
      i0 = (int)((xp[0] - lo[0]) * inv_h);
      j0 = (int)((xp[1] - lo[1]) * inv_h);
      k0 = (int)((xp[2] - lo[2]) * inv_h);
    } else {                                                    /* :204-208 */
      i0 = (int)((xp[0] - lo[0]) * inv_h - 1);
      j0 = (int)((xp[1] - lo[1]) * inv_h - 1);
      k0 = (int)((xp[2] - lo[2]) * inv_h - 1);
    }
    s->cell_ijk[3 * ip] = i0;
    s->cell_ijk[3 * ip + 1] = j0;
    s->cell_ijk[3 * ip + 2] = s->dim == 3 ? k0 : 0;
    int nc = 0;
    for (int i = i0; i < i0 + S1; i++) {                        /* :167-242 */
      for (int j = j0; j < j0 + S1; j++) {
        if (nz > 1) {
          for (int k = k0; k < k0 + S1; k++) {
            int64_t tag = (int64_t)nz * ny * i + (int64_t)nz * j + k;
            if (tag < nnodes && tag >= 0) cand[nc++] = tag;  /* negative tags: see note */
          }
        } else {
          int64_t tag = (int64_t)ny * i + j;
          if (tag < nnodes && tag >= 0) cand[nc++] = tag;
        }
      }
    }
    /* Note: the reference indexes map_ntag with any tag < nnodes; a negative
     * tag (a particle exactly on boxlo with cubic splines) would read before
     * the vector.  The restatement drops such tags instead. */
    int cnt = 0;
    for (int c = 0; c < nc; c++) {                              /* :244-316 */
      int64_t in = cand[c];
      double r[3], sv[3], sd[3] = {0, 0, 0}, wf;
      for (int d = 0; d < 3; d++) r[d] = (xp[d] - s->gx0[3 * in + d]) * inv_h;
      const int *nt = &s->gntype[3 * in];
      sv[0] = orc_basis(s->c.shape, 0, r[0], nt[0], inv_h);
      wf = sv[0];
      if (wf != 0) {
        if (s->dim >= 2) { sv[1] = orc_basis(s->c.shape, 0, r[1], nt[1], inv_h); wf *= sv[1]; }
        else sv[1] = 1;
        if (s->dim == 3 && wf != 0) { sv[2] = orc_basis(s->c.shape, 0, r[2], nt[2], inv_h); wf *= sv[2]; }
        else sv[2] = 1;
      }
      if (wf == 0) continue;
      sd[0] = orc_basis(s->c.shape, 1, r[0], nt[0], inv_h);
      if (s->dim >= 2) sd[1] = orc_basis(s->c.shape, 1, r[1], nt[1], inv_h);
      if (s->dim == 3) sd[2] = orc_basis(s->c.shape, 1, r[2], nt[2], inv_h);
      int64_t q = ip * s->S + cnt;
      s->neigh_pn[q] = in;
      s->wf_pn[q] = wf;
      if (s->dim == 3) {
        s->wfd_pn[3 * q] = sd[0] * sv[1] * sv[2];
        s->wfd_pn[3 * q + 1] = sv[0] * sd[1] * sv[2];
        s->wfd_pn[3 * q + 2] = sv[0] * sv[1] * sd[2];
      } else {
        s->wfd_pn[3 * q] = sd[0] * sv[1];
        s->wfd_pn[3 * q + 1] = sv[0] * sd[1];
        s->wfd_pn[3 * q + 2] = 0;
      }
      OMP_ATOMIC
      s->np_start[in + 1]++;
      cnt++;
    }
    s->nn_pn[ip] = cnt;
  }
  return node_lists(s);
}

/* ------------------------------------------------------------------------ */
/* P2G: ULMPM::particles_to_grid ulmpm.cpp:339-378 ->                        */
/*   Solid::compute_mass_nodes solid.cpp:317-335,                            */
/*   Solid::compute_velocity_nodes :337-390 (APIC: _APIC :392-426),          */
/*   Solid::compute_external_and_internal_forces_nodes_UL :482-522           */
/*   (MLS: _UL_MLS :524-574)                                                 */
/* then Modify::post_particles_to_grid -> FixBodyforce fix_body_force.cpp:106-179 */
/* USF splits it: mass and velocity (particles_to_grid_USF_1,                */
/* ulmpm.cpp:380-404), forces (particles_to_grid_USF_2, :406-431)            */
/* ------------------------------------------------------------------------ */
static void p2g_mass(orc_sim *s) {
  /* mass: solid 0 resets, every solid accumulates in ascending ip; TLMPM
   * computes it on the first step only (tlmpm.cpp:344-351) */
  if (!(s->tl && s->mass_done)) {
    OMP_FOR
    for (int64_t in = 0; in < s->nn; in++) {
      int64_t b = s->np_start[in], e = s->np_start[in + 1];
      s->gmass[in] = 0;
      for (int64_t o = b; o < e; o++) {
        int64_t ip = s->neigh_np[o];
        s->gmass[in] += s->wf_np[o] * s->mass[ip];
      }
    }
  }
  s->mass_done = 1;
}

/* node velocity, per solid (solid.cpp:357-378): v += (sum_solid (w m) v) /
 * mass; APIC (solid.cpp:419-424): v + C (x0_I - x_p), C = L (ULMPM) or Fdot
 * (TLMPM), x_p = x (ULMPM) or x0 (TLMPM) */
static void p2g_velocity(orc_sim *s) {
  const double *Cs = s->tl ? s->Fdot : s->L, *pos = s->tl ? s->x0 : s->x;
  OMP_FOR
  for (int64_t in = 0; in < s->nn; in++) {
    int64_t b = s->np_start[in], e = s->np_start[in + 1];
    double *gv = &s->gv[3 * in];
    gv[0] = gv[1] = gv[2] = 0;
    int64_t o = b;
    while (o < e) {
      int sol = s->solid[s->neigh_np[o]];
      double vt[3] = {0, 0, 0};
      int64_t o2 = o;
      for (; o2 < e && s->solid[s->neigh_np[o2]] == sol; o2++) {
        int64_t ip = s->neigh_np[o2];
        double wm = s->wf_np[o2] * s->mass[ip];
        if (s->apic) {
          const double *Cp = &Cs[9 * ip];
          double dx[3];
          for (int d = 0; d < 3; d++) dx[d] = s->gx0[3 * in + d] - pos[3 * ip + d];
          for (int d = 0; d < 3; d++)
            vt[d] += wm * (s->v[3 * ip + d] +
                           (Cp[3 * d] * dx[0] + Cp[3 * d + 1] * dx[1] + Cp[3 * d + 2] * dx[2]));
        } else {
          for (int d = 0; d < 3; d++) vt[d] += wm * s->v[3 * ip + d];
        }
      }
      if (s->gmass[in] > 0) {
        for (int d = 0; d < 3; d++) vt[d] /= s->gmass[in];
        for (int d = 0; d < 3; d++) gv[d] += vt[d];
      }
      o = o2;
    }
  }
}

/* internal force: UL f -= vol * (sigma * wfd) (:510-515); MLS f -= vol * w *
 * (sigma * Di * (x0_I - x)) (:560-565); TL f -= vol0PK1 * wfd0
 * (compute_internal_forces_nodes_TL :452-480); then the body force fix */
static void p2g_forces(orc_sim *s) {
  OMP_FOR
  for (int64_t in = 0; in < s->nn; in++) {
    int64_t b = s->np_start[in], e = s->np_start[in + 1];
    double *gf = &s->gf[3 * in], *gmb = &s->gmb[3 * in];
    gf[0] = gf[1] = gf[2] = 0;
    gmb[0] = gmb[1] = gmb[2] = 0;
    for (int64_t o = b; o < e; o++) {
      int64_t ip = s->neigh_np[o];
      const double *wd = &s->wfd_np[3 * o];
      if (s->tl) {
        const double *P = &s->PK1[9 * ip];
        for (int d = 0; d < 3; d++)
          gf[d] -= P[3 * d] * wd[0] + P[3 * d + 1] * wd[1] + P[3 * d + 2] * wd[2];
      } else if (s->mls) {
        const double *sg = &s->sig[9 * ip], *Di = s->Di;
        double M[9], dx[3];
        for (int i = 0; i < 3; i++)      /* sigma * Di, a 3x3 product */
          for (int j = 0; j < 3; j++)
            M[3 * i + j] = sg[3 * i] * Di[j] + sg[3 * i + 1] * Di[3 + j] + sg[3 * i + 2] * Di[6 + j];
        for (int d = 0; d < 3; d++) dx[d] = s->gx0[3 * in + d] - s->x[3 * ip + d];
        double vw = s->vol[ip] * s->wf_np[o];
        for (int d = 0; d < 3; d++)
          gf[d] -= vw * (M[3 * d] * dx[0] + M[3 * d + 1] * dx[1] + M[3 * d + 2] * dx[2]);
      } else {
        const double *sg = &s->sig[9 * ip];
        for (int d = 0; d < 3; d++) {
          double t = sg[3 * d] * wd[0] + sg[3 * d + 1] * wd[1] + sg[3 * d + 2] * wd[2];
          gf[d] -= s->vol[ip] * t;
        }
      }
      /* mb += wf * mbp with mbp == 0 (ULMPM::reset ulmpm.cpp:553-563,
       * TLMPM::reset tlmpm.cpp:516-525) */
    }
  }
  /* body force fix (fix_body_force.cpp:117-146) */
  for (int k = 0; k < s->nbf; k++) {
    fixbf_t *f = &s->bf[k];
    OMP_FOR
    for (int64_t in = 0; in < s->nn; in++) {
      if (s->gmass[in] > 0 && (s->gmask[in] & f->bit)) {
        double fv[3] = {0, 0, 0};
        for (int d = 0; d < 3; d++) if (f->set[d]) fv[d] = f->val[d];
        for (int d = 0; d < 3; d++) fv[d] *= s->gmass[in];
        for (int d = 0; d < 3; d++) s->gmb[3 * in + d] += fv[d];
      }
    }
  }
}

/* grid temperature, per solid (Solid::compute_temperature_nodes
 * solid.cpp:2743-2762): T += (sum_solid (w m) T) / mass where mass > 0 */
static void p2g_temperature(orc_sim *s) {
  OMP_FOR
  for (int64_t in = 0; in < s->nn; in++) {
    int64_t b = s->np_start[in], e = s->np_start[in + 1], o = b;
    s->gT[in] = 0;
    while (o < e) {
      int sol = s->solid[s->neigh_np[o]];
      double Tt = 0;
      int64_t o2 = o;
      for (; o2 < e && s->solid[s->neigh_np[o2]] == sol; o2++) {
        int64_t ip = s->neigh_np[o2];
        Tt += s->wf_np[o2] * s->mass[ip] * s->T[ip];
      }
      if (s->gmass[in] > 0) {
        Tt /= s->gmass[in];
        s->gT[in] += Tt;
      }
      o = o2;
    }
  }
}

/* thermal driving forces: Qext += sum w gamma where mass > 0
 * (compute_external_temperature_driving_forces_nodes solid.cpp:2764-2779),
 * Qint = sum wfd . q (compute_internal_temperature_driving_forces_nodes
 * :2781-2797, which resets Qint in every solid's call: the last solid's sum
 * remains) */
static void p2g_thermal_forces(orc_sim *s) {
  int last = 0;
  for (int64_t ip = 0; ip < s->np; ip++)
    if (s->solid[ip] > last) last = s->solid[ip];
  OMP_FOR
  for (int64_t in = 0; in < s->nn; in++) {
    int64_t b = s->np_start[in], e = s->np_start[in + 1];
    double qe = 0, qi = 0;
    for (int64_t o = b; o < e; o++) {
      int64_t ip = s->neigh_np[o];
      if (s->gmass[in] > 0) qe += s->wf_np[o] * s->gamma[ip];
      if (s->solid[ip] == last) {
        const double *wd = &s->wfd_np[3 * o], *qp = &s->q[3 * ip];
        qi += wd[0] * qp[0] + wd[1] * qp[1] + wd[2] * qp[2];
      }
    }
    s->gQext[in] = qe;
    s->gQint[in] = qi;
  }
}

int orc_phase_p2g(orc_sim *s) {
  p2g_mass(s);
  p2g_velocity(s);
  p2g_forces(s);
  if (s->temp) {                                  /* ulmpm.cpp:371-375, tlmpm.cpp:359-363 */
    p2g_temperature(s);
    p2g_thermal_forces(s);
  }
  return KMPM_OK;
}

/* FixVelocityNodes::post_update_grid_state fix_velocity_nodes.cpp:130-225:
 * fix after fix, nodes in order, v_update <- value, v <- previous value,
 * with the reaction force sum (1/dt) m (value - v_update) */
static void velocity_fixes_post_update(orc_sim *s) {
  double dt = s->dt;
  double inv_dt = 1.0 / s->dt;
  for (int k = 0; k < s->nfv; k++) {
    fixvel_t *f = &s->fv[k];
    int err = 0;
    for (int d = 0; d < 3; d++) {
      if (!f->set[d]) continue;
      f->cur[d] = f->expr ? ov_result(&f->val_v[d], s->time, dt_variable(s), &err) : f->val[d];
      f->cur_prev[d] = f->expr ? ov_result(&f->prev_v[d], s->time, dt_variable(s), &err)
                               : f->prev[d];
    }
    if (err) set_status(s, KMPM_ERR_ARG, -1);
    double ftot[3] = {0, 0, 0};
    for (int64_t in = 0; in < s->nn; in++) {
      if (s->gmask[in] & f->bit) {
        double Dv[3] = {0, 0, 0};
        for (int d = 0; d < 3; d++)
          if (f->set[d]) {
            Dv[d] = f->cur[d] - s->gvu[3 * in + d];
            s->gvu[3 * in + d] = f->cur[d];
            s->gv[3 * in + d] = f->cur_prev[d];
          }
        double c = inv_dt * s->gmass[in];
        for (int d = 0; d < 3; d++) ftot[d] += c * Dv[d];
      }
    }
    for (int d = 0; d < 3; d++) f->force[d] = ftot[d];
  }
  (void)dt;
}

/* ------------------------------------------------------------------------ */
/* grid update: Grid::update_grid_velocities grid.cpp:448-466, then          */
/* FixVelocityNodes::post_update_grid_state fix_velocity_nodes.cpp:130-225   */
/* ------------------------------------------------------------------------ */
int orc_phase_grid_update(orc_sim *s) {
  double dt = s->dt;
  OMP_FOR
  for (int64_t in = 0; in < s->nn; in++) {
    double m = s->gmass[in];
    for (int d = 0; d < 3; d++) {
      if (m != 0)
        s->gvu[3 * in + d] = s->gv[3 * in + d] + dt * (s->gf[3 * in + d] + s->gmb[3 * in + d]) / m;
      else
        s->gvu[3 * in + d] = s->gv[3 * in + d];
    }
  }
  if (s->temp) {                                  /* Grid::update_grid_temperature grid.cpp:1354-1362 */
    OMP_FOR
    for (int64_t in = 0; in < s->nn; in++) {
      if (s->gmass[in] != 0)
        s->gTu[in] = s->gT[in] + dt * (s->gQint[in] + s->gQext[in]) / s->gmass[in];
      else
        s->gTu[in] = s->gT[in];
    }
  }
  velocity_fixes_post_update(s);
  return KMPM_OK;
}

/* MUSL re-map: ULMPM::velocities_to_grid ulmpm.cpp:476-496 ->
 * compute_velocity_nodes (solid.cpp:337-390; APIC :392-426), then
 * FixVelocityNodes::post_velocities_to_grid (fix_velocity_nodes.cpp:227-271) */
static void phase_velocities_to_grid(orc_sim *s) {
  p2g_velocity(s);
  if (s->temp) p2g_temperature(s);                /* ulmpm.cpp:491-493 */
  for (int k = 0; k < s->nfv; k++) {
    fixvel_t *f = &s->fv[k];
    OMP_FOR
    for (int64_t in = 0; in < s->nn; in++)
      if (s->gmask[in] & f->bit)
        for (int d = 0; d < 3; d++) if (f->set[d]) s->gv[3 * in + d] = f->cur[d];
  }
}

/* Solid::compute_rate_deformation_gradient_UL solid.cpp:860-936 */
/* TLMPM: Solid::compute_rate_deformation_gradient_TL solid.cpp:798-858, the
 * same sums over the reference gradients into Fdot */
/* APIC: compute_rate_deformation_gradient_UL_APIC solid.cpp:1081-1153
 * (TL_APIC :1006-1079): L(a,b) += v_I[a] (x0_I - x_p)[b] w, then L *= Di */
static void phase_rate_of_deformation(orc_sim *s, int doublemapping) {
  const double *vn = doublemapping ? s->gv : s->gvu;
  const double *pos = s->tl ? s->x0 : s->x;
  OMP_FOR
  for (int64_t ip = 0; ip < s->np; ip++) {
    double *L = s->tl ? &s->Fdot[9 * ip] : &s->L[9 * ip];
    for (int e = 0; e < 9; e++) L[e] = 0;
    for (int c = 0; c < s->nn_pn[ip]; c++) {
      int64_t q = ip * s->S + c, in = s->neigh_pn[q];
      const double *wd = &s->wfd_pn[3 * q], *v = &vn[3 * in];
      if (s->apic) {
        double dx[3], w = s->wf_pn[q];
        for (int d = 0; d < 3; d++) dx[d] = s->gx0[3 * in + d] - pos[3 * ip + d];
        int nd = s->dim == 2 ? 2 : 3;
        for (int a = 0; a < nd; a++)
          for (int b = 0; b < nd; b++) L[3 * a + b] += v[a] * dx[b] * w;
      } else if (s->dim == 2) {
        L[0] += v[0] * wd[0];
        L[1] += v[0] * wd[1];
        L[3] += v[1] * wd[0];
        L[4] += v[1] * wd[1];
      } else {
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) L[3 * a + b] += v[a] * wd[b];
      }
    }
    if (s->apic) {                    /* L *= Di, a 3x3 product */
      double T[9];
      m3_mul(L, s->Di, T);
      memcpy(L, T, sizeof(T));
    }
  }
}

/* Solid::compute_particle_accelerations_velocities_and_positions
 * solid.cpp:576-635 (ASFLIP: compute_particle_accelerations_velocities
 * :637-685, no position update); Solid::update_particle_velocities :786-790
 * (ASFLIP: update_particle_velocities_and_positions :791-796, x += dt v) */
static int phase_g2p(orc_sim *s) {
  double inv_dt = 1.0 / s->dt;
  double alpha = s->alpha;
  const double *lo = s->c.boxlo, *hi = s->c.boxhi;
  OMP_FOR
  for (int64_t ip = 0; ip < s->np; ip++) {
    double *vu = &s->vu[3 * ip], *a = &s->a[3 * ip];
    double vc[4][3] = {{0}};
    const int q4 = s->cpdi == 1 + KMPM_CPDI_Q4;
    for (int d = 0; d < 3; d++) vu[d] = a[d] = 0;
    for (int c = 0; c < s->nn_pn[ip]; c++) {
      int64_t q = ip * s->S + c, in = s->neigh_pn[q];
      double w = s->wf_pn[q];
      for (int d = 0; d < 3; d++) vu[d] += w * s->gvu[3 * in + d];
      for (int d = 0; d < 3; d++) a[d] += w * (s->gvu[3 * in + d] - s->gv[3 * in + d]);
      /* CPDI (Solid::compute_particle_velocities_and_positions
       * solid.cpp:696-764): x moves node by node, Q4 corners by their own
       * weights */
      if (s->cpdi)
        for (int d = 0; d < 3; d++) s->x[3 * ip + d] += s->dt * w * s->gvu[3 * in + d];
      if (q4)
        for (int k = 0; k < 4; k++)
          for (int d = 0; d < 3; d++) vc[k][d] += s->wfc_pn[4 * q + k] * s->gvu[3 * in + d];
    }
    for (int d = 0; d < 3; d++) a[d] *= inv_dt;
    if (q4)
      for (int k = 0; k < 4; k++)
        for (int d = 0; d < 3; d++) s->xpc[12 * ip + 3 * k + d] += s->dt * vc[k][d];
    if (!s->asflip && !s->cpdi)
      for (int d = 0; d < 3; d++) s->x[3 * ip + d] += s->dt * vu[d];
    const double *x = &s->x[3 * ip];
    if (!s->tl &&        /* the box check is UL only (solid.cpp:617-627) */
        !(x[0] >= lo[0] && x[0] <= hi[0] && x[1] >= lo[1] && x[1] <= hi[1] &&
          x[2] >= (s->dim == 3 ? lo[2] : 0) && x[2] <= (s->dim == 3 ? hi[2] : 0))) {
      set_status(s, KMPM_ERR_PARTICLE_LEFT_BOX, s->ptag[ip]);
    }
  }
  OMP_FOR
  for (int64_t ip = 0; ip < s->np; ip++) {
    for (int d = 0; d < 3; d++)
      s->v[3 * ip + d] = (1 - alpha) * s->vu[3 * ip + d] +
                         alpha * (s->v[3 * ip + d] + s->dt * s->a[3 * ip + d]);
    if (s->asflip)
      for (int d = 0; d < 3; d++) s->x[3 * ip + d] += s->dt * s->v[3 * ip + d];
  }
  if (s->temp) {                                  /* update_particle_temperature solid.cpp:2799-2808 */
    OMP_FOR
    for (int64_t ip = 0; ip < s->np; ip++) {
      s->T[ip] = 0;
      for (int c = 0; c < s->nn_pn[ip]; c++) {
        int64_t q = ip * s->S + c;
        s->T[ip] += s->wf_pn[q] * s->gTu[s->neigh_pn[q]];
      }
    }
  }
  return s->status;
}

/* eye-based helpers mirroring Eigen expressions */
static double trace3(const double *M) { return M[0] + M[4] + M[8]; }
/* MPM_Math::Deviator mpm_math.h:28-33: M - eye*(trace/3) */
static void deviator(const double *M, double *out) {
  double t = trace3(M) / 3.0;
  for (int e = 0; e < 9; e++) out[e] = M[e] - ((e % 4 == 0) ? 1.0 : 0.0) * t;
}
/* Frobenius norm, row-major sequential sum */
static double norm9(const double *M) {
  double s = 0;
  for (int e = 0; e < 9; e++) s += M[e] * M[e];
  return sqrt(s);
}
/* 3x3 determinant, cofactor expansion along the first row */
static double det3(const double *m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

/* StrengthJohnsonCook::update_deviatoric_stress strength_jc.cpp:100-184 (T = 0) */
static void jc_strength_T(const kmpm_material *mt, const double sigma[9], const double D[9],
                          double dt, double eps_p, double epsdot, double damage, double T,
                          double out[9], double *dps) {
  const double *p = mt->strength_p;
  double G_ = p[0], A = p[1], B = p[2], n = p[3], epsdot0 = p[4], C = p[5], m = p[6],
         Tr = p[7], Tm = p[8];
  double Tmr = Tm - Tr;
  if (damage >= 1.0) {
    for (int e = 0; e < 9; e++) out[e] = 0;
    return;                          /* *dps left untouched: caller holds 0 */
  }
  double epsdot_ratio = epsdot / epsdot0;
  epsdot_ratio = dmax(epsdot_ratio, 1.0);
  double ys;
  if (eps_p < 1.0e-10) ys = A;
  else ys = A + B * pow(eps_p, n);
  if (C != 0) ys *= pow(1.0 + epsdot_ratio, C);
  if (T < Tm) {
    if (m != 0 && T >= Tr) ys *= 1.0 - pow((T - Tr) / Tmr, m);
  } else ys = 0;
  double Gd = G_;
  if (damage > 0) { Gd *= (1 - damage); ys *= (1 - damage); }
  double c = dt * 2.0 * Gd, trial[9], dev[9];
  for (int e = 0; e < 9; e++) trial[e] = sigma[e] + c * D[e];
  deviator(trial, dev);
  double J2 = SQRT_3_OVER_2 * norm9(dev);
  for (int e = 0; e < 9; e++) out[e] = dev[e];
  if (J2 < ys) {
    *dps = 0.0;
  } else {
    *dps = (J2 - ys) / (3.0 * Gd);
    for (int e = 0; e < 9; e++) out[e] *= (ys / J2);
  }
}

/* StrengthJohnsonCook::update_deviatoric_stress strength_jc.cpp:100-184,
 * the call without a temperature (T = 0) */
void orc_jc_strength(const kmpm_material *mt, const double sigma[9], const double D[9],
                     double dt, double eps_p, double epsdot, double damage,
                     double out[9], double *dps) {
  jc_strength_T(mt, sigma, D, dt, eps_p, epsdot, damage, 0.0, out, dps);
}

/* EOSShock::compute_pressure eos_shock.cpp:99-133 at temperature T */
static double shock_pressure_T(const kmpm_material *mt, double J, double rho, double damage,
                               const double D[9], double cellsize, double T) {
  const double *p = mt->eos_p;
  double rho0_ = p[0], c0 = p[2], S = p[3], Gamma = p[4], cv = p[5], Tr = p[6], Q1 = p[7],
         Q2 = p[8];
  double alpha = cv * rho0_, e0 = 0, e;
  double mu = rho / rho0_ - 1.0;
  double pH = rho0_ * square(c0) * mu * (1.0 + mu) / square(1.0 - (S - 1.0) * mu);
  if (T > Tr) e = alpha * (T - Tr);
  else e = 0;
  double pFinal = pH + Gamma * (e - e0);
  if (damage > 0.0) {
    if (pFinal < 0.0) {
      if (damage >= 1.0) pFinal = 0;
      else pFinal *= 1.0 - damage;
    }
  }
  if (Q1 != 0 || Q2 != 0) {
    double tr_eps = trace3(D);
    if (tr_eps < 0) {
      double q = rho * cellsize * (Q1 * cellsize * tr_eps * tr_eps - Q2 * c0 * sqrt(J) * tr_eps);
      pFinal += q;
    }
  }
  return pFinal;
}

/* the call without a temperature (T = 0) */
double orc_shock_pressure(const kmpm_material *mt, double J, double rho, double damage,
                          const double D[9], double cellsize) {
  return shock_pressure_T(mt, J, rho, damage, D, cellsize, 0.0);
}

/* DamageJohnsonCook::compute_damage damage_jc.cpp:91-143 at temperature T */
static void jc_damage_T(const kmpm_material *mt, double *damage_init, double *damage, double pH,
                        const double sdev[9], double epsdot, double dps, double T) {
  const double *p = mt->damage_p;
  double d1 = p[0], d2 = p[1], d3 = p[2], d4 = p[3], d5 = p[4], epsdot0 = p[5], Tr = p[6],
         Tm = p[7];
  double Tmr = Tm - Tr;
  if (dps == 0 && *damage >= 1.0) return;
  double vm = SQRT_3_OVER_2 * norm9(sdev);
  double triax = 0.0;
  if (pH != 0.0 && vm != 0.0) triax = -pH / (vm + 0.001 * fabs(pH));
  if (triax <= -3) { *damage_init = 0; return; }
  double jc = d1 + d2 * exp(d3 * triax);
  if (d4 > 0.0) {
    if (epsdot > epsdot0) {
      double r = epsdot / epsdot0;
      jc *= (1.0 + d4 * log(r));
    }
  }
  if (d5 > 0.0 && T >= Tr) jc *= 1 + d5 * (T - Tr) / Tmr;
  *damage_init += dps / jc;
  if (*damage_init >= 1.0) *damage = dmin((*damage_init - 1.0) * 10, 1.0);
}

/* the call without a temperature (T = 0) */
void orc_jc_damage(const kmpm_material *mt, double *damage_init, double *damage, double pH,
                   const double sdev[9], double epsdot, double dps) {
  jc_damage_T(mt, damage_init, damage, pH, sdev, epsdot, dps, 0.0);
}

/* Solid::update_stress solid.cpp:1246-1374 for one particle: the LINEAR
 * material (:1268-1280), or EOS + strength (+ damage) with the eff. plastic
 * strain, its rate filter (:1319-1329), the tension rule (:1353-1356) and
 * the elastic strain (:1358-1366).  h is the cell size (tav = 1000 h / c). */
void orc_update_stress_point(const kmpm_material *mt, double h, double dt, double J, double rho,
                             const double D[9], double sg[9], double eel[9], double *ep,
                             double *epdot, double *dinit, double *dmg) {
  double gamma = 0;
  orc_update_stress_point_T(mt, h, dt, J, rho, D, sg, eel, ep, epdot, dinit, dmg, 0.0, 0, 0.0,
                            &gamma);
}

/* The same with the particle's temperature T: a material with a temperature
 * law (cp != 0) evaluates its EOS and strength at T, adds the thermal
 * pressure alpha (T0 - T) (temperature_plastic_work.cpp:74-76) and sets the
 * heat source gamma = chi sigma_flow epsdot_p (0 from Tm on,
 * temperature_plastic_work.cpp:65-71) times vol / cp, vol being vol (ULMPM)
 * or vol0 (TLMPM) (solid.cpp:1301-1351); the damage sees T on
 * thermo-mechanical runs (method_temp, solid.cpp:1331-1341). */
void orc_update_stress_point_T(const kmpm_material *mt, double h, double dt, double J,
                               double rho, const double D[9], double sg[9], double eel[9],
                               double *ep, double *epdot, double *dinit, double *dmg, double T,
                               int method_temp, double vol, double *gamma) {
  const int has_temp = mt->temp == KMPM_TEMP_PLASTIC_WORK && mt->temp_p[1] != 0;
  const double Tc = has_temp ? T : 0.0;      /* the constitutive calls' temperature */
  if (mt->type == KMPM_MAT_LINEAR) {                           /* :1268-1280 */
    double si[9];
    for (int e = 0; e < 9; e++) si[e] = dt * D[e];
    for (int e = 0; e < 9; e++) eel[e] += si[e];
    double tr = trace3(si);
    for (int e = 0; e < 9; e++)
      sg[e] += 2 * mt->G * si[e] + mt->lambda * tr * ((e % 4 == 0) ? 1.0 : 0.0);
    return;
  }
  double pH = 0, dps = 0, sdev[9];                              /* :1292-1374 */
  /* EOS */
  if (mt->eos == KMPM_EOS_SHOCK) {
    pH = shock_pressure_T(mt, J, rho, *dmg, D, h, Tc);
  } else if (mt->eos == KMPM_EOS_FLUID) {                       /* eos_fluid.cpp:68-73 */
    double mu = rho / mt->eos_p[0];
    pH = mt->eos_p[1] * (pow(mu, mt->eos_p[2]) - 1.0);
  } else {                                                      /* eos_linear.cpp:71-74 */
    pH = mt->eos_p[1] * (1 - J) * (1 - *dmg);
  }
  if (has_temp) pH += mt->temp_p[3] * (mt->temp_p[4] - T);      /* :1305 */
  /* strength */
  if (mt->strength == KMPM_STRENGTH_JOHNSON_COOK) {
    jc_strength_T(mt, sg, D, dt, *ep, *epdot, *dmg, Tc, sdev, &dps);
  } else if (mt->strength == KMPM_STRENGTH_FLUID) {             /* strength_fluid.cpp:46-58 */
    double dv[9];
    deviator(D, dv);
    for (int e = 0; e < 9; e++) sdev[e] = 2.0 * mt->strength_p[0] * dv[e];
  } else {                                                      /* strength_linear.cpp:59-71 */
    double dv[9], ds[9];
    deviator(D, dv);
    deviator(sg, ds);
    double c2 = 2.0 * mt->strength_p[0] * (1 - *dmg);
    for (int e = 0; e < 9; e++) sdev[e] = ds[e] + dt * (c2 * dv[e]);
  }
  *ep += dps;                                                   /* :1319 */
  double tav = 1000 * h / mt->signal_velocity;                  /* :1324-1329 */
  *epdot -= *epdot * dt / tav;
  *epdot += dps / tav;
  *epdot = dmax(0.0, *epdot);
  if (mt->damage == KMPM_DAMAGE_JOHNSON_COOK)                   /* :1331-1341 */
    jc_damage_T(mt, dinit, dmg, pH, sdev, *epdot, dps, method_temp ? T : 0.0);
  if (has_temp) {                                               /* :1343-1351 */
    const double chi = mt->temp_p[0], invcp = 1.0 / mt->temp_p[1], Tm = mt->temp_p[5];
    const double flow_stress = SQRT_3_OVER_2 * norm9(sdev);
    *gamma = T < Tm ? chi * flow_stress * *epdot : 0;
    *gamma *= vol * invcp;
  }
  double dm = *dmg;
  if (dm == 0 || pH >= 0) {                                     /* :1353-1356 */
    for (int e = 0; e < 9; e++) sg[e] = -pH * ((e % 4 == 0) ? 1.0 : 0.0) + sdev[e];
  } else {
    for (int e = 0; e < 9; e++) sg[e] = -pH * (1.0 - dm) * ((e % 4 == 0) ? 1.0 : 0.0) + sdev[e];
  }
  double t3 = (dt * trace3(D) + trace3(eel)) / 3.0;             /* :1358-1366 */
  double Gd = dm > 1e-10 ? mt->G * (1 - dm) : mt->G;
  for (int e = 0; e < 9; e++) eel[e] = t3 * ((e % 4 == 0) ? 1.0 : 0.0) + sdev[e] / Gd;
}

/* Solid::update_deformation_gradient solid.cpp:1155-1244 (UL),
 * Solid::update_stress solid.cpp:1246-1438 */
/* Solid::update_heat_flux solid.cpp:2810-2839: q = -(vol0 | vol) kappa / cp
 * sum wfd T_n, T_n = T (doublemapping) or T_update */
static void phase_heat_flux(orc_sim *s, int doublemapping) {
  const double *Tn = doublemapping ? s->gT : s->gTu;
  OMP_FOR
  for (int64_t ip = 0; ip < s->np; ip++) {
    const kmpm_material *mt = &s->mats[s->mat[ip]];
    const int has = mt->temp == KMPM_TEMP_PLASTIC_WORK && mt->temp_p[1] != 0;
    const double invcp = has ? 1.0 / mt->temp_p[1] : 0.0, kappa = has ? mt->temp_p[2] : 0.0;
    double *qp = &s->q[3 * ip];
    qp[0] = qp[1] = qp[2] = 0;
    for (int c = 0; c < s->nn_pn[ip]; c++) {
      int64_t q = ip * s->S + c;
      const double *wd = &s->wfd_pn[3 * q], t = Tn[s->neigh_pn[q]];
      for (int d = 0; d < 3; d++) qp[d] -= wd[d] * t;
    }
    const double sc = (s->tl ? s->vol0[ip] : s->vol[ip]) * invcp * kappa;
    for (int d = 0; d < 3; d++) qp[d] *= sc;
  }
}

static int phase_F_and_stress(orc_sim *s, int doublemapping) {
  double dt = s->dt;
  double max_c = 0;
  OMP_FOR
  for (int64_t ip = 0; ip < s->np; ip++) {
    double *F = &s->F[9 * ip], *L = &s->L[9 * ip], *D = &s->D[9 * ip];
    if (s->tl) {                                                /* :1181-1182 */
      for (int e = 0; e < 9; e++) F[e] += dt * s->Fdot[9 * ip + e];
    } else {                                                    /* :1183-1184 */
      double M[9], Fn[9];
      for (int e = 0; e < 9; e++) M[e] = ((e % 4 == 0) ? 1.0 : 0.0) + dt * L[e];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          Fn[3 * i + j] = M[3 * i] * F[j] + M[3 * i + 1] * F[3 + j] + M[3 * i + 2] * F[6 + j];
      memcpy(F, Fn, sizeof(Fn));
    }
    if (s->tl) m3_inverse(F, &s->Finv[9 * ip]);                 /* :1186 */
    if (s->cpdi == 1 + KMPM_CPDI_Q4) {                          /* :1188-1200 */
      const double *c = &s->xpc[12 * ip];
      s->vol[ip] = 0.5 * (c[0] * c[4] - c[3] * c[1] + c[3] * c[7] - c[6] * c[4] +
                          c[6] * c[10] - c[9] * c[7] + c[9] * c[1] - c[0] * c[10]);
      s->J[ip] = s->vol[ip] / s->vol0[ip];
    } else {
      s->J[ip] = det3(F);
      s->vol[ip] = s->J[ip] * s->vol0[ip];
    }
    if (s->J[ip] <= 0.0 && s->dmg[ip] < 1.0) {
      set_status(s, KMPM_ERR_JACOBIAN, s->ptag[ip]);
    }
    s->rho[ip] = s->rho0[ip] / s->J[ip];
    if (s->tl) {                                                /* :1218-1234 */
      double *R = &s->R[9 * ip];
      if (!polar_rotation(F, R))                                /* PolDec failure aborts */
        set_status(s, KMPM_ERR_JACOBIAN, s->ptag[ip]);
      double Lt[9], Sym[9], Rt[9], T1[9], T2[9];
      m3_mul(&s->Fdot[9 * ip], &s->Finv[9 * ip], L);            /* L = Fdot F^-1 */
      m3_transpose(L, Lt);
      for (int e = 0; e < 9; e++) Sym[e] = L[e] + Lt[e];
      m3_transpose(R, Rt);
      m3_mul(Rt, Sym, T1);
      m3_mul(T1, R, T2);
      for (int e = 0; e < 9; e++) D[e] = 0.5 * T2[e];
    } else {
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) D[3 * i + j] = 0.5 * (L[3 * i + j] + L[3 * j + i]);
    }
  }
  if (s->cpdi == 1 + KMPM_CPDI_R4 && !s->tl) {    /* update_particle_domain solid.cpp:2338-2352 */
    OMP_FOR
    for (int64_t ip = 0; ip < s->np; ip++)
      for (int k = 0; k < 2; k++) {
        const double *F = &s->F[9 * ip], *r0 = &s->rp0[6 * ip + 3 * k];
        for (int i = 0; i < 3; i++)
          s->rp[6 * ip + 3 * k + i] = F[3 * i] * r0[0] + F[3 * i + 1] * r0[1] + F[3 * i + 2] * r0[2];
      }
  }
  OMP_FOR
  for (int64_t ip = 0; ip < s->np; ip++)
    orc_update_stress_point_T(&s->mats[s->mat[ip]], s->h, dt, s->J[ip], s->rho[ip],
                              &s->D[9 * ip], &s->sig[9 * ip], &s->eel[9 * ip], &s->ep[ip],
                              &s->epdot[ip], &s->dinit[ip], &s->dmg[ip], s->T[ip], s->temp,
                              s->tl ? s->vol0[ip] : s->vol[ip], &s->gamma[ip]);
  if (s->temp) phase_heat_flux(s, doublemapping);   /* ulmpm.cpp:515-521 */
  if (s->tl) {                                                  /* :1277-1280,1368-1372 */
    OMP_FOR
    for (int64_t ip = 0; ip < s->np; ip++)
      pk1(s->vol0[ip], s->J[ip], &s->R[9 * ip], &s->sig[9 * ip], &s->Finv[9 * ip],
          &s->PK1[9 * ip]);
  }
  /* CFL (:1376-1437); UL: min_h_ratio = 1, TL: the smallest |Re eig(F)| */
  double min_h_ratio = 1.0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(max : max_c) reduction(min : min_h_ratio)
#endif
  for (int64_t ip = 0; ip < s->np; ip++) {
    if (s->dmg[ip] >= 1.0) continue;
    const kmpm_material *mt = &s->mats[s->mat[ip]];
    const double *v = &s->v[3 * ip];
    max_c = dmax(max_c, sqrt((mt->K + FOUR_THIRD * mt->G) / s->rho[ip]) +
                            dmax(dmax(fabs(v[0]), fabs(v[1])), fabs(v[2])));
    if (isnan(max_c)) {
      set_status(s, KMPM_ERR_WAVE_SPEED, s->ptag[ip]);
    }
    if (s->tl) {
      double re[3];
      eig_real_parts(&s->F[9 * ip], re);
      for (int k = 0; k < 3; k++) min_h_ratio = dmin(min_h_ratio, fabs(re[k]));
    }
  }
  s->dtCFL = dmin(s->dtCFL, s->h * min_h_ratio / max_c);
  if (isnan(s->dtCFL) && s->status == KMPM_OK) s->status = KMPM_ERR_DTCFL;
  return s->status;
}

/* Update::update_time update.cpp:201-206; ULMPM::adjust_dt ulmpm.cpp:525-551 */
static int phase_time_and_dt(orc_sim *s) {
  s->time += s->dt;
  if (s->c.dt_constant) return s->status;
  if (s->dtCFL == 0 && s->status == KMPM_OK) s->status = KMPM_ERR_DTCFL;
  s->dt = s->dtCFL * s->c.dt_factor;
  return s->status;
}

/* USL::run usl.cpp:43-93, MUSL::run musl.cpp:43-94 and USF::run
 * usf.cpp:34-94, one step each */
int orc_step(orc_sim *s, int64_t nsteps) {
  for (int64_t k = 0; k < nsteps && s->status == KMPM_OK; k++) {
    s->step++;                                                  /* update_timestep */
    orc_phase_weights(s);
    s->dtCFL = 1.0e22;                                          /* reset */
    if (s->c.scheme == KMPM_SCHEME_USF) {
      p2g_mass(s);                                              /* particles_to_grid_USF_1 */
      p2g_velocity(s);
      if (s->temp) p2g_temperature(s);
      velocity_fixes_post_update(s);                            /* post_update_grid_state */
      phase_rate_of_deformation(s, 1);
      if (phase_F_and_stress(s, 1)) break;
      p2g_forces(s);                                            /* USF_2, post_particles_to_grid */
      if (s->temp) p2g_thermal_forces(s);
      orc_phase_grid_update(s);
      if (phase_g2p(s)) break;
      phase_time_and_dt(s);
      continue;
    }
    orc_phase_p2g(s);
    orc_phase_grid_update(s);
    if (s->c.scheme == KMPM_SCHEME_USL) {
      phase_rate_of_deformation(s, 0);
      if (phase_g2p(s)) break;
    } else {
      if (phase_g2p(s)) break;
      phase_velocities_to_grid(s);
      phase_rate_of_deformation(s, 1);
    }
    if (phase_F_and_stress(s, s->c.scheme != KMPM_SCHEME_USL)) break;
    phase_time_and_dt(s);
  }
  return s->status;
}

/* the heat flux q of every particle (3 per particle, particle order) */
int orc_get_heat_flux(orc_sim *s, double *q, int64_t np) {
  if (np < s->np) return KMPM_ERR_CAPACITY;
  memcpy(q, s->q, (size_t)s->np * 3 * sizeof(double));
  return KMPM_OK;
}

/* set the affine matrix (L or Fdot) of every particle, as a state the
 * reference reaches after a step (pins of the APIC transfers) */
int orc_set_affine(orc_sim *s, const double *C, int64_t np) {
  if (np != s->np) return KMPM_ERR_ARG;
  memcpy(s->tl ? s->Fdot : s->L, C, (size_t)np * 9 * sizeof(double));
  return KMPM_OK;
}

/* the affine matrix of the APIC transfers: L (ULMPM) or Fdot (TLMPM),
 * row-major, particle order */
int orc_get_affine(orc_sim *s, double *C, int64_t np) {
  if (np < s->np) return KMPM_ERR_CAPACITY;
  const double *src = s->tl ? s->Fdot : s->L;
  memcpy(C, src, (size_t)s->np * 9 * sizeof(double));
  return KMPM_OK;
}

int orc_get_time(orc_sim *s, double *dt, double *time, int64_t *step) {
  if (dt) *dt = s->dt;
  if (time) *time = s->time;
  if (step) *step = s->step;
  return KMPM_OK;
}

/* FixKineticEnergy::final_integrate fix_kinetic_energy.cpp:78-118,
 * FixStrainEnergy::final_integrate fix_strain_energy.cpp:79-133 */
int orc_energy(orc_sim *s, uint32_t bit, double *ek, double *es) {
  double Ek = 0, Es = 0;
  for (int64_t ip = 0; ip < s->np; ip++) {
    if (!(s->pmask[ip] & bit)) continue;
    const double *v = &s->v[3 * ip];
    Ek += 0.5 * s->mass[ip] * square(sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    const double *sg = &s->sig[9 * ip], *e = &s->eel[9 * ip];
    double t = sg[0] * e[0];
    for (int q = 1; q < 9; q++) t += sg[q] * e[q];
    Es += 0.5 * s->vol[ip] * t;
  }
  *ek = Ek;
  *es = Es;
  return KMPM_OK;
}

int orc_get_grid(orc_sim *s, double *mass, double *v, double *vu, double *f, int64_t nn) {
  if (nn != s->nn) return KMPM_ERR_ARG;
  if (mass) memcpy(mass, s->gmass, nn * sizeof(double));
  if (v) memcpy(v, s->gv, 3 * nn * sizeof(double));
  if (vu) memcpy(vu, s->gvu, 3 * nn * sizeof(double));
  if (f) memcpy(f, s->gf, 3 * nn * sizeof(double));
  return KMPM_OK;
}

int orc_get_neighbours(orc_sim *s, int32_t *numneigh, int64_t *tags, double *w, double *wd,
                       int64_t np, int S) {
  if (np != s->np || S < s->S) return KMPM_ERR_ARG;
  for (int64_t ip = 0; ip < np; ip++) {
    numneigh[ip] = s->nn_pn[ip];
    for (int c = 0; c < S; c++) {
      int64_t q = ip * s->S + c, o = ip * S + c;
      if (c < s->nn_pn[ip]) {
        tags[o] = s->neigh_pn[q];
        w[o] = s->wf_pn[q];
        for (int d = 0; d < 3; d++) wd[3 * o + d] = s->wfd_pn[3 * q + d];
      } else {
        tags[o] = -1;
        w[o] = 0;
        for (int d = 0; d < 3; d++) wd[3 * o + d] = 0;
      }
    }
  }
  return KMPM_OK;
}

int orc_cell_index(orc_sim *s, int32_t *ijk, int64_t np) {
  if (np != s->np) return KMPM_ERR_ARG;
  memcpy(ijk, s->cell_ijk, 3 * np * sizeof(int32_t));
  return KMPM_OK;
}

/* Cell of a particle for the sort: trunc((x - lo) * inv_h) per dimension
 * (the linear-shape base index of ulmpm.cpp:163-165; TLMPM: x0 against the
 * solid's grid), clamped to the node range; key is tile-major. */
static int64_t cell_key(orc_sim *s, int64_t ip, const int T[3]) {
  int c[3] = {0, 0, 0}, nt[3];
  const double *xp = s->tl ? s->x0 : s->x, *glo = grid_lo(&s->c);
  for (int d = 0; d < s->dim; d++) {
    double u = (xp[3 * ip + d] - glo[d]) * s->inv_h;
    int ci = (u >= 0 && u < 2147483647.0) ? (int)u : (u >= 0 ? s->n[d] - 1 : 0);
    if (ci > s->n[d] - 1) ci = s->n[d] - 1;
    c[d] = ci;
  }
  int Tz = s->dim == 3 ? T[2] : 1;
  int Tt[3] = {T[0], T[1], Tz};
  for (int d = 0; d < 3; d++) nt[d] = (s->n[d] + Tt[d] - 1) / Tt[d];
  int64_t tile = ((int64_t)(c[0] / Tt[0]) * nt[1] + c[1] / Tt[1]) * nt[2] + c[2] / Tt[2];
  int64_t local = ((int64_t)(c[0] % Tt[0]) * Tt[1] + c[1] % Tt[1]) * Tt[2] + c[2] % Tt[2];
  return tile * ((int64_t)Tt[0] * Tt[1] * Tt[2]) + local;
}

typedef struct { int64_t key, ptag; } kp_t;
static int kp_cmp(const void *a, const void *b) {
  const kp_t *x = (const kp_t *)a, *y = (const kp_t *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  if (x->ptag != y->ptag) return x->ptag < y->ptag ? -1 : 1;
  return 0;
}
int orc_sort_permutation(orc_sim *s, const int tile[3], int64_t *sorted, int64_t np) {
  if (np != s->np) return KMPM_ERR_ARG;
  kp_t *kp = (kp_t *)xcalloc(np, sizeof(kp_t));
  for (int64_t ip = 0; ip < np; ip++) { kp[ip].key = cell_key(s, ip, tile); kp[ip].ptag = s->ptag[ip]; }
  qsort(kp, np, sizeof(kp_t), kp_cmp);
  for (int64_t i = 0; i < np; i++) sorted[i] = kp[i].ptag;
  free(kp);
  return KMPM_OK;
}

int64_t orc_error_particle(orc_sim *s) { return s->err_ptag; }

int orc_polar_rotation(const double F[9], double R[9]) { return polar_rotation(F, R); }
void orc_eig_real_parts(const double F[9], double re[3]) { eig_real_parts(F, re); }

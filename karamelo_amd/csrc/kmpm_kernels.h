// kmpm_kernels.h -- device functions and tile-templated kernels of the
// explicit MPM step (the P2G, grid update and G2P kernels are compiled
// once per tile shape; see kmpm_tile_launch.h).
//
// Step structure (USL):
//   sort   : cell key per particle -> counting sort by tile-major cell key,
//            ties broken by particle id (deterministic permutation)
//   P2G    : one workgroup per active tile; lanes = stencil nodes; each lane
//            accumulates the particles of a cell in registers, flushes into an
//            LDS node block of the tile; the block is written to a per-tile
//            partial buffer (no global atomics)
//   grid   : each owned node of a touched tile sums the partial blocks that
//            cover it, then v = p/m, v_update = v + dt (f + mb)/m, fixes
//   G2P    : one workgroup per active tile; LDS block of node velocities;
//            lanes = particles: gather, L, x/v update, F, stress, CFL, next
//            cell key; state is written in sorted order (the permutation is
//            applied for free)
#pragma once

#include <type_traits>

#include "kmpm_constitutive.h"
#include "kmpm_ctx.h"
#include "kmpm_mat3.h"
#include "kmpm_surrogate.h"

namespace kmpm {

#define PF(base, f, i) ((base)[(size_t)(f) * (size_t)a.cap + (size_t)(i)])
#define IFLD(base, f, i) ((base)[(size_t)(f) * (size_t)a.cap + (size_t)(i)])

// Particle field access for both state layouts.  Values travel as double;
// the mixed layout widens its FP32 fields on load and rounds them on store.
template <int PREC>
__device__ inline double ldp(const KArgs &a, const double *d, const float *f, int fld, int i) {
  if (PREC == KMPM_PRECISION_FP64) return d[(size_t)fld * a.cap + i];
  if (mx_is_double(fld)) return d[(size_t)mx_dslot(fld) * a.cap + i];
  return (double)f[(size_t)mx_fslot(fld) * a.cap + i];
}
template <int PREC>
__device__ inline void stp(const KArgs &a, double *d, float *f, int fld, int i, double v) {
  if (PREC == KMPM_PRECISION_FP64) d[(size_t)fld * a.cap + i] = v;
  else if (mx_is_double(fld)) d[(size_t)mx_dslot(fld) * a.cap + i] = v;
  else f[(size_t)mx_fslot(fld) * a.cap + i] = (float)v;
}
#define LDA(f, i) ldp<PREC>(a, a.pa, a.fa, (f), (i))
#define LDB(f, i) ldp<PREC>(a, a.pb, a.fb, (f), (i))
#define STB(f, i, v) stp<PREC>(a, a.pb, a.fb, (f), (i), (v))

// Grid arrays hold R = Real<PREC>::T values in the double allocations.
template <int PREC> __device__ inline typename Real<PREC>::T *gptr(double *p) {
  return reinterpret_cast<typename Real<PREC>::T *>(p);
}

__device__ inline void set_error(DevScalars *sc, int code, int64_t pid) {
  if (atomicCAS(&sc->error, 0, code) == 0) sc->err_pid = pid;
}

// A step that raises a sticky error ends normally and k_adjust_dt then sets
// `frozen` (the reference aborts at that point, solid.cpp:617-627,
// :1208-1215).  Every later step still launches, as a captured graph must,
// but its kernels return at entry; `frozen` only changes between steps, so
// the whole grid of a kernel sees one value.  The kernels that write the
// next state copy the current one into it slot by slot instead, so the
// host's buffer swap keeps the state of the failing step.
__device__ inline bool frozen(const KArgs &a) { return a.sc->frozen != 0; }

// ---- velocity fixes: values, reaction force ---------------------------------
// A fix value program (kmpm_expr_compile) at time t with the input script's
// dt variable.
__device__ inline double expr_eval(const kmpm_expr &e, double t, double dt) {
  double st[KMPM_EXPR_STACK];
  int sp = 0;
  for (int q = 0; q < e.n && q < KMPM_EXPR_MAX; q++) {
    const int o = e.op[q];
    if (o == KMPM_OP_CONST || o == KMPM_OP_TIME || o == KMPM_OP_DT) {
      if (sp >= KMPM_EXPR_STACK) return NAN;
      st[sp++] = o == KMPM_OP_CONST ? e.c[q] : o == KMPM_OP_TIME ? t : dt;
    } else if (o <= KMPM_OP_ATAN2) {
      if (sp < 2) return NAN;
      const double y = st[--sp], x = st[sp - 1];
      double r;
      switch (o) {
        case KMPM_OP_ADD: r = x + y; break;
        case KMPM_OP_SUB: r = x - y; break;
        case KMPM_OP_MUL: r = x * y; break;
        case KMPM_OP_DIV: r = x / y; break;
        case KMPM_OP_POW: r = pow(x, y); break;
        case KMPM_OP_GT: r = x > y; break;
        case KMPM_OP_GE: r = x >= y; break;
        case KMPM_OP_LT: r = x < y; break;
        case KMPM_OP_LE: r = x <= y; break;
        case KMPM_OP_EQ: r = x == y; break;
        case KMPM_OP_NE: r = x != y; break;
        default: r = atan2(x, y); break;
      }
      st[sp - 1] = r;
    } else {
      if (sp < 1) return NAN;
      const double x = st[sp - 1];
      double r;
      switch (o) {
        case KMPM_OP_NEG: r = -x; break;
        case KMPM_OP_NOT: r = !x; break;
        case KMPM_OP_EXP: r = exp(x); break;
        case KMPM_OP_SQRT: r = sqrt(x); break;
        case KMPM_OP_COS: r = cos(x); break;
        case KMPM_OP_SIN: r = sin(x); break;
        case KMPM_OP_TAN: r = tan(x); break;
        default: r = log(x); break;
      }
      st[sp - 1] = r;
    }
  }
  return sp == 1 ? st[0] : NAN;
}

// The values every velocity fix imposes in this step, into LDS (val[3],
// prev[3] per fix; expression values from k_fix_values), and the fix force
// accumulators cleared; every thread of the block calls it (it
// synchronises).
__device__ inline void fix_values_to_lds(const KArgs &a, double (*fxv)[6], double (*fsum)[3]) {
  const int tid = threadIdx.x;
  if (tid < a.nfv * 6) {
    const int fi = tid / 6, w = tid % 6;
    const DevFixVel &f = a.fv[fi];
    double v;
    if (f.expr) {
      v = a.sc->fv_last[fi][w];
    } else {
      v = w < 3 ? f.val[w] : f.prev[w - 3];
      if (blockIdx.x == 0) a.sc->fv_last[fi][w] = v;
    }
    fxv[fi][w] = v;
  }
  if (tid < a.nfv * 3) fsum[tid / 3][tid % 3] = 0.0;
  __syncthreads();
}

// post_update_grid_state of the velocity fixes on one node, in fix order
// (fix_velocity_nodes.cpp:130-225): v_update <- value, v <- previous value,
// and the reaction sum (1/dt) m (value - v_update) into the block's LDS sums.
__device__ inline void apply_velocity_fixes(const KArgs &a, uint32_t mask, double m, double inv_dt,
                                            const double (*fxv)[6], double (*fsum)[3],
                                            double *v, double *vu) {
  for (int fi = 0; fi < a.nfv; fi++) {
    const DevFixVel &f = a.fv[fi];
    if (!(mask & f.bit)) continue;
    for (int d = 0; d < 3; d++) {
      if (!f.set[d]) continue;
      const double dv = fxv[fi][d] - vu[d];
      vu[d] = fxv[fi][d];
      v[d] = fxv[fi][3 + d];
      const double r = (inv_dt * m) * dv;
      if (r != 0.0) atomicAdd(&fsum[fi][d], r);
    }
  }
}

// End of a grid update: the block's reaction sums into the step's buffer.
__device__ inline void flush_fix_forces(const KArgs &a, double (*fsum)[3]) {
  __syncthreads();
  const int tid = threadIdx.x;
  if (tid < a.nfv * 3 && fsum[tid / 3][tid % 3] != 0.0)
    atomicAdd(&a.sc->fv_force[a.sc->step & 1][tid / 3][tid % 3], fsum[tid / 3][tid % 3]);
}

__device__ inline void frozen_copy(const KArgs &a) {
  const bool mixed = a.prec == KMPM_PRECISION_MIXED;
  const int64_t nd = (int64_t)state_nd(mixed, a.lay) * a.cap;
  const int64_t nf = (int64_t)state_nf(mixed, a.lay) * a.cap;
  const int64_t ni = (int64_t)NIF * a.cap;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t i = i0; i < nd; i += stride) a.pb[i] = a.pa[i];
  for (int64_t i = i0; i < nf; i += stride) a.fb[i] = a.fa[i];
  for (int64_t i = i0; i < ni; i += stride) a.ib[i] = a.ia[i];
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int cell_of_clamped(double x, double lo, double inv_h, int n) {
  double u = (x - lo) * inv_h;
  if (!(u >= 0.0)) return 0;   // negative or NaN
  if (u >= (double)(n - 1)) return n - 1;
  return (int)u;
}

// Cell of a particle relative to the origin of the tile it was sorted into.
// The sort guarantees 0 <= result < T; a value outside (keys out of date)
// would address outside the tile's LDS node block, so it is clamped and
// reported instead.
__device__ inline int tile_local(int c, int org, int T, DevScalars *sc) {
  const int l = c - org;
  if (l < 0 || l >= T) {
    set_error(sc, KMPM_ERR_INTERNAL, -1);
    return l < 0 ? 0 : T - 1;
  }
  return l;
}

// dt of the instance a tile belongs to (batched contexts) or of the context
// (dt_ctx, read once per kernel).
__device__ inline double tile_dt(const KArgs &a, int t, double dt_ctx) {
  if (!a.ninst) return dt_ctx;
  const int i = a.tinst[t];
  return i >= 0 ? a.inst[i].dt : dt_ctx;
}

__device__ inline int tile_key(const KArgs &a, int c0, int c1, int c2) {
  const TileGeom &g = a.tg;
  int t = ((c0 / g.T[0]) * g.nt[1] + c1 / g.T[1]) * g.nt[2] + c2 / g.T[2];
  int l = ((c0 % g.T[0]) * g.T[1] + c1 % g.T[1]) * g.T[2] + c2 % g.T[2];
  return t * g.cpt + l;
}

// Batched contexts: a particle's stencil nodes must lie inside the tile
// range of the instance that owns its tile, so every node a particle reaches
// is advanced with that instance's dt and no instance sees another's mass.
__device__ inline void check_instance(const KArgs &a, const int c[3], int t, int i) {
  const int ii = a.tinst[t];
  bool ok = ii >= 0;
  if (ok) {
    const InstScalars &s = a.inst[ii];
    const int lo_off = a.S1 == 4 ? 1 : 0, hi_off = a.S1 == 4 ? 2 : 1;
    for (int d = 0; d < a.dim; d++) {
      const int nlo = s.tlo[d] * a.tg.T[d];
      int nhi = s.thi[d] * a.tg.T[d];        // exclusive
      nhi = nhi < a.n[d] ? nhi : a.n[d];
      // nodes below 0 or at n and beyond carry no weight
      const int b0 = c[d] - lo_off < 0 ? 0 : c[d] - lo_off;
      const int b1 = c[d] + hi_off > a.n[d] - 1 ? a.n[d] - 1 : c[d] + hi_off;
      ok = ok && b0 >= nlo && b1 < nhi;
    }
  }
  if (!ok) set_error(a.sc, KMPM_ERR_INSTANCE, i);
}

// Sort key of a particle at (x, y, z).
__device__ inline int particle_key(const KArgs &a, double x, double y, double z) {
  const int c0 = cell_of_clamped(x, a.lo[0], a.inv_h, a.n[0]);
  const int c1 = cell_of_clamped(y, a.lo[1], a.inv_h, a.n[1]);
  const int c2 = a.dim == 3 ? cell_of_clamped(z, a.lo[2], a.inv_h, a.n[2]) : 0;
  return tile_key(a, c0, c1, c2);
}

// Cell (i, j, k) of a sort key (the inverse of tile_key).
__device__ inline void key_cell(const KArgs &a, int key, int c[3]) {
  const TileGeom &g = a.tg;
  const int t = key / g.cpt, l = key % g.cpt;
  const int tz = t % g.nt[2], ty = (t / g.nt[2]) % g.nt[1], tx = t / (g.nt[2] * g.nt[1]);
  const int lz = l % g.T[2], ly = (l / g.T[2]) % g.T[1], lx = l / (g.T[2] * g.T[1]);
  c[0] = tx * g.T[0] + lx;
  c[1] = ty * g.T[1] + ly;
  c[2] = tz * g.T[2] + lz;
}

// Weight positions: current x (ULMPM) or reference x0 (TLMPM).
// Index in the double block (FP64 fields, or the mixed layout's double slots).
__device__ inline int wpos(const KArgs &a) {
  return a.tl ? (a.prec == KMPM_PRECISION_MIXED ? mx_dslot(X0X) : X0X) : PX;
}

// ---------------------------------------------------------------------------
// P2G: staged particles, lane = stencil column, register sums per cell
// ---------------------------------------------------------------------------
constexpr int SCR_FIELDS = 7;   // scratch slot layout: mass, momentum[3], force[3]

// Grid storage is tile-major: node (i, j, k) lives at tile * cells-per-tile
// + its index inside the tile (x-major, z fastest), the same order as the
// cell sort key, so a tile's nodes are one contiguous run of every grid array.
template <int TX, int TY, int TZ>
__device__ inline int64_t node_sidx(const KArgs &a, int i, int j, int k) {
  const TileGeom &g = a.tg;
  const int t = ((i / TX) * g.nt[1] + j / TY) * g.nt[2] + k / TZ;
  return (int64_t)t * (TX * TY * TZ) + ((i % TX) * TY + j % TY) * TZ + k % TZ;
}

// A tile's partial node block is stored as sub-blocks, one per tile its
// nodes belong to.  Per axis the block's nodes split into classes: cubic,
// the node before the tile, the tile's own T nodes and the two after it;
// linear, the own nodes and the one after.  Sub-blocks are ordered by class
// (x, then y, then z) and each is stored (x, y, z)-major, so the grid update
// of a tile reads every source sub-block as one contiguous run.
template <int S1, int T, bool ONE = false> struct AxisCls {
  static constexpr int n = ONE ? 1 : (S1 == 4 ? 3 : 2);
  // first block offset of class k (the classes tile the block in order)
  __host__ __device__ static constexpr int lo_b(int k) {
    return ONE ? 0 : S1 == 4 ? (k == 0 ? 0 : k == 1 ? 1 : T + 1) : (k == 0 ? 0 : T);
  }
  __host__ __device__ static constexpr int size(int k) {
    return ONE ? 1 : S1 == 4 ? (k == 0 ? 1 : k == 1 ? T : 2) : (k == 0 ? T : 1);
  }
  // tile offset of the class's nodes from the block's tile, and the first
  // local index they have in that tile
  __host__ __device__ static constexpr int toff(int k) { return ONE ? 0 : S1 == 4 ? k - 1 : k; }
  __host__ __device__ static constexpr int lo_l(int k) {
    return ONE ? 0 : (S1 == 4 && k == 0) ? T - 1 : 0;
  }
  __host__ __device__ static constexpr int cls(int b) {
    return ONE ? 0 : S1 == 4 ? (b < 1 ? 0 : b < T + 1 ? 1 : 2) : (b < T ? 0 : 1);
  }
};

// Position of block node (bx, by, bz) in the sub-block layout.
template <int DIM, int S1, int TX, int TY, int TZ>
__device__ inline int subblock_pos(int bx, int by, int bz) {
  using CX = AxisCls<S1, TX>;
  using CY = AxisCls<S1, TY>;
  using CZ = AxisCls<S1, TZ, DIM == 2>;
  constexpr int EY = TY + S1 - 1, EZ = DIM == 3 ? TZ + S1 - 1 : 1;
  const int kx = CX::cls(bx), ky = CY::cls(by), kz = CZ::cls(bz);
  const int sx = CX::size(kx), sy = CY::size(ky), sz = CZ::size(kz);
  const int lx = CX::lo_b(kx), ly = CY::lo_b(ky), lz = CZ::lo_b(kz);
  return lx * EY * EZ + sx * ly * EZ + sx * sy * lz + ((bx - lx) * sy + (by - ly)) * sz + (bz - lz);
}

// Raw particle data a P2G lane loads for one particle (R: the transfer type).
// Stress terms a P2G particle carries: -vol sigma (symmetric: 6 in 3D,
// xx,yy,xy in 2D) or, TLMPM, -vol0 PK1 (row-major: 9 in 3D, 4 in 2D).
template <int DIM, bool TL> constexpr int p2g_nsig() {
  return TL ? (DIM == 3 ? 9 : 4) : (DIM == 3 ? 6 : 3);
}

// APIC (XF > 0): the affine term's position xa (x, or x0 on TLMPM; the
// advanced x in the MUSL re-map, whose weights are those of the step's
// start) and C, the affine matrix of the previous step (row-major, DIM x DIM).
template <int DIM, bool REMAP, bool TL, class R, int XF = 0>
struct P2GIn {
  double x[DIM];
  R v[DIM], m;
  R s[REMAP ? 1 : p2g_nsig<DIM, TL>()];
  double xa[XF ? DIM : 1];
  R c[XF ? DIM * DIM : 1];
};

template <int DIM, bool REMAP, bool TL, int PREC, class R, int XF>
__device__ inline void p2g_load(const KArgs &a, int j, P2GIn<DIM, REMAP, TL, R, XF> &in) {
  if constexpr (XF > 0) {
    const int slot = a.order[j];
    for (int r = 0; r < DIM; r++)
      for (int q = 0; q < DIM; q++) in.c[r * DIM + q] = (R)LDA(C00 + 3 * r + q, slot);
  }
  if (REMAP) {
    for (int d = 0; d < DIM; d++) {
      in.x[d] = TL ? LDB(X0X + d, j) : a.xo[(size_t)d * a.cap + j];
      in.v[d] = (R)LDB(VX + d, j);
      if constexpr (XF > 0) in.xa[d] = TL ? in.x[d] : LDB(PX + d, j);
    }
    in.m = (R)LDB(MASS, j);
  } else if (TL) {
    const int slot = a.order[j];
    for (int d = 0; d < DIM; d++) {
      in.x[d] = LDA(X0X + d, slot);
      in.v[d] = (R)LDA(VX + d, slot);
    }
    in.m = (R)LDA(MASS, slot);
    if (DIM == 3) {
      for (int e = 0; e < 9; e++) in.s[e] = (R)-LDA(P00 + e, slot);
    } else {
      in.s[0] = (R)-LDA(P00, slot);
      in.s[1] = (R)-LDA(P01, slot);
      in.s[2] = (R)-LDA(P10, slot);
      in.s[3] = (R)-LDA(P11, slot);
    }
  } else {
    const int slot = a.order[j];
    for (int d = 0; d < DIM; d++) {
      in.x[d] = LDA(PX + d, slot);
      in.v[d] = (R)LDA(VX + d, slot);
    }
    in.m = (R)LDA(MASS, slot);
    // MLS (solid.cpp:560-565): -vol sigma Di, Di = di I in the run's axes
    const R nvol = XF == 2 ? (R)(-LDA(VOL, slot) * a.di) : (R)-LDA(VOL, slot);
    if (DIM == 3) {
      for (int e = 0; e < 6; e++) in.s[e] = nvol * (R)LDA(SXX + e, slot);
    } else {
      in.s[0] = nvol * (R)LDA(SXX, slot);
      in.s[1] = nvol * (R)LDA(SYY, slot);
      in.s[2] = nvol * (R)LDA(SXY, slot);
    }
  }
  if constexpr (XF > 0)
    if (!REMAP)
      for (int d = 0; d < DIM; d++) in.xa[d] = in.x[d];
}

// Stress term S_de of a staged P2G payload: -vol sigma (symmetric, stored
// xx,yy,zz,xy,xz,yz in 3D and xx,yy,xy in 2D) or, TLMPM, -vol0 PK1 (row-major).
template <int DIM, bool TL, class R>
__device__ inline R p2g_stress(const R *s, int d, int e) {
  if (TL) return s[d * DIM + e];
  if (DIM == 3) {
    if (d == e) return s[d];
    const int o = d + e;   // (0,1) -> xy, (0,2) -> xz, (1,2) -> yz
    return s[o == 1 ? 3 : o == 2 ? 4 : 5];
  }
  return d == e ? s[d] : s[2];
}

// Scaled flush of a P2G register sum into a 64-bit fixed-point LDS node sum:
// q = val * scale is exact in FP32 (a power-of-two scale within the float
// range) and rounded to the nearest integer (ties to even, as in FP64).  Its
// two's-complement words are hi = floor(q / 2^32) and lo = q - hi 2^32, in
// [0, 2^32), formed exactly in FP64.  Returns false, and adds zero, when |q|
// reaches the flush limit (or val is NaN); the caller reports that once.
__device__ inline bool fx_add_f(unsigned long long *dst, float val, float scale) {
  const float q = rintf(val * scale);
  const bool ok = fabsf(q) < (float)FX_FLUSH_LIMIT;
  const float qc = ok ? q : 0.0f;
  const float hi = floorf(qc * 2.3283064365386963e-10f);   // |hi| < 2^22
  const double lo = fma(-(double)hi, 4294967296.0, (double)qc);
  const unsigned long long iv =
      ((unsigned long long)(unsigned)(int)hi << 32) | (unsigned long long)(unsigned)lo;
  atomicAdd(dst, iv);
  return ok;
}

// fx_add_f's word for |val * scale| < 2^51, in three instructions: t =
// val * scale + 1.5 * 2^52 is rounded once, to the nearest integer (ties to
// even, as rintf; val * scale is exact in FP64), and the bits of t are then
// those of 1.5 * 2^52 plus q = rint(val * scale) as a 64-bit integer.
constexpr double FX_MAGIC = 6755399441055744.0;              // 1.5 * 2^52
constexpr unsigned long long FX_MAGIC_BITS = 0x4338000000000000ull;
constexpr double FX_FAST_LIMIT = 2251799813685248.0;         // 2^51
__device__ inline unsigned long long fx_word_fast(float val, double scale) {
  const double t = fma((double)val, scale, FX_MAGIC);
  return (unsigned long long)__double_as_longlong(t) - FX_MAGIC_BITS;
}

// P2G over the active tiles, one workgroup per tile at a time.
//
// The tile's particles are staged in chunks of CHK (sorted by cell): one
// thread per particle evaluates its shape functions (FP64, the reference's
// expressions) and stores the per-axis (w, dw) pairs and its payload in LDS.
// Each run of particles of one cell (a segment) goes to a lane group: a
// lane owns one stencil column (the nodes that share all but the last axis
// offset) and accumulates its S1 nodes of that column in registers over the
// segment, so a node sum costs one multiply-add per field and particle; the
// group adds its sums to the tile's LDS node block once per segment.  The
// block is then written to the tile's partial buffer (no global atomics).
//
// FP64: double arithmetic and LDS sums.  Mixed: float arithmetic, register
// sums flushed into 64-bit fixed-point LDS sums (order-independent, so the
// scatter is deterministic).
//
// APIC (XF 1, solid.cpp:392-426): the momentum adds m w C (x_I - x_p).  Per
// axis the stage also holds g = (x_I - x_p) w, so sum_b C_db (x_I - x_p)_b w
// is a gradient-like contraction: the column product with one axis's w
// replaced by its g.  MLS (XF 2, solid.cpp:524-574): the force
// -vol w sigma Di (x_I - x_p) is the gradient form with dw replaced by g and
// sigma scaled by Di, so the stage holds (w, g) in place of (w, dw).
//
// TH: the thermal transfer of a thermo-mechanical run (FP64, XF 0) into its
// own partial blocks a.scratch_th [slot][3][bn]: sum w m T (the grid
// temperature's numerator, solid.cpp:2743-2762), sum w gamma (the external
// driving force, :2764-2779) and sum grad w . q (the internal one,
// :2781-2797); the MUSL re-map (REMAP) sums w m T of the updated
// temperatures only (ulmpm.cpp:491-493).
template <int DIM, int S1, int TX, int TY, int TZ, bool REMAP, int PREC, bool TL, int XF = 0,
          bool TH = false>
__global__ __launch_bounds__(256) void k_p2g(KArgs a) {
  if (frozen(a)) return;
  using R = typename Real<PREC>::T;
  using R2 = typename Real<PREC>::T2;
  constexpr bool FX = PREC == KMPM_PRECISION_MIXED;
  using BT = typename std::conditional<FX, unsigned long long, double>::type;
  constexpr int GL = DIM == 3 ? S1 * S1 : S1;   // lanes (stencil columns) per group
  constexpr int NGW = 64 / GL;                   // groups per wave
  constexpr int NG = 4 * NGW;                    // groups per block
  constexpr int CHK = XF ? 64 : 128;             // particles staged per chunk
  constexpr int EX = TX + S1 - 1, EY = TY + S1 - 1, EZ = DIM == 3 ? TZ + S1 - 1 : 1;
  constexpr int NSIG = p2g_nsig<DIM, TL>();
  static_assert(!TH || (XF == 0 && PREC == KMPM_PRECISION_FP64), "thermal transfer: FP64");
  constexpr int NF = TH ? (REMAP ? 1 : 3) : REMAP ? (1 + DIM) : (1 + 2 * DIM);
  constexpr int PC = 1 + DIM + (REMAP ? 0 : NSIG);         // payload offset of m C
  // m, m v, stress terms, m C; thermal: m T, gamma, q
  constexpr int NPAY = TH ? (REMAP ? 1 : 2 + DIM) : PC + (XF ? DIM * DIM : 0);
  // APIC with the gradient force needs dw and g: g in its own table
  constexpr bool GT = XF == 1 && !REMAP;
  constexpr int PSB = DIM == 3 ? EZ : 1;   // node block strides (last axis contiguous)
  constexpr int PSA = DIM == 3 ? EY * EZ : EY;
  constexpr int PBN = EX * PSA;
  static_assert(64 % GL == 0, "stencil shape");

  __shared__ BT blk[NF * PBN];
  // Rows of the stage tables are padded to an odd number of elements, so
  // the stage's per-particle row writes fall on distinct LDS banks.
  __shared__ R2 stw[CHK][DIM * S1 + 1];   // (w, dw) per axis and stencil offset; (w, g) for
                                          // the MLS force and the APIC re-map
  __shared__ R stg[GT ? CHK : 1][GT ? DIM * S1 + 1 : 1];   // APIC: g per axis and offset
  __shared__ R spay[CHK][NPAY + 1];     // +1: rows of distinct particles in distinct banks
  __shared__ int scl[CHK];              // block offset of the particle's cell
  __shared__ int seg[CHK + 1];          // segment starts; seg[nseg] = chunk size
  __shared__ int wheads[4];
  __shared__ int nseg_s;
  // sub-block layout position -> LDS node block offset, for the partial
  // block write (contiguous stores, gathered from LDS)
  constexpr int BNP = EX * EY * EZ;
  __shared__ short pinv[BNP];
  for (int o = threadIdx.x; o < BNP; o += 256) {
    const int bz = o % EZ, by = (o / EZ) % EY, bx = o / (EZ * EY);
    pinv[subblock_pos<DIM, S1, TX, TY, TZ>(bx, by, bz)] = (short)(bx * PSA + by * PSB + bz);
  }

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int nt1 = a.tg.nt[1], nt2 = a.tg.nt[2];
  const int off = S1 == 4 ? 1 : 0;
  const int ntile = a.sc->tile_count;

  // lane -> group and stencil column
  const int gl = lane % GL, q = wave * NGW + lane / GL;
  const int sa = DIM == 3 ? gl / S1 : gl;
  const int sb = DIM == 3 ? gl % S1 : 0;
  const int colofs = sa * PSA + sb * PSB;
  R fs[FX_N];
  for (int k = 0; k < FX_N; k++) fs[k] = (R)a.fxs[k];
  // |sum| < flim[k] implies |sum * fxs[k]| < 2^51 (a power of two, clamped
  // into the float range, which only sends more sums to the exact split)
  double fsd[FX_N];
  float flim[FX_N];
  for (int k = 0; k < FX_N; k++) {
    fsd[k] = a.fxs[k];
    const double l = FX_FAST_LIMIT / a.fxs[k];
    flim[k] = l > 0x1p126 ? 0x1p126f : (float)l;
  }

  for (int li = blockIdx.x; li < ntile; li += gridDim.x) {
    const int t = a.tile_list[li];
    const int tz = DIM == 3 ? t % nt2 : 0;
    const int ty = (t / nt2) % nt1;
    const int tx = t / (nt2 * nt1);
    const int org[3] = {tx * TX, ty * TY, DIM == 3 ? tz * TZ : 0};
    const int p0 = a.tile_p0[t], p1 = a.tile_p0[t + 1];
    for (int i = tid; i < NF * PBN; i += 256) blk[i] = 0;

    for (int c0 = p0; c0 < p1; c0 += CHK) {
      const int n = min(CHK, p1 - c0);
      __syncthreads();   // the previous chunk's segments are done
      // ---- stage: one thread per particle ----
      if (tid < n) {
        P2GIn<DIM, REMAP, TL, R, XF> in;
        double thp[TH ? 2 + DIM : 1];   // thermal payload: m T, gamma, q
        if constexpr (TH) {
          const int j = c0 + tid;
          if (REMAP) {   // the updated temperatures at the step's first positions
            for (int d = 0; d < DIM; d++)
              in.x[d] = TL ? LDB(X0X + d, j) : a.xo[(size_t)d * a.cap + j];
            thp[0] = LDB(MASS, j) * LDB(TEMP, j);
          } else {
            const int slot = a.order[j];
            for (int d = 0; d < DIM; d++) in.x[d] = LDA((TL ? X0X : PX) + d, slot);
            thp[0] = LDA(MASS, slot) * LDA(TEMP, slot);
            thp[1] = LDA(GAMMA, slot);
            for (int d = 0; d < DIM; d++) thp[2 + d] = LDA(QX + d, slot);
          }
        } else {
          p2g_load<DIM, REMAP, TL, PREC>(a, c0 + tid, in);
        }
        int cl[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; d++) {
          const int c = cell_of_clamped(in.x[d], a.lo[d], a.inv_h, a.n[d]);
          cl[d] = tile_local(c, org[d], d == 0 ? TX : d == 1 ? TY : TZ, a.sc);
          double w[S1], dw[S1];
          weights_1d<S1>(in.x[d], a.lo[d], a.h, a.inv_h, a.n[d], c - off, w, dw);
          if constexpr (XF > 0) {
#pragma unroll
            for (int k = 0; k < S1; k++) {
              // node coordinate as the weights compute it
              const double g = ((a.lo[d] + (0 + (c - off) + k) * a.h) - in.xa[d]) * w[k];
              if constexpr (GT) {
                stw[tid][d * S1 + k] = R2{(R)w[k], (R)dw[k]};
                stg[tid][d * S1 + k] = (R)g;
              } else {
                stw[tid][d * S1 + k] = R2{(R)w[k], (R)g};
              }
            }
          } else {
#pragma unroll
            for (int k = 0; k < S1; k++) stw[tid][d * S1 + k] = R2{(R)w[k], (R)dw[k]};
          }
        }
        const int cell = DIM == 3 ? cl[0] * PSA + cl[1] * PSB + cl[2] : cl[0] * PSA + cl[1];
        scl[tid] = cell;
        if constexpr (TH) {
          for (int e = 0; e < NPAY; e++) spay[tid][e] = thp[e];
        } else {
          spay[tid][0] = in.m;
          for (int d = 0; d < DIM; d++) spay[tid][1 + d] = in.m * in.v[d];
          if (!REMAP)
            for (int e = 0; e < NSIG; e++) spay[tid][1 + DIM + e] = in.s[e];
          if constexpr (XF > 0)
            for (int e = 0; e < DIM * DIM; e++) spay[tid][PC + e] = in.m * in.c[e];
        }
      }
      __syncthreads();
      // cells are sorted inside the tile: a segment starts where the cell changes
      const bool head = tid < n && (tid == 0 || scl[tid] != scl[tid - 1]);
      // ---- segment starts: ballot per wave, prefix over the waves ----
      const unsigned long long hb = __ballot(head);
      if (lane == 0) wheads[wave] = __popcll(hb);
      __syncthreads();
      int before = 0;
      for (int w = 0; w < wave; w++) before += wheads[w];
      if (head) seg[before + __popcll(hb & ((1ull << lane) - 1ull))] = tid;
      if (tid == 0) {
        const int tot = wheads[0] + wheads[1] + wheads[2] + wheads[3];
        nseg_s = tot;
        seg[tot] = n;
      }
      __syncthreads();
      const int nseg = nseg_s;
      // ---- accumulate: group q takes segments q, q + NG, ... ----
      for (int s = q; s < nseg; s += NG) {
        const int i0 = seg[s], i1 = seg[s + 1];
        R acc[S1][NF];
#pragma unroll
        for (int k = 0; k < S1; k++)
#pragma unroll
          for (int f = 0; f < NF; f++) acc[k][f] = 0;
        for (int i = i0; i < i1; i++) {
          // column factors: c = product of the non-last-axis weights, ce[e]
          // the same with axis e differentiated, cg[e] (APIC) with axis e's
          // w replaced by its g
          R c, ce[DIM - 1], cg[DIM - 1];
          if constexpr (DIM == 3) {
            const R2 px = stw[i][sa], py = stw[i][S1 + sb];
            c = px.x * py.x;
            ce[0] = px.y * py.x;
            ce[DIM - 2] = px.x * py.y;
            if constexpr (XF > 0) {
              cg[0] = (GT ? stg[i][sa] : px.y) * py.x;
              cg[DIM - 2] = px.x * (GT ? stg[i][S1 + sb] : py.y);
            }
          } else {
            const R2 px = stw[i][sa];
            c = px.x;
            ce[0] = px.y;
            if constexpr (XF > 0) cg[0] = GT ? stg[i][sa] : px.y;
          }
          const R *p = spay[i];
          if constexpr (TH) {
            // w m T, w gamma and grad w . q (the last axis' factor below)
            const R mtc = p[0] * c;
            R gc = 0, qa = 0, qb = 0;
            if constexpr (!REMAP) {
              gc = p[1] * c;
              qa = p[2] * ce[0];
              if constexpr (DIM == 3) qa = fma(p[3], ce[DIM - 2], qa);
              qb = p[1 + DIM] * c;
            }
#pragma unroll
            for (int k = 0; k < S1; k++) {
              const R2 wz = stw[i][(DIM - 1) * S1 + k];
              acc[k][0] = fma(mtc, wz.x, acc[k][0]);
              if constexpr (!REMAP) {
                acc[k][1] = fma(gc, wz.x, acc[k][1]);
                acc[k][2] = fma(qb, wz.y, fma(qa, wz.x, acc[k][2]));
              }
            }
          } else {
            const R mc = p[0] * c;
            R mvc[DIM], fa[DIM], fb[DIM], mvb[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) mvc[d] = p[1 + d] * c;
            if constexpr (XF > 0) {
              // m C (x_I - x_p) w: the non-last axes fold into the column's
              // momentum, the last axis pairs with its g below
              const R *mC = p + PC;
#pragma unroll
              for (int d = 0; d < DIM; d++) {
                R u = mC[d * DIM] * cg[0];
                if constexpr (DIM == 3) u = fma(mC[d * DIM + 1], cg[DIM - 2], u);
                mvc[d] += u;
                mvb[d] = mC[d * DIM + DIM - 1] * c;
              }
            }
            if (!REMAP) {
              const R *sg = p + 1 + DIM;
#pragma unroll
              for (int d = 0; d < DIM; d++) {
                R u = p2g_stress<DIM, TL>(sg, d, 0) * ce[0];
                if constexpr (DIM == 3) u = fma(p2g_stress<DIM, TL>(sg, d, 1), ce[DIM - 2], u);
                fa[d] = u;
                fb[d] = p2g_stress<DIM, TL>(sg, d, DIM - 1) * c;
              }
            }
#pragma unroll
            for (int k = 0; k < S1; k++) {
              const R2 wz = stw[i][(DIM - 1) * S1 + k];
              acc[k][0] = fma(mc, wz.x, acc[k][0]);
#pragma unroll
              for (int d = 0; d < DIM; d++) acc[k][1 + d] = fma(mvc[d], wz.x, acc[k][1 + d]);
              if constexpr (XF > 0) {
                const R gz = GT ? stg[i][(DIM - 1) * S1 + k] : wz.y;
#pragma unroll
                for (int d = 0; d < DIM; d++) acc[k][1 + d] = fma(mvb[d], gz, acc[k][1 + d]);
              }
              if (!REMAP)
#pragma unroll
                for (int d = 0; d < DIM; d++)
                  acc[k][1 + DIM + d] = fma(fb[d], wz.y, fma(fa[d], wz.x, acc[k][1 + DIM + d]));
            }
          }
        }
        const int base = scl[i0] + colofs;
        bool ok = true;
        if constexpr (FX) {
          // the three-instruction words unless a sum of the wave is beyond
          // their range (then the exact split, which also reports overflow).
          // The largest magnitude per scale class, on the bit patterns of
          // |sum|: they order as the values do, with infinities and NaNs
          // above every finite value, so those take the exact split too.
          uint32_t mb[3] = {0u, 0u, 0u};
#pragma unroll
          for (int k = 0; k < S1; k++)
#pragma unroll
            for (int f = 0; f < NF; f++) {
              const int c = f == 0 ? 0 : f <= DIM ? 1 : 2;
              mb[c] = max(mb[c], __float_as_uint(acc[k][f]) & 0x7fffffffu);
            }
          const bool fast = mb[0] < __float_as_uint(flim[FX_MASS]) &&
                            mb[1] < __float_as_uint(flim[FX_MOM]) &&
                            mb[2] < __float_as_uint(flim[FX_FORCE]);
          if (__ballot(!fast) == 0) {
#pragma unroll
            for (int k = 0; k < S1; k++)
#pragma unroll
              for (int f = 0; f < NF; f++)
                atomicAdd(&blk[f * PBN + base + k],
                          fx_word_fast(acc[k][f],
                                       fsd[f == 0 ? FX_MASS : f <= DIM ? FX_MOM : FX_FORCE]));
          } else {
#pragma unroll
            for (int k = 0; k < S1; k++)
#pragma unroll
              for (int f = 0; f < NF; f++)
                ok &= fx_add_f(&blk[f * PBN + base + k], acc[k][f],
                               fs[f == 0 ? FX_MASS : f <= DIM ? FX_MOM : FX_FORCE]);
          }
        } else {
#pragma unroll
          for (int k = 0; k < S1; k++)
#pragma unroll
            for (int f = 0; f < NF; f++) atomicAdd(&blk[f * PBN + base + k], acc[k][f]);
        }
        if (!ok) set_error(a.sc, KMPM_ERR_FIXED_POINT, -1);
      }
    }
    __syncthreads();
    constexpr int BN = EX * EY * EZ;
    if constexpr (TH) {
      double *tdst = a.scratch_th + (size_t)li * 3 * BN;
      for (int qq = tid; qq < NF * BN; qq += 256) {
        const int f = qq / BN, p = qq % BN;
        tdst[qq] = blk[f * PBN + pinv[p]];
      }
      __syncthreads();
      continue;
    }
    BT *dst = reinterpret_cast<BT *>(a.scratch) + (size_t)li * SCR_FIELDS * BN;
    // blk fields: mass, mom[DIM], force[DIM] -> scratch fields mass, mom[3],
    // force[3]; components a 2D run does not have are written as zeros.
    const int f_lo = REMAP ? 1 : 0, f_hi = REMAP ? 4 : SCR_FIELDS;
    for (int qq = tid + f_lo * BN; qq < f_hi * BN; qq += 256) {
      const int sf = qq / BN, p = qq % BN;   // p: position in the sub-block layout
      int bf;
      if (sf == 0) bf = 0;
      else if (sf <= 3) bf = sf <= DIM ? sf : -1;
      else bf = (sf - 3) <= DIM ? DIM + (sf - 3) : -1;
      dst[qq] = bf >= 0 ? blk[bf * PBN + pinv[p]] : (BT)0;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// grid: sum partial blocks, v = p/m, v_update, fixes (grid.cpp:448-466,
// fix_body_force.cpp:117-146, fix_velocity_nodes.cpp:130-271)
// ---------------------------------------------------------------------------
// The node fields a grid pass (GridMode) reads from the partial blocks:
// [f0, f1) of mass, momentum[3], force[3].
__host__ __device__ constexpr int gm_f0(int mode) {
  return mode == GM_REMAP ? 1 : mode == GM_USF2 ? 4 : 0;
}
__host__ __device__ constexpr int gm_f1(int mode) {
  return mode == GM_REMAP || mode == GM_USF1 ? 4 : 7;
}

// One owned node of a grid pass, from its summed node fields `sum` (FP64):
//   GM_STEP   v = p/m, v_update = v + dt (f + mb)/m (grid.cpp:448-466), then
//             the velocity fixes' post_update_grid_state
//   GM_REMAP  MUSL re-map (ulmpm.cpp:476-496): v = p/m with the step's mass,
//             then post_velocities_to_grid, v <- value
//             (fix_velocity_nodes.cpp:227-271)
//   GM_USF1   USF after particles_to_grid_USF_1 (usf.cpp:53-55): m, v = p/m,
//             then post_update_grid_state's v <- previous value (its
//             v_update and reaction are replaced by GM_USF2's)
//   GM_USF2   USF update_grid_state (usf.cpp:63-67): GM_STEP from GM_USF1's
//             m and v with the forces of particles_to_grid_USF_2
template <class R>
__device__ inline void grid_node(const KArgs &a, int mode, int64_t sn, const double *sum,
                                 uint32_t mask, double dt, R *gmass, R *gv, R *gvu,
                                 const double (*fxv)[6], double (*fsum)[3]) {
  if (mode == GM_REMAP) {
    const double m = gmass[sn];
    double v[3];
    for (int d = 0; d < 3; d++) v[d] = m > 0 ? sum[1 + d] / m : 0.0;
    for (int fi = 0; fi < a.nfv; fi++)
      if (mask & a.fv[fi].bit)
        for (int d = 0; d < 3; d++)
          if (a.fv[fi].set[d]) v[d] = fxv[fi][d];
    for (int d = 0; d < 3; d++) gv[d * a.ns + sn] = (R)v[d];
    return;
  }
  if (mode == GM_USF1) {
    const double m = sum[0];
    double v[3];
    for (int d = 0; d < 3; d++) v[d] = m > 0 ? sum[1 + d] / m : 0.0;
    for (int fi = 0; fi < a.nfv; fi++)
      if (mask & a.fv[fi].bit)
        for (int d = 0; d < 3; d++)
          if (a.fv[fi].set[d]) v[d] = fxv[fi][3 + d];
    gmass[sn] = (R)m;
    for (int d = 0; d < 3; d++) gv[d * a.ns + sn] = (R)v[d];
    return;
  }
  const bool usf = mode == GM_USF2;
  const double m = usf ? (double)gmass[sn] : sum[0];
  double v[3], vu[3], mb[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++)
    v[d] = usf ? (double)gv[d * a.ns + sn] : m > 0 ? sum[1 + d] / m : 0.0;
  for (int fi = 0; fi < a.nbf; fi++)
    if (m > 0 && (mask & a.bf[fi].bit))
      for (int d = 0; d < 3; d++)
        if (a.bf[fi].set[d]) mb[d] += a.bf[fi].val[d] * m;
  for (int d = 0; d < 3; d++) vu[d] = m != 0 ? v[d] + dt * (sum[4 + d] + mb[d]) / m : v[d];
  apply_velocity_fixes(a, mask, m, 1.0 / dt, fxv, fsum, v, vu);
  gmass[sn] = (R)m;
  for (int d = 0; d < 3; d++) {
    gv[d * a.ns + sn] = (R)v[d];
    gvu[d * a.ns + sn] = (R)vu[d];
  }
}
template <int DIM, int S1, int TX, int TY, int TZ, int PREC>
__global__ __launch_bounds__(256) void k_grid_update(KArgs a, int mode) {
  if (frozen(a)) return;
  using R = typename Real<PREC>::T;
  constexpr bool FX = PREC == KMPM_PRECISION_MIXED;
  using BT = typename std::conditional<FX, long long, double>::type;
  using CX = AxisCls<S1, TX>;
  using CY = AxisCls<S1, TY>;
  using CZ = AxisCls<S1, TZ, DIM == 2>;
  R *gmass = gptr<PREC>(a.gmass), *gv = gptr<PREC>(a.gv), *gvu = gptr<PREC>(a.gvu);
  constexpr int EX = TX + S1 - 1, EY = TY + S1 - 1, EZ = DIM == 3 ? TZ + S1 - 1 : 1;
  constexpr int BN = EX * EY * EZ;
  constexpr int TZZ = DIM == 3 ? TZ : 1;
  constexpr int NO = TX * TY * TZZ;
  __shared__ double fxv[MAX_FIX][6];
  __shared__ double fsum[MAX_FIX][3];
  fix_values_to_lds(a, fxv, fsum);
  if constexpr (FX) {
    // Mixed: one group of 256 / TPB threads per dilated tile.  The tile's
    // nodes lie in the partial blocks of the tiles around it, one contiguous
    // sub-block per source tile at the same block position in every source;
    // the group streams each source's run (coalesced) and adds every word
    // into the tile's node sums in LDS with 64-bit integer atomics (exact, so
    // the order of the adds does not matter).
    constexpr int TPB = NO >= 256 ? 1 : 256 / NO;   // tiles per workgroup
    constexpr int GS = 256 / TPB;                     // threads per tile
    constexpr int NC = CX::n * CY::n * CZ::n;
    __shared__ unsigned long long acc[TPB][SCR_FIELDS][NO];
    __shared__ unsigned short otab[BN];   // block position -> node of the gathering tile
    __shared__ unsigned char ctab[BN];    // block position -> source class combo
    __shared__ int sslot[TPB][NC];
    const int tid = threadIdx.x, grp = tid / GS, gl = tid % GS;
    for (int o = tid; o < BN; o += 256) {
      const int bz = o % EZ, by = (o / EZ) % EY, bx = o / (EZ * EY);
      const int kx = CX::cls(bx), ky = CY::cls(by), kz = CZ::cls(bz);
      const int lx = bx - CX::lo_b(kx) + CX::lo_l(kx), ly = by - CY::lo_b(ky) + CY::lo_l(ky),
                lz = bz - CZ::lo_b(kz) + CZ::lo_l(kz);
      const int p = subblock_pos<DIM, S1, TX, TY, TZ>(bx, by, bz);
      ctab[p] = (unsigned char)((kx * CY::n + ky) * CZ::n + kz);
      otab[p] = (unsigned short)((lx * TY + ly) * TZZ + lz);
    }
    const int nt0 = a.tg.nt[0], nt1 = a.tg.nt[1], nt2 = a.tg.nt[2];
    const int ndil = a.sc->dil_count;
    const double dt_ctx = a.sc->dt;
    const unsigned long long *scr = reinterpret_cast<const unsigned long long *>(a.scratch);
    const int f0 = gm_f0(mode), f1 = gm_f1(mode);
    for (int l0 = blockIdx.x * TPB; l0 < ndil; l0 += gridDim.x * TPB) {
      const int li = l0 + grp;
      const bool have = li < ndil;
      const int t = have ? a.dil_list[li] : 0;
      const int tz = t % nt2, ty = (t / nt2) % nt1, tx = t / (nt2 * nt1);
      __syncthreads();   // tables written; the previous tiles' nodes are done
      for (int q = gl; q < SCR_FIELDS * NO; q += GS) acc[grp][q / NO][q % NO] = 0;
      if (gl < NC) {
        const int kx = gl / (CY::n * CZ::n), ky = (gl / CZ::n) % CY::n, kz = gl % CZ::n;
        const int ux = tx - CX::toff(kx), uy = ty - CY::toff(ky), uz = tz - CZ::toff(kz);
        const bool in = have && ux >= 0 && ux < nt0 && uy >= 0 && uy < nt1 && uz >= 0 &&
                        uz < nt2;
        sslot[grp][gl] = in ? a.tile_slot[(ux * nt1 + uy) * nt2 + uz] : -1;
      }
      __syncthreads();
      for (int p = gl; p < BN; p += GS) {
        const int sl = sslot[grp][ctab[p]];
        if (sl < 0) continue;
        const unsigned long long *src = scr + (size_t)sl * SCR_FIELDS * BN + p;
        unsigned long long v[SCR_FIELDS];
#pragma unroll
        for (int f = 0; f < SCR_FIELDS; f++)
          if (f >= f0 && f < f1) v[f] = src[f * BN];
        const int o = otab[p];
#pragma unroll
        for (int f = 0; f < SCR_FIELDS; f++)
          if (f >= f0 && f < f1) atomicAdd(&acc[grp][f][o], v[f]);
      }
      __syncthreads();
      if (!have) continue;
      const double dt = tile_dt(a, t, dt_ctx);
      for (int o = gl; o < NO; o += GS) {
        const int lz = o % TZZ, ly = (o / TZZ) % TY, lx = o / (TZZ * TY);
        const int i = tx * TX + lx, j = ty * TY + ly, k = DIM == 3 ? tz * TZ + lz : 0;
        if (i >= a.n[0] || j >= a.n[1] || k >= a.n[2]) continue;
        const int64_t sn = (int64_t)t * NO + o;   // tile-major storage index
        BT isum[SCR_FIELDS];
        for (int f = 0; f < SCR_FIELDS; f++) isum[f] = (BT)acc[grp][f][o];
        double sum[SCR_FIELDS];
        for (int f = 0; f < SCR_FIELDS; f++)
          sum[f] = FX ? (double)isum[f] * a.fxi[f == 0 ? FX_MASS : f <= 3 ? FX_MOM : FX_FORCE]
                      : (double)isum[f];
        grid_node<R>(a, mode, sn, sum, a.gmask[sn], dt, gmass, gv, gvu, fxv, fsum);
      }
    }
  } else {
    const int nt0 = a.tg.nt[0], nt1 = a.tg.nt[1], nt2 = a.tg.nt[2];
    const int total = a.sc->dil_count * NO;
    const double dt_ctx = a.sc->dt;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
      const int t = a.dil_list[q / NO], o = q % NO;
      const double dt = tile_dt(a, t, dt_ctx);
      const int tz = t % nt2, ty = (t / nt2) % nt1, tx = t / (nt2 * nt1);
      const int lz = o % TZZ, ly = (o / TZZ) % TY, lx = o / (TZZ * TY);
      const int i = tx * TX + lx, j = ty * TY + ly, k = DIM == 3 ? tz * TZ + lz : 0;
      if (i >= a.n[0] || j >= a.n[1] || k >= a.n[2]) continue;
      const int64_t sn = (int64_t)t * NO + o;   // tile-major storage index
      BT isum[SCR_FIELDS];
      for (int f = 0; f < SCR_FIELDS; f++) isum[f] = 0;
      // the sub-blocks of the source tiles whose blocks cover this node
  #pragma unroll
      for (int kx = 0; kx < CX::n; kx++) {
        const int ux = tx - CX::toff(kx), rx = lx - CX::lo_l(kx);
        if (rx < 0 || rx >= CX::size(kx) || ux < 0 || ux >= nt0) continue;
  #pragma unroll
        for (int ky = 0; ky < CY::n; ky++) {
          const int uy = ty - CY::toff(ky), ry = ly - CY::lo_l(ky);
          if (ry < 0 || ry >= CY::size(ky) || uy < 0 || uy >= nt1) continue;
  #pragma unroll
          for (int kz = 0; kz < CZ::n; kz++) {
            const int uz = tz - CZ::toff(kz), rz = lz - CZ::lo_l(kz);
            if (rz < 0 || rz >= CZ::size(kz) || uz < 0 || uz >= nt2) continue;
            const int slot = a.tile_slot[(ux * nt1 + uy) * nt2 + uz];
            if (slot < 0) continue;
            const int sx = CX::size(kx), sy = CY::size(ky), sz = CZ::size(kz);
            const int pos = CX::lo_b(kx) * EY * EZ + sx * CY::lo_b(ky) * EZ +
                            sx * sy * CZ::lo_b(kz) + (rx * sy + ry) * sz + rz;
            const BT *src = reinterpret_cast<const BT *>(a.scratch) + (size_t)slot * SCR_FIELDS * BN + pos;
            const int f0 = gm_f0(mode), f1 = gm_f1(mode);
#pragma unroll
            for (int f = 0; f < SCR_FIELDS; f++)
              if (f >= f0 && f < f1) isum[f] += src[f * BN];
          }
        }
      }
      double sum[SCR_FIELDS];
      for (int f = 0; f < SCR_FIELDS; f++)
        sum[f] = FX ? (double)isum[f] * a.fxi[f == 0 ? FX_MASS : f <= 3 ? FX_MOM : FX_FORCE]
                    : (double)isum[f];
      grid_node<R>(a, mode, sn, sum, a.gmask[sn], dt, gmass, gv, gvu, fxv, fsum);
    }
  }
  if (mode != GM_USF1) flush_fix_forces(a, fsum);   // GM_USF2 records the step's reaction
}

// The sub-blocks of the source tiles' partial blocks that hold node
// (tile t = (tx, ty, tz), local (lx, ly, lz)): fn(slot, pos) per source,
// pos the node's offset inside a field of the source's block.
template <int DIM, int S1, int TX, int TY, int TZ, class Fn>
__device__ inline void for_node_sources(const KArgs &a, int tx, int ty, int tz, int lx, int ly,
                                        int lz, Fn &&fn) {
  using CX = AxisCls<S1, TX>;
  using CY = AxisCls<S1, TY>;
  using CZ = AxisCls<S1, TZ, DIM == 2>;
  constexpr int EY = TY + S1 - 1, EZ = DIM == 3 ? TZ + S1 - 1 : 1;
  const int nt0 = a.tg.nt[0], nt1 = a.tg.nt[1], nt2 = a.tg.nt[2];
#pragma unroll
  for (int kx = 0; kx < CX::n; kx++) {
    const int ux = tx - CX::toff(kx), rx = lx - CX::lo_l(kx);
    if (rx < 0 || rx >= CX::size(kx) || ux < 0 || ux >= nt0) continue;
#pragma unroll
    for (int ky = 0; ky < CY::n; ky++) {
      const int uy = ty - CY::toff(ky), ry = ly - CY::lo_l(ky);
      if (ry < 0 || ry >= CY::size(ky) || uy < 0 || uy >= nt1) continue;
#pragma unroll
      for (int kz = 0; kz < CZ::n; kz++) {
        const int uz = tz - CZ::toff(kz), rz = lz - CZ::lo_l(kz);
        if (rz < 0 || rz >= CZ::size(kz) || uz < 0 || uz >= nt2) continue;
        const int slot = a.tile_slot[(ux * nt1 + uy) * nt2 + uz];
        if (slot < 0) continue;
        const int sx = CX::size(kx), sy = CY::size(ky), sz = CZ::size(kz);
        fn(slot, CX::lo_b(kx) * EY * EZ + sx * CY::lo_b(ky) * EZ + sx * sy * CZ::lo_b(kz) +
                     (rx * sy + ry) * sz + rz);
      }
    }
  }
}

// Grid temperature of a thermo-mechanical run (FP64), after the grid
// update of the same P2G: the thermal partial blocks summed per owned node,
// then
//   T = sum w m T / m where m > 0 (solid.cpp:2743-2762),
//   Qext = sum w gamma where m > 0 (:2764-2779), Qint = sum grad w . q
//   (:2781-2797), and T_update = T + dt (Qint + Qext) / m, or T where
//   m = 0 (Grid::update_grid_temperature grid.cpp:1354-1362);
// TM_T: T alone (the MUSL re-map, ulmpm.cpp:491-493; USF's first pass).
template <int DIM, int S1, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void k_grid_thermal(KArgs a, int mode) {
  if (frozen(a)) return;
  constexpr int EX = TX + S1 - 1, EY = TY + S1 - 1, EZ = DIM == 3 ? TZ + S1 - 1 : 1;
  constexpr int BN = EX * EY * EZ;
  constexpr int TZZ = DIM == 3 ? TZ : 1;
  constexpr int NO = TX * TY * TZZ;
  const int nt1 = a.tg.nt[1], nt2 = a.tg.nt[2];
  const int total = a.sc->dil_count * NO;
  const double dt_ctx = a.sc->dt;
  const bool full = mode == TM_FULL;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
    const int t = a.dil_list[q / NO], o = q % NO;
    const int tz = t % nt2, ty = (t / nt2) % nt1, tx = t / (nt2 * nt1);
    const int lz = o % TZZ, ly = (o / TZZ) % TY, lx = o / (TZZ * TY);
    const int i = tx * TX + lx, j = ty * TY + ly, k = DIM == 3 ? tz * TZ + lz : 0;
    if (i >= a.n[0] || j >= a.n[1] || k >= a.n[2]) continue;
    const int64_t sn = (int64_t)t * NO + o;   // tile-major storage index
    double s[3] = {0, 0, 0};
    for_node_sources<DIM, S1, TX, TY, TZ>(a, tx, ty, tz, lx, ly, lz, [&](int slot, int pos) {
      const double *src = a.scratch_th + (size_t)slot * 3 * BN + pos;
      s[0] += src[0];
      if (full) {
        s[1] += src[BN];
        s[2] += src[2 * BN];
      }
    });
    const double m = a.gmass[sn];
    const double T = m > 0 ? s[0] / m : 0.0;
    a.gT[sn] = T;
    if (!full) continue;
    const double qext = m > 0 ? s[1] : 0.0;
    const double dt = tile_dt(a, t, dt_ctx);
    a.gTu[sn] = m != 0 ? T + dt * (s[2] + qext) / m : T;
  }
}

// Grid state for download: nodes outside every touched tile hold zeros
// (and their velocity-fix values), as the reference's reset leaves them.
template <int DIM, int TX, int TY, int TZ, int PREC>
__global__ void k_grid_clean(KArgs a) {
  using R = typename Real<PREC>::T;
  R *gmass = gptr<PREC>(a.gmass), *gv = gptr<PREC>(a.gv), *gvu = gptr<PREC>(a.gvu);
  constexpr int TZZ = DIM == 3 ? TZ : 1;
  constexpr int NO = TX * TY * TZZ;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int nt1 = a.tg.nt[1], nt2 = a.tg.nt[2];
  const int S1 = a.S1;
  const int up = S1 == 4 ? 1 : 0;
  for (int64_t sn = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; sn < a.ns; sn += stride) {
    const int t = (int)(sn / NO), o = (int)(sn % NO);
    const int tz = t % nt2, ty = (t / nt2) % nt1, tx = t / (nt2 * nt1);
    const int lz = o % TZZ, ly = (o / TZZ) % TY, lx = o / (TZZ * TY);
    const int i = tx * TX + lx, j = ty * TY + ly, k = DIM == 3 ? tz * TZ + lz : 0;
    if (i >= a.n[0] || j >= a.n[1] || k >= a.n[2]) continue;   // padding of edge tiles
    bool touched = false;
    for (int dx = -1; dx <= up && !touched; dx++)
      for (int dy = -1; dy <= up && !touched; dy++)
        for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? up : 0) && !touched; dz++) {
          const int ux = tx + dx, uy = ty + dy, uz = tz + dz;
          if (ux < 0 || uy < 0 || uz < 0 || ux >= a.tg.nt[0] || uy >= nt1 || uz >= nt2) continue;
          if (a.tile_slot[(ux * nt1 + uy) * nt2 + uz] >= 0) touched = true;
        }
    if (touched) continue;
    double v[3] = {0, 0, 0}, vu[3] = {0, 0, 0};
    const uint32_t mask = a.gmask[sn];
    for (int fi = 0; fi < a.nfv; fi++)
      if (mask & a.fv[fi].bit)
        for (int d = 0; d < 3; d++)
          if (a.fv[fi].set[d]) {
            vu[d] = a.sc->fv_last[fi][d];      // the values of the last grid update
            v[d] = a.sc->fv_last[fi][3 + d];
          }
    gmass[sn] = 0;
    for (int d = 0; d < 3; d++) {
      gv[d * a.ns + sn] = (R)v[d];
      gvu[d * a.ns + sn] = (R)vu[d];
    }
    if (a.thermo) a.gT[sn] = a.gTu[sn] = 0;
  }
}

// ---------------------------------------------------------------------------
// G2P helpers
// ---------------------------------------------------------------------------
__device__ inline void block_max_to(double val, unsigned long long *dst) {
  // val >= 0 (or NaN handled by caller); wave max then LDS then atomic max
  __shared__ double red[4];
  for (int o = 32; o > 0; o >>= 1) {
    double other = __shfl_xor(val, o, 64);
    val = other > val ? other : val;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) red[wave] = val;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = red[w] > m ? red[w] : m;
    if (m > 0) atomicMax(dst, (unsigned long long)__double_as_longlong(m));
  }
}

__device__ inline void block_min_ratio(double val, unsigned long long *dst) {
  __shared__ double red[4];
  for (int o = 32; o > 0; o >>= 1) {
    const double other = __shfl_xor(val, o, 64);
    val = other < val ? other : val;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) red[wave] = val;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = red[w] < m ? red[w] : m;
    if (m < 1.0) atomicMin(dst, (unsigned long long)__double_as_longlong(m));
  }
}

// CFL of a batched context: the block's maximum wave speed (and, TLMPM,
// smallest |Re eig(F)|) of tile t go to the tile's instance; the
// accumulators restart.  Block-uniform (it synchronises the block).
template <bool TL>
__device__ inline void flush_tile_cfl(const KArgs &a, int t, double &cmax, double &rmin) {
  __shared__ unsigned long long sink[2];   // for tiles no instance owns
  const int i = a.tinst[t];
  block_max_to(cmax, i >= 0 ? &a.inst[i].max_c_bits : &sink[0]);
  if (TL) block_min_ratio(rmin, i >= 0 ? &a.inst[i].min_ratio_bits : &sink[1]);
  cmax = 0;
  rmin = 1.0;
  __syncthreads();   // thread 0 has read the reduction slots before they are reused
}

// End of a CFL kernel: the block's reductions into the context's scalars
// (batched contexts have flushed per tile).
template <bool TL>
__device__ inline void finish_cfl(const KArgs &a, double cmax, double rmin) {
  if (a.ninst) return;
  block_max_to(cmax, &a.sc->max_c_bits);
  if (TL) block_min_ratio(rmin, &a.sc->min_ratio_bits);
}

// Load the tile's node block (velocities) into LDS: comps [nc][BN].
template <int DIM, int S1, int TX, int TY, int TZ, int MODE, int PREC>
__device__ inline void load_node_block(const KArgs &a, typename Real<PREC>::T *nb, int ox, int oy,
                                       int oz) {
  using R = typename Real<PREC>::T;
  const R *gv = gptr<PREC>(a.gv), *gvu = gptr<PREC>(a.gvu);
  // MODE 0: v_update and v_update - v (USL / MUSL part A); MODE 1: v (MUSL part B)
  constexpr int EX = TX + S1 - 1, EY = TY + S1 - 1, EZ = DIM == 3 ? TZ + S1 - 1 : 1;
  constexpr int BN = EX * EY * EZ;
  const int off = S1 == 4 ? 1 : 0;
  for (int q = threadIdx.x; q < BN; q += blockDim.x) {
    const int bz = q % EZ, by = (q / EZ) % EY, bx = q / (EZ * EY);
    const int i = ox - off + bx, j = oy - off + by, k = DIM == 3 ? oz - off + bz : 0;
    const bool ok = i >= 0 && j >= 0 && k >= 0 && i < a.n[0] && j < a.n[1] && k < a.n[2];
    const int64_t sn = ok ? node_sidx<TX, TY, DIM == 3 ? TZ : 1>(a, i, j, k) : 0;
    for (int d = 0; d < DIM; d++) {
      if (MODE == 0) {
        const R vu = ok ? gvu[d * a.ns + sn] : (R)0;
        const R v = ok ? gv[d * a.ns + sn] : (R)0;
        nb[d * BN + q] = vu;
        nb[(DIM + d) * BN + q] = vu - v;
      } else {
        nb[d * BN + q] = ok ? gv[d * a.ns + sn] : (R)0;
      }
    }
  }
}

// Load a scalar node field of the grid (FP64, tile-major [ns]: the grid
// temperature) into the tile's LDS block [BN].
template <int DIM, int S1, int TX, int TY, int TZ>
__device__ inline void load_node_scalar(const KArgs &a, const double *g, double *nf, int ox, int oy,
                                        int oz) {
  constexpr int EX = TX + S1 - 1, EY = TY + S1 - 1, EZ = DIM == 3 ? TZ + S1 - 1 : 1;
  constexpr int BN = EX * EY * EZ;
  const int off = S1 == 4 ? 1 : 0;
  for (int q = threadIdx.x; q < BN; q += blockDim.x) {
    const int bz = q % EZ, by = (q / EZ) % EY, bx = q / (EZ * EY);
    const int i = ox - off + bx, j = oy - off + by, k = DIM == 3 ? oz - off + bz : 0;
    const bool ok = i >= 0 && j >= 0 && k >= 0 && i < a.n[0] && j < a.n[1] && k < a.n[2];
    nf[q] = ok ? g[node_sidx<TX, TY, DIM == 3 ? TZ : 1>(a, i, j, k)] : 0.0;
  }
}

// Gather of a scalar node field from its LDS block: s = sum w f and
// g = sum grad w f, contracted one axis at a time as gather_stencil does.
template <int DIM, int S1, int EY, int EZ, class R>
__device__ inline void gather_scalar(const double *nf, const int *cl, const R (*w)[S1],
                                     const R (*dw)[S1], double &s, double *g) {
  s = 0;
  g[0] = g[1] = g[2] = 0;
#pragma unroll
  for (int ia = 0; ia < S1; ia++) {
    double uw = 0, ud1 = 0, ud2 = 0;
#pragma unroll
    for (int ib = 0; ib < S1; ib++) {
      if (DIM == 3) {
        const int row = ((cl[0] + ia) * EY + (cl[1] + ib)) * EZ + cl[2];
        double tw = 0, tz = 0;
#pragma unroll
        for (int ic = 0; ic < S1; ic++) {
          const double f = nf[row + ic];
          tw = fma((double)w[2][ic], f, tw);
          tz = fma((double)dw[2][ic], f, tz);
        }
        uw = fma((double)w[1][ib], tw, uw);
        ud1 = fma((double)dw[1][ib], tw, ud1);
        ud2 = fma((double)w[1][ib], tz, ud2);
      } else {
        const double f = nf[(cl[0] + ia) * EY + (cl[1] + ib)];
        uw = fma((double)w[1][ib], f, uw);
        ud1 = fma((double)dw[1][ib], f, ud1);
      }
    }
    s = fma((double)w[0][ia], uw, s);
    g[0] = fma((double)dw[0][ia], uw, g[0]);
    g[1] = fma((double)w[0][ia], ud1, g[1]);
    if (DIM == 3) g[2] = fma((double)w[0][ia], ud2, g[2]);
  }
}

// Gather over the stencil: vu, acc (MODE 0 / A), L (USL / B).
// The weights are products of per-axis factors, so the sums are contracted
// one axis at a time (last axis first): per node only the multiply-adds of
// the innermost contraction remain, and the weight products are never
// formed per node.
//   vu_d  = sum_abc wx_a wy_b wz_c V_d          acc_d = sum wx wy wz A_d
//   L_d0  = sum dwx wy wz V_d   L_d1 = sum wx dwy wz V_d   L_d2 = sum wx wy dwz V_d
template <int DIM, int S1, int EY, int EZ, int BN, bool WANT_VA, bool WANT_L, bool L_FROM_V,
          class R>
__device__ inline void gather_stencil(const R *nb, const int *cl, const R (*w)[S1],
                                      const R (*dw)[S1], R *vu, R *acc, R *L) {
  (void)L_FROM_V;
#pragma unroll
  for (int ia = 0; ia < S1; ia++) {
    if (WANT_VA)
      for (int d = 0; d < DIM; d++) asm volatile("" : "+v"(vu[d]), "+v"(acc[d]));
    if (WANT_L)
      for (int d = 0; d < 9; d++) asm volatile("" : "+v"(L[d]));
    // contractions over the axes after the first, for this x offset
    R uw[DIM], ua[DIM], ud1[DIM], ud2[DIM];   // sum wy wz V, wy wz A, dwy wz V, wy dwz V
#pragma unroll
    for (int d = 0; d < DIM; d++) uw[d] = ua[d] = ud1[d] = ud2[d] = 0;
#pragma unroll
    for (int ib = 0; ib < S1; ib++) {
      // Opaque barrier: the previous row's accumulations complete here and
      // this row's LDS loads do not move above it (keeps registers bounded).
#pragma unroll
      for (int d = 0; d < DIM; d++) asm volatile("" : "+v"(uw[d]), "+v"(ua[d]), "+v"(ud1[d]), "+v"(ud2[d])::"memory");
      if (DIM == 3) {
        const int row = ((cl[0] + ia) * EY + (cl[1] + ib)) * EZ + cl[2];
        R tw[DIM], ta[DIM], tz[DIM];   // sum_c wz V, wz A, dwz V
#pragma unroll
        for (int d = 0; d < DIM; d++) tw[d] = ta[d] = tz[d] = 0;
#pragma unroll
        for (int ic = 0; ic < S1; ic++) {
#pragma unroll
          for (int d = 0; d < DIM; d++) {
            const R vn = nb[d * BN + row + ic];
            tw[d] = fma(w[2][ic], vn, tw[d]);
            if (WANT_L) tz[d] = fma(dw[2][ic], vn, tz[d]);
            if (WANT_VA) ta[d] = fma(w[2][ic], nb[(DIM + d) * BN + row + ic], ta[d]);
          }
        }
#pragma unroll
        for (int d = 0; d < DIM; d++) {
          uw[d] = fma(w[1][ib], tw[d], uw[d]);
          if (WANT_VA) ua[d] = fma(w[1][ib], ta[d], ua[d]);
          if (WANT_L) {
            ud1[d] = fma(dw[1][ib], tw[d], ud1[d]);
            ud2[d] = fma(w[1][ib], tz[d], ud2[d]);
          }
        }
      } else {
        const int idx = (cl[0] + ia) * EY + (cl[1] + ib);
#pragma unroll
        for (int d = 0; d < DIM; d++) {
          const R vn = nb[d * BN + idx];
          uw[d] = fma(w[1][ib], vn, uw[d]);
          if (WANT_L) ud1[d] = fma(dw[1][ib], vn, ud1[d]);
          if (WANT_VA) ua[d] = fma(w[1][ib], nb[(DIM + d) * BN + idx], ua[d]);
        }
      }
    }
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      if (WANT_VA) {
        vu[d] = fma(w[0][ia], uw[d], vu[d]);
        acc[d] = fma(w[0][ia], ua[d], acc[d]);
      }
      if (WANT_L) {
        L[3 * d + 0] = fma(dw[0][ia], uw[d], L[3 * d + 0]);
        L[3 * d + 1] = fma(w[0][ia], ud1[d], L[3 * d + 1]);
        if (DIM == 3) L[3 * d + 2] = fma(w[0][ia], ud2[d], L[3 * d + 2]);
      }
    }
  }
}

// 1D weights of one axis in FP64 (reference rounding), stored as R.
template <int S1, class R>
__device__ inline void weights_1d_as(double x, double lo, double h, double inv_h, int n, int base,
                                     R *w, R *dw) {
  double wd[S1], dwd[S1];
  weights_1d<S1>(x, lo, h, inv_h, n, base, wd, dwd);
#pragma unroll
  for (int q = 0; q < S1; q++) {
    w[q] = (R)wd[q];
    dw[q] = (R)dwd[q];
  }
}

// Per-axis weights of a G2P gather in the path's type R.  For the APIC
// family (apic) the derivative slot holds Di (x_I - xa) w instead, so the
// gather's gradient contraction yields the APIC rate of deformation
// Di sum_I v_I (x_I - xa)^T w (solid.cpp:1006-1153), xa being the particle
// position of that rate: x (USL, USF), the advanced x (MUSL) or x0 (TLMPM).
template <int S1, class R>
__device__ inline void g2p_weights_1d(const KArgs &a, int d, double x, int base, bool apic,
                                      double xa, R *w, R *dw) {
  double wd[S1], dwd[S1];
  weights_1d<S1>(x, a.lo[d], a.h, a.inv_h, a.n[d], base, wd, dwd);
#pragma unroll
  for (int q = 0; q < S1; q++) {
    w[q] = (R)wd[q];
    // node coordinate as the weights compute it
    dw[q] = apic ? (R)(a.di * (((a.lo[d] + (0 + base + q) * a.h) - xa) * wd[q])) : (R)dwd[q];
  }
}

// F update (solid.cpp:1181-1216, UL): F = (I + dt L) F, J = det F.
__device__ inline double update_F(double dt, const double *L, double *F) {
  double M[9], Fn[9];
  for (int e = 0; e < 9; e++) M[e] = ((e % 4 == 0) ? 1.0 : 0.0) + dt * L[e];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++)
      Fn[3 * r + c] = M[3 * r] * F[c] + M[3 * r + 1] * F[3 + c] + M[3 * r + 2] * F[6 + c];
  for (int e = 0; e < 9; e++) F[e] = Fn[e];
  return det3(F);
}

// The constitutive update in the arithmetic of the path: FP64, or FP32 on
// the mixed path (whose state is stored in FP32), with J and rho in FP64.
// The stress state (sig, eel, ep, epdot, dinit, dmg) is held in the type of
// the path's storage, R.
// th: the thermal state of a thermo-mechanical context (FP64, the generic
// material class), else null.
template <int MC, int PREC, class R = typename Real<PREC>::T>
__device__ inline void stress_update(const DevMat &m, double dt, double h, double J, double rho,
                                     const double *D, R *sig, R *eel, R &ep, R &epdot, R &dinit,
                                     R &dmg, ThermalPt *th = nullptr) {
  if constexpr (PREC == KMPM_PRECISION_MIXED) {
    float Df[6];
    for (int e = 0; e < 6; e++) Df[e] = (float)D[e];
    update_stress_point<MC, float>(m, (float)dt, (float)h, J, rho, Df, sig, eel, ep, epdot, dinit,
                                   dmg);
  } else if constexpr (MC == MC_GENERIC) {
    update_stress_point<MC>(m, dt, h, J, rho, D, sig, eel, ep, epdot, dinit, dmg, th);
  } else {
    update_stress_point<MC>(m, dt, h, J, rho, D, sig, eel, ep, epdot, dinit, dmg);
  }
}

// D = sym(L) (solid.cpp:1239), stress update (solid.cpp:1246-1374) and the
// particle's CFL wave speed (:1376-1385); returns -1 for fully damaged
// particles (skipped by the CFL maximum).
template <int MC, int PREC, class R = typename Real<PREC>::T>
__device__ inline double stress_and_speed(const KArgs &a, double dt, const double *L, double J,
                                          R *sig, R *eel, R &ep, R &epdot, R &dinit, R &dmg,
                                          int mat, const double *v, int pid,
                                          ThermalPt *th = nullptr) {
  const DevMat &m = a.mats[mat];
  const double rho = m.rho0 / J;
  double D[6];
  D[0] = 0.5 * (L[0] + L[0]);
  D[1] = 0.5 * (L[4] + L[4]);
  D[2] = 0.5 * (L[8] + L[8]);
  D[3] = 0.5 * (L[1] + L[3]);
  D[4] = 0.5 * (L[2] + L[6]);
  D[5] = 0.5 * (L[5] + L[7]);
  stress_update<MC, PREC>(m, dt, a.h, J, rho, D, sig, eel, ep, epdot, dinit, dmg, th);
  if (dmg >= 1.0) return -1.0;
  const double vm = fmax(fmax(fabs(v[0]), fabs(v[1])), fabs(v[2]));
  const double c = sqrt(m.cfl_k43g / rho) + vm;
  if (isnan(c)) set_error(a.sc, KMPM_ERR_WAVE_SPEED, pid);
  return c;
}

__device__ inline bool outside_box(const KArgs &a, const double *x) {
  // Domain::inside (domain.cpp:174-182); in 2D z == boxlo[2] == boxhi[2] == 0
  for (int d = 0; d < a.dim; d++)
    if (!(x[d] >= a.lo[d] && x[d] <= a.hi[d])) return true;
  return false;
}

// Where a stress kernel writes a particle: the next state at its sorted
// position, or (USF, whose stress update precedes the G2P that carries the
// state over) the current state's own slot.
struct OutSlot {
  double *d;
  float *f;
  int i;
};
__device__ inline OutSlot out_slot(const KArgs &a, bool in_place, int slot, int j) {
  return in_place ? OutSlot{a.pa, a.fa, slot} : OutSlot{a.pb, a.fb, j};
}
#define STO(fld, val) stp<PREC>(a, o.d, o.f, (fld), o.i, (val))

// The heat source gamma and the heat flux q = -vol kappa / cp grad T_n of a
// thermo-mechanical particle (solid.cpp:1343-1351, 2810-2839), vol being
// the current volume (ULMPM) or vol0 (TLMPM).
template <int PREC>
__device__ inline void store_heat(const KArgs &a, const OutSlot &o, int mat, double gamma,
                                  double vol, const double *gradT) {
  const DevMat &m = a.mats[mat];
  const double sc = vol * m.invcp * m.kappa;
  STO(GAMMA, gamma);
  for (int d = 0; d < 3; d++) STO(QX + d, -gradT[d] * sc);
}

// ---------------------------------------------------------------------------
// USL fused G2P: L(v_update), G2P, advance, F, stress, CFL, next key
// ---------------------------------------------------------------------------
// STAGE: a learned constitutive update (the surrogate's networks or the cut
// network) follows in k_stress_learned; this kernel leaves the stress state
// alone and stages its inputs (the rate of deformation, the EOS pressure and
// the wave speed) in a.cut_stage.
// XF: the APIC family (the rate of deformation from the affine moments, L
// kept for the next step's P2G; ASFLIP's position update when a.asflip).
template <int DIM, int S1, int TX, int TY, int TZ, int MC, int PREC, bool STAGE, bool XF = false>
// Mixed linear and JC classes fit four waves per SIMD without spilling (the
// fluid and generic classes would spill at that bound); the plain mixed
// kernel takes 0.44 ms at four waves, 0.54 at three and 0.57 at five on the
// 4.27M bar.  The staging kernel is bounded for three waves: it still fits
// four (126 VGPRs), and the schedule made for three takes 0.31 ms against
// 0.35 for the one made for four.
__global__ __launch_bounds__(256, STAGE ? 3
                                        : (PREC == KMPM_PRECISION_MIXED && !XF &&
                                                   (MC == MC_LINEAR || MC == MC_JC_SHOCK)
                                               ? 4
                                               : 1)) void k_g2p_usl(KArgs a) {
  if (frozen(a)) {
    frozen_copy(a);
    return;
  }
  using R = typename Real<PREC>::T;
  constexpr int EX = TX + S1 - 1, EY = TY + S1 - 1, EZ = DIM == 3 ? TZ + S1 - 1 : 1;
  constexpr int BN = EX * EY * EZ;
  __shared__ R nb[2 * DIM * BN];
  // thermo-mechanical runs (FP64, generic class): T_update's node block
  constexpr bool THK = MC == MC_GENERIC && PREC == KMPM_PRECISION_FP64 && !STAGE && !XF;
  __shared__ double nT[THK ? BN : 1];
  const bool th = THK && a.thermo;
  const int nt1 = a.tg.nt[1], nt2 = a.tg.nt[2];
  const int off = S1 == 4 ? 1 : 0;
  const int ntile = a.sc->tile_count;
  const double dt_ctx = a.sc->dt;
  const double alpha = a.flip;
  double cmax = 0, rmin = 1.0;
  for (int li = blockIdx.x; li < ntile; li += gridDim.x) {
    const int t = a.tile_list[li];
    const double dt = tile_dt(a, t, dt_ctx);
    const double inv_dt = 1.0 / dt;
    const int tz = DIM == 3 ? t % nt2 : 0, ty = (t / nt2) % nt1, tx = t / (nt2 * nt1);
    const int org[3] = {tx * TX, ty * TY, DIM == 3 ? tz * TZ : 0};
    const int p0 = a.tile_p0[t], p1 = a.tile_p0[t + 1];
    __syncthreads();
    load_node_block<DIM, S1, TX, TY, TZ, 0, PREC>(a, nb, org[0], org[1], org[2]);
    if (th) load_node_scalar<DIM, S1, TX, TY, TZ>(a, a.gTu, nT, org[0], org[1], org[2]);
    __syncthreads();
    auto particle = [&](const int j) {
      const int slot = a.order[j];
      double x[3], v[3];
      for (int d = 0; d < 3; d++) {
        x[d] = LDA(PX + d, slot);
        v[d] = LDA(VX + d, slot);
      }
      R w[3][S1], dw[3][S1];
      int cl[3] = {0, 0, 0};
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const int c = cell_of_clamped(x[d], a.lo[d], a.inv_h, a.n[d]);
        cl[d] = tile_local(c, org[d], d == 0 ? TX : d == 1 ? TY : TZ, a.sc);
        if constexpr (XF)
          g2p_weights_1d<S1>(a, d, x[d], c - off, true, x[d], w[d], dw[d]);
        else
          weights_1d_as<S1>(x[d], a.lo[d], a.h, a.inv_h, a.n[d], c - off, w[d], dw[d]);
      }
      double vu[3], acc[3], L[9];
      {
        R vur[3] = {0, 0, 0}, accr[3] = {0, 0, 0}, Lr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        gather_stencil<DIM, S1, EY, EZ, BN, true, true, false>(nb, cl, w, dw, vur, accr, Lr);
        for (int d = 0; d < 3; d++) {
          vu[d] = vur[d];
          acc[d] = accr[d];
        }
        for (int e = 0; e < 9; e++) L[e] = Lr[e];
      }
      if constexpr (XF)
        for (int e = 0; e < 9; e++) STB(C00 + e, j, L[e]);
      // the particle's temperature from T_update (solid.cpp:2799-2808) and
      // the gradient the heat flux takes (:2810-2839)
      double Tp = 0, gradT[3] = {0, 0, 0};
      if (th) {
        gather_scalar<DIM, S1, EY, EZ>(nT, cl, w, dw, Tp, gradT);
        STB(TEMP, j, Tp);
      }
      const int pid = IFLD(a.ia, PID, slot);
      for (int d = 0; d < 3; d++) acc[d] *= inv_dt;
      if (XF && a.asflip) {
        // ASFLIP (solid.cpp:637-685, 791-796): the box check sees x before
        // the update, which then moves x by the blended velocity
        if (outside_box(a, x)) set_error(a.sc, KMPM_ERR_PARTICLE_LEFT_BOX, pid);
        for (int d = 0; d < 3; d++) v[d] = (1 - alpha) * vu[d] + alpha * (v[d] + dt * acc[d]);
        for (int d = 0; d < DIM; d++) x[d] += dt * v[d];
      } else {
        for (int d = 0; d < DIM; d++) x[d] += dt * vu[d];
        if (outside_box(a, x)) set_error(a.sc, KMPM_ERR_PARTICLE_LEFT_BOX, pid);
        for (int d = 0; d < 3; d++) v[d] = (1 - alpha) * vu[d] + alpha * (v[d] + dt * acc[d]);
      }
      // State is written in sorted order as soon as each part is final, so
      // it does not stay live across the constitutive update.
      const int mat = IFLD(a.ia, MATI, slot);
      {
        for (int d = 0; d < 3; d++) {
          STB(PX + d, j, x[d]);
          STB(VX + d, j, v[d]);
        }
        a.key[j] = particle_key(a, x[0], x[1], x[2]);
        IFLD(a.ib, PID, j) = pid;
        IFLD(a.ib, MATI, j) = mat;
      }
      const double vol0 = LDA(VOL0, slot);
      {
        STB(MASS, j, LDA(MASS, slot));
        STB(VOL0, j, vol0);
      }
      double J;
      {
        double F[9];
        for (int e = 0; e < 9; e++) F[e] = LDA(F00 + e, slot);
        J = update_F(dt, L, F);
        for (int e = 0; e < 9; e++) STB(F00 + e, j, F[e]);
      }
      STB(VOL, j, J * vol0);
      R dmg = (R)LDA(DMG, slot);
      if (J <= 0.0 && dmg < 1.0) set_error(a.sc, KMPM_ERR_JACOBIAN, pid);
      if constexpr (STAGE) {
        // the cut's inputs: D = sym(L) (solid.cpp:1239), the EOS pressure
        // (eos_shock.cpp:99-133, FP64 from J as on the plain mixed path) and
        // the wave speed before the damage test (solid.cpp:1376-1385)
        const DevMat &m = a.mats[mat];
        const double rho = m.rho0 / J;
        double D[6];
        D[0] = L[0];
        D[1] = L[4];
        D[2] = L[8];
        D[3] = 0.5 * (L[1] + L[3]);
        D[4] = 0.5 * (L[2] + L[6]);
        D[5] = 0.5 * (L[5] + L[7]);
        const double pH = shock_pressure<double>(m, J, rho, (double)dmg, D, a.h);
        const double vm = fmax(fmax(fabs(v[0]), fabs(v[1])), fabs(v[2]));
        const double c = sqrt(m.cfl_k43g / rho) + vm;
        float *st = a.cut_stage;
        const size_t cs = (size_t)a.cap;
        for (int e = 0; e < 6; e++) st[e * cs + j] = (float)D[e];
        st[6 * cs + j] = (float)pH;
        st[7 * cs + j] = (float)c;
        return;
      }
      R sig[6], eel[6];
      for (int e = 0; e < 6; e++) {
        sig[e] = (R)LDA(SXX + e, slot);
        eel[e] = (R)LDA(EXX + e, slot);
      }
      R ep = (R)LDA(EP, slot), epdot = (R)LDA(EPDOT, slot);
      R dinit = (R)LDA(DINIT, slot);
      ThermalPt tp{Tp, 1, J * vol0, th ? LDA(GAMMA, slot) : 0.0};
      const double c = stress_and_speed<MC, PREC>(a, dt, L, J, sig, eel, ep, epdot, dinit, dmg,
                                                  mat, v, pid, th ? &tp : nullptr);
      if (c > cmax) cmax = c;
      if (th) store_heat<PREC>(a, OutSlot{a.pb, a.fb, j}, mat, tp.gamma, J * vol0, gradT);
      for (int e = 0; e < 6; e++) {
        STB(SXX + e, j, sig[e]);
        STB(EXX + e, j, eel[e]);
      }
      STB(EP, j, ep);
      STB(EPDOT, j, epdot);
      STB(DINIT, j, dinit);
      STB(DMG, j, dmg);
    };
    for (int j = p0 + (int)threadIdx.x; j < p1; j += blockDim.x) particle(j);
    if (a.ninst) flush_tile_cfl<false>(a, t, cmax, rmin);
  }
  finish_cfl<false>(a, cmax, rmin);
}

// ---------------------------------------------------------------------------
// TLMPM particle update after the gather (FP64): solid.cpp:1155-1244 (TL
// branch) and 1246-1437.  F += dt Fdot, F^-1, J, vol, rho, R = polar(F),
// D = sym(R^T (L + L^T) R) / 2 with L = Fdot F^-1, the stress update,
// vol0 PK1, the wave speed and the smallest |Re eig(F)| (both -1 for fully
// damaged particles, which the CFL skips).  Writes slot j of the next state.
// ---------------------------------------------------------------------------


template <int MC, int PREC>
__device__ inline void tl_update(const KArgs &a, double dt, const double *Fdot, int slot,
                                 const OutSlot &o, int pid, int mat, const double *v,
                                 double &cmax, double &rmin, ThermalPt *th = nullptr,
                                 const double *jv = nullptr) {
  double F[9], Finv[9], R[9];
  for (int e = 0; e < 9; e++) F[e] = LDA(F00 + e, slot) + dt * Fdot[e];   // :1181-1182
  m3_inverse(F, Finv);                                                   // :1186
  // jv: J and vol given (CPDI-Q4: the corners' polygon, :1188-1200)
  const double J = jv ? jv[0] : m3_det(F);                               // :1203
  const double vol0 = LDA(VOL0, slot);
  using RS = typename Real<PREC>::T;   // the stored state's type
  RS dmg = (RS)LDA(DMG, slot);
  if (J <= 0.0 && dmg < 1.0) set_error(a.sc, KMPM_ERR_JACOBIAN, pid);    // :1208-1215
  if (!polar_rotation(F, R)) set_error(a.sc, KMPM_ERR_JACOBIAN, pid);    // :1222-1232
  double D[6];
  tl_rate(Fdot, Finv, R, D);                                             // :1225-1226
  const DevMat &m = a.mats[mat];
  const double rho = m.rho0 / J;
  RS sig[6], eel[6];
  for (int e = 0; e < 6; e++) {
    sig[e] = (RS)LDA(SXX + e, slot);
    eel[e] = (RS)LDA(EXX + e, slot);
  }
  RS ep = (RS)LDA(EP, slot), epdot = (RS)LDA(EPDOT, slot), dinit = (RS)LDA(DINIT, slot);
  stress_update<MC, PREC>(m, dt, a.h, J, rho, D, sig, eel, ep, epdot, dinit, dmg, th);
  double S[9] = {sig[0], sig[3], sig[4], sig[3], sig[1], sig[5], sig[4], sig[5], sig[2]};
  double P[9];
  pk1(vol0, J, R, S, Finv, P);
  for (int e = 0; e < 9; e++) {
    STO(F00 + e, F[e]);
    STO(P00 + e, P[e]);
  }
  STO(VOL0, vol0);
  STO(VOL, jv ? jv[1] : J * vol0);
  for (int e = 0; e < 6; e++) {
    STO(SXX + e, sig[e]);
    STO(EXX + e, eel[e]);
  }
  STO(EP, ep);
  STO(EPDOT, epdot);
  STO(DINIT, dinit);
  STO(DMG, dmg);
  if (dmg >= 1.0) return;                                                // :1378-1379
  const double vm = fmax(fmax(fabs(v[0]), fabs(v[1])), fabs(v[2]));
  const double c = sqrt(m.cfl_k43g / rho) + vm;
  if (isnan(c)) set_error(a.sc, KMPM_ERR_WAVE_SPEED, pid);
  if (c > cmax) cmax = c;
  double re[3];
  eig_real_parts(F, re);                                                 // :1400-1408
  for (int k = 0; k < 3; k++) rmin = fabs(re[k]) < rmin ? fabs(re[k]) : rmin;
}

// TLMPM fused USL G2P: weights of x0 on the solid's grid, Fdot(v_update),
// G2P, advance (no box check), then tl_update.
template <int DIM, int S1, int TX, int TY, int TZ, int MC, int PREC>
__global__ __launch_bounds__(256) void k_g2p_usl_tl(KArgs a) {
  if (frozen(a)) {
    frozen_copy(a);
    return;
  }
  using R = typename Real<PREC>::T;
  constexpr int EX = TX + S1 - 1, EY = TY + S1 - 1, EZ = DIM == 3 ? TZ + S1 - 1 : 1;
  constexpr int BN = EX * EY * EZ;
  __shared__ R nb[2 * DIM * BN];
  constexpr bool THK = MC == MC_GENERIC && PREC == KMPM_PRECISION_FP64;
  __shared__ double nT[THK ? BN : 1];   // T_update (thermo-mechanical runs)
  const bool th = THK && a.thermo;
  const int nt1 = a.tg.nt[1], nt2 = a.tg.nt[2];
  const int off = S1 == 4 ? 1 : 0;
  const int ntile = a.sc->tile_count;
  const double dt_ctx = a.sc->dt;
  const double alpha = a.flip;
  double cmax = 0, rmin = 1.0;
  for (int li = blockIdx.x; li < ntile; li += gridDim.x) {
    const int t = a.tile_list[li];
    const double dt = tile_dt(a, t, dt_ctx);
    const double inv_dt = 1.0 / dt;
    const int tz = DIM == 3 ? t % nt2 : 0, ty = (t / nt2) % nt1, tx = t / (nt2 * nt1);
    const int org[3] = {tx * TX, ty * TY, DIM == 3 ? tz * TZ : 0};
    const int p0 = a.tile_p0[t], p1 = a.tile_p0[t + 1];
    __syncthreads();
    load_node_block<DIM, S1, TX, TY, TZ, 0, PREC>(a, nb, org[0], org[1], org[2]);
    if (th) load_node_scalar<DIM, S1, TX, TY, TZ>(a, a.gTu, nT, org[0], org[1], org[2]);
    __syncthreads();
    for (int j = p0 + (int)threadIdx.x; j < p1; j += blockDim.x) {
      const int slot = a.order[j];
      double x0[3], x[3], v[3];
      for (int d = 0; d < 3; d++) {
        x0[d] = LDA(X0X + d, slot);
        x[d] = LDA(PX + d, slot);
        v[d] = LDA(VX + d, slot);
      }
      R w[3][S1], dw[3][S1];
      int cl[3] = {0, 0, 0};
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const int c = cell_of_clamped(x0[d], a.lo[d], a.inv_h, a.n[d]);
        cl[d] = tile_local(c, org[d], d == 0 ? TX : d == 1 ? TY : TZ, a.sc);
        g2p_weights_1d<S1>(a, d, x0[d], c - off, a.xf != 0, x0[d], w[d], dw[d]);
      }
      double vu[3], acc[3], Fdot[9];
      {
        R vur[3] = {0, 0, 0}, accr[3] = {0, 0, 0}, Fr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        gather_stencil<DIM, S1, EY, EZ, BN, true, true, false>(nb, cl, w, dw, vur, accr, Fr);
        for (int d = 0; d < 3; d++) {
          vu[d] = vur[d];
          acc[d] = accr[d];
        }
        for (int e = 0; e < 9; e++) Fdot[e] = Fr[e];
      }
      if (a.xf)   // APIC: Fdot is the next step's affine matrix
        for (int e = 0; e < 9; e++) STB(C00 + e, j, Fdot[e]);
      const int pid = IFLD(a.ia, PID, slot);
      const int mat = IFLD(a.ia, MATI, slot);
      for (int d = 0; d < 3; d++) acc[d] *= inv_dt;
      for (int d = 0; d < DIM; d++) x[d] += dt * vu[d];                 // solid.cpp:616
      for (int d = 0; d < 3; d++) v[d] = (1 - alpha) * vu[d] + alpha * (v[d] + dt * acc[d]);
      for (int d = 0; d < 3; d++) {
        STB(PX + d, j, x[d]);
        STB(VX + d, j, v[d]);
        STB(X0X + d, j, x0[d]);
      }
      a.key[j] = particle_key(a, x0[0], x0[1], x0[2]);
      IFLD(a.ib, PID, j) = pid;
      IFLD(a.ib, MATI, j) = mat;
      STB(MASS, j, LDA(MASS, slot));
      double Tp = 0, gradT[3] = {0, 0, 0};
      if (th) {   // the particle's temperature and its gradient, from T_update
        gather_scalar<DIM, S1, EY, EZ>(nT, cl, w, dw, Tp, gradT);
        STB(TEMP, j, Tp);
      }
      const double vol0 = LDA(VOL0, slot);
      ThermalPt tp{Tp, 1, vol0, th ? LDA(GAMMA, slot) : 0.0};
      const OutSlot o{a.pb, a.fb, j};
      tl_update<MC, PREC>(a, dt, Fdot, slot, o, pid, mat, v, cmax, rmin, th ? &tp : nullptr);
      if (th) store_heat<PREC>(a, o, mat, tp.gamma, vol0, gradT);
    }
    if (a.ninst) flush_tile_cfl<true>(a, t, cmax, rmin);
  }
  finish_cfl<true>(a, cmax, rmin);
}

// ---------------------------------------------------------------------------
// MUSL part A: G2P + advance (musl.cpp:61-67); keeps start-of-step x
// ---------------------------------------------------------------------------
// TL: weights of x0 on the solid's grid, no box check, x0 carried along.
// USF (a.scheme): the G2P and advance of usf.cpp:69-75, after the stress
// kernel updated the current state in place; every other field is carried
// into the next state here.  ASFLIP (a.asflip): the box check sees x before
// the update, which moves x by the blended velocity (solid.cpp:637-685,
// 791-796).
template <int DIM, int S1, int TX, int TY, int TZ, int PREC, bool TL>
__global__ __launch_bounds__(256) void k_g2p_musl_a(KArgs a) {
  if (frozen(a)) return;
  using R = typename Real<PREC>::T;
  constexpr int EX = TX + S1 - 1, EY = TY + S1 - 1, EZ = DIM == 3 ? TZ + S1 - 1 : 1;
  constexpr int BN = EX * EY * EZ;
  __shared__ R nb[2 * DIM * BN];
  constexpr bool THK = PREC == KMPM_PRECISION_FP64;
  __shared__ double nT[THK ? BN : 1];   // T_update (thermo-mechanical runs)
  const bool th = THK && a.thermo;
  const int nt1 = a.tg.nt[1], nt2 = a.tg.nt[2];
  const int off = S1 == 4 ? 1 : 0;
  const int ntile = a.sc->tile_count;
  const double dt_ctx = a.sc->dt;
  const double alpha = a.flip;
  for (int li = blockIdx.x; li < ntile; li += gridDim.x) {
    const int t = a.tile_list[li];
    const double dt = tile_dt(a, t, dt_ctx);
    const double inv_dt = 1.0 / dt;
    const int tz = DIM == 3 ? t % nt2 : 0, ty = (t / nt2) % nt1, tx = t / (nt2 * nt1);
    const int org[3] = {tx * TX, ty * TY, DIM == 3 ? tz * TZ : 0};
    const int p0 = a.tile_p0[t], p1 = a.tile_p0[t + 1];
    __syncthreads();
    load_node_block<DIM, S1, TX, TY, TZ, 0, PREC>(a, nb, org[0], org[1], org[2]);
    if (th) load_node_scalar<DIM, S1, TX, TY, TZ>(a, a.gTu, nT, org[0], org[1], org[2]);
    __syncthreads();
    for (int j = p0 + (int)threadIdx.x; j < p1; j += blockDim.x) {
      const int slot = a.order[j];
      double x[3], v[3];
      for (int d = 0; d < 3; d++) {
        x[d] = LDA(PX + d, slot);
        v[d] = LDA(VX + d, slot);
      }
      double xw[3];
      for (int d = 0; d < 3; d++) xw[d] = TL ? LDA(X0X + d, slot) : x[d];
      R w[3][S1], dw[3][S1];
      int cl[3] = {0, 0, 0};
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const int c = cell_of_clamped(xw[d], a.lo[d], a.inv_h, a.n[d]);
        cl[d] = tile_local(c, org[d], d == 0 ? TX : d == 1 ? TY : TZ, a.sc);
        weights_1d_as<S1>(xw[d], a.lo[d], a.h, a.inv_h, a.n[d], c - off, w[d], dw[d]);
      }
      double vu[3], acc[3];
      {
        R vur[3] = {0, 0, 0}, accr[3] = {0, 0, 0};
        gather_stencil<DIM, S1, EY, EZ, BN, true, false, false>(nb, cl, w, dw, vur, accr,
                                                                (R *)nullptr);
        for (int d = 0; d < 3; d++) {
          vu[d] = vur[d];
          acc[d] = accr[d];
        }
      }
      const int pid = IFLD(a.ia, PID, slot);
      if (TL) {
        for (int d = 0; d < 3; d++) STB(X0X + d, j, xw[d]);
      } else {
        for (int d = 0; d < 3; d++) a.xo[(size_t)d * a.cap + j] = x[d];
      }
      for (int d = 0; d < 3; d++) acc[d] *= inv_dt;
      if (!TL && a.asflip) {
        if (outside_box(a, x)) set_error(a.sc, KMPM_ERR_PARTICLE_LEFT_BOX, pid);
        for (int d = 0; d < 3; d++) v[d] = (1 - alpha) * vu[d] + alpha * (v[d] + dt * acc[d]);
        for (int d = 0; d < DIM; d++) x[d] += dt * v[d];
      } else {
        for (int d = 0; d < DIM; d++) x[d] += dt * vu[d];
        if (!TL && outside_box(a, x)) set_error(a.sc, KMPM_ERR_PARTICLE_LEFT_BOX, pid);
        for (int d = 0; d < 3; d++) v[d] = (1 - alpha) * vu[d] + alpha * (v[d] + dt * acc[d]);
      }
      for (int d = 0; d < 3; d++) {
        STB(PX + d, j, x[d]);
        STB(VX + d, j, v[d]);
      }
      STB(MASS, j, LDA(MASS, slot));
      if (a.scheme == KMPM_SCHEME_USF) {
        const int nfld = state_nd(false, a.lay);
        for (int f = F00; f < nfld; f++)
          if (f != MASS && !(TL && f >= X0X && f <= X0Z)) STB(f, j, LDA(f, slot));
        IFLD(a.ib, PID, j) = pid;
        IFLD(a.ib, MATI, j) = IFLD(a.ia, MATI, slot);
      }
      if (th) {   // update_particle_temperature (solid.cpp:2799-2808)
        double Tp, gradT[3];
        gather_scalar<DIM, S1, EY, EZ>(nT, cl, w, dw, Tp, gradT);
        STB(TEMP, j, Tp);
      }
      // next step's cell: of the advanced x (ULMPM) or of x0 (TLMPM)
      const double *xk = TL ? xw : x;
      a.key[j] = particle_key(a, xk[0], xk[1], xk[2]);
    }
  }
}

// ---------------------------------------------------------------------------
// MUSL part B: L from re-mapped v (old weights), F, stress, CFL (musl.cpp:75-77)
// ---------------------------------------------------------------------------
// USF (a.scheme): the rate of deformation from the first P2G's grid velocity
// with the current positions, F and the stress (usf.cpp:58-60), written in
// place: the USF forces then come from the updated stress and the G2P
// kernel carries the state over.  APIC (a.xf): L from the affine moments
// about the advanced x (MUSL) or the current x (USF), kept as the next
// step's affine matrix.
template <int DIM, int S1, int TX, int TY, int TZ, int MC, int PREC>
__global__ __launch_bounds__(256) void k_stress_musl_b(KArgs a) {
  if (frozen(a)) {
    frozen_copy(a);
    return;
  }
  using R = typename Real<PREC>::T;
  constexpr int EX = TX + S1 - 1, EY = TY + S1 - 1, EZ = DIM == 3 ? TZ + S1 - 1 : 1;
  constexpr int BN = EX * EY * EZ;
  __shared__ R nb[DIM * BN];
  // thermo-mechanical runs (FP64, generic class): the re-mapped grid
  // temperature (MUSL) or USF's first (ulmpm.cpp:515-521, doublemapping)
  constexpr bool THK = MC == MC_GENERIC && PREC == KMPM_PRECISION_FP64;
  __shared__ double nT[THK ? BN : 1];
  const bool th = THK && a.thermo;
  const int nt1 = a.tg.nt[1], nt2 = a.tg.nt[2];
  const int off = S1 == 4 ? 1 : 0;
  const int ntile = a.sc->tile_count;
  const double dt_ctx = a.sc->dt;
  double cmax = 0, rmin = 1.0;
  for (int li = blockIdx.x; li < ntile; li += gridDim.x) {
    const int t = a.tile_list[li];
    const double dt = tile_dt(a, t, dt_ctx);
    const int tz = DIM == 3 ? t % nt2 : 0, ty = (t / nt2) % nt1, tx = t / (nt2 * nt1);
    const int org[3] = {tx * TX, ty * TY, DIM == 3 ? tz * TZ : 0};
    const int p0 = a.tile_p0[t], p1 = a.tile_p0[t + 1];
    __syncthreads();
    load_node_block<DIM, S1, TX, TY, TZ, 1, PREC>(a, nb, org[0], org[1], org[2]);
    if (th) load_node_scalar<DIM, S1, TX, TY, TZ>(a, a.gT, nT, org[0], org[1], org[2]);
    __syncthreads();
    const bool usf = a.scheme == KMPM_SCHEME_USF;
    for (int j = p0 + (int)threadIdx.x; j < p1; j += blockDim.x) {
      const int slot = a.order[j];
      double x[3], xa[3];
      for (int d = 0; d < 3; d++) {
        x[d] = usf ? LDA(PX + d, slot) : a.xo[(size_t)d * a.cap + j];
        xa[d] = usf || !a.xf ? x[d] : LDB(PX + d, j);
      }
      R w[3][S1], dw[3][S1];
      int cl[3] = {0, 0, 0};
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const int c = cell_of_clamped(x[d], a.lo[d], a.inv_h, a.n[d]);
        cl[d] = tile_local(c, org[d], d == 0 ? TX : d == 1 ? TY : TZ, a.sc);
        g2p_weights_1d<S1>(a, d, x[d], c - off, a.xf != 0, xa[d], w[d], dw[d]);
      }
      double L[9];
      {
        R Lr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        gather_stencil<DIM, S1, EY, EZ, BN, false, true, true>(nb, cl, w, dw, (R *)nullptr,
                                                               (R *)nullptr, Lr);
        for (int e = 0; e < 9; e++) L[e] = Lr[e];
      }
      const OutSlot o = out_slot(a, usf, slot, j);
      if (a.xf)
        for (int e = 0; e < 9; e++) STO(C00 + e, L[e]);
      const int pid = IFLD(a.ia, PID, slot);
      double v[3];
      for (int d = 0; d < 3; d++) v[d] = usf ? LDA(VX + d, slot) : LDB(VX + d, j);
      const int mat = IFLD(a.ia, MATI, slot);
      if (!usf) {
        IFLD(a.ib, PID, j) = pid;
        IFLD(a.ib, MATI, j) = mat;
      }
      const double vol0 = LDA(VOL0, slot);
      STO(VOL0, vol0);
      double J;
      {
        double F[9];
        for (int e = 0; e < 9; e++) F[e] = LDA(F00 + e, slot);
        J = update_F(dt, L, F);
        for (int e = 0; e < 9; e++) STO(F00 + e, F[e]);
      }
      STO(VOL, J * vol0);
      R dmg = (R)LDA(DMG, slot);
      if (J <= 0.0 && dmg < 1.0) set_error(a.sc, KMPM_ERR_JACOBIAN, pid);
      R sig[6], eel[6];
      for (int e = 0; e < 6; e++) {
        sig[e] = (R)LDA(SXX + e, slot);
        eel[e] = (R)LDA(EXX + e, slot);
      }
      R ep = (R)LDA(EP, slot), epdot = (R)LDA(EPDOT, slot);
      R dinit = (R)LDA(DINIT, slot);
      // the temperature: MUSL's updated one, USF's of the step's start
      double Tn, gradT[3] = {0, 0, 0};
      if (th) gather_scalar<DIM, S1, EY, EZ>(nT, cl, w, dw, Tn, gradT);
      ThermalPt tp{th ? (usf ? LDA(TEMP, slot) : LDB(TEMP, j)) : 0.0, 1, J * vol0,
                   th ? LDA(GAMMA, slot) : 0.0};
      const double c = stress_and_speed<MC, PREC>(a, dt, L, J, sig, eel, ep, epdot, dinit, dmg,
                                                  mat, v, pid, th ? &tp : nullptr);
      if (c > cmax) cmax = c;
      if (th) store_heat<PREC>(a, o, mat, tp.gamma, J * vol0, gradT);
      for (int e = 0; e < 6; e++) {
        STO(SXX + e, sig[e]);
        STO(EXX + e, eel[e]);
      }
      STO(EP, ep);
      STO(EPDOT, epdot);
      STO(DINIT, dinit);
      STO(DMG, dmg);
    }
    if (a.ninst) flush_tile_cfl<false>(a, t, cmax, rmin);
  }
  finish_cfl<false>(a, cmax, rmin);
}

// TLMPM MUSL part B: Fdot from the re-mapped v with the x0 weights
// (compute_rate_deformation_gradient_TL(true), musl.cpp:75-77), then
// tl_update.
template <int DIM, int S1, int TX, int TY, int TZ, int MC, int PREC>
__global__ __launch_bounds__(256) void k_stress_musl_b_tl(KArgs a) {
  if (frozen(a)) {
    frozen_copy(a);
    return;
  }
  using R = typename Real<PREC>::T;
  constexpr int EX = TX + S1 - 1, EY = TY + S1 - 1, EZ = DIM == 3 ? TZ + S1 - 1 : 1;
  constexpr int BN = EX * EY * EZ;
  __shared__ R nb[DIM * BN];
  constexpr bool THK = MC == MC_GENERIC && PREC == KMPM_PRECISION_FP64;
  __shared__ double nT[THK ? BN : 1];   // the grid temperature (thermo-mechanical runs)
  const bool th = THK && a.thermo;
  const int nt1 = a.tg.nt[1], nt2 = a.tg.nt[2];
  const int off = S1 == 4 ? 1 : 0;
  const int ntile = a.sc->tile_count;
  const double dt_ctx = a.sc->dt;
  double cmax = 0, rmin = 1.0;
  for (int li = blockIdx.x; li < ntile; li += gridDim.x) {
    const int t = a.tile_list[li];
    const double dt = tile_dt(a, t, dt_ctx);
    const int tz = DIM == 3 ? t % nt2 : 0, ty = (t / nt2) % nt1, tx = t / (nt2 * nt1);
    const int org[3] = {tx * TX, ty * TY, DIM == 3 ? tz * TZ : 0};
    const int p0 = a.tile_p0[t], p1 = a.tile_p0[t + 1];
    __syncthreads();
    load_node_block<DIM, S1, TX, TY, TZ, 1, PREC>(a, nb, org[0], org[1], org[2]);
    if (th) load_node_scalar<DIM, S1, TX, TY, TZ>(a, a.gT, nT, org[0], org[1], org[2]);
    __syncthreads();
    const bool usf = a.scheme == KMPM_SCHEME_USF;
    for (int j = p0 + (int)threadIdx.x; j < p1; j += blockDim.x) {
      const int slot = a.order[j];
      double x0[3];
      for (int d = 0; d < 3; d++) x0[d] = LDA(X0X + d, slot);
      R w[3][S1], dw[3][S1];
      int cl[3] = {0, 0, 0};
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const int c = cell_of_clamped(x0[d], a.lo[d], a.inv_h, a.n[d]);
        cl[d] = tile_local(c, org[d], d == 0 ? TX : d == 1 ? TY : TZ, a.sc);
        g2p_weights_1d<S1>(a, d, x0[d], c - off, a.xf != 0, x0[d], w[d], dw[d]);
      }
      double Fdot[9];
      {
        R Fr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        gather_stencil<DIM, S1, EY, EZ, BN, false, true, true>(nb, cl, w, dw, (R *)nullptr,
                                                               (R *)nullptr, Fr);
        for (int e = 0; e < 9; e++) Fdot[e] = Fr[e];
      }
      const int pid = IFLD(a.ia, PID, slot);
      const int mat = IFLD(a.ia, MATI, slot);
      if (!usf) {
        IFLD(a.ib, PID, j) = pid;
        IFLD(a.ib, MATI, j) = mat;
      }
      double v[3];
      for (int d = 0; d < 3; d++) v[d] = usf ? LDA(VX + d, slot) : LDB(VX + d, j);
      const OutSlot o = out_slot(a, usf, slot, j);
      if (a.xf)
        for (int e = 0; e < 9; e++) STO(C00 + e, Fdot[e]);
      double Tn, gradT[3] = {0, 0, 0};
      if (th) gather_scalar<DIM, S1, EY, EZ>(nT, cl, w, dw, Tn, gradT);
      const double vol0 = LDA(VOL0, slot);
      ThermalPt tp{th ? (usf ? LDA(TEMP, slot) : LDB(TEMP, j)) : 0.0, 1, vol0,
                   th ? LDA(GAMMA, slot) : 0.0};
      tl_update<MC, PREC>(a, dt, Fdot, slot, o, pid, mat, v, cmax, rmin, th ? &tp : nullptr);
      if (th) store_heat<PREC>(a, o, mat, tp.gamma, vol0, gradT);
    }
    if (a.ninst) flush_tile_cfl<true>(a, t, cmax, rmin);
  }
  finish_cfl<true>(a, cmax, rmin);
}

// ---------------------------------------------------------------------------
// tile shapes and launch helpers
// ---------------------------------------------------------------------------
// Tile shapes (cells per axis) the tile kernels are compiled for, CH the
// particles a P2G wave stages at once.  The node arrays are z-contiguous in
// 3D (tag = (i ny + j) nz + k) and y-contiguous in 2D, so the longer extents
// go there: they lengthen the contiguous runs of the grid update and of the
// node block loads.  kmpm_autotune picks the variant per context.
constexpr int NTILE_VARIANTS = 3;
template <int DIM, int S1, int V> struct Tiles;
template <> struct Tiles<3, 4, 0> { static constexpr int TX = 4, TY = 4, TZ = 4, CH = 24; };
template <> struct Tiles<3, 4, 1> { static constexpr int TX = 4, TY = 4, TZ = 8, CH = 24; };
template <> struct Tiles<3, 4, 2> { static constexpr int TX = 4, TY = 8, TZ = 8, CH = 24; };
template <> struct Tiles<3, 2, 0> { static constexpr int TX = 4, TY = 4, TZ = 4, CH = 32; };
template <> struct Tiles<3, 2, 1> { static constexpr int TX = 4, TY = 4, TZ = 8, CH = 32; };
template <> struct Tiles<3, 2, 2> { static constexpr int TX = 8, TY = 8, TZ = 8, CH = 32; };
template <> struct Tiles<2, 4, 0> { static constexpr int TX = 8, TY = 8, TZ = 1, CH = 64; };
template <> struct Tiles<2, 4, 1> { static constexpr int TX = 8, TY = 16, TZ = 1, CH = 64; };
template <> struct Tiles<2, 4, 2> { static constexpr int TX = 16, TY = 16, TZ = 1, CH = 64; };
template <> struct Tiles<2, 2, 0> { static constexpr int TX = 8, TY = 8, TZ = 1, CH = 64; };
template <> struct Tiles<2, 2, 1> { static constexpr int TX = 8, TY = 16, TZ = 1, CH = 64; };
template <> struct Tiles<2, 2, 2> { static constexpr int TX = 16, TY = 16, TZ = 1, CH = 64; };

static inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }


static inline int persistent_grid(kmpm_ctx *c, int per_cu) {
  int64_t g = (int64_t)c->cu_count * per_cu;
  int64_t maxt = c->a.tg.ntiles < c->a.np ? c->a.tg.ntiles : c->a.np;
  if (maxt < 1) maxt = 1;
  return (int)(g < maxt ? g : maxt);
}

#define PREC_DISPATCH(c, ...)                                                             \
  do {                                                                                    \
    if ((c)->a.prec == KMPM_PRECISION_MIXED) { constexpr int PREC = KMPM_PRECISION_MIXED; __VA_ARGS__; } \
    else { constexpr int PREC = KMPM_PRECISION_FP64; __VA_ARGS__; }                        \
  } while (0)

#define MC_DISPATCH(c, ...)                                                         \
  do {                                                                              \
    switch ((c)->a.mat_class) {                                                     \
      case MC_LINEAR: { constexpr int MC = MC_LINEAR; __VA_ARGS__; } break;         \
      case MC_JC_SHOCK: { constexpr int MC = MC_JC_SHOCK; __VA_ARGS__; } break;     \
      case MC_FLUID: { constexpr int MC = MC_FLUID; __VA_ARGS__; } break;           \
      default: { constexpr int MC = MC_GENERIC; __VA_ARGS__; } break;               \
    }                                                                               \
  } while (0)

// Launchers of the tile kernels for one (dimension, shape, tile variant);
// defined in kmpm_tile_launch.h and instantiated once per variant
// (kmpm_tiles.hip), so the variants compile in parallel.
template <int DIM, int S1, int V> struct TileLaunch {
  static hipError_t p2g(kmpm_ctx *c, int remap);
  static hipError_t grid_update(kmpm_ctx *c, int mode);   // GridMode
  static hipError_t grid_clean(kmpm_ctx *c);
  static hipError_t g2p_usl(kmpm_ctx *c);
  static hipError_t g2p_musl_a(kmpm_ctx *c);
  static hipError_t stress_musl_b(kmpm_ctx *c);
  static hipError_t p2g_thermal(kmpm_ctx *c, int remap);   // variant 0 only
  static hipError_t grid_thermal(kmpm_ctx *c, int mode);   // ThermalMode; variant 0 only
};

}  // namespace kmpm

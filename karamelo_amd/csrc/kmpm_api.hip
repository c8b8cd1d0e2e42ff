// kmpm_api.hip -- C ABI of the MI355X explicit MPM step: context, buffers,
// phases and the stream-ordered step loop.
#include <algorithm>
#include <cstddef>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kmpm_ctx.h"
#include "kmpm_mat3.h"
#include "kmpm_cutnet.h"
#include "kmpm_surrogate.h"

using namespace kmpm;

namespace kmpm {
void tile_shape(int dim, int S1, int variant, int T[3]);
}

#define HIPCHK(x)                          \
  do {                                     \
    hipError_t e__ = (x);                  \
    if (e__ != hipSuccess) return KMPM_ERR_HIP; \
  } while (0)

static int status_from_device(kmpm_ctx *c, int64_t *ptag_out) {
  DevScalars sc;
  if (hipMemcpyAsync(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    return KMPM_ERR_HIP;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return KMPM_ERR_HIP;
  if (ptag_out) *ptag_out = (sc.err_pid >= 0 && sc.err_pid < c->a.np) ? c->h_ptag[sc.err_pid] : -1;
  return sc.error;
}

template <class T>
static hipError_t dalloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  return hipMalloc((void **)p, n * sizeof(T));
}

static void free_particle_buffers(kmpm_ctx *c) {
  void *ptrs[] = {c->d_pa, c->d_pb, c->d_fa, c->d_fb, c->d_xo, c->d_ia, c->d_ib, c->d_key,
                  c->d_order_tmp, c->d_sort_tmp, c->d_order, c->d_pmask, c->d_solid};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  c->d_pa = c->d_pb = c->d_xo = nullptr;
  c->d_fa = c->d_fb = nullptr;
  c->d_ia = c->d_ib = c->d_key = c->d_order = nullptr;
  c->d_order_tmp = nullptr;
  c->d_sort_tmp = nullptr;
  c->d_pmask = nullptr;
  c->d_solid = nullptr;
  free(c->h_ptag);
  free(c->h_solid);
  free(c->h_pmask);
  free(c->h_x0);
  c->h_ptag = nullptr;
  c->h_solid = nullptr;
  c->h_pmask = nullptr;
  c->h_x0 = nullptr;
}

static void refresh_args(kmpm_ctx *c) {
  KArgs &a = c->a;
  a.pa = c->d_pa;
  a.pb = c->d_pb;
  a.fa = c->d_fa;
  a.fb = c->d_fb;
  a.ia = c->d_ia;
  a.ib = c->d_ib;
  a.xo = c->d_xo;
  a.key = c->d_key;
  a.order_tmp = c->d_order_tmp;
  a.sort_tmp = c->d_sort_tmp;
  a.order = c->d_order;
  a.solid_of = c->d_solid;
}

static void swap_state(kmpm_ctx *c) {
  std::swap(c->d_pa, c->d_pb);
  std::swap(c->d_fa, c->d_fb);
  std::swap(c->d_ia, c->d_ib);
  refresh_args(c);
}

// Captured step graphs hold the buffers and library workspaces of the
// moment they were captured; anything that reallocates them drops the graphs.
static void drop_graphs(kmpm_ctx *c) {
  for (auto &g : c->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
    g.steps = 0;
  }
}

// Phase timing: a start/stop HIP event pair recorded on the context stream
// around each launch of a phase while timing is enabled.
static void tbegin(kmpm_ctx *c, int phase) {
  if (!c->timing) return;
  std::vector<hipEvent_t> &pool = *c->ev_pool;
  const size_t k = c->ev_phase->size() * 2;
  while (pool.size() < k + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    pool.push_back(e);
  }
  (void)hipEventRecord(pool[k], c->stream);
  c->ev_phase->push_back(phase);
}
static void tend(kmpm_ctx *c) {
  if (!c->timing || c->ev_phase->empty()) return;
  const size_t k = c->ev_phase->size() * 2 - 1;
  (void)hipEventRecord((*c->ev_pool)[k], c->stream);
}

// Mixed layout <-> the FP64 field layout used on the host.
static void split_mixed(const std::vector<double> &hd, int cap, int lay,
                        std::vector<double> &dd, std::vector<float> &ff) {
  for (int f = 0; f < state_nd(false, lay); f++)
    for (int i = 0; i < cap; i++) {
      const double v = hd[(size_t)f * cap + i];
      if (mx_is_double(f)) dd[(size_t)mx_dslot(f) * cap + i] = v;
      else ff[(size_t)mx_fslot(f) * cap + i] = (float)v;
    }
}
static void join_mixed(const std::vector<double> &dd, const std::vector<float> &ff, int cap,
                       int lay, std::vector<double> &hd) {
  for (int f = 0; f < state_nd(false, lay); f++)
    for (int i = 0; i < cap; i++)
      hd[(size_t)f * cap + i] = mx_is_double(f) ? dd[(size_t)mx_dslot(f) * cap + i]
                                                : (double)ff[(size_t)mx_fslot(f) * cap + i];
}

// Fixed-point scales of the mixed path's node sums: the largest power of
// two that keeps one particle's contribution below 2^FX_CONTRIB_BITS, from
// bounds taken at upload with a factor 4 of headroom:
//   mass      m_max                      (weights are at most 1)
//   momentum  m_max * v_ref,  v_ref = 4 max(|v|_max, max wave speed of the materials)
//   force     vol_ref * sigma_ref / h,   vol_ref = 4 vol_max,
//             sigma_ref = 4 max(|sigma|_max, max (K + 4/3 G))
// (|grad w| <= 1/h for both shape families; 8/h for the MLS force).
// Power-of-two scales of the fixed-point node sums: one particle's
// contribution stays below 2^FX_CONTRIB_BITS.  The P2G kernels scale in FP32,
// so a scale must lie inside the float range (false otherwise).
static bool fixed_point_scales(const kmpm_ctx *c, const std::vector<double> &hd, int cap,
                               int64_t np, double fxs[FX_N], double fxi[FX_N]) {
  double mmax = 0, vmax = 0, smax = 0, volmax = 0;
  for (int64_t i = 0; i < np; i++) {
    mmax = std::max(mmax, std::fabs(hd[(size_t)MASS * cap + i]));
    volmax = std::max(volmax, std::fabs(hd[(size_t)VOL * cap + i]));
    for (int d = 0; d < 3; d++) vmax = std::max(vmax, std::fabs(hd[(size_t)(VX + d) * cap + i]));
    for (int e = 0; e < 6; e++) smax = std::max(smax, std::fabs(hd[(size_t)(SXX + e) * cap + i]));
  }
  double cmax = 0, modmax = 0;
  for (int k = 0; k < c->nmat; k++) {
    const kmpm_material &m = c->mats_host[k];
    cmax = std::max(cmax, m.signal_velocity);
    modmax = std::max(modmax, std::fabs(m.K) + FOUR_THIRD * std::fabs(m.G));
  }
  const double vref = 4 * std::max(vmax, cmax);
  const double sref = 4 * std::max(smax, modmax);
  // the MLS force's factor Di (x_I - x_p) reaches 3/h^2 * 2h = 6/h
  const double fmls = c->a.xf == 2 ? 8.0 : 1.0;
  const double bound[FX_N] = {mmax, mmax * vref, 4 * volmax * sref * c->a.inv_h * fmls};
  for (int k = 0; k < FX_N; k++) {
    int e = 0;
    std::frexp(bound[k] > 0 && std::isfinite(bound[k]) ? bound[k] : 1.0, &e);   // bound < 2^e
    if (FX_CONTRIB_BITS - e > 120 || FX_CONTRIB_BITS - e < -120) return false;
    fxs[k] = std::ldexp(1.0, FX_CONTRIB_BITS - e);
    fxi[k] = std::ldexp(1.0, e - FX_CONTRIB_BITS);
  }
  return true;
}

// Device record of a material (Mat constructors, material.cpp:720-769,
// with the plugin parameters in constructor order).
DevMat kmpm::make_devmat(const kmpm_material &m) {
  DevMat d;
  memset(&d, 0, sizeof(d));
  d.type = m.type;
  d.eos = m.eos;
  d.strength = m.strength;
  d.damage = m.damage;
  d.rho0 = m.rho0;
  d.K = m.K;
  d.G = m.G;
  d.lambda = m.lambda;
  d.signal_velocity = m.signal_velocity;
  d.cfl_k43g = m.K + FOUR_THIRD * m.G;
  {  // JC thermal factor at T = 0 (strength_jc.cpp:127-133)
    const double mm = m.strength_p[6], Tr = m.strength_p[7], Tm = m.strength_p[8], T = 0;
    double f = 1.0;
    if (T < Tm) {
      if (mm != 0 && T >= Tr) f = 1.0 - std::pow((T - Tr) / (Tm - Tr), mm);
    } else {
      f = 0.0;
    }
    d.jc_thermal = f;
  }
  memcpy(d.eos_p, m.eos_p, sizeof(d.eos_p));
  memcpy(d.strength_p, m.strength_p, sizeof(d.strength_p));
  memcpy(d.damage_p, m.damage_p, sizeof(d.damage_p));
  if (m.type == KMPM_MAT_EOS_STRENGTH && m.temp == KMPM_TEMP_PLASTIC_WORK && m.temp_p[1] != 0) {
    d.temp_on = 1;
    d.chi = m.temp_p[0];
    d.invcp = 1.0 / m.temp_p[1];
    d.kappa = m.temp_p[2];
    d.talpha = m.temp_p[3];
    d.T0 = m.temp_p[4];
    d.Tmelt = m.temp_p[5];
  }
  return d;
}

// Tile ranges of instances given as cell ranges (multiples of the tile
// shape); false when a range is not tile-aligned or leaves the grid.
static bool instance_tiles(const TileGeom &g, int dim, const int32_t *clo, const int32_t *chi,
                           int tlo[3], int thi[3]) {
  for (int d = 0; d < 3; d++) {
    int lo = clo[d], hi = chi[d];
    if (d >= dim) {
      if (lo != 0 || hi != 1) return false;
      tlo[d] = 0;
      thi[d] = 1;
      continue;
    }
    if (lo < 0 || hi <= lo || lo % g.T[d] || hi % g.T[d]) return false;
    tlo[d] = lo / g.T[d];
    thi[d] = hi / g.T[d];
    if (thi[d] > g.nt[d]) return false;
  }
  return true;
}

// Tile-major storage index of node `tag` (the device's node_sidx).
static int64_t host_sidx(const KArgs &a, int64_t tag) {
  const TileGeom &g = a.tg;
  const int k = (int)(tag % a.n[2]), j = (int)((tag / a.n[2]) % a.n[1]);
  const int i = (int)(tag / ((int64_t)a.n[2] * a.n[1]));
  const int t = ((i / g.T[0]) * g.nt[1] + j / g.T[1]) * g.nt[2] + k / g.T[2];
  return (int64_t)t * g.cpt + ((i % g.T[0]) * g.T[1] + j % g.T[1]) * g.T[2] + k % g.T[2];
}

// The node group masks (kept on the host in tag order) in storage order.
static int upload_gmask(kmpm_ctx *c) {
  const KArgs &a = c->a;
  std::vector<uint32_t> m((size_t)a.ns, 0u);
  const std::vector<uint32_t> &h = *c->h_gmask;
  for (int64_t tag = 0; tag < a.nn; tag++) m[host_sidx(a, tag)] = h[tag];
  HIPCHK(hipMemcpy(c->d_gmask, m.data(), a.ns * sizeof(uint32_t), hipMemcpyHostToDevice));
  return KMPM_OK;
}

// Tile geometry of a tile variant and every buffer sized by it (cell lists,
// tile lists, scan workspace, P2G partial blocks, the instance tile map).
// The state keeps its values; its cell keys and sort are redone on the next
// step.
static int setup_tiles(kmpm_ctx *c, int variant) {
  KArgs &a = c->a;
  TileGeom g;
  tile_shape(a.dim, a.S1, variant, g.T);
  if (variant < 0 || variant >= KMPM_NTILE_VARIANTS || g.T[0] == 0) return KMPM_ERR_ARG;
  // the APIC transfers and the USF step are compiled for the first tile
  // shape of each dimension and shape function only
  if (variant != 0 && (a.xf || a.thermo || a.scheme == KMPM_SCHEME_USF)) return KMPM_ERR_ARG;
  g.ntiles = 1;
  g.cpt = 1;
  g.bn = 1;
  g.off = a.S1 == 4 ? 1 : 0;
  for (int d = 0; d < 3; d++) {
    g.nt[d] = (a.n[d] + g.T[d] - 1) / g.T[d];
    g.E[d] = d < a.dim ? g.T[d] + a.S1 - 1 : 1;
    g.ntiles *= g.nt[d];
    g.cpt *= g.T[d];
    g.bn *= g.E[d];
  }
  const int64_t ncells = (int64_t)g.ntiles * g.cpt;
  if (ncells >= 2147483647LL) return KMPM_ERR_ARG;
  // the per-tile cell sort's LDS table; cell keys split into tile and cell by shift and mask
  if (g.cpt > TILE_SORT_MAX_CPT || (g.cpt & (g.cpt - 1)) != 0) return KMPM_ERR_ARG;
  // instances keep their cell ranges, which must be multiples of the tiles
  const int ninst = a.ninst;
  std::vector<int> map;
  std::vector<InstScalars> is(ninst);
  if (ninst) {
    map.assign(g.ntiles, -1);
    HIPCHK(hipMemcpyAsync(is.data(), c->d_inst, ninst * sizeof(InstScalars),
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int32_t *cells = c->inst_cells->data();
    for (int i = 0; i < ninst; i++) {
      InstScalars &q = is[i];
      if (!instance_tiles(g, a.dim, cells + 6 * i, cells + 6 * i + 3, q.tlo, q.thi))
        return KMPM_ERR_ARG;
      for (int x = q.tlo[0]; x < q.thi[0]; x++)
        for (int y = q.tlo[1]; y < q.thi[1]; y++)
          for (int z = q.tlo[2]; z < q.thi[2]; z++)
            map[((int64_t)x * g.nt[1] + y) * g.nt[2] + z] = i;
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  void *old[] = {c->d_tile_np, c->d_tile_p0, c->d_tile_cur, c->d_tile_list, c->d_tile_slot,
                 c->d_dil_list, c->d_scan_tmp, c->d_scratch, c->d_tinst,
                 c->d_gmass, c->d_gv, c->d_gvu, c->d_gmask, c->d_gT, c->d_gTu,
                 c->d_scratch_th, c->d_cacc};
  for (void *p : old)
    if (p) (void)hipFree(p);
  c->d_cacc = nullptr;
  c->d_tile_np = c->d_tile_p0 = c->d_tile_cur = c->d_tile_list = c->d_tile_slot = c->d_dil_list = nullptr;
  c->d_scan_tmp = nullptr;
  c->d_scratch = nullptr;
  c->d_tinst = nullptr;
  c->d_gmass = c->d_gv = c->d_gvu = nullptr;
  c->d_gT = c->d_gTu = c->d_scratch_th = nullptr;
  c->d_gmask = nullptr;
  // the grid, tile-major over whole tiles
  const int64_t ns = ncells;
  HIPCHK(dalloc(&c->d_gmass, ns));
  HIPCHK(dalloc(&c->d_gv, 3 * ns));
  HIPCHK(dalloc(&c->d_gvu, 3 * ns));
  HIPCHK(dalloc(&c->d_gmask, ns));
  HIPCHK(hipMemset(c->d_gmass, 0, ns * sizeof(double)));
  HIPCHK(hipMemset(c->d_gv, 0, 3 * ns * sizeof(double)));
  HIPCHK(hipMemset(c->d_gvu, 0, 3 * ns * sizeof(double)));
  HIPCHK(dalloc(&c->d_tile_np, g.ntiles + 1));
  HIPCHK(dalloc(&c->d_tile_p0, g.ntiles + 1));
  HIPCHK(dalloc(&c->d_tile_cur, g.ntiles));
  HIPCHK(hipMemset(c->d_tile_np, 0, (g.ntiles + 1) * sizeof(int)));   // zero between sorts
  HIPCHK(dalloc(&c->d_tile_list, g.ntiles));
  HIPCHK(dalloc(&c->d_tile_slot, g.ntiles));
  HIPCHK(dalloc(&c->d_dil_list, g.ntiles));
  HIPCHK(hipMemset(c->d_tile_slot, 0xff, g.ntiles * sizeof(int)));
  c->scan_tmp_bytes = scan_temp_bytes(g.ntiles + 1);
  HIPCHK(hipMalloc(&c->d_scan_tmp, c->scan_tmp_bytes ? c->scan_tmp_bytes : 16));
  // P2G partial blocks: one per active tile, at most min(tiles, particles)
  const int64_t slots = std::max<int64_t>(1, std::min<int64_t>(g.ntiles, a.np));
  HIPCHK(dalloc(&c->d_scratch, (size_t)slots * 7 * g.bn));
  if (a.thermo) {   // grid temperatures and the thermal P2G's partial blocks
    HIPCHK(dalloc(&c->d_gT, ns));
    HIPCHK(dalloc(&c->d_gTu, ns));
    HIPCHK(hipMemset(c->d_gT, 0, ns * sizeof(double)));
    HIPCHK(hipMemset(c->d_gTu, 0, ns * sizeof(double)));
    HIPCHK(dalloc(&c->d_scratch_th, (size_t)slots * 3 * g.bn));
  }
  // CPDI node sums: mass, momentum[3], force[3]
  if (a.cpdi) HIPCHK(dalloc(&c->d_cacc, (size_t)7 * ns));
  if (ninst) {
    HIPCHK(dalloc(&c->d_tinst, g.ntiles));
    HIPCHK(hipMemcpy(c->d_tinst, map.data(), g.ntiles * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_inst, is.data(), ninst * sizeof(InstScalars), hipMemcpyHostToDevice));
  }
  a.tg = g;
  a.ns = ns;
  a.gmass = c->d_gmass;
  a.gv = c->d_gv;
  a.gvu = c->d_gvu;
  a.gmask = c->d_gmask;
  c->ncells = ncells;
  c->tile_variant = variant;
  {
    const int st = upload_gmask(c);
    if (st != KMPM_OK) return st;
  }
  a.tile_np = c->d_tile_np;
  a.tile_p0 = c->d_tile_p0;
  a.tile_cur = c->d_tile_cur;
  a.tile_list = c->d_tile_list;
  a.tile_slot = c->d_tile_slot;
  a.dil_list = c->d_dil_list;
  a.scratch = c->d_scratch;
  a.scratch_slots = (int)slots;
  a.gT = c->d_gT;
  a.gTu = c->d_gTu;
  a.scratch_th = c->d_scratch_th;
  a.cacc = c->d_cacc;
  a.tinst = c->d_tinst;
  c->keys_valid = false;
  c->sorted = false;
  c->tl_identity = false;
  return KMPM_OK;
}

extern "C" {

int kmpm_ctx_create(const kmpm_config *cfg, const kmpm_material *mats, int nmat,
                    kmpm_ctx **out) {
  if (!cfg || !out || nmat < 1 || nmat > KMPM_MAX_MATERIALS || !mats) return KMPM_ERR_ARG;
  *out = nullptr;
  if (cfg->method < KMPM_METHOD_ULMPM || cfg->method > KMPM_METHOD_TLCPDI) return KMPM_ERR_ARG;
  if (cfg->scheme != KMPM_SCHEME_USL && cfg->scheme != KMPM_SCHEME_MUSL &&
      cfg->scheme != KMPM_SCHEME_USF)
    return KMPM_ERR_ARG;
  if (cfg->shape != KMPM_SHAPE_LINEAR && cfg->shape != KMPM_SHAPE_CUBIC_SPLINE) return KMPM_ERR_ARG;
  if (cfg->sub_method < KMPM_SUB_PIC || cfg->sub_method > KMPM_SUB_MLS) return KMPM_ERR_ARG;
  // velocity-update sub-method: ULMPM::setup (ulmpm.cpp:79-85), TLMPM::setup
  // (tlmpm.cpp:82-84); the inertia tensor of Solid::compute_inertia_tensor
  // (solid.cpp:1440-1470) needs cubic splines on ULMPM and 1 or 2 particles
  // per cell with linear shape functions on TLMPM
  const bool tl_cfg = cfg->method == KMPM_METHOD_TLMPM || cfg->method == KMPM_METHOD_TLCPDI;
  const bool cpdi = cfg->method == KMPM_METHOD_ULCPDI || cfg->method == KMPM_METHOD_TLCPDI;
  const int sub = cfg->sub_method;
  // CPDI (ulcpdi.cpp:115-118): 2D, FP64, the fast order, PIC/FLIP, one
  // context per run (no batched instances or learned updates, below)
  if (cpdi && (cfg->dimension != 2 || cfg->precision != KMPM_PRECISION_FP64 ||
               cfg->order != KMPM_ORDER_FAST || sub >= KMPM_SUB_APIC || cfg->thermo ||
               (cfg->cpdi_style != KMPM_CPDI_R4 && cfg->cpdi_style != KMPM_CPDI_Q4)))
    return KMPM_ERR_ARG;
  const bool apic = tl_cfg ? sub == KMPM_SUB_APIC : sub >= KMPM_SUB_APIC;
  if (apic && cfg->shape == KMPM_SHAPE_LINEAR &&
      (!tl_cfg || (cfg->np_per_cell != 1 && cfg->np_per_cell != 2)))
    return KMPM_ERR_ARG;
  // the reference-order kernels implement the PIC/FLIP transfers of USL
  // and MUSL
  if (cfg->order == KMPM_ORDER_REFERENCE && (apic || cfg->scheme == KMPM_SCHEME_USF))
    return KMPM_ERR_ARG;
  // thermo-mechanical runs: FP64, the fast order, the PIC/FLIP transfers
  if (cfg->thermo && (cfg->precision != KMPM_PRECISION_FP64 ||
                      cfg->order != KMPM_ORDER_FAST || apic))
    return KMPM_ERR_ARG;
  if (cfg->precision != KMPM_PRECISION_FP64 && cfg->precision != KMPM_PRECISION_MIXED)
    return KMPM_ERR_ARG;
  if (cfg->dimension != 2 && cfg->dimension != 3) return KMPM_ERR_ARG;
  if (cfg->order != KMPM_ORDER_FAST && cfg->order != KMPM_ORDER_REFERENCE) return KMPM_ERR_ARG;
  if (cfg->order == KMPM_ORDER_REFERENCE &&
      (cfg->method != KMPM_METHOD_ULMPM || cfg->precision != KMPM_PRECISION_FP64))
    return KMPM_ERR_ARG;
  int64_t n3[3];
  int st = kmpm_grid_size(cfg, n3);
  if (st != KMPM_OK) return st;
  const int S1 = cfg->shape == KMPM_SHAPE_LINEAR ? 2 : 4;
  for (int d = 0; d < cfg->dimension; d++)
    if (n3[d] < S1) return KMPM_ERR_GRID;  // node tags must not wrap inside one stencil
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return KMPM_ERR_NO_DEVICE;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return KMPM_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return KMPM_ERR_NO_DEVICE;
  if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) return KMPM_ERR_NO_DEVICE;

  kmpm_ctx *c = (kmpm_ctx *)calloc(1, sizeof(kmpm_ctx));
  c->cfg = *cfg;
  c->device = dev;
  c->cu_count = prop.multiProcessorCount;
  c->nmat = nmat;
  c->mats_host = (kmpm_material *)malloc(nmat * sizeof(kmpm_material));
  memcpy(c->mats_host, mats, nmat * sizeof(kmpm_material));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    free(c);
    return KMPM_ERR_HIP;
  }
  KArgs &a = c->a;
  a.dim = cfg->dimension;
  a.tl = tl_cfg;
  a.cpdi = cpdi ? 1 + cfg->cpdi_style : 0;
  a.ref_order = cfg->order == KMPM_ORDER_REFERENCE;
  a.prec = cfg->precision;
  a.S1 = S1;
  a.scheme = cfg->scheme;
  for (int d = 0; d < 3; d++) {
    a.n[d] = (int)n3[d];
    // the grid's bounds: the box (UL) or the solid's limits (TL)
    a.lo[d] = d < cfg->dimension ? (tl_cfg ? cfg->solidlo[d] : cfg->boxlo[d]) : 0.0;
    a.hi[d] = d < cfg->dimension ? (tl_cfg ? cfg->solidhi[d] : cfg->boxhi[d]) : 0.0;
    a.blo[d] = d < cfg->dimension ? cfg->boxlo[d] : 0.0;
  }
  a.nn = n3[0] * n3[1] * n3[2];
  a.h = cfg->cellsize;
  a.inv_h = 1.0 / cfg->cellsize;  // ulmpm.cpp:122
  a.flip = cfg->pic_flip;
  if (sub == KMPM_SUB_APIC || (!a.tl && sub == KMPM_SUB_MLS)) a.flip = 0.0;   // ulmpm.cpp:81
  a.xf = !apic ? 0 : (!a.tl && sub == KMPM_SUB_MLS) ? 2 : 1;
  a.asflip = !a.tl && sub == KMPM_SUB_ASFLIP;
  a.thermo = cfg->thermo != 0;
  a.lay = a.cpdi ? LAY_CP : a.thermo ? LAY_TH : apic ? LAY_AP : a.tl ? LAY_TL : LAY_UL;
  if (apic) {
    const double inv2 = 1.0 / (cfg->cellsize * cfg->cellsize);
    a.di = cfg->shape == KMPM_SHAPE_CUBIC_SPLINE ? 3.0 * inv2
           : cfg->np_per_cell == 1               ? 16.0 / 4.0 * inv2
                                                 : 16.0 / 3.0 * inv2;
  }
  a.dt_factor = cfg->dt_factor;
  a.dt_constant = cfg->dt_constant;
  a.nmat = nmat;
  c->inst_cells = new std::vector<int32_t>();
  // material class of the stress kernels
  {
    bool all_lin = true, all_jc = true, all_fluid = true;
    for (int i = 0; i < nmat; i++) {
      const kmpm_material &m = mats[i];
      all_lin = all_lin && m.type == KMPM_MAT_LINEAR;
      all_jc = all_jc && m.type == KMPM_MAT_EOS_STRENGTH && m.eos == KMPM_EOS_SHOCK &&
               m.strength == KMPM_STRENGTH_JOHNSON_COOK;
      all_fluid = all_fluid && m.type == KMPM_MAT_EOS_STRENGTH && m.eos == KMPM_EOS_FLUID &&
                  m.strength == KMPM_STRENGTH_FLUID && m.damage == KMPM_DAMAGE_NONE;
    }
    a.mat_class = all_lin ? MC_LINEAR : all_jc ? MC_JC_SHOCK : all_fluid ? MC_FLUID : MC_GENERIC;
    if (a.thermo) a.mat_class = MC_GENERIC;   // the temperature-aware update is the generic one
  }
  // material table
  std::vector<DevMat> dm(nmat);
  for (int i = 0; i < nmat; i++) dm[i] = make_devmat(mats[i]);
  bool ok = dalloc(&c->d_mats, nmat) == hipSuccess &&
            hipMemcpy(c->d_mats, dm.data(), nmat * sizeof(DevMat), hipMemcpyHostToDevice) ==
                hipSuccess;
  ok = ok && dalloc(&c->d_sc, 1) == hipSuccess;
  ok = ok && dalloc(&c->d_red, 2 + 2 * ENERGY_BLOCKS) == hipSuccess;
  ok = ok && dalloc(&c->d_fexpr, 6 * MAX_FIX) == hipSuccess &&
       hipMemset(c->d_fexpr, 0, 6 * MAX_FIX * sizeof(kmpm_expr)) == hipSuccess;
  if (ok) {
    c->h_gmask = new std::vector<uint32_t>(a.nn, 1u);   // Grid::grow sets mask = 1 (grid.cpp:429)
    DevScalars sc;
    memset(&sc, 0, sizeof(sc));
    sc.dt = cfg->dt;
    sc.err_pid = -1;
    const double one = 1.0;
    memcpy(&sc.min_ratio_bits, &one, sizeof(one));
    ok = ok && hipMemcpy(c->d_sc, &sc, sizeof(sc), hipMemcpyHostToDevice) == hipSuccess;
  }
  a.mats = c->d_mats;
  a.sc = c->d_sc;
  a.fexpr = c->d_fexpr;
  if (!ok) {
    kmpm_ctx_destroy(c);
    return KMPM_ERR_HIP;
  }
  if ((st = setup_tiles(c, 0)) != KMPM_OK) {
    kmpm_ctx_destroy(c);
    return st;
  }
  *out = c;
  return KMPM_OK;
}

int kmpm_ctx_destroy(kmpm_ctx *c) {
  if (!c) return KMPM_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_particle_buffers(c);
  drop_graphs(c);
  void *ptrs[] = {c->d_mats, c->d_tile_np, c->d_tile_p0, c->d_tile_cur, c->d_tile_list, c->d_tile_slot,
                  c->d_dil_list, c->d_scratch, c->d_gmass, c->d_gv, c->d_gvu, c->d_gmask,
                  c->d_sc, c->d_scan_tmp, c->d_surr, c->d_inst, c->d_tinst, c->d_red,
                  c->d_fexpr, c->d_cut, c->d_cut_stage, c->d_gT, c->d_gTu,
                  c->d_scratch_th, c->d_cacc};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  free(c->mats_host);
  delete c->inst_cells;
  delete c->h_gmask;
  if (c->ev_pool) {
    for (hipEvent_t e : *c->ev_pool) (void)hipEventDestroy(e);
    delete c->ev_pool;
    delete c->ev_phase;
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  free(c);
  return KMPM_OK;
}

int kmpm_upload_particles(kmpm_ctx *c, const kmpm_particles *p) {
  if (!c || !p || p->np < 0 || p->np >= 2147483647LL) return KMPM_ERR_ARG;
  if (!p->ptag || !p->solid || !p->mat || !p->x || !p->v || !p->F || !p->sigma ||
      !p->strain_el || !p->mass || !p->vol0 || !p->eff_plastic_strain ||
      !p->eff_plastic_strain_rate || !p->damage || !p->damage_init)
    return KMPM_ERR_ARG;
  const int64_t np = p->np;
  for (int64_t i = 0; i < np; i++) {
    if (p->mat[i] < 0 || p->mat[i] >= c->nmat) return KMPM_ERR_ARG;
    if (i > 0 && (p->ptag[i] <= p->ptag[i - 1] || p->solid[i] < p->solid[i - 1]))
      return KMPM_ERR_ARG;  // ascending tags, ordered by solid
  }
  KArgs &a = c->a;
  if (a.tl || a.thermo)   // one solid: one TL grid; the thermal sums of one solid
    for (int64_t i = 1; i < np; i++)
      if (p->solid[i] != p->solid[0]) return KMPM_ERR_ARG;
  const int cap = (int)std::max<int64_t>(np, 1);
  const bool mixed = a.prec == KMPM_PRECISION_MIXED;
  const int ndf = state_nd(mixed, a.lay);

  // Host images of the new state.  Everything that can reject the upload is
  // checked here, while the context still holds its previous particles.
  // The APIC layout's affine matrices start at zero, as the reference's
  // L and Fdot do after populate or a restart.
  std::vector<double> hd((size_t)state_nd(false, a.lay) * cap, 0.0);
  std::vector<int> hi((size_t)NIF * cap, 0);
  auto D = [&](int f, int64_t i) -> double & { return hd[(size_t)f * cap + i]; };
  for (int64_t i = 0; i < np; i++) {
    for (int d = 0; d < 3; d++) {
      D(PX + d, i) = p->x[3 * i + d];
      D(VX + d, i) = p->v[3 * i + d];
    }
    for (int e = 0; e < 9; e++) D(F00 + e, i) = p->F[9 * i + e];
    const double *s = &p->sigma[9 * i], *e = &p->strain_el[9 * i];
    const int sym[6] = {0, 4, 8, 1, 2, 5};
    for (int k = 0; k < 6; k++) {
      D(SXX + k, i) = s[sym[k]];
      D(EXX + k, i) = e[sym[k]];
    }
    D(EP, i) = p->eff_plastic_strain[i];
    D(EPDOT, i) = p->eff_plastic_strain_rate[i];
    D(DINIT, i) = p->damage_init[i];
    D(DMG, i) = p->damage[i];
    D(MASS, i) = p->mass[i];
    D(VOL0, i) = p->vol0[i];
    D(VOL, i) = p->vol ? p->vol[i] : p->vol0[i];
    if (a.tl) {
      // reference position and vol0 PK1 = vol0 J (R sigma R^T) F^-T of the
      // uploaded F and sigma (zero for an unstressed solid)
      double F[9], Finv[9], R[9], P[9];
      for (int e = 0; e < 9; e++) F[e] = p->F[9 * i + e];
      for (int d = 0; d < 3; d++) D(X0X + d, i) = p->x0 ? p->x0[3 * i + d] : p->x[3 * i + d];
      m3_inverse(F, Finv);
      if (!polar_rotation(F, R)) return KMPM_ERR_JACOBIAN;
      pk1(p->vol0[i], m3_det(F), R, s, Finv, P);
      for (int e = 0; e < 9; e++) D(P00 + e, i) = P[e];
    }
    if (a.thermo) D(TEMP, i) = p->T ? p->T[i] : 0.0;   // gamma = q = 0 (solid.cpp:2305-2309)
    if (a.cpdi == 1 + KMPM_CPDI_R4) {   // r1x, r1y, r2x, r2y of the [2][3] vectors
      if (!p->rp0) return KMPM_ERR_ARG;
      const double *r0 = &p->rp0[6 * i], *r = p->rp ? &p->rp[6 * i] : r0;
      const int xy[4] = {0, 1, 3, 4};
      for (int e = 0; e < 4; e++) {
        D(REF0 + e, i) = r0[xy[e]];
        D(DOM0 + e, i) = r[xy[e]];
      }
    } else if (a.cpdi == 1 + KMPM_CPDI_Q4) {   // x, y of the four [3] corners
      if (!p->xpc0) return KMPM_ERR_ARG;
      const double *c0 = &p->xpc0[12 * i], *cc = p->xpc ? &p->xpc[12 * i] : c0;
      for (int k = 0; k < 4; k++)
        for (int d = 0; d < 2; d++) {
          D(REF0 + 2 * k + d, i) = c0[3 * k + d];
          D(DOM0 + 2 * k + d, i) = cc[3 * k + d];
        }
    }
    hi[(size_t)PID * cap + i] = (int)i;
    hi[(size_t)MATI * cap + i] = p->mat[i];
  }
  double fxs[FX_N], fxi[FX_N];
  if (mixed && !fixed_point_scales(c, hd, cap, np, fxs, fxi)) return KMPM_ERR_FIXED_POINT;

  // The previous particles go.  From here a failure leaves the context
  // without particles and rejects every step until an upload succeeds.
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  free_particle_buffers(c);
  c->upload_failed = true;
  a.np = 0;
  a.cap = 0;
  refresh_args(c);
  HIPCHK(dalloc(&c->d_pa, (size_t)ndf * cap));
  HIPCHK(dalloc(&c->d_pb, (size_t)ndf * cap));
  if (mixed) {
    HIPCHK(dalloc(&c->d_fa, (size_t)state_nf(true, a.lay) * cap));
    HIPCHK(dalloc(&c->d_fb, (size_t)state_nf(true, a.lay) * cap));
  }
  HIPCHK(dalloc(&c->d_xo, (size_t)3 * cap));
  HIPCHK(dalloc(&c->d_ia, (size_t)NIF * cap));
  HIPCHK(dalloc(&c->d_ib, (size_t)NIF * cap));
  HIPCHK(dalloc(&c->d_key, cap));
  HIPCHK(dalloc(&c->d_order_tmp, cap));
  HIPCHK(dalloc(&c->d_sort_tmp, cap));
  HIPCHK(dalloc(&c->d_order, cap));
  HIPCHK(dalloc(&c->d_pmask, cap));
  HIPCHK(dalloc(&c->d_solid, cap));
  a.np = (int)np;
  a.cap = cap;
  if (c->d_cut_stage && c->cut_stage_cap < cap) {   // the cut network's staging, [8][cap]
    const int was_on = a.cut_on;
    (void)hipFree(c->d_cut_stage);
    c->d_cut_stage = nullptr;
    c->cut_stage_cap = 0;
    a.cut_stage = nullptr;
    a.cut_on = 0;
    HIPCHK(dalloc(&c->d_cut_stage, (size_t)8 * cap));
    c->cut_stage_cap = cap;
    a.cut_stage = c->d_cut_stage;
    a.cut_on = was_on;
  }
  {
    const int st = setup_tiles(c, c->tile_variant);   // P2G partial blocks for np particles
    if (st != KMPM_OK) {
      a.np = 0;
      return st;
    }
  }
  c->h_ptag = (int64_t *)malloc(sizeof(int64_t) * cap);
  c->h_solid = (int32_t *)malloc(sizeof(int32_t) * cap);
  c->h_pmask = (uint32_t *)malloc(sizeof(uint32_t) * cap);
  c->h_x0 = (double *)malloc(sizeof(double) * 3 * cap);
  if (!c->h_ptag || !c->h_solid || !c->h_pmask || !c->h_x0) {
    a.np = 0;
    return KMPM_ERR_ARG;
  }
  for (int64_t i = 0; i < np; i++) {
    c->h_ptag[i] = p->ptag[i];
    c->h_solid[i] = p->solid[i];
    c->h_pmask[i] = p->mask ? p->mask[i] : 1u;
    for (int d = 0; d < 3; d++) c->h_x0[3 * i + d] = p->x0 ? p->x0[3 * i + d] : p->x[3 * i + d];
  }
  refresh_args(c);
  auto fail = [&](int st) {
    a.np = 0;
    return st;
  };
  if (mixed) {
    std::vector<double> dd((size_t)state_nd(true, a.lay) * cap);
    std::vector<float> ff((size_t)state_nf(true, a.lay) * cap);
    split_mixed(hd, cap, a.lay, dd, ff);
    if (hipMemcpy(c->d_pa, dd.data(), dd.size() * sizeof(double), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(c->d_fa, ff.data(), ff.size() * sizeof(float), hipMemcpyHostToDevice) !=
            hipSuccess)
      return fail(KMPM_ERR_HIP);
    for (int k = 0; k < FX_N; k++) {
      a.fxs[k] = fxs[k];
      a.fxi[k] = fxi[k];
    }
  } else if (hipMemcpy(c->d_pa, hd.data(), hd.size() * sizeof(double), hipMemcpyHostToDevice) !=
             hipSuccess) {
    return fail(KMPM_ERR_HIP);
  }
  if (hipMemcpy(c->d_ia, hi.data(), hi.size() * sizeof(int), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(c->d_pmask, c->h_pmask, np * sizeof(uint32_t), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(c->d_solid, c->h_solid, np * sizeof(int32_t), hipMemcpyHostToDevice) !=
          hipSuccess ||
      launch_iota(c->d_order, (int)np, c->stream) != hipSuccess)
    return fail(KMPM_ERR_HIP);
  c->keys_valid = false;
  c->tl_identity = false;
  c->sorted = false;
  c->upload_failed = false;
  return KMPM_OK;
}

// The current state in the FP64 field layout (hd) and the int fields (hi),
// in storage order; waits for the stream.
static int fetch_state(kmpm_ctx *c, std::vector<double> &hd, std::vector<int> &hi) {
  const int cap = c->a.cap;
  hd.assign((size_t)state_nd(false, c->a.lay) * cap, 0.0);
  hi.assign((size_t)NIF * cap, 0);
  std::vector<double> dd;
  std::vector<float> ff;
  const bool mixed = c->a.prec == KMPM_PRECISION_MIXED;
  if (mixed) {
    dd.resize((size_t)state_nd(true, c->a.lay) * cap);
    ff.resize((size_t)state_nf(true, c->a.lay) * cap);
    HIPCHK(hipMemcpyAsync(dd.data(), c->d_pa, dd.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemcpyAsync(ff.data(), c->d_fa, ff.size() * sizeof(float), hipMemcpyDeviceToHost,
                          c->stream));
  } else {
    HIPCHK(hipMemcpyAsync(hd.data(), c->d_pa, hd.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  }
  HIPCHK(hipMemcpyAsync(hi.data(), c->d_ia, hi.size() * sizeof(int), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (mixed) join_mixed(dd, ff, cap, c->a.lay, hd);
  return KMPM_OK;
}

int kmpm_download_affine(kmpm_ctx *c, double *C, int64_t np) {
  if (!c || !C || c->a.lay != LAY_AP || c->upload_failed) return KMPM_ERR_ARG;
  if (np < c->a.np) return KMPM_ERR_CAPACITY;
  std::vector<double> hd;
  std::vector<int> hi;
  const int st = fetch_state(c, hd, hi);
  if (st) return st;
  const int cap = c->a.cap;
  for (int64_t i = 0; i < c->a.np; i++) {
    const int64_t q = hi[(size_t)PID * cap + i];
    if (q < 0 || q >= c->a.np) return KMPM_ERR_HIP;
    for (int e = 0; e < 9; e++) C[9 * q + e] = hd[(size_t)(C00 + e) * cap + i];
  }
  return KMPM_OK;
}

int kmpm_download_particles(kmpm_ctx *c, kmpm_particles *p) {
  if (!c || !p) return KMPM_ERR_ARG;
  const int64_t np = c->a.np;
  if (p->np < np) return KMPM_ERR_CAPACITY;
  const int cap = c->a.cap;
  std::vector<double> hd;
  std::vector<int> hi;
  const int st = fetch_state(c, hd, hi);
  if (st) return st;
  auto D = [&](int f, int64_t i) { return hd[(size_t)f * cap + i]; };
  for (int64_t i = 0; i < np; i++) {
    const int64_t q = hi[(size_t)PID * cap + i];  // output index = upload index
    if (q < 0 || q >= np) return KMPM_ERR_HIP;
    const int mat = hi[(size_t)MATI * cap + i];
    if (p->ptag) p->ptag[q] = c->h_ptag[q];
    if (p->solid) p->solid[q] = c->h_solid[q];
    if (p->mat) p->mat[q] = mat;
    if (p->mask) p->mask[q] = c->h_pmask[q];
    for (int d = 0; d < 3; d++) {
      if (p->x0) p->x0[3 * q + d] = c->h_x0[3 * q + d];
      if (p->x) p->x[3 * q + d] = D(PX + d, i);
      if (p->v) p->v[3 * q + d] = D(VX + d, i);
    }
    if (p->F)
      for (int e = 0; e < 9; e++) p->F[9 * q + e] = D(F00 + e, i);
    const int full[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    for (int e = 0; e < 9; e++) {
      if (p->sigma) p->sigma[9 * q + e] = D(SXX + full[e], i);
      if (p->strain_el) p->strain_el[9 * q + e] = D(EXX + full[e], i);
    }
    const double *F = nullptr;
    double Fl[9];
    for (int e = 0; e < 9; e++) Fl[e] = D(F00 + e, i);
    F = Fl;
    // CPDI-Q4: J = vol / vol0 of the corners' polygon (solid.cpp:1188-1200)
    const double J = c->a.cpdi == 1 + KMPM_CPDI_Q4
                         ? D(VOL, i) / D(VOL0, i)
                         : F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                               F[2] * (F[3] * F[7] - F[4] * F[6]);
    const double rho0 = c->mats_host[mat].rho0;
    if (p->mass) p->mass[q] = D(MASS, i);
    if (p->vol0) p->vol0[q] = D(VOL0, i);
    if (p->vol) p->vol[q] = D(VOL, i);
    if (p->rho0) p->rho0[q] = rho0;
    if (p->rho) p->rho[q] = rho0 / J;
    if (p->J) p->J[q] = J;
    if (p->eff_plastic_strain) p->eff_plastic_strain[q] = D(EP, i);
    if (p->eff_plastic_strain_rate) p->eff_plastic_strain_rate[q] = D(EPDOT, i);
    if (p->damage) p->damage[q] = D(DMG, i);
    if (p->damage_init) p->damage_init[q] = D(DINIT, i);
    if (p->T) p->T[q] = c->a.thermo ? D(TEMP, i) : 0.0;
    if (p->gamma) p->gamma[q] = c->a.thermo ? D(GAMMA, i) : 0.0;
    if (c->a.cpdi == 1 + KMPM_CPDI_R4) {
      const int xy[4] = {0, 1, 3, 4};
      for (int e = 0; e < 6; e++) {
        if (p->rp0) p->rp0[6 * q + e] = 0.0;
        if (p->rp) p->rp[6 * q + e] = 0.0;
      }
      for (int e = 0; e < 4; e++) {
        if (p->rp0) p->rp0[6 * q + xy[e]] = D(REF0 + e, i);
        if (p->rp) p->rp[6 * q + xy[e]] = D(DOM0 + e, i);
      }
    } else if (c->a.cpdi == 1 + KMPM_CPDI_Q4) {
      for (int k = 0; k < 4; k++)
        for (int d = 0; d < 3; d++) {
          if (p->xpc0) p->xpc0[12 * q + 3 * k + d] = d < 2 ? D(REF0 + 2 * k + d, i) : 0.0;
          if (p->xpc) p->xpc[12 * q + 3 * k + d] = d < 2 ? D(DOM0 + 2 * k + d, i) : 0.0;
        }
    }
  }
  return KMPM_OK;
}

int kmpm_set_node_mask(kmpm_ctx *c, const uint32_t *mask, int64_t nn) {
  if (!c || !mask || nn != c->a.nn) return KMPM_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->h_gmask->assign(mask, mask + nn);
  return upload_gmask(c);
}

int kmpm_fix_velocity_nodes(kmpm_ctx *c, uint32_t bit, const int set[3], const double val[3],
                            const double prev[3]) {
  if (!c || !set || !val || !prev || c->a.nfv >= MAX_FIX) return KMPM_ERR_ARG;
  DevFixVel &f = c->a.fv[c->a.nfv++];
  f.bit = bit;
  f.expr = 0;
  for (int d = 0; d < 3; d++) {
    f.set[d] = set[d] && d < c->a.dim;  // fix_velocity_nodes.cpp:81-116
    f.val[d] = val[d];
    f.prev[d] = prev[d];
  }
  return KMPM_OK;
}

int kmpm_fix_velocity_nodes_expr(kmpm_ctx *c, uint32_t bit, const kmpm_expr value[3],
                                 const kmpm_expr previous[3]) {
  if (!c || !value || !previous || c->a.nfv >= MAX_FIX || c->a.ninst > 0) return KMPM_ERR_ARG;
  kmpm_expr prog[6];
  memset(prog, 0, sizeof(prog));
  int set[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++) {
    const kmpm_expr &v = value[d], &p = previous[d];
    if (v.n < 0 || v.n > KMPM_EXPR_MAX || p.n < 0 || p.n > KMPM_EXPR_MAX) return KMPM_ERR_ARG;
    if ((v.n == 0) != (p.n == 0)) return KMPM_ERR_ARG;
    set[d] = v.n > 0 && d < c->a.dim;     // NULL, or a component the dimension lacks
    if (set[d]) {
      prog[d] = v;
      prog[3 + d] = p;
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  const int idx = c->a.nfv;
  HIPCHK(hipMemcpy(c->d_fexpr + 6 * idx, prog, sizeof(prog), hipMemcpyHostToDevice));
  DevFixVel &f = c->a.fv[c->a.nfv++];
  memset(&f, 0, sizeof(f));
  f.bit = bit;
  f.expr = 1;
  for (int d = 0; d < 3; d++) f.set[d] = set[d];
  return KMPM_OK;
}

int kmpm_fix_velocity_nodes_force(kmpm_ctx *c, int fix, double force[3]) {
  if (!c || !force || fix < 0 || fix >= c->a.nfv) return KMPM_ERR_ARG;
  DevScalars sc;
  HIPCHK(hipMemcpyAsync(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int d = 0; d < 3; d++) force[d] = sc.step > 0 ? sc.fv_force[(sc.step - 1) & 1][fix][d] : 0.0;
  return KMPM_OK;
}

int kmpm_fix_body_force(kmpm_ctx *c, uint32_t bit, const int set[3], const double val[3]) {
  if (!c || !set || !val || c->a.nbf >= MAX_FIX) return KMPM_ERR_ARG;
  DevFixBF &f = c->a.bf[c->a.nbf++];
  f.bit = bit;
  for (int d = 0; d < 3; d++) {
    f.set[d] = set[d];
    f.val[d] = val[d];
  }
  return KMPM_OK;
}

// ---- phases ---------------------------------------------------------------
int kmpm_compute_grid_weight_functions_and_gradients(kmpm_ctx *c) {
  if (!c) return KMPM_ERR_ARG;
  if (c->upload_failed) return KMPM_ERR_ARG;
  if (c->a.ref_order) {
    // the reference-order kernels evaluate each pair's weights where they
    // use them, over the state in upload order
    c->sorted = true;
    return KMPM_OK;
  }
  if (c->tl_identity) {
    // TLMPM weights are those of x0 (computed once, tlmpm.cpp:89,339): after
    // the first step the state is stored in its sorted order and the cell
    // lists stay valid
    c->sorted = true;
    return KMPM_OK;
  }
  tbegin(c, KMPM_PHASE_SORT);
  if (!c->keys_valid) {
    HIPCHK(launch_keys(c));
    c->keys_valid = true;
  }
  HIPCHK(launch_sort(c));
  tend(c);
  c->sorted = true;
  return KMPM_OK;
}

int kmpm_particles_to_grid(kmpm_ctx *c) {
  if (!c || !c->sorted) return KMPM_ERR_ARG;
  if (c->upload_failed) return KMPM_ERR_ARG;
  tbegin(c, KMPM_PHASE_P2G);
  if (c->a.cpdi)
    HIPCHK(launch_cpdi_p2g(c, 0, 0));
  else
    HIPCHK(c->a.ref_order ? launch_ref_p2g(c, 0) : launch_p2g(c, 0));
  // the grid temperature's sums and the thermal driving forces
  // (ulmpm.cpp:371-375, tlmpm.cpp:359-363)
  if (c->a.thermo) HIPCHK(launch_p2g_thermal(c, 0));
  tend(c);
  return KMPM_OK;
}

int kmpm_update_grid_state(kmpm_ctx *c) {
  if (!c || !c->sorted) return KMPM_ERR_ARG;
  if (c->upload_failed) return KMPM_ERR_ARG;
  tbegin(c, KMPM_PHASE_GRID);
  // USF: the first pass's mass and velocity with particles_to_grid_USF_2's forces
  const int mode = c->cfg.scheme == KMPM_SCHEME_USF ? GM_USF2 : GM_STEP;
  if (c->a.cpdi)
    HIPCHK(launch_cpdi_grid(c, mode));
  else
    HIPCHK(c->a.ref_order ? launch_ref_grid(c, 0) : launch_grid_update(c, mode));
  // Grid::update_grid_temperature (grid.cpp:1354-1362), with the node mass
  // of this pass
  if (c->a.thermo) HIPCHK(launch_grid_thermal(c, TM_FULL));
  tend(c);
  return KMPM_OK;
}

// USF (usf.cpp:53-55): particles_to_grid_USF_1 (node mass and velocity,
// ulmpm.cpp:380-404) and the fixes' post_update_grid_state.  The P2G
// kernel also sums the forces of the current stress, which this pass leaves
// unused.
int kmpm_particles_to_grid_usf_1(kmpm_ctx *c) {
  if (!c || !c->sorted || c->cfg.scheme != KMPM_SCHEME_USF) return KMPM_ERR_ARG;
  if (c->upload_failed) return KMPM_ERR_ARG;
  tbegin(c, KMPM_PHASE_P2G);
  HIPCHK(c->a.cpdi ? launch_cpdi_p2g(c, 0, 0) : launch_p2g(c, 0));
  if (c->a.thermo) HIPCHK(launch_p2g_thermal(c, 0));   // usf.cpp:53 (grid temperature)
  tend(c);
  tbegin(c, KMPM_PHASE_GRID_B);
  HIPCHK(c->a.cpdi ? launch_cpdi_grid(c, GM_USF1) : launch_grid_update(c, GM_USF1));
  if (c->a.thermo) HIPCHK(launch_grid_thermal(c, TM_T));
  tend(c);
  return KMPM_OK;
}

// USF (usf.cpp:62-64): particles_to_grid_USF_2, the forces of the updated
// stress (ulmpm.cpp:406-431), and post_particles_to_grid's body forces
// (applied by the grid update, fix_body_force.cpp:117-146).
int kmpm_particles_to_grid_usf_2(kmpm_ctx *c) {
  if (!c || !c->sorted || c->cfg.scheme != KMPM_SCHEME_USF) return KMPM_ERR_ARG;
  if (c->upload_failed) return KMPM_ERR_ARG;
  tbegin(c, KMPM_PHASE_P2G_B);
  HIPCHK(c->a.cpdi ? launch_cpdi_p2g(c, 0, 1) : launch_p2g(c, 0));
  // the thermal driving forces of the updated heat source and flux
  if (c->a.thermo) HIPCHK(launch_p2g_thermal(c, 0));
  tend(c);
  return KMPM_OK;
}

// After the G2P that wrote the state in sorted order: TLMPM keeps that
// order for good (its keys come from x0), with the identity permutation.
static int after_g2p_state_swap(kmpm_ctx *c) {
  swap_state(c);
  c->sorted = false;
  c->keys_valid = true;
  if (c->a.tl && !c->tl_identity) {
    HIPCHK(launch_iota(c->d_order, c->a.np, c->stream));
    c->tl_identity = true;
  }
  return KMPM_OK;
}

int kmpm_grid_to_points_and_stress(kmpm_ctx *c) {
  if (!c || !c->sorted) return KMPM_ERR_ARG;
  if (c->upload_failed) return KMPM_ERR_ARG;
  tbegin(c, KMPM_PHASE_G2P);
  if (c->a.cpdi)
    HIPCHK(launch_cpdi_g2p(c, CG_USL));
  else
    HIPCHK(c->a.ref_order ? launch_ref_g2p_usl(c) : launch_g2p_usl(c));
  tend(c);
  if (learned_constitutive(c)) {
    tbegin(c, KMPM_PHASE_LEARNED);
    HIPCHK(launch_learned_stress(c));
    tend(c);
  }
  return after_g2p_state_swap(c);
}

int kmpm_grid_to_points(kmpm_ctx *c) {
  if (!c || !c->sorted) return KMPM_ERR_ARG;
  if (c->upload_failed) return KMPM_ERR_ARG;
  tbegin(c, KMPM_PHASE_G2P_A);
  if (c->a.cpdi)
    HIPCHK(launch_cpdi_g2p(c, c->cfg.scheme == KMPM_SCHEME_USF ? CG_USF : CG_MUSL));
  else
    HIPCHK(c->a.ref_order ? launch_ref_g2p_musl_a(c) : launch_g2p_musl_a(c));
  tend(c);
  // USF: the G2P carries the (already updated) state into the next buffer
  if (c->cfg.scheme == KMPM_SCHEME_USF) return after_g2p_state_swap(c);
  return KMPM_OK;
}

int kmpm_velocities_to_grid(kmpm_ctx *c) {
  if (!c || !c->sorted) return KMPM_ERR_ARG;
  if (c->upload_failed) return KMPM_ERR_ARG;
  tbegin(c, KMPM_PHASE_P2G_B);
  if (c->a.cpdi)
    HIPCHK(launch_cpdi_p2g(c, 1, 0));
  else
    HIPCHK(c->a.ref_order ? launch_ref_p2g(c, 1) : launch_p2g(c, 1));
  // the grid temperature of the updated particle temperatures
  // (ulmpm.cpp:491-493)
  if (c->a.thermo) HIPCHK(launch_p2g_thermal(c, 1));
  tend(c);
  tbegin(c, KMPM_PHASE_GRID_B);
  if (c->a.cpdi)
    HIPCHK(launch_cpdi_grid(c, GM_REMAP));
  else
    HIPCHK(c->a.ref_order ? launch_ref_grid(c, 1) : launch_grid_update(c, 1));
  if (c->a.thermo) HIPCHK(launch_grid_thermal(c, TM_T));
  tend(c);
  return KMPM_OK;
}

int kmpm_update_stress(kmpm_ctx *c) {
  if (!c || !c->sorted) return KMPM_ERR_ARG;
  if (c->upload_failed) return KMPM_ERR_ARG;
  tbegin(c, KMPM_PHASE_STRESS_B);
  if (c->a.cpdi)
    HIPCHK(launch_cpdi_stress(c));
  else
    HIPCHK(c->a.ref_order ? launch_ref_stress_musl_b(c) : launch_stress_musl_b(c));
  tend(c);
  // USF updates the current state in place, before its forces and G2P
  if (c->cfg.scheme == KMPM_SCHEME_USF) return KMPM_OK;
  return after_g2p_state_swap(c);
}

int kmpm_adjust_dt(kmpm_ctx *c) {
  if (!c) return KMPM_ERR_ARG;
  if (c->upload_failed) return KMPM_ERR_ARG;
  tbegin(c, KMPM_PHASE_DT);
  HIPCHK(launch_adjust_dt(c));
  tend(c);
  return KMPM_OK;
}

int kmpm_step(kmpm_ctx *c, int64_t nsteps) {
  if (!c || nsteps < 0) return KMPM_ERR_ARG;
  for (int64_t k = 0; k < nsteps; k++) {
    int s = kmpm_compute_grid_weight_functions_and_gradients(c);
    if (s) return s;
    if (c->cfg.scheme == KMPM_SCHEME_USF) {   // USF::run usf.cpp:48-83
      if ((s = kmpm_particles_to_grid_usf_1(c))) return s;
      if ((s = kmpm_update_stress(c))) return s;
      if ((s = kmpm_particles_to_grid_usf_2(c))) return s;
      if ((s = kmpm_update_grid_state(c))) return s;
      if ((s = kmpm_grid_to_points(c))) return s;
      if ((s = kmpm_adjust_dt(c))) return s;
      continue;
    }
    if ((s = kmpm_particles_to_grid(c))) return s;
    if ((s = kmpm_update_grid_state(c))) return s;
    if (c->cfg.scheme == KMPM_SCHEME_USL) {
      if ((s = kmpm_grid_to_points_and_stress(c))) return s;
    } else {
      if ((s = kmpm_grid_to_points(c))) return s;
      if ((s = kmpm_velocities_to_grid(c))) return s;
      if ((s = kmpm_update_stress(c))) return s;
    }
    if ((s = kmpm_adjust_dt(c))) return s;
  }
  return KMPM_OK;
}

int kmpm_synchronize(kmpm_ctx *c, int64_t *ptag_out) {
  if (!c) return KMPM_ERR_ARG;
  return status_from_device(c, ptag_out);
}

// ---- batched instances ------------------------------------------------------
int kmpm_set_instances(kmpm_ctx *c, int ninst, const int32_t *clo, const int32_t *chi) {
  if (!c || ninst < 0 || (ninst > 0 && (!clo || !chi))) return KMPM_ERR_ARG;
  if (ninst > 0 && (c->a.tl || c->a.ref_order || c->a.surr_on || c->a.cut_on || c->a.xf ||
                    c->a.thermo || c->a.cpdi || c->a.scheme == KMPM_SCHEME_USF))
    return KMPM_ERR_ARG;
  for (int q = 0; q < c->a.nfv; q++)
    if (ninst > 0 && c->a.fv[q].expr) return KMPM_ERR_ARG;   // values of the context's time
  const TileGeom &g = c->a.tg;
  std::vector<int> map(g.ntiles, -1);
  std::vector<InstScalars> is(ninst);
  for (int i = 0; i < ninst; i++) {
    InstScalars &q = is[i];
    memset(&q, 0, sizeof(q));
    if (!instance_tiles(g, c->a.dim, clo + 3 * i, chi + 3 * i, q.tlo, q.thi)) return KMPM_ERR_ARG;
    for (int x = q.tlo[0]; x < q.thi[0]; x++)
      for (int y = q.tlo[1]; y < q.thi[1]; y++)
        for (int z = q.tlo[2]; z < q.thi[2]; z++) {
          int &m = map[((int64_t)x * g.nt[1] + y) * g.nt[2] + z];
          if (m >= 0) return KMPM_ERR_ARG;   // overlapping ranges
          m = i;
        }
    q.dt = c->cfg.dt;
    const double one = 1.0;
    memcpy(&q.min_ratio_bits, &one, sizeof(one));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->d_inst) (void)hipFree(c->d_inst);
  if (c->d_tinst) (void)hipFree(c->d_tinst);
  c->d_inst = nullptr;
  c->d_tinst = nullptr;
  c->a.ninst = 0;
  c->a.inst = nullptr;
  c->a.tinst = nullptr;
  if (ninst == 0) return KMPM_OK;
  HIPCHK(dalloc(&c->d_inst, ninst));
  HIPCHK(dalloc(&c->d_tinst, g.ntiles));
  HIPCHK(hipMemcpy(c->d_inst, is.data(), ninst * sizeof(InstScalars), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_tinst, map.data(), g.ntiles * sizeof(int), hipMemcpyHostToDevice));
  c->a.ninst = ninst;
  c->a.inst = c->d_inst;
  c->a.tinst = c->d_tinst;
  c->inst_cells->resize(6 * (size_t)ninst);
  for (int i = 0; i < ninst; i++)
    for (int d = 0; d < 3; d++) {
      (*c->inst_cells)[6 * i + d] = clo[3 * i + d];
      (*c->inst_cells)[6 * i + 3 + d] = chi[3 * i + d];
    }
  return KMPM_OK;
}

int kmpm_instance_time(kmpm_ctx *c, int ninst, double *dt, double *time) {
  if (!c || ninst != c->a.ninst) return KMPM_ERR_ARG;
  if (ninst == 0) return KMPM_OK;
  std::vector<InstScalars> is(ninst);
  HIPCHK(hipMemcpyAsync(is.data(), c->d_inst, ninst * sizeof(InstScalars), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < ninst; i++) {
    if (dt) dt[i] = is[i].dt;
    if (time) time[i] = is[i].time;
  }
  return KMPM_OK;
}

// ---- tile variants and autotuning ---------------------------------------------
int kmpm_set_tile_variant(kmpm_ctx *c, int variant) {
  if (!c) return KMPM_ERR_ARG;
  if (variant == c->tile_variant) return KMPM_OK;
  return setup_tiles(c, variant);
}

// Device copy of the particle state and scalars that a step changes.
struct SavedState {
  double *pa = nullptr;
  float *fa = nullptr;
  int *ia = nullptr;
  InstScalars *inst = nullptr;
  DevScalars sc;
  size_t nd = 0, nf = 0, ni = 0;
};

static int save_particles(kmpm_ctx *c, SavedState &s) {
  const KArgs &a = c->a;
  const bool mixed = a.prec == KMPM_PRECISION_MIXED;
  s.nd = (size_t)state_nd(mixed, a.lay) * a.cap;
  s.nf = (size_t)state_nf(mixed, a.lay) * a.cap;
  s.ni = (size_t)NIF * a.cap;
  HIPCHK(dalloc(&s.pa, s.nd));
  if (s.nf) HIPCHK(dalloc(&s.fa, s.nf));
  HIPCHK(dalloc(&s.ia, s.ni));
  if (a.ninst) HIPCHK(dalloc(&s.inst, a.ninst));
  HIPCHK(hipMemcpyAsync(s.pa, c->d_pa, s.nd * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  if (s.nf)
    HIPCHK(hipMemcpyAsync(s.fa, c->d_fa, s.nf * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(s.ia, c->d_ia, s.ni * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
  if (a.ninst)
    HIPCHK(hipMemcpyAsync(s.inst, c->d_inst, a.ninst * sizeof(InstScalars),
                          hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(&s.sc, c->d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return KMPM_OK;
}

// Back to the saved state (in the buffers that are current now); its cell
// keys and sort are redone on the next step.
static int restore_particles(kmpm_ctx *c, const SavedState &s) {
  HIPCHK(hipMemcpyAsync(c->d_pa, s.pa, s.nd * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  if (s.nf)
    HIPCHK(hipMemcpyAsync(c->d_fa, s.fa, s.nf * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_ia, s.ia, s.ni * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
  if (c->a.ninst)
    HIPCHK(hipMemcpyAsync(c->d_inst, s.inst, c->a.ninst * sizeof(InstScalars),
                          hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_sc, &s.sc, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->keys_valid = false;
  c->sorted = false;
  c->tl_identity = false;
  return KMPM_OK;
}

static void free_saved(SavedState &s) {
  void *p[] = {s.pa, s.fa, s.ia, s.inst};
  for (void *q : p)
    if (q) (void)hipFree(q);
}

int kmpm_autotune(kmpm_ctx *c, int trial_steps, int *variant, double ms[KMPM_NTILE_VARIANTS]) {
  if (!c || trial_steps < 1 || c->timing || c->a.np == 0) return KMPM_ERR_ARG;
  int st = status_from_device(c, nullptr);
  if (st != KMPM_OK) return st;
  SavedState saved;
  if ((st = save_particles(c, saved)) != KMPM_OK) {
    free_saved(saved);
    return st;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int best = c->tile_variant;
  double best_ms = 1e300;
  const int start_variant = c->tile_variant;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = KMPM_ERR_HIP;
  for (int v = 0; v < KMPM_NTILE_VARIANTS && st == KMPM_OK; v++) {
    double t = -1;
    // the saved instance scalars come back first: setup_tiles recomputes
    // their tile ranges for this variant
    if ((st = restore_particles(c, saved)) != KMPM_OK) break;
    const int sv = setup_tiles(c, v);
    if (sv == KMPM_ERR_ARG) {         // instances not aligned with these tiles
      if (ms) ms[v] = -1;
      continue;
    }
    if ((st = sv) != KMPM_OK) break;
    if ((st = kmpm_step(c, 1)) != KMPM_OK) break;   // first step: cell keys of this variant
    if (hipEventRecord(e0, c->stream) != hipSuccess) st = KMPM_ERR_HIP;
    if (st == KMPM_OK) st = kmpm_step(c, trial_steps);
    if (st == KMPM_OK && hipEventRecord(e1, c->stream) != hipSuccess) st = KMPM_ERR_HIP;
    if (st == KMPM_OK) st = status_from_device(c, nullptr);
    float f = 0;
    if (st == KMPM_OK && hipEventElapsedTime(&f, e0, e1) != hipSuccess) st = KMPM_ERR_HIP;
    if (st != KMPM_OK) break;
    t = (double)f / trial_steps;
    if (ms) ms[v] = t;
    if (t < best_ms) {
      best_ms = t;
      best = v;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  const int keep = st == KMPM_OK ? best : start_variant;
  int st2 = restore_particles(c, saved);
  if (st2 == KMPM_OK) st2 = setup_tiles(c, keep);
  free_saved(saved);
  if (variant) *variant = c->tile_variant;
  return st != KMPM_OK ? st : st2;
}

// ---- captured steps -----------------------------------------------------------
static HostStepState save_state(const kmpm_ctx *c) {
  HostStepState s;
  memcpy(&s.a, &c->a, sizeof(KArgs));
  s.keys_valid = c->keys_valid;
  s.sorted = c->sorted;
  s.tl_identity = c->tl_identity;
  return s;
}

static void load_state(kmpm_ctx *c, const HostStepState &s) {
  memcpy(&c->a, &s.a, sizeof(KArgs));
  c->d_pa = s.a.pa;
  c->d_pb = s.a.pb;
  c->d_fa = s.a.fa;
  c->d_fb = s.a.fb;
  c->d_ia = s.a.ia;
  c->d_ib = s.a.ib;
  c->keys_valid = s.keys_valid;
  c->sorted = s.sorted;
  c->tl_identity = s.tl_identity;
}

static bool same_state(const HostStepState &x, const HostStepState &y) {
  return memcmp(&x.a, &y.a, sizeof(KArgs)) == 0 && x.keys_valid == y.keys_valid &&
         x.sorted == y.sorted && x.tl_identity == y.tl_identity;
}

// Captures `steps` steps of kmpm_step from the current host state into a
// graph (nothing runs); leaves the host state at the end of those steps.
static int capture_steps(kmpm_ctx *c, int64_t steps, kmpm_ctx::StepGraph &out) {
  out.start = save_state(c);
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  int st = kmpm_step(c, steps);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (st != KMPM_OK || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    load_state(c, out.start);
    return st != KMPM_OK ? st : KMPM_ERR_HIP;
  }
  out.end = save_state(c);
  const hipError_t ei = hipGraphInstantiate(&out.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    out.exec = nullptr;
    load_state(c, out.start);
    return KMPM_ERR_HIP;
  }
  out.steps = steps;
  load_state(c, out.start);
  return KMPM_OK;
}

int kmpm_step_graph(kmpm_ctx *c, int64_t nsteps, int64_t graph_steps) {
  if (!c || nsteps < 0 || graph_steps < 1 || c->timing) return KMPM_ERR_ARG;
  int st;
  // steps whose launches differ from the steady state (first cell keys,
  // TLMPM's first sort) run eagerly
  while (nsteps > 0 && (!c->keys_valid || (c->a.tl && !c->tl_identity))) {
    if ((st = kmpm_step(c, 1))) return st;
    nsteps--;
  }
  while (nsteps >= graph_steps) {
    const HostStepState now = save_state(c);
    kmpm_ctx::StepGraph *g = nullptr;
    for (auto &q : c->graphs)
      if (q.exec && q.steps == graph_steps && same_state(q.start, now)) g = &q;
    if (!g) {
      // replace the entry that does not end where this state starts (an
      // odd step count alternates between two start states)
      g = &c->graphs[0];
      for (auto &q : c->graphs)
        if (!q.exec) g = &q;
      if (g->exec && c->graphs[1].exec) {
        g = same_state(c->graphs[0].end, now) ? &c->graphs[1] : &c->graphs[0];
      }
      if (g->exec) {
        (void)hipGraphExecDestroy(g->exec);
        g->exec = nullptr;
      }
      if ((st = capture_steps(c, graph_steps, *g))) return st;
    }
    HIPCHK(hipGraphLaunch(g->exec, c->stream));
    load_state(c, g->end);
    nsteps -= graph_steps;
  }
  return kmpm_step(c, nsteps);
}

int kmpm_run_time(kmpm_ctx *c, double duration, int64_t max_steps, int64_t *steps_done) {
  if (steps_done) *steps_done = 0;
  if (!c || max_steps < 0 || c->a.ninst > 0) return KMPM_ERR_ARG;
  DevScalars sc;
  auto read = [&]() -> int {
    HIPCHK(hipMemcpyAsync(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return KMPM_OK;
  };
  int st = read();
  if (st) return st;
  const double time_end = duration + sc.time;          // run_time.cpp:43
  char lit[512];                                       // std::to_string(double): "%f"
  std::snprintf(lit, sizeof(lit), "%f", time_end);
  const double cond = (double)std::strtof(lit, nullptr);   // input.cpp:407-411 (stof)
  int64_t n = 0;
  while (n < max_steps && sc.time < cond && sc.error == 0) {
    if ((st = kmpm_step(c, 1))) break;
    n++;
    if ((st = read())) break;
    if (sc.time > time_end) break;                     // usl.cpp:84-88
  }
  if (steps_done) *steps_done = n;
  return st;
}

int kmpm_get_time(kmpm_ctx *c, double *dt, double *time, int64_t *step) {
  if (!c) return KMPM_ERR_ARG;
  DevScalars sc;
  HIPCHK(hipMemcpyAsync(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (dt) *dt = sc.dt;
  if (time) *time = sc.time;
  if (step) *step = sc.step;
  return KMPM_OK;
}

int kmpm_reduce_energy(kmpm_ctx *c, uint32_t bit, double *ek, double *es) {
  if (!c || !ek || !es) return KMPM_ERR_ARG;
  hipError_t e = launch_energy(c, bit, c->d_red);
  double h[2] = {0, 0};
  if (e == hipSuccess)
    e = hipMemcpyAsync(h, c->d_red, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return KMPM_ERR_HIP;
  *ek = h[0];
  *es = h[1];
  return KMPM_OK;
}

int kmpm_download_grid(kmpm_ctx *c, double *mass, double *v, double *vu, int64_t nn) {
  if (!c || nn != c->a.nn) return KMPM_ERR_ARG;
  // the reference order and CPDI write every node each step
  if (!c->a.ref_order && !c->a.cpdi) HIPCHK(launch_grid_clean(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  // device layout [comp][node] (FP32 values in the mixed path) -> reference
  // layout [node][comp]
  const bool mixed = c->a.prec == KMPM_PRECISION_MIXED;
  const size_t esz = mixed ? sizeof(float) : sizeof(double);
  const int64_t ns = c->a.ns;
  std::vector<char> raw(3 * ns * esz);
  auto val = [&](size_t k) {
    return mixed ? (double)reinterpret_cast<const float *>(raw.data())[k]
                 : reinterpret_cast<const double *>(raw.data())[k];
  };
  std::vector<int64_t> sidx(nn);   // tile-major storage -> tag order
  for (int64_t i = 0; i < nn; i++) sidx[i] = host_sidx(c->a, i);
  if (mass) {
    HIPCHK(hipMemcpy(raw.data(), c->d_gmass, ns * esz, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nn; i++) mass[i] = val(sidx[i]);
  }
  double *dsts[2] = {v, vu};
  double *srcs[2] = {c->d_gv, c->d_gvu};
  for (int k = 0; k < 2; k++) {
    if (!dsts[k]) continue;
    HIPCHK(hipMemcpy(raw.data(), srcs[k], 3 * ns * esz, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nn; i++)
      for (int d = 0; d < 3; d++) dsts[k][3 * i + d] = val((size_t)d * ns + sidx[i]);
  }
  return KMPM_OK;
}

}  // extern "C"

// The grid temperature in tag order (thermo-mechanical contexts; zeros
// otherwise), for the grid dump's T column.
int kmpm::download_grid_temperature(kmpm_ctx *c, double *T, int64_t nn) {
  if (!c || !T || nn != c->a.nn) return KMPM_ERR_ARG;
  if (!c->a.thermo) {
    std::fill(T, T + nn, 0.0);
    return KMPM_OK;
  }
  HIPCHK(launch_grid_clean(c));
  std::vector<double> raw(c->a.ns);
  HIPCHK(hipMemcpyAsync(raw.data(), c->d_gT, c->a.ns * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < nn; i++) T[i] = raw[host_sidx(c->a, i)];
  return KMPM_OK;
}

extern "C" {

int kmpm_download_cells(kmpm_ctx *c, int32_t *cell_ijk, int64_t *sorted_ptags, int64_t np) {
  if (!c || np != c->a.np) return KMPM_ERR_ARG;
  if (cell_ijk) {
    int *d = nullptr;
    HIPCHK(hipMalloc(&d, 3 * std::max<int64_t>(np, 1) * sizeof(int)));
    hipError_t e = launch_cells(c, d);
    if (e == hipSuccess)
      e = hipMemcpyAsync(cell_ijk, d, 3 * np * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return KMPM_ERR_HIP;
  }
  if (sorted_ptags) {
    if (!c->sorted) return KMPM_ERR_ARG;
    std::vector<int> order(np), pid((size_t)c->a.cap);
    HIPCHK(hipMemcpyAsync(order.data(), c->d_order, np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(pid.data(), c->d_ia + (size_t)PID * c->a.cap, c->a.cap * sizeof(int),
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t j = 0; j < np; j++) sorted_ptags[j] = c->h_ptag[pid[order[j]]];
  }
  return KMPM_OK;
}

int kmpm_timing_enable(kmpm_ctx *c, int enable) {
  if (!c) return KMPM_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!c->ev_pool) {
    c->ev_pool = new std::vector<hipEvent_t>();
    c->ev_phase = new std::vector<int>();
  }
  if (enable) c->ev_phase->clear();
  c->timing = enable != 0;
  return KMPM_OK;
}

int kmpm_timing_read(kmpm_ctx *c, double ms[KMPM_NPHASE], int64_t launches[KMPM_NPHASE]) {
  if (!c || !ms || !launches) return KMPM_ERR_ARG;
  for (int k = 0; k < KMPM_NPHASE; k++) {
    ms[k] = 0;
    launches[k] = 0;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!c->ev_phase) return KMPM_OK;
  for (size_t i = 0; i < c->ev_phase->size(); i++) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, (*c->ev_pool)[2 * i], (*c->ev_pool)[2 * i + 1]));
    const int ph = (*c->ev_phase)[i];
    ms[ph] += t;
    launches[ph] += 1;
  }
  return KMPM_OK;
}

// The staging buffer of a learned constitutive update ([8][cap] floats: the
// rate of deformation, the EOS pressure and the wave speed of each particle).
static int ensure_learned_stage(kmpm_ctx *c) {
  if (c->d_cut_stage && c->cut_stage_cap >= c->a.cap) return KMPM_OK;
  if (c->d_cut_stage) (void)hipFree(c->d_cut_stage);
  c->d_cut_stage = nullptr;
  c->cut_stage_cap = 0;
  c->a.cut_stage = nullptr;
  const int cap = c->a.cap > 0 ? c->a.cap : 1;
  HIPCHK(dalloc(&c->d_cut_stage, (size_t)8 * cap));
  c->cut_stage_cap = cap;
  c->a.cut_stage = c->d_cut_stage;
  return KMPM_OK;
}

int kmpm_ctx_set_surrogate(kmpm_ctx *c, const kmpm_surrogate *s) {
  if (!c) return KMPM_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  if (!s) {
    c->a.surr_on = 0;
    return KMPM_OK;
  }
  // one network pair serves the whole wave, so every material must be the
  // surrogate's JC + shock material; one instance (the constitutive kernel
  // reduces the context's CFL)
  if (c->a.prec != KMPM_PRECISION_MIXED || c->a.mat_class != MC_JC_SHOCK || c->a.tl ||
      c->a.ninst || c->a.xf || c->a.thermo || c->a.cpdi)
    return KMPM_ERR_ARG;
  for (int i = 0; i < c->nmat; i++)
    if (!surrogate_matches(s, c->mats_host[i])) return KMPM_ERR_ARG;
  SurrogateDev img;
  int st = surrogate_device_image(s, &img);
  if (st) return st;
  st = ensure_learned_stage(c);
  if (st) return st;
  if (!c->d_surr) HIPCHK(dalloc(&c->d_surr, 1));
  HIPCHK(hipMemcpy(c->d_surr, &img, sizeof(img), hipMemcpyHostToDevice));
  c->a.surr = c->d_surr;
  c->a.surr_on = 1;
  return KMPM_OK;
}

int kmpm_ctx_set_cutnet(kmpm_ctx *c, const kmpm_cutnet *s) {
  if (!c) return KMPM_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  if (!s) {
    c->a.cut_on = 0;
    return KMPM_OK;
  }
  // one network serves every particle of the wave: one material, the cut
  // network's; USL on one instance of the mixed ULMPM path
  if (c->a.prec != KMPM_PRECISION_MIXED || c->a.mat_class != MC_JC_SHOCK || c->a.tl ||
      c->a.ref_order || c->a.ninst || c->a.xf || c->a.thermo || c->a.cpdi ||
      c->cfg.scheme != KMPM_SCHEME_USL ||
      c->nmat != 1 ||
      !cutnet_matches(s, c->mats_host[0]))
    return KMPM_ERR_ARG;
  if (!c->d_cut) HIPCHK(dalloc(&c->d_cut, 1));
  int st = ensure_learned_stage(c);
  if (st) return st;
  st = cutnet_device_image(s, c->d_cut, c->stream);
  if (st) return st;
  c->a.cut = c->d_cut;
  c->a.cut_on = 1;
  return KMPM_OK;
}

int kmpm_cutnet_stats(kmpm_ctx *c, int64_t *evals, int64_t *fallbacks, int64_t *judged,
                      double *judge_err2, double *judge_ref2) {
  if (!c || !evals || !fallbacks || !judged || !judge_err2 || !judge_ref2) return KMPM_ERR_ARG;
  DevScalars sc;
  HIPCHK(hipMemcpyAsync(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *evals = (int64_t)sc.surr_evals;
  *fallbacks = (int64_t)sc.surr_fallbacks;
  *judged = (int64_t)sc.judge_n;
  *judge_err2 = sc.judge_err2;
  *judge_ref2 = sc.judge_ref2;
  const unsigned long long zc[2] = {0, 0};
  HIPCHK(hipMemcpyAsync((char *)c->d_sc + offsetof(DevScalars, surr_evals), zc, sizeof(zc),
                        hipMemcpyHostToDevice, c->stream));
  struct {
    unsigned long long n;
    double e, r;
  } zj = {0, 0.0, 0.0};
  static_assert(offsetof(DevScalars, judge_err2) == offsetof(DevScalars, judge_n) + 8 &&
                    offsetof(DevScalars, judge_ref2) == offsetof(DevScalars, judge_n) + 16,
                "judge counters are contiguous");
  HIPCHK(hipMemcpyAsync((char *)c->d_sc + offsetof(DevScalars, judge_n), &zj, 24,
                        hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return KMPM_OK;
}

int kmpm_surrogate_stats(kmpm_ctx *c, int64_t *evals, int64_t *fallbacks) {
  if (!c || !evals || !fallbacks) return KMPM_ERR_ARG;
  DevScalars sc;
  HIPCHK(hipMemcpyAsync(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *evals = (int64_t)sc.surr_evals;
  *fallbacks = (int64_t)sc.surr_fallbacks;
  sc.surr_evals = 0;
  sc.surr_fallbacks = 0;
  HIPCHK(hipMemcpyAsync((char *)c->d_sc + offsetof(DevScalars, surr_evals), &sc.surr_evals,
                        2 * sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return KMPM_OK;
}

int kmpm_copy_probe(int64_t bytes, int reps, double *gbps) {
  if (bytes < 1024 || reps < 1 || !gbps) return KMPM_ERR_ARG;
  return kmpm::copy_probe(bytes, reps, gbps) == hipSuccess ? KMPM_OK : KMPM_ERR_HIP;
}

int kmpm_get_tile_shape(kmpm_ctx *c, int tile[3]) {
  if (!c || !tile) return KMPM_ERR_ARG;
  for (int d = 0; d < 3; d++) tile[d] = c->a.tg.T[d];
  return KMPM_OK;
}

int kmpm_debug_weights(kmpm_ctx *c, int64_t *tag, double *w, double *wd, int64_t np) {
  if (!c || !tag || !w || !wd || np != c->a.np) return KMPM_ERR_ARG;
  const int S = c->a.dim == 3 ? c->a.S1 * c->a.S1 * c->a.S1 : c->a.S1 * c->a.S1;
  const size_t n = (size_t)std::max<int64_t>(np, 1) * S;
  int64_t *dt = nullptr;
  double *dw = nullptr, *dwd = nullptr;
  HIPCHK(hipMalloc(&dt, n * sizeof(int64_t)));
  if (hipMalloc(&dw, n * sizeof(double)) != hipSuccess ||
      hipMalloc(&dwd, 3 * n * sizeof(double)) != hipSuccess) {
    (void)hipFree(dt);
    if (dw) (void)hipFree(dw);
    return KMPM_ERR_HIP;
  }
  hipError_t e = launch_debug_weights(c, dt, dw, dwd);
  const size_t m = (size_t)np * S;
  if (e == hipSuccess) e = hipMemcpyAsync(tag, dt, m * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(w, dw, m * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(wd, dwd, 3 * m * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dt);
  (void)hipFree(dw);
  (void)hipFree(dwd);
  return e == hipSuccess ? KMPM_OK : KMPM_ERR_HIP;
}

}  // extern "C"

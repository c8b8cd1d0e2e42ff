// kmpm_ctx.h -- the device context behind the C ABI.
#pragma once

#include <vector>

#include "kmpm_device.h"

namespace kmpm {

struct SurrogateDev;
struct CutNetDev;

// Tile geometry shared by the sort key, P2G, grid update and G2P kernels.
struct TileGeom {
  int T[3];        // tile size in cells (T[2] == 1 in 2D)
  int nt[3];       // tiles per axis
  int ntiles;      // nt[0]*nt[1]*nt[2]
  int cpt;         // cells per tile
  int E[3];        // node block extent per tile: T + S1 - 1 (1 in unused z)
  int bn;          // nodes per block
  int off;         // block origin = tile origin - off (1 cubic, 0 linear)
};

// Pointers and sizes handed to every kernel by value.
struct KArgs {
  // problem
  int dim, S1, scheme;
  int n[3];
  int64_t nn;
  double lo[3], hi[3];
  double h, inv_h, flip, dt_factor;
  int dt_constant;
  int nmat;
  int mat_class;    // MatClass the stress kernels are specialised for
  const DevMat *mats;
  TileGeom tg;
  // particles
  int tl;           // TLMPM: weights from x0 on the solid's grid, F += dt Fdot
  int lay;          // state Layout (LAY_UL, LAY_TL, LAY_AP)
  // APIC family (kmpm_config.sub_method): xf = 1 APIC transfers (affine
  // P2G momentum, rate of deformation from the affine moments), 2 APIC with
  // the MLS force; asflip: x advanced by the blended velocity; di: the
  // diagonal of the inertia tensor Di (solid.cpp:1440-1470) in the axes of
  // the dimension
  int xf, asflip;
  double di;
  int ref_order;    // KMPM_ORDER_REFERENCE: the reference-order kernels
  const int *solid_of;   // solid index by pid
  int prec;         // KMPM_PRECISION_*
  int np, cap;
  double *pa;       // current state: state_nd(mixed, lay) * cap doubles
  float *fa;        // current FP32 state (mixed: state_nf(true, lay) * cap)
  int *ia;          // current int state (NIF * cap)
  double *pb;       // next state
  float *fb;
  int *ib;
  double fxs[FX_N], fxi[FX_N];  // fixed-point scales (powers of two) and inverses
  double *xo;       // MUSL: positions at the start of the step, sorted slots (3*cap)
  int *key;         // cell key per current slot
  int *order_tmp;   // counting-sort scatter: slots grouped by tile, in arrival order
  int *sort_tmp;    // cell-ordered slots of tiles too large for LDS
  int *order;       // sorted position -> current slot
  // tiles
  int *tile_np;     // [ntiles + 1] particles per tile
  int *tile_p0;     // [ntiles + 1] first sorted position of each tile
  int *tile_cur;    // [ntiles] scatter cursor of each tile
  int *tile_list;   // active tiles (by list slot)
  int *tile_slot;   // tile -> list slot or -1
  int *dil_list;    // tiles whose owned nodes get contributions
  double *scratch;  // partial node blocks [slot][field][bn] (int64 in the mixed path)
  int scratch_slots;
  // grid, tile-major (node_sidx), SoA per component, ns entries per
  // component (whole tiles: edge tiles are padded); the mixed path stores
  // FP32 values in the same allocations
  int64_t ns;
  double *gmass, *gv, *gvu;   // gv/gvu: 3 * ns
  // thermo-mechanical (FP64): grid T and T_update [ns], and the thermal P2G's
  // partial blocks [slot][3][bn] (sum w m T, sum w gamma, sum grad w . q)
  int thermo;
  double *gT, *gTu, *scratch_th;
  // CPDI (FP64, 2D): 1 + KMPM_CPDI_* or 0; the box's lower corner (the
  // reference takes the candidate nodes from it on the solid's grid too);
  // the node sums of the particle-major P2G [7][ns] (mass, momentum, force)
  int cpdi;
  double blo[3];
  double *cacc;
  uint32_t *gmask;
  // fixes
  int nfv, nbf;
  DevFixVel fv[MAX_FIX];
  DevFixBF bf[MAX_FIX];
  const kmpm_expr *fexpr;   // [MAX_FIX][6] programs of expression-valued velocity fixes
  // batched instances (0: one instance, whose dt/time are in sc)
  int ninst;
  InstScalars *inst;     // [ninst]
  const int *tinst;      // tile -> instance, -1 for tiles no instance reaches
  // learned constitutive surrogate (JC + shock class, mixed path)
  const SurrogateDev *surr;
  int surr_on;
  // cut network (JC + shock + JC damage class, mixed path, USL): the fused
  // G2P stages the cut's inputs (rate of deformation, pressure, wave speed)
  // in cut_stage [8][cap] and a separate kernel runs the network
  const CutNetDev *cut;
  float *cut_stage;
  int cut_on;
  // scalars
  DevScalars *sc;
};

// Host-side state that selects what a step launches: the kernel arguments
// (with the current/next buffers) and the phase flags.
struct HostStepState {
  KArgs a;
  bool keys_valid, sorted, tl_identity;
};

}  // namespace kmpm

struct kmpm_ctx {
  kmpm_config cfg;
  kmpm::KArgs a;
  int device;
  int cu_count;
  hipStream_t stream;
  int nmat;
  kmpm_material *mats_host;   // [nmat]
  // static per-particle data in upload order (index = pid)
  int64_t *h_ptag;
  int32_t *h_solid;
  uint32_t *h_pmask;
  double *h_x0;
  // device allocations
  kmpm::DevMat *d_mats;
  double *d_pa, *d_pb, *d_xo;
  float *d_fa, *d_fb;
  int *d_ia, *d_ib;
  int *d_key, *d_order;
  int *d_order_tmp;
  int *d_sort_tmp;
  int *d_tile_np, *d_tile_p0, *d_tile_cur;
  int *d_tile_list, *d_tile_slot, *d_dil_list;
  double *d_scratch;
  double *d_gmass, *d_gv, *d_gvu;
  double *d_gT, *d_gTu, *d_scratch_th;
  double *d_cacc;
  uint32_t *d_gmask;
  uint32_t *d_pmask;     // particle group masks (by pid)
  kmpm_expr *d_fexpr;    // [MAX_FIX][6]
  int *d_solid;          // solid index (by pid)
  kmpm::DevScalars *d_sc;
  kmpm::SurrogateDev *d_surr;
  kmpm::CutNetDev *d_cut;
  float *d_cut_stage;     // [8][cut_stage_cap]
  int cut_stage_cap;
  kmpm::InstScalars *d_inst;
  int *d_tinst;
  std::vector<int32_t> *inst_cells;   // per instance: cell_lo[3], cell_hi[3]
  std::vector<uint32_t> *h_gmask;     // node group masks in tag order
  int tile_variant;                   // Tiles<dim, S1, variant> of the tile kernels
  // captured step graphs: each replays `steps` steps from the host state
  // it was captured in (buffer ping-pong and flags) to its end state
  struct StepGraph {
    hipGraphExec_t exec;
    int64_t steps;
    kmpm::HostStepState start, end;
  } graphs[2];
  void *d_scan_tmp;
  size_t scan_tmp_bytes;
  int64_t ncells;
  // phase timing (HIP events on `stream`)
  bool timing;
  std::vector<hipEvent_t> *ev_pool;   // recorded pairs: start, stop
  std::vector<int> *ev_phase;         // phase of each pair
  bool upload_failed;    // an upload failed after freeing the previous particles:
                         // steps are rejected until an upload succeeds
  double *d_red;         // energy reduction: 2 results + 2 per reduction block
  bool keys_valid;       // d_key holds keys of the current positions
  bool sorted;           // order/tiles valid for the current step
  bool tl_identity;      // TLMPM: the state is in its final sorted order (keys
                         // come from x0 and never change)
};

namespace kmpm {
// The USL step's constitutive update is a learned one (the surrogate's
// networks or the cut network, JC + shock class on the mixed ULMPM path),
// run by k_stress_learned after the G2P kernel staged its inputs.
static inline bool learned_constitutive(const kmpm_ctx *c) {
  return (c->a.cut_on || c->a.surr_on) && c->a.mat_class == MC_JC_SHOCK &&
         c->a.prec == KMPM_PRECISION_MIXED && !c->a.tl;
}

// Passes of the grid update kernel: the step's update (GM_STEP), the MUSL
// re-map of velocities (GM_REMAP), and USF's two passes, node mass and
// velocity after its first P2G (GM_USF1) and the update with the forces of
// its second P2G (GM_USF2).
enum GridMode { GM_STEP = 0, GM_REMAP = 1, GM_USF1 = 2, GM_USF2 = 3 };
// Passes of the grid temperature kernel (thermo-mechanical runs): T, the
// driving forces and T_update (TM_FULL), or T alone (TM_T: the MUSL re-map
// and USF's first pass).
enum ThermalMode { TM_FULL = 0, TM_T = 1 };
// CPDI G2P kernels: USL's fused G2P, F and stress; MUSL's part A (G2P and
// advance); USF's G2P (after its in-place stress pass)
enum CpdiG2P { CG_USL = 0, CG_MUSL = 1, CG_USF = 2 };

// launchers (kmpm_kernels.hip); all stream-ordered, return hipError_t
hipError_t launch_keys(kmpm_ctx *c);
hipError_t launch_sort(kmpm_ctx *c);
hipError_t launch_p2g(kmpm_ctx *c, int momentum_only);
hipError_t launch_grid_update(kmpm_ctx *c, int mode);   // GridMode (kmpm_kernels.h)
hipError_t launch_p2g_thermal(kmpm_ctx *c, int remap);
hipError_t launch_grid_thermal(kmpm_ctx *c, int mode);  // ThermalMode
// CPDI (kmpm_cpdi.hip): P2G (remap: the MUSL momentum re-map; saved: USF's
// second, whose weights divide by the volume its stress pass kept), the
// grid pass (GridMode), the G2P (CpdiG2P) and the MUSL / USF stress pass
hipError_t launch_cpdi_p2g(kmpm_ctx *c, int remap, int saved);
hipError_t launch_cpdi_grid(kmpm_ctx *c, int mode);
hipError_t launch_cpdi_g2p(kmpm_ctx *c, int kind);
hipError_t launch_cpdi_stress(kmpm_ctx *c);
hipError_t launch_fix_values(kmpm_ctx *c);   // expression-valued fixes, before a grid pass
hipError_t launch_g2p_usl(kmpm_ctx *c);
hipError_t launch_g2p_musl_a(kmpm_ctx *c);
hipError_t launch_stress_musl_b(kmpm_ctx *c);
hipError_t launch_learned_stress(kmpm_ctx *c);   // kmpm_cutnet.hip
hipError_t launch_adjust_dt(kmpm_ctx *c);
hipError_t launch_grid_clean(kmpm_ctx *c);
int download_grid_temperature(kmpm_ctx *c, double *T, int64_t nn);   // tag order
hipError_t launch_energy(kmpm_ctx *c, uint32_t bit, double *d_out2);
hipError_t launch_debug_weights(kmpm_ctx *c, int64_t *d_tag, double *d_w, double *d_wd);
hipError_t launch_cells(kmpm_ctx *c, int *d_ijk);
hipError_t launch_iota(int *d, int n, hipStream_t stream);
// reference-order FP64 kernels (kmpm_reforder.hip)
hipError_t launch_ref_p2g(kmpm_ctx *c, int remap);
hipError_t launch_ref_grid(kmpm_ctx *c, int remap);
hipError_t launch_ref_g2p_usl(kmpm_ctx *c);
hipError_t launch_ref_g2p_musl_a(kmpm_ctx *c);
hipError_t launch_ref_stress_musl_b(kmpm_ctx *c);
size_t scan_temp_bytes(int64_t n);
DevMat make_devmat(const kmpm_material &m);
int surrogate_device_image(const kmpm_surrogate *s, SurrogateDev *out);
bool surrogate_matches(const kmpm_surrogate *s, const kmpm_material &m);
bool cutnet_matches(const kmpm_cutnet *s, const kmpm_material &m);
int cutnet_device_image(const kmpm_cutnet *s, CutNetDev *dst_dev, hipStream_t stream);
hipError_t copy_probe(int64_t bytes, int reps, double *gbps);
}  // namespace kmpm

/*
 * kmpm.h -- C ABI of the MI355X-native explicit MPM step (libkmpm.so).
 *
 * Plain C types only: pointers, sizes, POD structs.  Every entry point
 * returns an int status (KMPM_OK == 0); kmpm_status_string() maps it to the
 * reference's error text.  A context owns every device buffer; the host
 * passes copies in and receives copies out and no host pointer is retained.
 *
 * The entry points replace, one for one, the virtual step interface of the
 * reference's Method class and the per-particle constitutive plugin calls:
 *   Method virtuals ............ src/method.h:33-48
 *   ULMPM implementation ....... src/ulmpm.cpp:88-563
 *   Scheme loops ............... src/usl.cpp:43-93, src/musl.cpp:43-94,
 *                                src/usf.cpp:34-94
 *   constitutive plugins ....... src/eos.h:40, src/strength.h:38-45,
 *                                src/damage.h:35-41
 *   plugin registry keys ....... src/material.cpp:40-66, src/style_*.h
 * File:line citations are relative to the reference repository root.
 */
#ifndef KMPM_H
#define KMPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  KMPM_OK = 0,
  KMPM_ERR_ARG = 1,            /* invalid argument / unsupported option */
  KMPM_ERR_HIP = 2,            /* HIP runtime error */
  KMPM_ERR_NO_DEVICE = 3,      /* no usable gfx950 device */
  KMPM_ERR_CAPACITY = 4,       /* output buffer too small */
  KMPM_ERR_PARTICLE_LEFT_BOX = 10,  /* solid.cpp:617-627 */
  KMPM_ERR_JACOBIAN = 11,           /* J<=0 with damage<1, solid.cpp:1208-1215 */
  KMPM_ERR_DTCFL = 12,              /* dtCFL NaN or 0, ulmpm.cpp:535-544 */
  KMPM_ERR_WAVE_SPEED = 13,         /* NaN wave speed, solid.cpp:1387-1398 */
  KMPM_ERR_GRID = 14,               /* grid too small for the stencil */
  KMPM_ERR_FIXED_POINT = 15,        /* mixed path: a node sum left the range of
                                       its 64-bit fixed-point scale (no
                                       counterpart in the reference, which
                                       accumulates in FP64) */
  KMPM_ERR_INTERNAL = 16,           /* a particle was found outside the tile it
                                       was sorted into (an internal
                                       inconsistency; its access is clamped) */
  KMPM_ERR_INSTANCE = 17,           /* batched context: a particle's stencil
                                       reaches outside its instance's tiles
                                       (no counterpart in the reference) */
  KMPM_ERR_DOMAIN = 18              /* CPDI: a particle domain spans more than
                                       KMPM_CPDI_MAX_SPAN nodes on an axis (no
                                       counterpart in the reference, whose
                                       neighbour lists grow without bound) */
};

/* ---- options ----------------------------------------------------------- */
/* method(ulmpm|tlmpm, ...) -> update.cpp:108-196, style_method.h */
/* ULCPDI and TLCPDI: the CPDI methods (ulcpdi.cpp, tlcpdi.cpp; 2D, the
 * style in kmpm_config.cpdi_style) */
enum { KMPM_METHOD_ULMPM = 0, KMPM_METHOD_TLMPM = 1, KMPM_METHOD_ULCPDI = 2,
       KMPM_METHOD_TLCPDI = 3 };
/* CPDI styles, ulcpdi.h:59 known_styles {"R4", "Q4"}: R4 particle domains
 * are parallelograms spanned by two vectors carried by F; Q4 domains are
 * quadrilaterals whose corners move with the grid velocity */
enum { KMPM_CPDI_R4 = 0, KMPM_CPDI_Q4 = 1 };
/* Nodes per axis a CPDI particle domain's corner stencils may span (the
 * union of the corners' stencils); wider domains stop the step with
 * KMPM_ERR_DOMAIN */
enum { KMPM_CPDI_MAX_SPAN = 8 };
/* scheme(usl|musl|usf) -> update.cpp:83-101, style_scheme.h (usf.cpp:34-94) */
enum { KMPM_SCHEME_USL = 0, KMPM_SCHEME_MUSL = 1, KMPM_SCHEME_USF = 2 };
/* velocity-update sub-method, the second word of method(...) (update.cpp:
 * 129-146; values as Update::SubMethodType, update.h:62-69).
 *   PIC, FLIP  blend v by pic_flip (PIC is pic_flip 0)
 *   APIC       affine P2G momentum v + C (x_I - x_p) with C = L (ULMPM) or
 *              Fdot (TLMPM), rate of deformation from the affine moments
 *              D^-1 sum v_I (x_I - x_p)^T w; pic_flip forced to 0
 *              (ulmpm.cpp:79-85, tlmpm.cpp:82-84, solid.cpp:392-426,1006-1153)
 *   AFLIP      APIC transfers with FLIP blending (ULMPM)
 *   ASFLIP     AFLIP with x advanced by the blended v (solid.cpp:637-685,
 *              791-796; ULMPM)
 *   MLS        APIC with the MLS internal force -vol w sigma D^-1 (x_I - x_p)
 *              (solid.cpp:524-574; ULMPM)
 * TLMPM implements PIC/FLIP and APIC only; the other values run as FLIP
 * there, as in the reference.  The APIC family needs cubic splines on
 * ULMPM (solid.cpp:1448-1452) and, with linear shape functions on TLMPM,
 * np_per_cell 1 or 2 (solid.cpp:1454-1462). */
enum { KMPM_SUB_PIC = 0, KMPM_SUB_FLIP = 1, KMPM_SUB_APIC = 2, KMPM_SUB_AFLIP = 3,
       KMPM_SUB_ASFLIP = 4, KMPM_SUB_MLS = 5 };
/* shape functions -> basis_functions.h:21-104 */
enum { KMPM_SHAPE_LINEAR = 0, KMPM_SHAPE_CUBIC_SPLINE = 1 };
/* arithmetic of the device step: FP64 throughout, or mixed (FP32 state,
 * transfers and constitutive update; FP64 position, deformation gradient,
 * J and pressure; 64-bit fixed-point node mass, momentum and force sums),
 * for both methods */
enum { KMPM_PRECISION_FP64 = 0, KMPM_PRECISION_MIXED = 1 };
/* summation order of the FP64 step: the fast order of the tiled kernels, or
 * the reference's order (node-major P2G over each node's particles in
 * ascending index with the per-solid velocity sums, solid.cpp:317-522;
 * per-particle gathers in the candidate order of ulmpm.cpp:167-265), which
 * reproduces the reference's roundings.  The reference order costs
 * O(nodes x particles) per P2G: it is for parity runs of small problems
 * (ULMPM, FP64, one instance). */
enum { KMPM_ORDER_FAST = 0, KMPM_ORDER_REFERENCE = 1 };

typedef struct kmpm_config {
  int dimension;          /* 2 or 3 (domain.cpp:496-551) */
  int method;             /* KMPM_METHOD_* */
  int scheme;             /* KMPM_SCHEME_* (reference default: MUSL, update.cpp:42-45) */
  int shape;              /* KMPM_SHAPE_* */
  int precision;          /* KMPM_PRECISION_* */
  int order;              /* KMPM_ORDER_* (FP64 only) */
  double boxlo[3], boxhi[3];
  double cellsize;        /* grid cell size h */
  double pic_flip;        /* FLIP blending alpha (update.cpp:163-166) */
  double dt;              /* dt of the first step (1e-16 unless set_dt, update.cpp:38) */
  int    dt_constant;     /* nonzero after set_dt (update.cpp:63-70) */
  int    sub_method;      /* KMPM_SUB_* (PIC and FLIP both blend by pic_flip) */
  double dt_factor;       /* dt_factor (default 0.9, update.cpp:40) */
  /* TLMPM: limits of the solid's region (Region::limits, see
   * kmpm_region_limits), the extent of the solid's own grid
   * (Grid::init(solidlo, solidhi), grid.cpp:91-97, solid.cpp:1833-1838,1927-1930).
   * Unused by ULMPM, whose grid covers the box. */
  double solidlo[3], solidhi[3];
  /* particles per cell per axis of the solid (solid.cpp:2027), read only by
   * the APIC inertia of linear shape functions (solid.cpp:1454-1462) */
  int    np_per_cell;
  /* method(..., thermo-mechanical) (update.cpp:168-180): particle and grid
   * temperatures, heat sources and fluxes (solid.cpp:2743-2839,
   * grid.cpp:1354-1362); FP64, one solid per context */
  int    thermo;
  /* ULCPDI / TLCPDI: KMPM_CPDI_R4 or KMPM_CPDI_Q4 (ulcpdi.cpp:49-108) */
  int    cpdi_style;
} kmpm_config;

/* ---- materials: material(...) + eos/strength/damage styles ------------- */
/* material types, material.cpp:250-340 */
enum { KMPM_MAT_LINEAR = 0, KMPM_MAT_EOS_STRENGTH = 1 };
/* eos(id, <key>, ...) keys: linear, shock, fluid (style_eos.h) */
enum { KMPM_EOS_NONE = 0, KMPM_EOS_LINEAR = 1, KMPM_EOS_SHOCK = 2, KMPM_EOS_FLUID = 3 };
/* strength(id, <key>, ...) keys: linear, johnson_cook, fluid (style_strength.h) */
enum { KMPM_STRENGTH_NONE = 0, KMPM_STRENGTH_LINEAR = 1,
       KMPM_STRENGTH_JOHNSON_COOK = 2, KMPM_STRENGTH_FLUID = 3 };
/* damage(id, <key>, ...) keys: damage_johnson_cook (style_damage.h) */
enum { KMPM_DAMAGE_NONE = 0, KMPM_DAMAGE_JOHNSON_COOK = 1 };
/* temperature(id, <key>, ...) keys: plastic_work (style_temperature.h) */
enum { KMPM_TEMP_NONE = 0, KMPM_TEMP_PLASTIC_WORK = 1 };

typedef struct kmpm_material {
  int type;               /* KMPM_MAT_* */
  int eos, strength, damage;
  /* Parameters in the order the reference constructors parse them:
   *   linear material   lin_p      = rho0, E, nu               (material.cpp:279-289)
   *   eos linear        eos_p      = rho0, K                   (eos_linear.cpp:52-53)
   *   eos shock         eos_p      = rho0, K, c0, S, Gamma, cv, Tr, Q1, Q2 (eos_shock.cpp:53-61)
   *   eos fluid         eos_p      = rho0, K, Gamma            (eos_fluid.cpp:43-45)
   *   strength linear   strength_p = G                         (strength_linear.cpp:50)
   *   strength JC       strength_p = G, A, B, n, epsdot0, C, m, Tr, Tm (strength_jc.cpp:55-63)
   *   strength fluid    strength_p = mu                        (strength_fluid.cpp:35)
   *   damage JC         damage_p   = d1, d2, d3, d4, d5, epsdot0, Tr, Tm (damage_jc.cpp:54-61)
   */
  double lin_p[3];
  double eos_p[9];
  double strength_p[9];
  double damage_p[8];
  /* Derived moduli, filled by kmpm_material_init (Mat ctors, material.cpp:720-769). */
  double rho0, K, G, E, nu, lambda, signal_velocity;
  /* The material's temperature law, the last argument of material(...)
   * (material.cpp:310-358); temp = KMPM_TEMP_* and
   *   temperature plastic_work  temp_p = chi, cp, kappa, alpha, T0, Tm
   *                             (temperature_plastic_work.cpp:50-55)
   * A law makes the stress update temperature dependent (solid.cpp:1301-1351). */
  int    temp, reserved_t;
  double temp_p[6];
} kmpm_material;

/* ---- regions used to populate solids and groups ------------------------ */
enum { KMPM_REGION_BLOCK = 0, KMPM_REGION_CYLINDER = 1 };
typedef struct kmpm_region {
  int kind;               /* KMPM_REGION_* */
  int axis;               /* cylinder axis: 0=x, 1=y, 2=z (2D disks use z) */
  int interior;           /* 1: inside matches (default); 0: outside matches (region.cpp:65-72) */
  int reserved;
  /* block: xlo, xhi, ylo, yhi, zlo, zhi (region_block.cpp:51-135)
   * cylinder: c1, c2, R, lo, hi (region_cylinder.cpp:60-110; 2D: lo=hi=0) */
  double p[6];
} kmpm_region;

/* ---- particle interchange (host side, reference layouts) --------------- */
/* Vectors are [np][3], tensors [np][3][3] row-major, as Eigen::Vector3d /
 * Matrix3d per particle in solid.h:50-97.  On download any NULL pointer is
 * skipped, so zero-initialise the struct and set the fields wanted (fields
 * may be appended at its end).  Particles are returned in ascending ptag
 * order. */
typedef struct kmpm_particles {
  int64_t np;
  int64_t *ptag;
  int32_t *solid;         /* solid index in creation order */
  int32_t *mat;           /* material index */
  uint32_t *mask;         /* group bitmask (group.cpp:65-234) */
  double *x0, *x, *v;
  double *F, *sigma, *strain_el;
  double *mass, *vol0, *vol, *rho0, *rho, *J;
  double *eff_plastic_strain, *eff_plastic_strain_rate;
  double *damage, *damage_init;
  /* thermo-mechanical contexts: temperature T (the solid's T0 at populate,
   * solid.cpp:2305; kmpm_populate writes T0 = 0 and the caller sets it) and,
   * on download, the heat source gamma (solid.cpp:1343-1351); either may be
   * NULL; other contexts ignore T and download zeros */
  double *T, *gamma;
  /* CPDI contexts (solid.h): R4 domain vectors rp0 (reference) and rp
   * (current, F rp0 on ULCPDI, solid.cpp:2338-2352), [np][2][3]; Q4 corners
   * xpc0 and xpc, [np][4][3], counter-clockwise from (-, -)
   * (solid.cpp:2208-2222).  On upload a NULL current field starts from its
   * reference one; other contexts ignore them */
  double *rp0, *rp, *xpc0, *xpc;
} kmpm_particles;

/* ---- host-side set-up (no device needed) -------------------------------- */
const char *kmpm_status_string(int status);
/* Derive rho0/K/G/E/nu/lambda/signal_velocity (Mat ctors, material.cpp:720-769). */
int kmpm_material_init(kmpm_material *mat);
/* Grid node counts per dimension (Grid::init, grid.cpp:151-173, one rank):
 * over the box for ULMPM, over cfg->solidlo/solidhi for TLMPM. */
int kmpm_grid_size(const kmpm_config *cfg, int64_t n_out[3]);
/* Lattice population of a region (Solid::populate, solid.cpp:1810-2336, one rank);
 * the lattice starts at the box for ULMPM and at the region's limits for
 * TLMPM (solid.cpp:1927-1936).
 * Writes up to `capacity` particles into `out` (arrays must be allocated);
 * *np_out receives the count (also when it exceeds capacity -> KMPM_ERR_CAPACITY).
 * ptag = ptag_offset + i + 1 (solid.cpp:2320). */
int kmpm_populate(const kmpm_config *cfg, const kmpm_region *region, int ppc1d,
                  const kmpm_material *mat, int mat_index, int solid_index,
                  int64_t ptag_offset, kmpm_particles *out, int64_t capacity,
                  int64_t *np_out);
/* Region::match for one point (region.cpp:65-72 + the style's inside()). */
int kmpm_region_match(const kmpm_region *region, int dimension, double x, double y, double z);
/* Region::limits: bounding box lo/hi of a region (region_block.cpp:159-168,
 * region_cylinder.cpp:113-137,173-182). */
int kmpm_region_limits(const kmpm_region *region, double lo[3], double hi[3]);

/* ---- device context ------------------------------------------------------ */
typedef struct kmpm_ctx kmpm_ctx;

/* Materials per context (each instance of a batched context may have its own). */
enum { KMPM_MAX_MATERIALS = 1024 };
/* Creates the context on the current HIP device: grid (Grid::init), material
 * table, streams.  Replaces Update::create_method/create_scheme + Domain::set_dimension. */
int kmpm_ctx_create(const kmpm_config *cfg, const kmpm_material *mats, int nmat,
                    kmpm_ctx **out);
int kmpm_ctx_destroy(kmpm_ctx *ctx);
/* Upload all particles (all solids, ordered by solid).  Replaces the Solid
 * particle vectors (solid.h:50-112). */
int kmpm_upload_particles(kmpm_ctx *ctx, const kmpm_particles *p);
int kmpm_download_particles(kmpm_ctx *ctx, kmpm_particles *p);
/* Node group bitmask, one uint32 per node in tag order (Grid::mask, grid.h:94). */
int kmpm_set_node_mask(kmpm_ctx *ctx, const uint32_t *mask, int64_t nnodes);
/* fix(id, velocity_nodes, group, vx, vy, vz): v_update <- value(t),
 * v <- value(t-dt) on the group's nodes (fix_velocity_nodes.cpp:130-271). */
int kmpm_fix_velocity_nodes(kmpm_ctx *ctx, uint32_t groupbit, const int set[3],
                            const double value[3], const double value_prev[3]);
/* fix(id, body_force, group, gx, gy, gz): mb += m*g on nodes with m>0
 * (fix_body_force.cpp:106-179). */
int kmpm_fix_body_force(kmpm_ctx *ctx, uint32_t groupbit, const int set[3],
                        const double value[3]);

/* ---- fix values that depend on time --------------------------------------
 * The input script gives each component of fix velocity_nodes as an
 * expression (fix_velocity_nodes.cpp:64-116) that the reference parses
 * again at every evaluation (Var::result, var.cpp:52-59, Input::parsev,
 * input.cpp:374-735).  kmpm_expr_compile turns such a text into a program
 * of `time` (the time at the start of the step) and `dt` (the input
 * script's dt variable: 0 in the first step unless set_dt, then the step's
 * dt, input.cpp:55, ulmpm.cpp:550) that the device evaluates every step,
 * following the parser's rules: literals read by stof, its operator
 * precedence and unary-minus handling, constant folding, and constants of
 * a time-dependent expression re-read from their "%.15f" text.  previous
 * != 0 first replaces the text "time" by "time - dt" (fix_velocity_nodes.cpp:
 * 69-77), which gives the value the fix imposes on v.  Variables: time, dt,
 * PI and the nvar constants names[]/values[] (user variables of the script;
 * the last definition of a name wins).  Operators + - * / ^ ** e E > >= < <=
 * == != !; functions exp sqrt cos sin tan log atan2. */
enum { KMPM_EXPR_MAX = 32, KMPM_EXPR_STACK = 8 };
enum {
  KMPM_OP_CONST = 0, KMPM_OP_TIME, KMPM_OP_DT,
  KMPM_OP_ADD, KMPM_OP_SUB, KMPM_OP_MUL, KMPM_OP_DIV, KMPM_OP_POW,
  KMPM_OP_GT, KMPM_OP_GE, KMPM_OP_LT, KMPM_OP_LE, KMPM_OP_EQ, KMPM_OP_NE,
  KMPM_OP_ATAN2,
  KMPM_OP_NEG, KMPM_OP_NOT, KMPM_OP_EXP, KMPM_OP_SQRT, KMPM_OP_COS, KMPM_OP_SIN,
  KMPM_OP_TAN, KMPM_OP_LOG
};
typedef struct kmpm_expr {
  int32_t n;                      /* operations in postfix order; 0: not set */
  int32_t op[KMPM_EXPR_MAX];      /* KMPM_OP_* */
  double c[KMPM_EXPR_MAX];        /* value of each KMPM_OP_CONST */
} kmpm_expr;
/* KMPM_ERR_ARG for text the reference's parser rejects or that uses
 * anything beyond the list above; KMPM_ERR_CAPACITY beyond KMPM_EXPR_MAX
 * operations or KMPM_EXPR_STACK stack levels. */
int kmpm_expr_compile(const char *text, int previous, int nvar, const char *const *names,
                      const double *values, kmpm_expr *out);
/* fix velocity_nodes with time-dependent components: value[d] gives
 * v_update(t) and previous[d] the v it sets in the same step (compile both
 * from the same text, previous with previous = 1); n == 0 leaves the
 * component free (NULL in the script).  Not for batched contexts. */
int kmpm_fix_velocity_nodes_expr(kmpm_ctx *ctx, uint32_t groupbit, const kmpm_expr value[3],
                                 const kmpm_expr previous[3]);
/* The reaction force of velocity fix `fix` (0-based, in the order the
 * velocity_nodes fixes were added) in the last completed step:
 * sum over its nodes of m (v_imposed - v_update) / dt, the Vars <id>_x,
 * <id>_y, <id>_z the reference publishes (fix_velocity_nodes.cpp:161-224).
 * Waits for the stream. */
int kmpm_fix_velocity_nodes_force(kmpm_ctx *ctx, int fix, double force[3]);

/* ---- step phases (Method virtuals, method.h:33-48) ------------------------ */
/* compute_grid_weight_functions_and_gradients (ulmpm.cpp:88-337): particle
 * cell indices + deterministic cell sort; weights are evaluated in the
 * kernels that use them. */
int kmpm_compute_grid_weight_functions_and_gradients(kmpm_ctx *ctx);
/* particles_to_grid + post_particles_to_grid fixes (ulmpm.cpp:339-378). */
int kmpm_particles_to_grid(kmpm_ctx *ctx);
/* update_grid_state + post_update_grid_state fixes (ulmpm.cpp:436-440,
 * grid.cpp:448-466). */
int kmpm_update_grid_state(kmpm_ctx *ctx);
/* USL: compute_rate_deformation_gradient(false) + grid_to_points +
 * advance_particles + update_deformation_gradient + update_stress fused
 * (usl.cpp:59-74). */
int kmpm_grid_to_points_and_stress(kmpm_ctx *ctx);
/* MUSL: grid_to_points + advance_particles (musl.cpp:61-67). */
int kmpm_grid_to_points(kmpm_ctx *ctx);
/* MUSL: velocities_to_grid + post_velocities_to_grid (ulmpm.cpp:476-496). */
int kmpm_velocities_to_grid(kmpm_ctx *ctx);
/* MUSL: compute_rate_deformation_gradient(true) + update_deformation_gradient
 * + update_stress (musl.cpp:75-77); USF: the same from the first P2G's grid
 * velocity (usf.cpp:58-60). */
int kmpm_update_stress(kmpm_ctx *ctx);
/* USF: particles_to_grid_USF_1 (node mass and velocity, ulmpm.cpp:380-404)
 * + the post_update_grid_state fixes that follow it (usf.cpp:53-55). */
int kmpm_particles_to_grid_usf_1(kmpm_ctx *ctx);
/* USF: particles_to_grid_USF_2 (the forces of the updated stress,
 * ulmpm.cpp:406-431) + post_particles_to_grid (usf.cpp:62-64); then
 * kmpm_update_grid_state and kmpm_grid_to_points complete the step. */
int kmpm_particles_to_grid_usf_2(kmpm_ctx *ctx);
/* The affine matrix of the APIC transfers (L on ULMPM, Fdot on TLMPM; the
 * state compute_velocity_nodes_APIC reads, solid.cpp:392-426), row-major,
 * 9 per particle in upload order.  KMPM_ERR_ARG for contexts without the
 * APIC transfers.  Waits for the stream. */
int kmpm_download_affine(kmpm_ctx *ctx, double *C, int64_t np);
/* update_time + adjust_dt (update.cpp:201-206, ulmpm.cpp:525-551). */
int kmpm_adjust_dt(kmpm_ctx *ctx);
/* nsteps complete steps of the configured scheme (USL::run / MUSL::run /
 * USF::run),
 * stream-ordered, no host synchronisation inside. */
int kmpm_step(kmpm_ctx *ctx, int64_t nsteps);
/* RunTime::command (run_time.cpp:31-49) with the scheme loop's end test
 * (usl.cpp:43-93, musl.cpp:43-94): time_end = time + duration (duration as
 * the input script's parser gives it); steps run while the time is below
 * the loop condition "time<" + to_string(time_end), whose literal the
 * script parser reads with stof, and the loop breaks after a step whose
 * time exceeds time_end.  At most max_steps steps; *steps_done (may be
 * NULL) receives their number.  The time is read after every step (one
 * host synchronisation per step); a device error ends the loop and is
 * reported by kmpm_synchronize.  Not for batched contexts. */
int kmpm_run_time(kmpm_ctx *ctx, double duration, int64_t max_steps, int64_t *steps_done);
/* Waits for the stream and returns the sticky device error (first offending
 * particle tag in *ptag_out, may be NULL). */
int kmpm_synchronize(kmpm_ctx *ctx, int64_t *ptag_out);

/* ---- batched instances and captured steps ----------------------------------
 * No counterpart in the reference, which runs one simulation per process.
 * A batched context packs many independent simulations ("instances") side
 * by side on one grid: instance i owns the cells [cell_lo, cell_hi) per axis
 * (3 ints each, the z range [0, 1) in 2D), bounds that are multiples of the
 * tile shape (kmpm_get_tile_shape); ranges must not overlap, and every
 * particle's stencil must stay in the cells of its instance
 * (KMPM_ERR_INSTANCE otherwise).  Each instance
 * has its own dt (from its own particles' CFL condition, the rule of
 * ulmpm.cpp:521-544) and time; results equal those of the instance run alone
 * on the same grid.  ULMPM only.  ninst = 0 returns to one simulation.
 * Resets every instance's time to 0 and dt to config.dt. */
int kmpm_set_instances(kmpm_ctx *ctx, int ninst, const int32_t *cell_lo,
                       const int32_t *cell_hi);
/* dt and time per instance (arrays of ninst; NULL skipped); waits. */
int kmpm_instance_time(kmpm_ctx *ctx, int ninst, double *dt, double *time);
/* nsteps steps as replays of one HIP graph that holds `graph_steps`
 * captured steps (captured on the first call and whenever graph_steps or
 * the buffers change), then the remainder eagerly.  Same results as
 * kmpm_step; phase timing must be off. */
int kmpm_step_graph(kmpm_ctx *ctx, int64_t nsteps, int64_t graph_steps);

/* ---- phase timing ---------------------------------------------------------- */
/* Phases whose launches can be timed with HIP events recorded on the
 * context's stream (no counterpart in the reference; measurement only). */
enum {
  KMPM_PHASE_SORT = 0,      /* cell keys + counting sort + tile lists */
  KMPM_PHASE_P2G = 1,       /* particles_to_grid */
  KMPM_PHASE_GRID = 2,      /* update_grid_state */
  KMPM_PHASE_G2P = 3,       /* USL fused G2P + F + stress */
  KMPM_PHASE_G2P_A = 4,     /* MUSL grid_to_points */
  KMPM_PHASE_P2G_B = 5,     /* MUSL velocities_to_grid (P2G of momentum) */
  KMPM_PHASE_GRID_B = 6,    /* MUSL velocities_to_grid (node velocities) */
  KMPM_PHASE_STRESS_B = 7,  /* MUSL rate of deformation + F + stress */
  KMPM_PHASE_DT = 8,        /* update_time + adjust_dt */
  KMPM_PHASE_LEARNED = 9,   /* USL with the surrogate or the cut network: the
                               constitutive kernel after G2P */
  KMPM_NPHASE = 10
};
/* enable != 0 starts recording (and clears totals); 0 stops. */
int kmpm_timing_enable(kmpm_ctx *ctx, int enable);
/* Waits for the stream; total milliseconds and launch counts per phase. */
int kmpm_timing_read(kmpm_ctx *ctx, double ms[KMPM_NPHASE], int64_t launches[KMPM_NPHASE]);
/* Streaming copy of `bytes` bytes between two device buffers, `reps` times,
 * on a private stream; returns GB/s (10^9 B/s) of read+write traffic.  Used
 * to scale hardware byte counters; no counterpart in the reference. */
int kmpm_copy_probe(int64_t bytes, int reps, double *gbps);

/* ---- observables ---------------------------------------------------------- */
/* dt, time, step counter (update.h) */
int kmpm_get_time(kmpm_ctx *ctx, double *dt, double *time, int64_t *step);
/* Ek = sum 1/2 m |v|^2, Es = sum 1/2 vol sigma:strain_el over particles whose
 * mask has groupbit (compute_kinetic_energy.cpp:63-102, compute_strain_energy.cpp:64-117). */
int kmpm_reduce_energy(kmpm_ctx *ctx, uint32_t groupbit, double *ek, double *es);
/* Grid node state in tag order; NULL pointers skipped.  mass [nn], v/v_update [nn][3]. */
int kmpm_download_grid(kmpm_ctx *ctx, double *mass, double *v, double *v_update,
                       int64_t nnodes);
/* Particle-cell indices (i0,j0,k0) of the last weight pass, [np][3] in ptag
 * order, as ulmpm.cpp:163-165 / 206-208 computes them, and the cell sort
 * permutation as the list of ptags in sorted order [np]. */
int kmpm_download_cells(kmpm_ctx *ctx, int32_t *cell_ijk, int64_t *sorted_ptags,
                        int64_t np);
/* Tile shape (cells per tile per axis) of the cell sort key: particles are
 * ordered by tile, then by cell inside the tile (x-major), then by ptag. */
int kmpm_get_tile_shape(kmpm_ctx *ctx, int tile[3]);

/* ---- tile shapes and autotuning --------------------------------------------
 * No counterpart in the reference (its loops have no tiles).  The tile
 * kernels are compiled for KMPM_NTILE_VARIANTS tile shapes per dimension and
 * shape function (variant 0, the default: 4x4x4 cells in 3D, 8x8 in 2D;
 * the others lengthen the z (3D) or y (2D) extent, the contiguous axis of
 * the node arrays).  Changing the variant keeps the particle state; the next
 * step redoes its cell keys and sort.  Batched instances must stay aligned
 * to the new tiles (KMPM_ERR_ARG otherwise). */
enum { KMPM_NTILE_VARIANTS = 3 };
int kmpm_set_tile_variant(kmpm_ctx *ctx, int variant);
/* Times `trial_steps` steps (after one warm-up step) of every variant from
 * the current state, restores that state, and keeps the fastest variant.
 * ms_per_step[v] is -1 for a variant the instances do not align with. */
int kmpm_autotune(kmpm_ctx *ctx, int trial_steps, int *variant,
                  double ms_per_step[KMPM_NTILE_VARIANTS]);
/* Dense stencil weights of the last weight pass for parity checks:
 * per particle (ptag order) S = 2^d (linear) or 4^d (cubic) entries of
 * node tag (-1 when the pair is dropped), w and grad w [3].  Arrays of
 * np*S, np*S, np*S*3. */
int kmpm_debug_weights(kmpm_ctx *ctx, int64_t *node_tag, double *w, double *wd,
                       int64_t np);

/* ---- on-disk formats ---------------------------------------------------------
 * LAMMPS text dumps, as the reference's `dump(id, all, particle|grid, N,
 * name, vars...)` commands write them: the '*' of the name becomes the
 * step (dump_particle.cpp:54-68; a name without '*' becomes
 * <name><step><name>, as the reference's unsigned test makes it), header
 * "ITEM: TIMESTEP 0 ...", one line per particle / node "id type tag vars...",
 * values in the stream's default %g format.  vars: names from
 * dump_particle.h:32-43 / dump_grid.h:32-37 (an unknown name is
 * KMPM_ERR_ARG, the reference's "output variable ... is unknown").
 * Quantities this library does not carry print as the reference prints
 * them without them: T, gamma, ienergy and rigid 0; the particle body force
 * (set only by contact fixes) 0.  TLMPM particle stresses are rotated by R
 * (dump_particle.cpp:100-103). */
int kmpm_dump_particles(kmpm_ctx *ctx, const char *filename, const char *const *vars, int nvars);
int kmpm_dump_grid(kmpm_ctx *ctx, const char *filename, const char *const *vars, int nvars);
/* Restart file: the reference's update block (Update::write_restart,
 * update.cpp:220-280) and one solid block per solid (Solid::write_restart,
 * solid.cpp:2841-2889: Eigen matrices column-major, TLMPM vol0 PK1
 * recomputed from F and sigma), framed by the magic "KMPMRST1", the
 * dimension, the box and the solid count.  The regions, materials, groups
 * and fixes the reference's file also holds stay with the caller (the
 * input script re-creates them).  kmpm_read_restart checks the method,
 * scheme, shape and FLIP blend against the context, uploads the particles
 * (vol = J vol0, mass = rho0 vol0, as Solid::read_restart,
 * solid.cpp:2931-2936) and restores dt, time and step. */
int kmpm_write_restart(kmpm_ctx *ctx, const char *filename);
int kmpm_read_restart(kmpm_ctx *ctx, const char *filename);
/* Set dt, the elapsed time and the step count (what Update::read_restart
 * restores, update.cpp:362-385). */
int kmpm_set_time(kmpm_ctx *ctx, double dt, double time, int64_t step);

/* ---- learned constitutive surrogate ----------------------------------------
 * Replaces, for a johnson_cook (+ damage_johnson_cook) material, the two
 * transcendental material functions of
 *   StrengthJohnsonCook::update_deviatoric_stress (strength_jc.cpp:100-184)
 *   DamageJohnsonCook::compute_damage             (damage_jc.cpp:91-143)
 * with two 2 -> 32 -> 32 -> 1 tanh MLPs run on the BF16 matrix cores:
 *   net 0: (eps_p^(1/4), ln epsdot*)              -> sigma_y / A
 *   net 1: (triaxiality, ln(epsdot / epsdot0_d))  -> ln eps_f
 * The radial return, damage accumulation and every branch of the reference
 * stay exact; states outside the training envelope use the plain FP64
 * functions.  Training runs on the device on tuples the plain functions
 * generate.  No counterpart in the reference. */
typedef struct kmpm_surrogate_config {
  double eps_p_max;       /* envelope of net 0: eps_p in [1e-10, eps_p_max] */
  double rate_max;        /* envelope of both: epsdot / epsdot0 in [1, rate_max] */
  double triax_max;       /* envelope of net 1: triaxiality in (-3, triax_max] */
  int64_t samples;        /* training tuples per network (device-generated) */
  int64_t iterations;     /* Adam iterations per network */
  int32_t batch;          /* tuples per iteration, a multiple of 256 */
  int32_t reserved;
  double learning_rate;   /* initial Adam step; cosine decay to 1% of it */
  uint64_t seed;
} kmpm_surrogate_config;

typedef struct kmpm_surrogate kmpm_surrogate;
/* number of FP32 parameters of one network (flat order: w1[32][2], b1[32],
 * w2[32][32] ([out][in]), b2[32], w3[32], b3) */
enum { KMPM_SURROGATE_NPARAM = 1185 };

/* Generates the training tuples and initialises both networks. */
int kmpm_surrogate_create(const kmpm_material *mat, const kmpm_surrogate_config *cfg,
                          kmpm_surrogate **out);
int kmpm_surrogate_destroy(kmpm_surrogate *s);
/* Trains both networks (fused forward/backward kernels + Adam); returns the
 * final mean squared error of each on its tuple set (normalised units). */
int kmpm_surrogate_train(kmpm_surrogate *s, double final_loss[2]);
/* n fresh tuples of network `net` from the plain functions:
 * features [n][2], targets [n]. */
int kmpm_surrogate_tuples(kmpm_surrogate *s, int net, int64_t n, uint64_t seed,
                          float *features, float *targets);
/* Tuples at the states a context's simulation visits (the outputs of a
 * batched run streamed as training data): slot k * stride of the context's
 * current state, k < n = ceil(np / stride), gives the features of network
 * `net` (net 0: its eps_p and plastic strain rate; net 1: the triaxiality
 * of its stress and its rate), labelled by the plain functions of this
 * surrogate's material.  Rows without such a state are NaN (net 0: eps_p
 * below 1e-10; net 1: zero pressure or deviator, triaxiality <= -3).
 * features [n][2], targets [n].  No counterpart in the reference. */
int kmpm_surrogate_tuples_from_ctx(kmpm_surrogate *s, kmpm_ctx *ctx, int net, int64_t stride,
                                   float *features, float *targets, int64_t n);
/* Replaces rows [offset, offset + n) of network net's training set (all
 * finite values); a following kmpm_surrogate_train continues from the
 * current parameters on the updated set. */
int kmpm_surrogate_set_tuples(kmpm_surrogate *s, int net, int64_t offset, const float *features,
                              const float *targets, int64_t n);
/* Network outputs (target units) for features [n][2]. */
int kmpm_surrogate_eval(kmpm_surrogate *s, int net, const float *features, int64_t n,
                        float *out);
/* Gradient of L = (1/n) sum (z - (t - base(x) - out_shift)/out_scale)^2
 * with respect to the FP32 parameters (flat order above), z the raw
 * network output. */
int kmpm_surrogate_gradient(kmpm_surrogate *s, int net, const float *features,
                            const float *targets, int64_t n, float *grad, double *loss);
/* Read (set = 0) or replace (set = 1) the FP32 parameters of a network. */
int kmpm_surrogate_params(kmpm_surrogate *s, int net, float *params, int nparam, int set);
/* The maps of a network: in_scale[2], in_shift[2], out_scale, out_shift and
 * the quadratic baseline base[6] (coefficients of 1, x0, x1, x0^2, x0 x1,
 * x1^2 in the inputs x = f * in_scale + in_shift); output =
 * base(x) + out_shift + out_scale * z. */
int kmpm_surrogate_normalization(kmpm_surrogate *s, int net, double map[12]);
/* Feature ranges of the envelope: lo/hi of (net 0 f0, f1, net 1 f0, f1). */
int kmpm_surrogate_envelope(kmpm_surrogate *s, double lo[4], double hi[4]);
/* Use the surrogate in the step of a mixed-precision context whose
 * materials are all this JC + shock material (NULL: back to the plain
 * functions).  The context copies the networks. */
int kmpm_ctx_set_surrogate(kmpm_ctx *ctx, const kmpm_surrogate *s);
/* Particle updates since the last call that used a network, and that fell
 * back to the plain functions (outside the envelope); resets both. */
int kmpm_surrogate_stats(kmpm_ctx *ctx, int64_t *evaluations, int64_t *fallbacks);

/* ---- the cut network ----------------------------------------------------------
 * One network in place of both constitutive calls of Solid::update_stress
 * for a johnson_cook + shock (+ damage_johnson_cook) material: the strength
 * call (solid.cpp:1314-1316, StrengthJohnsonCook::update_deviatoric_stress,
 * strength_jc.cpp:100-184) and the damage call (solid.cpp:1337-1339,
 * DamageJohnsonCook::compute_damage, damage_jc.cpp:91-143).  The EOS call
 * before, the eps_p accumulation and rate filter between them and the
 * stress / strain assembly after stay the plain code.
 *   18 features -> 64 -> 64 -> 64 -> 64 -> 64 -> 9, GELU (tanh form), every
 *   product on the BF16 matrix cores with FP32 accumulation, over an FP32
 *   linear skip of the inputs.
 * Features (A = JC A, G = shear modulus, Gd = G (1 - damage), epsdot0 =
 * JC reference rate): t / A [6] with t = dev(sigma_n + 2 Gd dt D) the
 * trial deviator, 2 G dt dev(D) / A [6], eps_p^(1/4),
 * ln max(epsdot_p / epsdot0, 1), damage, damage_init, pH / A, q / A with
 * q the trial von Mises stress.
 * Outputs: dev(sigma_{n+1}) / A [6], the signed plastic increment
 *   G (q - sigma_y) / (Gd A) (= 3 G dps / A where positive),
 *   3 G (damage_init_{n+1} - damage_init_n) / A, damage_{n+1}.
 * In the step the plastic branch is the network's where its dps output is
 * positive; the elastic branch keeps the trial deviator exactly, the
 * damage_init reset at triaxiality <= -3 and the damage closed form are
 * applied exactly.  States outside the training envelope use the plain FP64
 * functions, which also judge a sample of the network's updates.  Training
 * runs on the device on tuples the plain functions generate.  No
 * counterpart in the reference. */
typedef struct kmpm_cutnet_config {
  double eps_p_max;        /* envelope and sampling: eps_p in [0, eps_p_max] */
  double rate_min;         /* sampled epsdot_p / epsdot0, log-uniform in [rate_min, rate_max] */
  double rate_max;         /* (the envelope's bound) */
  double damage_max;       /* damage in [0, damage_max), damage_max < 1 */
  double damage_init_max;  /* damage_init in [0, damage_init_max] */
  double pressure_max;     /* |pH| / A <= pressure_max */
  double increment_min;    /* 2 G dt vm(dev D) / A, log-uniform in [min, max] */
  double increment_max;
  double dt;               /* time step of the sampled states (the rate filter) */
  double cellsize;         /* h of the simulations (the filter's 1000 h / c_signal) */
  int64_t samples;         /* training tuples generated on the device */
  int64_t iterations;      /* default Adam iterations (kmpm_cutnet_train takes its own) */
  int32_t batch;           /* tuples per iteration, a multiple of 128 */
  int32_t judge_stride;    /* in the step, one wave in judge_stride also runs the plain
                              strength and sums the squared error (0: no judge) */
  double learning_rate;    /* initial Adam step; cosine decay to 1% of it */
  uint64_t seed;
} kmpm_cutnet_config;

typedef struct kmpm_cutnet kmpm_cutnet;
/* FP32 parameters (flat order: w1[64][18], b1[64], then w_l[64][64], b_l[64]
 * for l = 2..5, w6[9][64], b6[9]; weights [out][in]) */
enum { KMPM_CUTNET_NPARAM = 18441, KMPM_CUTNET_NI = 18, KMPM_CUTNET_NO = 9 };

/* Generates the training tuples, fits the input maps and the linear skip,
 * and initialises the network (JC strength + shock EOS material). */
int kmpm_cutnet_create(const kmpm_material *mat, const kmpm_cutnet_config *cfg,
                       kmpm_cutnet **out);
int kmpm_cutnet_destroy(kmpm_cutnet *net);
/* `iterations` Adam iterations (forward + backward on the matrix cores);
 * final_loss: mean loss of the last (up to 16) iterations; seconds: wall time. */
int kmpm_cutnet_train(kmpm_cutnet *net, int64_t iterations, double *final_loss, double *seconds);
/* n fresh tuples from the plain functions: features [n][18], targets [n][9]. */
int kmpm_cutnet_tuples(kmpm_cutnet *net, int64_t n, uint64_t seed, float *features,
                       float *targets);
/* Network outputs (target units) for features [n][18] -> outputs [n][9]. */
int kmpm_cutnet_eval(kmpm_cutnet *net, const float *features, int64_t n, float *outputs);
/* Gradient of L = (1/n) sum_p sum_k (z_k - (y_k - skip_k(x)) / out_scale_k)^2
 * with respect to the FP32 parameters (flat order), z the raw network output. */
int kmpm_cutnet_gradient(kmpm_cutnet *net, const float *features, const float *targets,
                         int64_t n, float *grad, double *loss);
/* kmpm_cutnet_gradient without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight training loops, harnesses that forbid hipMalloc).
 * features [n][18], targets [n][9], grad [KMPM_CUTNET_NPARAM] and loss [1]
 * are DEVICE pointers on the network's device, which must be current.  The
 * block partial sums live in a fixed file-scope device array (64 partial
 * gradients, 4.7 MB per GPU, held until the process ends), so n may not
 * exceed KMPM_CUTNET_NOALLOC_MAX_TUPLES (each block's FP32 sums then cover
 * at most 2^20 tuples): larger n returns KMPM_ERR_CAPACITY.  Stream-ordered
 * on the network's stream and synchronised before returning; calls that
 * overlap on one GPU share the array and race. */
enum { KMPM_CUTNET_NOALLOC_MAX_TUPLES = 1 << 26 };
int kmpm_cutnet_gradient_noalloc(kmpm_cutnet *net, const float *features, const float *targets,
                                 int64_t n, float *grad, float *loss);
/* Read (set = 0) or replace (set = 1) the FP32 parameters. */
int kmpm_cutnet_params(kmpm_cutnet *net, float *params, int nparam, int set);
/* The maps: in_scale[18], in_shift[18] (x = f in_scale + in_shift),
 * out_scale[9], skip[9][19] (output = skip[k][0..17] . x + skip[k][18] +
 * out_scale[k] z_k); set = 1 replaces them. */
int kmpm_cutnet_maps(kmpm_cutnet *net, float *in_scale, float *in_shift, float *out_scale,
                     float *skip, int set);
/* The envelope: eps_p, epsdot/epsdot0, damage, damage_init, |pH|/A, trial
 * von Mises / A, relative increment (upper bounds). */
int kmpm_cutnet_envelope(kmpm_cutnet *net, double env[7]);
/* Use the network in the USL step of a mixed-precision ULMPM context with
 * one material, this network's (NULL: back to the plain functions); the
 * context copies it. */
int kmpm_ctx_set_cutnet(kmpm_ctx *ctx, const kmpm_cutnet *net);
/* Since the last call: particle updates by the network, by the plain
 * fallback, updates judged, and the judged sums of the squared Frobenius
 * error of the deviatoric stress and of the plain deviatoric stress. */
int kmpm_cutnet_stats(kmpm_ctx *ctx, int64_t *evaluations, int64_t *fallbacks, int64_t *judged,
                      double *judge_err2, double *judge_ref2);

#ifdef __cplusplus
}
#endif
#endif /* KMPM_H */

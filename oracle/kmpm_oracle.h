/*
 * kmpm_oracle.h -- CPU FP64 restatement of Karamelo's explicit MPM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; nothing in the product (libkmpm.so, karamelo_amd/) may link,
 * load or call it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * It follows the reference's loops and summation order (node-major gather
 * P2G over per-node neighbour lists in ascending particle order, per-solid
 * velocity normalisation, Eigen expression order written out) and is built
 * with -ffp-contract=off.  Inputs use the same POD problem description as
 * the product's C ABI (include/kmpm.h).
 */
#ifndef KMPM_ORACLE_H
#define KMPM_ORACLE_H

#include "../include/kmpm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* set-up restatements (independent of the product's host code) */
int orc_material_init(kmpm_material *mat);
int orc_grid_size(const kmpm_config *cfg, int64_t n_out[3]);
int orc_region_match(const kmpm_region *region, int dimension, double x, double y, double z);
int orc_region_limits(const kmpm_region *region, double lo[3], double hi[3]);
int orc_populate(const kmpm_config *cfg, const kmpm_region *region, int ppc1d,
                 const kmpm_material *mat, int mat_index, int solid_index,
                 int64_t ptag_offset, kmpm_particles *out, int64_t capacity,
                 int64_t *np_out);

/* simulation */
orc_sim *orc_create(const kmpm_config *cfg, const kmpm_material *mats, int nmat);
void orc_destroy(orc_sim *s);
int orc_set_particles(orc_sim *s, const kmpm_particles *p);
int orc_get_particles(orc_sim *s, kmpm_particles *p);
int orc_set_node_mask(orc_sim *s, const uint32_t *mask, int64_t nnodes);
int orc_fix_velocity_nodes(orc_sim *s, uint32_t groupbit, const int set[3],
                           const double value[3], const double value_prev[3]);
int orc_fix_body_force(orc_sim *s, uint32_t groupbit, const int set[3],
                       const double value[3]);
/* velocity_nodes with the script's expression texts (NULL / "NULL": free),
 * evaluated every step as the reference's Vars are; user variables are
 * constants names[]/values[] */
int orc_fix_velocity_nodes_expr(orc_sim *s, uint32_t groupbit, const char *const texts[3],
                                int nvar, const char *const *names, const double *values);
/* a script expression (or its "time - dt" form) evaluated as a Var at
 * (time, dt) */
int orc_var_eval(const char *text, int previous, double time, double dt, int nvar,
                 const char *const *names, const double *values, double *out);
/* reaction force of velocity fix `fix` in the last step */
int orc_fix_velocity_force(orc_sim *s, int fix, double out[3]);
int orc_step(orc_sim *s, int64_t nsteps);
/* L (ULMPM) or Fdot (TLMPM) of the last step, row-major, 9 per particle */
int orc_get_affine(orc_sim *s, double *C, int64_t np);
int orc_set_affine(orc_sim *s, const double *C, int64_t np);
/* thermo-mechanical runs: the heat flux q, 3 per particle */
int orc_get_heat_flux(orc_sim *s, double *q, int64_t np);
int orc_get_time(orc_sim *s, double *dt, double *time, int64_t *step);
int orc_energy(orc_sim *s, uint32_t groupbit, double *ek, double *es);

/* phases (for phase-level parity tests) */
int orc_phase_weights(orc_sim *s);
int orc_phase_p2g(orc_sim *s);          /* particles_to_grid + post_particles_to_grid */
int orc_phase_grid_update(orc_sim *s);  /* update_grid_state + post_update_grid_state */
int orc_get_grid(orc_sim *s, double *mass, double *v, double *v_update, double *f,
                 int64_t nnodes);
/* neighbour lists of the last weight pass, particle-major in list order:
 * numneigh[np], tags[np*S], w[np*S], wd[np*S*3] (S = max stencil) */
int orc_get_neighbours(orc_sim *s, int32_t *numneigh, int64_t *tags, double *w,
                       double *wd, int64_t np, int S);
/* reference particle-cell indices (i0,j0,k0), ulmpm.cpp:163-165 / 206-208 */
int orc_cell_index(orc_sim *s, int32_t *ijk, int64_t np);
/* stable sort of particles by tile-major cell key, ties by ptag; returns
 * the ptags in sorted order */
int orc_sort_permutation(orc_sim *s, const int tile[3], int64_t *sorted_ptags, int64_t np);
/* last error: particle tag of the offending particle */
int64_t orc_error_particle(orc_sim *s);

/* constitutive point functions (for unit-level known-answer tests) */
void orc_jc_strength(const kmpm_material *m, const double sigma[9], const double D[9],
                     double dt, double eps_p, double epsdot, double damage,
                     double sigma_dev_out[9], double *dps_out);
double orc_shock_pressure(const kmpm_material *m, double J, double rho, double damage,
                          const double D[9], double cellsize);
void orc_jc_damage(const kmpm_material *m, double *damage_init, double *damage,
                   double pH, const double sdev[9], double epsdot, double dps);
double orc_basis(int shape, int derivative, double r, int ntype, double inv_h);
/* Solid::update_stress for one particle (solid.cpp:1246-1374): sigma,
 * strain_el (row-major 3x3), eff. plastic strain, its rate, damage_init and
 * damage are updated in place from D (3x3), J, rho, dt and the cell size */
void orc_update_stress_point(const kmpm_material *m, double cellsize, double dt, double J,
                             double rho, const double D[9], double sigma[9], double strain_el[9],
                             double *eps_p, double *epsdot, double *damage_init, double *damage);
/* the same with the particle temperature T (a material with a temperature
 * law), the thermo-mechanical flag of the method (the damage's T), the
 * volume that scales the heat source, and the heat source out */
void orc_update_stress_point_T(const kmpm_material *m, double cellsize, double dt, double J,
                               double rho, const double D[9], double sigma[9],
                               double strain_el[9], double *eps_p, double *epsdot,
                               double *damage_init, double *damage, double T, int method_temp,
                               double vol, double *gamma);
/* TLMPM matrix functions: rotation of the polar decomposition of F (row-major
 * 3x3; returns 0 when det F <= 0) and the real parts of F's eigenvalues */
int orc_polar_rotation(const double F[9], double R[9]);
void orc_eig_real_parts(const double F[9], double re[3]);

#ifdef __cplusplus
}
#endif
#endif

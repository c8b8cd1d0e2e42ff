// kmpm_host.cpp -- host-side set-up of the C ABI: material moduli, grid node
// counts, regions and lattice population.  No device is touched here.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/kmpm.h"

namespace {

struct Axis {
  double lo, hi;
};

// Node count along one axis (Grid::init, grid.cpp:151-171).  The y and z
// rules truncate the length before dividing, as the reference does.
int global_nodes(double L, double h, bool truncate_length) {
  int n = truncate_length ? (int)((double)((int)L) / h) + 1 : ((int)(L / h)) + 1;
  while (n * h <= L + 0.5 * h) n++;
  return n;
}

// Nodes a single rank creates along one axis (grid.cpp:177-206 with no
// neighbouring rank).
int local_nodes(double L, double lo, double hi, double h) {
  int n = std::max(0, (int)std::ceil(L / h));
  while (lo + h * (n - 0.5) < hi) n++;
  return n;
}

bool inside_region(const kmpm_region *r, double x, double y, double z) {
  const double *p = r->p;
  if (r->kind == KMPM_REGION_BLOCK)  // region_block.cpp:148-154
    return x >= p[0] && x <= p[1] && y >= p[2] && y <= p[3] && z >= p[4] && z <= p[5];
  // region_cylinder.cpp:147-168: axis coordinate in [lo, hi], radial distance^2 <= R^2
  const double c1 = p[0], c2 = p[1], R = p[2], lo = p[3], hi = p[4];
  double along, a, b;
  if (r->axis == 0) { along = x; a = y; b = z; }
  else if (r->axis == 1) { along = y; a = x; b = z; }
  else { along = z; a = x; b = y; }
  const double da = a - c1, db = b - c2;
  const double dSq = da * da + db * db;
  return along >= lo && along <= hi && dSq <= R * R;
}

}  // namespace

extern "C" {

const char *kmpm_status_string(int s) {
  switch (s) {
    case KMPM_OK: return "ok";
    case KMPM_ERR_ARG: return "invalid argument or unsupported option";
    case KMPM_ERR_HIP: return "HIP runtime error";
    case KMPM_ERR_NO_DEVICE: return "no usable gfx950 device";
    case KMPM_ERR_CAPACITY: return "output buffer too small";
    case KMPM_ERR_PARTICLE_LEFT_BOX: return "Error: Particle left the domain";
    case KMPM_ERR_JACOBIAN: return "Error: J<=0.0";
    case KMPM_ERR_DTCFL: return "Error: dtCFL == 0 or NaN";
    case KMPM_ERR_WAVE_SPEED: return "Error: max_p_wave_speed is nan";
    case KMPM_ERR_GRID: return "grid does not support the stencil on one device";
    case KMPM_ERR_FIXED_POINT: return "node sum exceeded the fixed-point range of the mixed path";
    case KMPM_ERR_INTERNAL: return "internal error: particle outside its sorted tile";
    case KMPM_ERR_INSTANCE: return "a particle reaches outside the tiles of its instance";
    default: return "unknown status";
  }
}

int kmpm_material_init(kmpm_material *m) {
  if (!m) return KMPM_ERR_ARG;
  if (m->type == KMPM_MAT_LINEAR) {
    // Mat(id, type, rho0, E, nu, cp, kappa), material.cpp:745-764
    m->rho0 = m->lin_p[0];
    m->E = m->lin_p[1];
    m->nu = m->lin_p[2];
    m->G = m->E / (2 * (1 + m->nu));
    m->lambda = m->E * m->nu / ((1 + m->nu) * (1 - 2 * m->nu));
    m->K = m->E / (3 * (1 - 2 * m->nu));
    m->signal_velocity = std::sqrt(m->K / m->rho0);
    return KMPM_OK;
  }
  if (m->type != KMPM_MAT_EOS_STRENGTH) return KMPM_ERR_ARG;
  if (m->eos < KMPM_EOS_LINEAR || m->eos > KMPM_EOS_FLUID) return KMPM_ERR_ARG;
  if (m->strength < KMPM_STRENGTH_LINEAR || m->strength > KMPM_STRENGTH_FLUID)
    return KMPM_ERR_ARG;
  if (m->damage != KMPM_DAMAGE_NONE && m->damage != KMPM_DAMAGE_JOHNSON_COOK)
    return KMPM_ERR_ARG;
  // Mat(id, type, eos, strength, damage, temp), material.cpp:720-743:
  // rho0 and K from the EOS, G from the strength style (viscosity for fluid).
  m->rho0 = m->eos_p[0];
  m->K = m->eos_p[1];
  m->G = m->strength_p[0];
  m->E = 9 * m->K * m->G / (3 * m->K + m->G);
  m->nu = (3 * m->K - 2 * m->G) / (2 * (3 * m->K + m->G));
  m->lambda = m->K - 2 * m->G / 3;
  m->signal_velocity = std::sqrt((m->lambda + 2 * m->G) / m->rho0);
  // Tr == Tm is rejected by both JC styles (strength_jc.cpp:89-93, damage_jc.cpp:80-84)
  if (m->strength == KMPM_STRENGTH_JOHNSON_COOK && m->strength_p[7] == m->strength_p[8])
    return KMPM_ERR_ARG;
  if (m->damage == KMPM_DAMAGE_JOHNSON_COOK && m->damage_p[6] == m->damage_p[7])
    return KMPM_ERR_ARG;
  return KMPM_OK;
}

int kmpm_grid_size(const kmpm_config *c, int64_t out[3]) {
  if (!c || !(c->cellsize > 0) || c->dimension < 2 || c->dimension > 3) return KMPM_ERR_ARG;
  const double h = c->cellsize;
  const bool tl = c->method == KMPM_METHOD_TLMPM || c->method == KMPM_METHOD_TLCPDI;
  for (int d = 0; d < 3; d++) {
    if (d >= c->dimension) {
      out[d] = 1;
      continue;
    }
    // the bounds of the grid: the box (UL) or the solid (TL), grid.cpp:91-97
    const double lo = tl ? c->solidlo[d] : c->boxlo[d];
    const double hi = tl ? c->solidhi[d] : c->boxhi[d];
    if (tl && (lo < c->boxlo[d] || hi > c->boxhi[d] || !(hi > lo))) return KMPM_ERR_ARG;
    const double L = hi - lo;
    const int g = global_nodes(L, h, d > 0);
    const int l = local_nodes(L, lo, hi, h);
    if (g != l) return KMPM_ERR_GRID;
    out[d] = g;
  }
  return KMPM_OK;
}

int kmpm_region_limits(const kmpm_region *r, double lo[3], double hi[3]) {
  if (!r || !lo || !hi) return KMPM_ERR_ARG;
  const double *p = r->p;
  if (r->kind == KMPM_REGION_BLOCK) {   // region_block.cpp:159-168
    for (int d = 0; d < 3; d++) {
      lo[d] = p[2 * d];
      hi[d] = p[2 * d + 1];
    }
    return KMPM_OK;
  }
  if (r->kind != KMPM_REGION_CYLINDER) return KMPM_ERR_ARG;
  // region_cylinder.cpp:113-137: the axis spans [lo, hi], the two others
  // c1 -+ R and c2 -+ R
  const double c1 = p[0], c2 = p[1], R = p[2], alo = p[3], ahi = p[4];
  const int ax = r->axis;
  const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
  lo[ax] = alo;
  hi[ax] = ahi;
  lo[o1] = c1 - R;
  hi[o1] = c1 + R;
  lo[o2] = c2 - R;
  hi[o2] = c2 + R;
  return KMPM_OK;
}

int kmpm_region_match(const kmpm_region *r, int dimension, double x, double y, double z) {
  (void)dimension;
  const bool in = inside_region(r, x, y, z);
  return r->interior ? (in ? 1 : 0) : (in ? 0 : 1);  // region.cpp:65-72
}

int kmpm_populate(const kmpm_config *c, const kmpm_region *region, int ppc1d,
                  const kmpm_material *mat, int mat_index, int solid_index, int64_t ptag_offset,
                  kmpm_particles *out, int64_t capacity, int64_t *np_out) {
  if (!c || !region || !mat || !out || !np_out) return KMPM_ERR_ARG;
  if (c->method < KMPM_METHOD_ULMPM || c->method > KMPM_METHOD_TLCPDI) return KMPM_ERR_ARG;
  const bool tl = c->method == KMPM_METHOD_TLMPM || c->method == KMPM_METHOD_TLCPDI;
  const bool cpdi = c->method == KMPM_METHOD_ULCPDI || c->method == KMPM_METHOD_TLCPDI;
  const bool q4 = cpdi && c->cpdi_style == KMPM_CPDI_Q4;
  const int dim = c->dimension;
  const double delta = c->cellsize;
  // CPDI: the half-width of a particle's domain (solid.cpp:2029-2106)
  double lp = delta;
  if (ppc1d == 1) lp *= 0.5;
  else if (ppc1d == 2) lp *= 0.25;
  else lp *= 1.0 / (2 * ppc1d);
  // Lattice of cells over the bounds (Solid::populate, solid.cpp:1927-2002,
  // one rank: sub-domain == box, no neighbour adjustments).  The bounds are
  // the box for ULMPM and the region's limits for TLMPM.
  double blo[3], bhi[3];
  if (tl) {
    kmpm_region_limits(region, blo, bhi);
  } else {
    for (int d = 0; d < 3; d++) {
      blo[d] = c->boxlo[d];
      bhi[d] = c->boxhi[d];
    }
  }
  int ncell[3] = {1, 1, 1}, noff[3] = {0, 0, 0};
  for (int d = 0; d < dim; d++) {
    const double Llo = std::max(0.0, c->boxlo[d] - blo[d]);
    const double Lhi = std::max(0.0, std::min(c->boxhi[d], bhi[d]) - blo[d]);
    noff[d] = (int)std::floor(Llo / delta);
    int n = std::max(0, (int)std::ceil(Lhi / delta) - noff[d]);
    while (blo[d] + delta * (noff[d] + n - 0.5) < std::min(c->boxhi[d], bhi[d])) n++;
    ncell[d] = n;
  }
  // Integration-point offsets inside a cell (solid.cpp:2026-2121).
  std::vector<double> ip;
  if (ppc1d == 1) {
    ip = {0, 0, 0};
  } else if (ppc1d == 2) {
    const double xi = 0.25;
    const double t[24] = {-xi, -xi, -xi, -xi, xi, -xi, xi, -xi, -xi, xi, xi, -xi,
                          -xi, -xi, xi,  -xi, xi, xi,  xi, -xi, xi,  xi, xi, xi};
    const int nip = dim == 2 ? 4 : 8;
    ip.assign(t, t + 3 * nip);
  } else if (ppc1d >= 4) {
    const double dd = 1.0 / ppc1d;
    for (int k = 0; k < ppc1d; k++)
      for (int i = 0; i < ppc1d; i++)
        for (int j = 0; j < ppc1d; j++) {
          ip.push_back((i + 0.5) * dd - 0.5);
          ip.push_back((j + 0.5) * dd - 0.5);
          ip.push_back((k + 0.5) * dd - 0.5);
        }
    const int nip = dim == 2 ? ppc1d * ppc1d : ppc1d * ppc1d * ppc1d;
    ip.resize(3 * nip);
  } else {
    return KMPM_ERR_ARG;
  }
  const int nip = (int)ip.size() / 3;
  double vol_ = dim == 2 ? delta * delta : delta * delta * delta;
  double mass_ = mat->rho0 * vol_;
  mass_ /= (double)nip;
  vol_ /= (double)nip;

  int64_t count = 0;
  for (int i = 0; i < ncell[0]; i++)
    for (int j = 0; j < ncell[1]; j++)
      for (int k = 0; k < ncell[2]; k++)
        for (int q = 0; q < nip; q++) {
          const double x = blo[0] + delta * (noff[0] + i + 0.5 + ip[3 * q]);
          const double y = blo[1] + delta * (noff[1] + j + 0.5 + ip[3 * q + 1]);
          const double z = dim == 3 ? blo[2] + delta * (noff[2] + k + 0.5 + ip[3 * q + 2]) : 0;
          // Domain::inside_subdomain (domain.cpp:187-195) then Region::inside
          bool in_box = x >= c->boxlo[0] && x <= c->boxhi[0] && y >= c->boxlo[1] &&
                        y <= c->boxhi[1];
          if (dim == 3) in_box = in_box && z >= c->boxlo[2] && z <= c->boxhi[2];
          if (!in_box || !inside_region(region, x, y, z)) continue;
          if (count < capacity) {
            const int64_t p = count;
            const double xs[3] = {x, y, z};
            for (int d = 0; d < 3; d++) {
              out->x0[3 * p + d] = xs[d];
              out->x[3 * p + d] = xs[d];
              out->v[3 * p + d] = 0;
            }
            for (int e = 0; e < 9; e++) {
              out->F[9 * p + e] = (e == 0 || e == 4 || e == 8) ? 1.0 : 0.0;
              out->sigma[9 * p + e] = 0;
              out->strain_el[9 * p + e] = 0;
            }
            out->mass[p] = mass_;
            out->vol0[p] = vol_;
            if (out->vol) out->vol[p] = vol_;
            if (out->rho0) out->rho0[p] = mat->rho0;
            if (out->rho) out->rho[p] = mat->rho0;
            if (out->J) out->J[p] = 1;
            out->eff_plastic_strain[p] = 0;
            out->eff_plastic_strain_rate[p] = 0;
            out->damage[p] = 0;
            out->damage_init[p] = 0;
            out->ptag[p] = ptag_offset + p + 1;
            out->solid[p] = solid_index;
            out->mat[p] = mat_index;
            if (out->mask) out->mask[p] = 1;
            if (cpdi && !q4) {   // R4 vectors (lp, 0, 0), (0, lp, 0), solid.cpp:2188-2205
              for (int e = 0; e < 6; e++) {
                const double r = e == 0 || e == 4 ? lp : 0.0;
                if (out->rp0) out->rp0[6 * p + e] = r;
                if (out->rp) out->rp[6 * p + e] = r;
              }
            } else if (q4) {     // Q4 corners from (-, -), counter-clockwise, :2208-2222
              const double cx[4] = {x - lp, x + lp, x + lp, x - lp};
              const double cy[4] = {y - lp, y - lp, y + lp, y + lp};
              for (int k = 0; k < 4; k++)
                for (int d = 0; d < 3; d++) {
                  const double v = d == 0 ? cx[k] : d == 1 ? cy[k] : 0.0;
                  if (out->xpc0) out->xpc0[12 * p + 3 * k + d] = v;
                  if (out->xpc) out->xpc[12 * p + 3 * k + d] = v;
                }
            }
          }
          count++;
        }
  *np_out = count;
  return count > capacity ? KMPM_ERR_CAPACITY : KMPM_OK;
}

}  // extern "C"

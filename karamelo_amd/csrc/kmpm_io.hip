// kmpm_io.hip -- the reference's on-disk formats for a context's state:
// the LAMMPS text dumps of particles (DumpParticle::write,
// dump_particle.cpp:52-161) and of grid nodes (DumpGrid::write,
// dump_grid.cpp:50-139), and a restart file made of the reference's update
// block (Update::write_restart, update.cpp:220-280) and solid blocks
// (Solid::write_restart, solid.cpp:2841-2889), with their readers.  Host
// code above the context: it downloads the state and formats it.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "kmpm_ctx.h"
#include "kmpm_mat3.h"

using namespace kmpm;

#define HIPCHK(x)                                 \
  do {                                            \
    if ((x) != hipSuccess) return KMPM_ERR_HIP;   \
  } while (0)

namespace {

// Dump::write's file name: the '*' is replaced by the step.  Without a '*'
// the reference's unsigned test (`pos_asterisk >= 0`) still takes that
// branch and its size arithmetic wraps, so the name becomes
// <name><step><name> (dump_particle.cpp:54-68, dump_grid.cpp:53-67).
std::string dump_name(const std::string &name, int64_t step) {
  const size_t pos = name.find('*');
  std::string out = name.substr(0, pos);
  out += std::to_string(step);
  if (name.size() - pos - 1 > 0) out += name.substr(pos + 1, name.size() - pos - 1);
  return out;
}

// Scheme style key (update.cpp:83-101).
const char *scheme_key(int scheme) {
  return scheme == KMPM_SCHEME_USL ? "usl" : scheme == KMPM_SCHEME_MUSL ? "musl" : "usf";
}

// Update::sub_method_type as the restart stores it (update.h:62-69): the
// configured value for the APIC family; PIC (0) or FLIP (1) otherwise, PIC
// being FLIP with PIC_FLIP = 0 (update.cpp:135-136).
int32_t restart_sub_method(const kmpm_config &cf) {
  if (cf.sub_method >= KMPM_SUB_APIC) return cf.sub_method;
  return cf.pic_flip == 0.0 ? KMPM_SUB_PIC : KMPM_SUB_FLIP;
}

// ostream << double with the default flags and precision 6 is %g.
void put(std::string &s, double v) {
  char b[40];
  std::snprintf(b, sizeof b, "%g ", v);
  s += b;
}
void put_i(std::string &s, int64_t v) {
  char b[32];
  std::snprintf(b, sizeof b, "%" PRId64 " ", v);
  s += b;
}

struct Host {
  std::vector<int64_t> ptag;
  std::vector<int32_t> solid, mat;
  std::vector<uint32_t> mask;
  std::vector<double> x0, x, v, F, sigma, strain_el, mass, vol0, vol, rho0, rho, J, ep, epdot,
      damage, damage_init, T, gamma;
  kmpm_particles p{};
  explicit Host(int64_t n) {
    auto z = [n](std::vector<double> &a, int c) { a.assign((size_t)(n * c), 0.0); };
    ptag.assign(n, 0);
    solid.assign(n, 0);
    mat.assign(n, 0);
    mask.assign(n, 0);
    z(x0, 3); z(x, 3); z(v, 3); z(F, 9); z(sigma, 9); z(strain_el, 9);
    z(mass, 1); z(vol0, 1); z(vol, 1); z(rho0, 1); z(rho, 1); z(J, 1);
    z(ep, 1); z(epdot, 1); z(damage, 1); z(damage_init, 1); z(T, 1); z(gamma, 1);
    p.np = n;
    p.ptag = ptag.data(); p.solid = solid.data(); p.mat = mat.data(); p.mask = mask.data();
    p.x0 = x0.data(); p.x = x.data(); p.v = v.data(); p.F = F.data(); p.sigma = sigma.data();
    p.strain_el = strain_el.data(); p.mass = mass.data(); p.vol0 = vol0.data();
    p.vol = vol.data(); p.rho0 = rho0.data(); p.rho = rho.data(); p.J = J.data();
    p.eff_plastic_strain = ep.data(); p.eff_plastic_strain_rate = epdot.data();
    p.damage = damage.data(); p.damage_init = damage_init.data();
    p.T = T.data(); p.gamma = gamma.data();
  }
};

bool in_list(const char *v, const char *const *list, int n) {
  for (int i = 0; i < n; i++)
    if (std::strcmp(v, list[i]) == 0) return true;
  return false;
}

const char *const PARTICLE_VARS[] = {"x", "y", "z", "x0", "y0", "z0", "vx", "vy", "vz",
                                     "s11", "s22", "s33", "s12", "s13", "s23", "e11", "e22",
                                     "e33", "e12", "e13", "e23", "seq", "volume", "mass",
                                     "damage", "damage_init", "bx", "by", "bz", "ep", "epdot",
                                     "T", "ienergy", "gamma"};   // dump_particle.h:32-43
const char *const GRID_VARS[] = {"x", "y", "z", "vx", "vy", "vz", "bx", "by", "bz", "mass",
                                 "mask", "rigid", "T", "ntypex", "ntypey", "ntypez"};
                                                                // dump_grid.h:32-37

void header(std::string &s, const kmpm_ctx *c, int64_t n, const char *const *vars, int nvars) {
  s += "ITEM: TIMESTEP\n0\nITEM: NUMBER OF ATOMS\n";
  s += std::to_string(n) + "\n";
  s += "ITEM: BOX BOUNDS sm sm sm\n";
  for (int d = 0; d < 3; d++) {
    char b[64];
    std::snprintf(b, sizeof b, "%g %g\n", c->cfg.boxlo[d], c->cfg.boxhi[d]);
    s += b;
  }
  s += "ITEM: ATOMS id type tag ";
  for (int k = 0; k < nvars; k++) s += std::string(vars[k]) + " ";
  s += "\n";
}

int write_file(const std::string &name, const std::string &text) {
  std::FILE *f = std::fopen(name.c_str(), "w");
  if (!f) return KMPM_ERR_ARG;
  const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
  return (std::fclose(f) == 0 && ok) ? KMPM_OK : KMPM_ERR_ARG;
}

template <class T> void wr(std::ofstream &o, const T &v) {
  o.write(reinterpret_cast<const char *>(&v), sizeof(T));
}
template <class T> bool rd(std::ifstream &i, T &v) {
  return (bool)i.read(reinterpret_cast<char *>(&v), sizeof(T));
}
void wr_str(std::ofstream &o, const std::string &s) {
  const size_t n = s.size();
  wr(o, n);
  o.write(s.data(), (std::streamsize)n);
}
bool rd_str(std::ifstream &i, std::string &s) {
  size_t n = 0;
  if (!rd(i, n) || n > 4096) return false;
  s.assign(n, '\0');
  return (bool)i.read(&s[0], (std::streamsize)n);
}
// Eigen::Matrix3d is stored column-major; the context's arrays are row-major.
void wr_mat(std::ofstream &o, const double *m) {
  for (int col = 0; col < 3; col++)
    for (int row = 0; row < 3; row++) wr(o, m[3 * row + col]);
}
bool rd_mat(std::ifstream &i, double *m) {
  for (int col = 0; col < 3; col++)
    for (int row = 0; row < 3; row++)
      if (!rd(i, m[3 * row + col])) return false;
  return true;
}

const char MAGIC[8] = {'K', 'M', 'P', 'M', 'R', 'S', 'T', '1'};

}  // namespace

extern "C" {

int kmpm_set_time(kmpm_ctx *c, double dt, double time, int64_t step) {
  if (!c || !(dt > 0) || step < 0) return KMPM_ERR_ARG;
  DevScalars sc;
  HIPCHK(hipMemcpyAsync(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  sc.dt = dt;
  sc.time = time;
  sc.step = step;
  HIPCHK(hipMemcpyAsync(c->d_sc, &sc, sizeof(sc), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return KMPM_OK;
}

int kmpm_dump_particles(kmpm_ctx *c, const char *filename, const char *const *vars, int nvars) {
  if (!c || !filename || nvars < 0 || (nvars > 0 && !vars)) return KMPM_ERR_ARG;
  for (int k = 0; k < nvars; k++)
    if (!vars[k] || !in_list(vars[k], PARTICLE_VARS, (int)(sizeof PARTICLE_VARS / sizeof *PARTICLE_VARS)))
      return KMPM_ERR_ARG;   // "Error: output variable ... is unknown!" (dump_particle.cpp:24-36)
  const int64_t n = c->a.np;
  Host h(n);
  int st = kmpm_download_particles(c, &h.p);
  if (st) return st;
  int64_t step = 0;
  if ((st = kmpm_get_time(c, nullptr, nullptr, &step))) return st;
  std::string s;
  header(s, c, n, vars, nvars);
  const bool tl = (c->cfg.method == KMPM_METHOD_TLMPM || c->cfg.method == KMPM_METHOD_TLCPDI);
  // solids in order, each in its particles' order (the upload order)
  int32_t nsol = 0;
  for (int64_t i = 0; i < n; i++) nsol = std::max(nsol, h.solid[i] + 1);
  for (int32_t is = 0; is < nsol; is++)
    for (int64_t i = 0; i < n; i++) {
      if (h.solid[i] != is) continue;
      double sg[9];
      std::memcpy(sg, &h.sigma[9 * i], sizeof sg);
      if (tl) {   // the TL stress is unrotated: R sigma R^T (dump_particle.cpp:100-103)
        double R[9], t[9];
        if (!polar_rotation(&h.F[9 * i], R)) return KMPM_ERR_JACOBIAN;
        for (int r = 0; r < 3; r++)
          for (int q = 0; q < 3; q++) {
            double acc = 0;
            for (int k = 0; k < 3; k++) acc += R[3 * r + k] * h.sigma[9 * i + 3 * k + q];
            t[3 * r + q] = acc;
          }
        for (int r = 0; r < 3; r++)
          for (int q = 0; q < 3; q++) {
            double acc = 0;
            for (int k = 0; k < 3; k++) acc += t[3 * r + k] * R[3 * q + k];
            sg[3 * r + q] = acc;
          }
      }
      put_i(s, h.ptag[i]);
      put_i(s, is + 1);
      put_i(s, h.ptag[i]);
      const double *e = &h.strain_el[9 * i];
      for (int k = 0; k < nvars; k++) {
        const std::string v = vars[k];
        if (v == "x") put(s, h.x[3 * i]);
        else if (v == "y") put(s, h.x[3 * i + 1]);
        else if (v == "z") put(s, h.x[3 * i + 2]);
        else if (v == "x0") put(s, h.x0[3 * i]);
        else if (v == "y0") put(s, h.x0[3 * i + 1]);
        else if (v == "z0") put(s, h.x0[3 * i + 2]);
        else if (v == "vx") put(s, h.v[3 * i]);
        else if (v == "vy") put(s, h.v[3 * i + 1]);
        else if (v == "vz") put(s, h.v[3 * i + 2]);
        else if (v == "s11") put(s, sg[0]);
        else if (v == "s22") put(s, sg[4]);
        else if (v == "s33") put(s, sg[8]);
        else if (v == "s12") put(s, sg[1]);
        else if (v == "s13") put(s, sg[2]);
        else if (v == "s23") put(s, sg[5]);
        else if (v == "seq") {
          // sqrt(3/2) |dev sigma| (Frobenius, the reference's row-major sum)
          const double t3 = (sg[0] + sg[4] + sg[8]) / 3.0;
          double dv[9];
          std::memcpy(dv, sg, sizeof dv);
          dv[0] -= t3;
          dv[4] -= t3;
          dv[8] -= t3;
          double n2 = 0;
          for (int q = 0; q < 9; q++) n2 += dv[q] * dv[q];
          put(s, std::sqrt(3. / 2.) * std::sqrt(n2));
        }
        else if (v == "e11") put(s, e[0]);
        else if (v == "e22") put(s, e[4]);
        else if (v == "e33") put(s, e[8]);
        else if (v == "e12") put(s, e[1]);
        else if (v == "e13") put(s, e[2]);
        else if (v == "e23") put(s, e[5]);
        else if (v == "damage") put(s, h.damage[i]);
        else if (v == "damage_init") put(s, h.damage_init[i]);
        else if (v == "volume") put(s, h.vol[i]);
        else if (v == "mass") put(s, h.mass[i]);
        // the particle body force is only set by contact fixes (zero here)
        else if (v == "bx" || v == "by" || v == "bz") put(s, 0.0);
        else if (v == "ep") put(s, h.ep[i]);
        else if (v == "epdot") put(s, h.epdot[i]);
        // T and gamma: 0 without a temperature law; the internal energy is
        // not integrated (0, as the reference leaves it without one)
        else if (v == "ienergy") put(s, 0.0);
        else if (v == "T") put(s, h.T[i]);
        else if (v == "gamma") put(s, h.gamma[i]);
      }
      s += "\n";
    }
  return write_file(dump_name(filename, step), s);
}

int kmpm_dump_grid(kmpm_ctx *c, const char *filename, const char *const *vars, int nvars) {
  if (!c || !filename || nvars < 0 || (nvars > 0 && !vars)) return KMPM_ERR_ARG;
  for (int k = 0; k < nvars; k++)
    if (!vars[k] || !in_list(vars[k], GRID_VARS, (int)(sizeof GRID_VARS / sizeof *GRID_VARS)))
      return KMPM_ERR_ARG;   // dump_grid.cpp:30-42
  const int64_t nn = c->a.nn;
  std::vector<double> mass(nn), v(3 * nn), T(nn);
  int st = kmpm_download_grid(c, mass.data(), v.data(), nullptr, nn);
  if (st) return st;
  if ((st = download_grid_temperature(c, T.data(), nn))) return st;
  int64_t step = 0;
  if ((st = kmpm_get_time(c, nullptr, nullptr, &step))) return st;
  const bool tl = (c->cfg.method == KMPM_METHOD_TLMPM || c->cfg.method == KMPM_METHOD_TLCPDI);
  int64_t n[3];
  if ((st = kmpm_grid_size(&c->cfg, n))) return st;
  const double *lo = tl ? c->cfg.solidlo : c->cfg.boxlo;
  const double h = c->cfg.cellsize;
  const bool cubic = c->cfg.shape == KMPM_SHAPE_CUBIC_SPLINE;
  const std::vector<uint32_t> *gm = c->h_gmask;
  std::string s;
  header(s, c, nn, vars, nvars);
  for (int64_t tag = 0; tag < nn; tag++) {
    // tag = nz ny i + nz j + k (grid.cpp:251)
    const int64_t ijk[3] = {tag / (n[1] * n[2]), (tag / n[2]) % n[1], tag % n[2]};
    put_i(s, tag);
    s += "1 ";
    put_i(s, tag);
    const uint32_t mask = (gm && (int64_t)gm->size() == nn) ? (*gm)[tag] : 1u;
    for (int k = 0; k < nvars; k++) {
      const std::string vv = vars[k];
      int d = -1;
      if (vv == "x" || vv == "vx" || vv == "bx" || vv == "ntypex") d = 0;
      if (vv == "y" || vv == "vy" || vv == "by" || vv == "ntypey") d = 1;
      if (vv == "z" || vv == "vz" || vv == "bz" || vv == "ntypez") d = 2;
      if (vv == "x" || vv == "y" || vv == "z") {
        // node positions (grid.cpp:222-226); TL nodes keep their reference positions
        put(s, d < c->cfg.dimension ? lo[d] + (double)ijk[d] * h : 0.0);
      } else if (vv[0] == 'v') {
        put(s, v[3 * tag + d]);
      } else if (vv[0] == 'b') {
        // the body force fixes' m g on the nodes of their group with mass
        // (fix_body_force.cpp:120-140)
        double f = 0.0;
        if (mass[tag] > 0)
          for (int q = 0; q < c->a.nbf; q++)
            if ((mask & c->a.bf[q].bit) && c->a.bf[q].set[d]) f += c->a.bf[q].val[d] * mass[tag];
        put(s, f);
      } else if (vv == "mass") {
        put(s, mass[tag]);
      } else if (vv == "mask") {
        put_i(s, (int64_t)mask);
      } else if (vv == "T") {
        put(s, T[tag]);
      } else if (vv == "rigid") {
        s += "0 ";
      } else {   // ntype (grid.cpp:228-240)
        int64_t t = 0;
        if (cubic && d < c->cfg.dimension)
          t = std::min<int64_t>(2, ijk[d]) - std::min<int64_t>(n[d] - 1 - ijk[d], 2);
        put_i(s, t);
      }
    }
    s += "\n";
  }
  return write_file(dump_name(filename, step), s);
}

int kmpm_write_restart(kmpm_ctx *c, const char *filename) {
  // the reference's restart blocks hold no CPDI domains (solid.cpp:2841-2889)
  if (!c || !filename || c->a.cpdi) return KMPM_ERR_ARG;
  const int64_t n = c->a.np;
  Host h(n);
  int st = kmpm_download_particles(c, &h.p);
  if (st) return st;
  double dt = 0, time = 0;
  int64_t step = 0;
  if ((st = kmpm_get_time(c, &dt, &time, &step))) return st;
  std::ofstream o(filename, std::ios::out | std::ios::binary);
  if (!o.is_open()) return KMPM_ERR_ARG;
  const kmpm_config &cf = c->cfg;
  const bool tl = (cf.method == KMPM_METHOD_TLMPM || cf.method == KMPM_METHOD_TLCPDI);
  o.write(MAGIC, sizeof MAGIC);
  wr(o, (int32_t)cf.dimension);
  // ---- Update::write_restart (update.cpp:220-280) ----
  wr_str(o, tl ? "tlmpm" : "ulmpm");
  wr(o, (bool)(cf.thermo != 0));                       // method->temp
  wr_str(o, scheme_key(cf.scheme));
  wr(o, restart_sub_method(cf));
  wr(o, c->a.flip);                                    // PIC_FLIP after the method's setup
  wr(o, (int32_t)(cf.shape == KMPM_SHAPE_CUBIC_SPLINE ? 1 : 0));   // ShapeFunctions
  wr(o, (size_t)0);                                    // additional_args
  wr(o, time);
  wr(o, step);
  wr(o, dt);
  wr(o, cf.dt_factor);
  wr(o, (bool)(cf.dt_constant != 0));
  // ---- the box and the solids ----
  for (int d = 0; d < 3; d++) wr(o, cf.boxlo[d]);
  for (int d = 0; d < 3; d++) wr(o, cf.boxhi[d]);
  int32_t nsol = 0;
  for (int64_t i = 0; i < n; i++) nsol = std::max(nsol, h.solid[i] + 1);
  wr(o, nsol);
  for (int32_t is = 0; is < nsol; is++) {
    int64_t ns = 0;
    int32_t imat = 0;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < n; i++)
      if (h.solid[i] == is) {
        imat = h.mat[i];
        ns++;
      }
    // the solid's bounds: its grid's for TLMPM, the box for ULMPM
    for (int d = 0; d < 3; d++) {
      lo[d] = tl ? cf.solidlo[d] : cf.boxlo[d];
      hi[d] = tl ? cf.solidhi[d] : cf.boxhi[d];
    }
    wr_str(o, "solid" + std::to_string(is + 1));
    // ---- Solid::write_restart (solid.cpp:2841-2889) ----
    for (int d = 0; d < 3; d++) wr(o, lo[d]);
    for (int d = 0; d < 3; d++) wr(o, hi[d]);
    for (int d = 0; d < 3; d++) wr(o, lo[d]);   // solidsublo (one rank)
    for (int d = 0; d < 3; d++) wr(o, hi[d]);
    wr(o, ns);                                   // np
    wr(o, (int32_t)ns);                          // np_local
    wr(o, (int32_t)(cf.dimension == 3 ? 8 : 4)); // nc: corners per particle
    wr(o, imat);
    wr(o, cf.cellsize);
    for (int64_t i = 0; i < n; i++) {
      if (h.solid[i] != is) continue;
      wr(o, h.ptag[i]);
      for (int d = 0; d < 3; d++) wr(o, h.x0[3 * i + d]);
      for (int d = 0; d < 3; d++) wr(o, h.x[3 * i + d]);
      for (int d = 0; d < 3; d++) wr(o, h.v[3 * i + d]);
      wr_mat(o, &h.sigma[9 * i]);
      wr_mat(o, &h.strain_el[9 * i]);
      if (tl) {
        // vol0 PK1 = vol0 J (R sigma R^T) F^-T, recomputed from F and sigma
        double Finv[9], R[9], P[9];
        m3_inverse(&h.F[9 * i], Finv);
        if (!polar_rotation(&h.F[9 * i], R)) return KMPM_ERR_JACOBIAN;
        pk1(h.vol0[i], m3_det(&h.F[9 * i]), R, &h.sigma[9 * i], Finv, P);
        wr_mat(o, P);
      }
      wr_mat(o, &h.F[9 * i]);
      wr(o, h.J[i]);
      wr(o, h.vol0[i]);
      wr(o, h.rho0[i]);
      wr(o, h.ep[i]);
      wr(o, h.epdot[i]);
      wr(o, h.damage[i]);
      wr(o, h.damage_init[i]);
      if (cf.thermo) wr(o, h.T[i]);
      wr(o, 0.0);                                // ienergy (not integrated)
      wr(o, (int32_t)h.mask[i]);
    }
  }
  o.close();
  return o.good() ? KMPM_OK : KMPM_ERR_ARG;
}

int kmpm_read_restart(kmpm_ctx *c, const char *filename) {
  if (!c || !filename || c->a.cpdi) return KMPM_ERR_ARG;
  std::ifstream in(filename, std::ios::in | std::ios::binary);
  if (!in.is_open()) return KMPM_ERR_ARG;
  char magic[8];
  if (!in.read(magic, 8) || std::memcmp(magic, MAGIC, 8) != 0) return KMPM_ERR_ARG;
  const kmpm_config &cf = c->cfg;
  const bool tl = (cf.method == KMPM_METHOD_TLMPM || cf.method == KMPM_METHOD_TLCPDI);
  int32_t dim = 0, sub = 0, shape = 0;
  std::string method, scheme;
  bool temp = false, dtc = false;
  size_t nargs = 0;
  double flip = 0, time = 0, dt = 0, dtf = 0;
  int64_t step = 0;
  if (!rd(in, dim) || !rd_str(in, method) || !rd(in, temp) || !rd_str(in, scheme) ||
      !rd(in, sub) || !rd(in, flip) || !rd(in, shape) || !rd(in, nargs))
    return KMPM_ERR_ARG;
  for (size_t k = 0; k < nargs; k++) {
    std::string a;
    if (!rd_str(in, a)) return KMPM_ERR_ARG;
  }
  if (!rd(in, time) || !rd(in, step) || !rd(in, dt) || !rd(in, dtf) || !rd(in, dtc))
    return KMPM_ERR_ARG;
  // the file must describe this context's method, scheme and grid
  if (dim != cf.dimension || method != (tl ? "tlmpm" : "ulmpm") ||
      scheme != scheme_key(cf.scheme) || temp != (cf.thermo != 0) ||
      sub != restart_sub_method(cf) ||
      shape != (cf.shape == KMPM_SHAPE_CUBIC_SPLINE ? 1 : 0) || flip != c->a.flip)
    return KMPM_ERR_ARG;
  double box[6];
  for (double &b : box)
    if (!rd(in, b)) return KMPM_ERR_ARG;
  int32_t nsol = 0;
  if (!rd(in, nsol) || nsol < 0) return KMPM_ERR_ARG;
  struct Row {
    int64_t ptag;
    int32_t solid, mat;
    uint32_t mask;
    double x0[3], x[3], v[3], sig[9], eel[9], F[9], J, vol0, rho0, ep, epdot, dmg, dinit, T;
  };
  std::vector<Row> rows;
  for (int32_t is = 0; is < nsol; is++) {
    std::string id;
    double bounds[12], h = 0;
    int64_t np = 0;
    int32_t npl = 0, nc = 0, imat = 0;
    if (!rd_str(in, id)) return KMPM_ERR_ARG;
    for (double &b : bounds)
      if (!rd(in, b)) return KMPM_ERR_ARG;
    if (!rd(in, np) || !rd(in, npl) || !rd(in, nc) || !rd(in, imat) || !rd(in, h))
      return KMPM_ERR_ARG;
    if (npl < 0 || imat < 0 || imat >= c->nmat || h != cf.cellsize) return KMPM_ERR_ARG;
    for (int32_t i = 0; i < npl; i++) {
      Row r{};
      double ien = 0, pk[9];
      int32_t mask = 0;
      r.solid = is;
      r.mat = imat;
      bool ok = rd(in, r.ptag);
      for (int d = 0; d < 3 && ok; d++) ok = rd(in, r.x0[d]);
      for (int d = 0; d < 3 && ok; d++) ok = rd(in, r.x[d]);
      for (int d = 0; d < 3 && ok; d++) ok = rd(in, r.v[d]);
      ok = ok && rd_mat(in, r.sig) && rd_mat(in, r.eel);
      if (tl) ok = ok && rd_mat(in, pk);
      ok = ok && rd_mat(in, r.F) && rd(in, r.J) && rd(in, r.vol0) && rd(in, r.rho0) &&
           rd(in, r.ep) && rd(in, r.epdot) && rd(in, r.dmg) && rd(in, r.dinit);
      if (temp) ok = ok && rd(in, r.T);
      ok = ok && rd(in, ien) && rd(in, mask);
      if (!ok) return KMPM_ERR_ARG;
      r.mask = (uint32_t)mask;
      rows.push_back(r);
    }
  }
  const int64_t n = (int64_t)rows.size();
  Host hb(n);
  for (int64_t i = 0; i < n; i++) {
    const Row &r = rows[i];
    hb.ptag[i] = r.ptag;
    hb.solid[i] = r.solid;
    hb.mat[i] = r.mat;
    hb.mask[i] = r.mask;
    for (int d = 0; d < 3; d++) {
      hb.x0[3 * i + d] = r.x0[d];
      hb.x[3 * i + d] = r.x[d];
      hb.v[3 * i + d] = r.v[d];
    }
    std::memcpy(&hb.F[9 * i], r.F, sizeof r.F);
    std::memcpy(&hb.sigma[9 * i], r.sig, sizeof r.sig);
    std::memcpy(&hb.strain_el[9 * i], r.eel, sizeof r.eel);
    // Solid::read_restart: vol = J vol0, mass = rho0 vol0 (solid.cpp:2931-2936)
    hb.vol0[i] = r.vol0;
    hb.vol[i] = r.J * r.vol0;
    hb.mass[i] = r.rho0 * r.vol0;
    hb.rho0[i] = r.rho0;
    hb.rho[i] = r.rho0 / r.J;
    hb.J[i] = r.J;
    hb.ep[i] = r.ep;
    hb.epdot[i] = r.epdot;
    hb.damage[i] = r.dmg;
    hb.damage_init[i] = r.dinit;
    hb.T[i] = r.T;   // gamma and the heat flux restart from 0
  }
  int st = kmpm_upload_particles(c, &hb.p);
  if (st) return st;
  return kmpm_set_time(c, dt, time, step);
}

}  // extern "C"

// kmpm_abi_driver.cpp -- a C++ program that uses libkmpm.so through
// include/kmpm.h only, the way a Method subclass in the reference would:
// the set-up calls (materials, grid, lattice population, fix expressions),
// then a run loop that calls the phase entry points in the order of the
// reference's scheme loops (usl.cpp:43-93, musl.cpp:43-94), then the error
// path.  Built and run by tests/test_abi_driver.py:
//   kmpm_abi_driver host     set-up calls only (no device)
//   kmpm_abi_driver device   set-up, 200 USL and 200 MUSL steps of the two
//                            disks, the reaction force of a time-dependent
//                            fix, a particle leaving the box
// Prints one line per check and exits 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kmpm.h"

static int failures = 0;

#define CHECK(cond, ...)                      \
  do {                                        \
    if (!(cond)) {                            \
      std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);               \
      std::printf("\n");                      \
      failures++;                             \
    }                                         \
  } while (0)

#define OK(call)                                                              \
  do {                                                                        \
    const int st_ = (call);                                                   \
    CHECK(st_ == KMPM_OK, "%s -> %d (%s)", #call, st_, kmpm_status_string(st_)); \
  } while (0)

// Particle arrays in the reference's per-particle layouts.
struct Arrays {
  std::vector<int64_t> ptag;
  std::vector<int32_t> solid, mat;
  std::vector<uint32_t> mask;
  std::vector<double> x0, x, v, F, sigma, eel, mass, vol0, vol, rho0, rho, J, ep, epdot, dmg, dinit;
  void resize(size_t n) {
    ptag.resize(n); solid.resize(n); mat.resize(n); mask.resize(n);
    for (auto *a : {&x0, &x, &v}) a->assign(3 * n, 0.0);
    for (auto *a : {&F, &sigma, &eel}) a->assign(9 * n, 0.0);
    for (auto *a : {&mass, &vol0, &vol, &rho0, &rho, &J, &ep, &epdot, &dmg, &dinit}) a->assign(n, 0.0);
  }
  kmpm_particles view(size_t n, size_t off = 0) {
    kmpm_particles p{};   // fields this driver does not use stay null
    p.np = (int64_t)n;
    p.ptag = ptag.data() + off; p.solid = solid.data() + off; p.mat = mat.data() + off;
    p.mask = mask.data() + off;
    p.x0 = x0.data() + 3 * off; p.x = x.data() + 3 * off; p.v = v.data() + 3 * off;
    p.F = F.data() + 9 * off; p.sigma = sigma.data() + 9 * off; p.strain_el = eel.data() + 9 * off;
    p.mass = mass.data() + off; p.vol0 = vol0.data() + off; p.vol = vol.data() + off;
    p.rho0 = rho0.data() + off; p.rho = rho.data() + off; p.J = J.data() + off;
    p.eff_plastic_strain = ep.data() + off; p.eff_plastic_strain_rate = epdot.data() + off;
    p.damage = dmg.data() + off; p.damage_init = dinit.data() + off;
    return p;
  }
};

// The two-disk problem (2D, unit square, h = 1/20, disks R = 0.2, 2x2
// particles per cell, linear elastic E = 1000, nu = 0.3, rho = 1000, +-0.1
// along the diagonal, set_dt(0.001)).
struct Setup {
  kmpm_config cfg;
  kmpm_material mat;
  Arrays p;
  size_t np = 0;
};

static void setup(Setup &s, int scheme) {
  std::memset(&s.cfg, 0, sizeof(s.cfg));
  s.cfg.dimension = 2;
  s.cfg.method = KMPM_METHOD_ULMPM;
  s.cfg.scheme = scheme;
  s.cfg.shape = KMPM_SHAPE_LINEAR;
  s.cfg.precision = KMPM_PRECISION_FP64;
  s.cfg.order = KMPM_ORDER_FAST;
  s.cfg.boxhi[0] = s.cfg.boxhi[1] = 1.0;
  s.cfg.cellsize = 1.0 / 20.0;
  s.cfg.pic_flip = 0.99;
  s.cfg.dt = (double)1.0f * std::pow(10.0, -3);    // "1e-3" as the script parser reads it
  s.cfg.dt_constant = 1;
  s.cfg.dt_factor = 0.9;
  std::memset(&s.mat, 0, sizeof(s.mat));
  s.mat.type = KMPM_MAT_LINEAR;
  s.mat.lin_p[0] = 1000.0;
  s.mat.lin_p[1] = 1000.0;
  s.mat.lin_p[2] = (double)0.3f;
  OK(kmpm_material_init(&s.mat));
  CHECK(std::fabs(s.mat.G - 1000.0 / (2 * (1 + s.mat.nu))) < 1e-12, "shear modulus %g", s.mat.G);
  int64_t n[3];
  OK(kmpm_grid_size(&s.cfg, n));
  CHECK(n[0] == 21 && n[1] == 21 && n[2] == 1, "grid %lld %lld %lld", (long long)n[0],
        (long long)n[1], (long long)n[2]);
  const double c[2][2] = {{0.25, 0.25}, {0.75, 0.75}};
  s.p.resize(4096);
  for (int d = 0; d < 2; d++) {
    kmpm_region r;
    std::memset(&r, 0, sizeof(r));
    r.kind = KMPM_REGION_CYLINDER;
    r.axis = 2;
    r.interior = 1;
    r.p[0] = c[d][0];
    r.p[1] = c[d][1];
    r.p[2] = (double)0.2f;
    int64_t got = 0;
    kmpm_particles out = s.p.view(4096 - s.np, s.np);
    OK(kmpm_populate(&s.cfg, &r, 2, &s.mat, 0, d, (int64_t)s.np, &out, 4096 - (int64_t)s.np, &got));
    for (int64_t i = 0; i < got; i++)
      for (int k = 0; k < 2; k++) s.p.v[3 * (s.np + i) + k] = d == 0 ? (double)0.1f : -(double)0.1f;
    s.np += (size_t)got;
  }
  CHECK(s.np == 416, "populated %zu particles", s.np);
  std::printf("set-up: %zu particles, G = %.6f\n", s.np, s.mat.G);
}

static int host_checks() {
  Setup s;
  setup(s, KMPM_SCHEME_USL);
  // a time-dependent value and the previous value of its "time - dt" text
  const char *names[1] = {"v"};
  const double values[1] = {10.0};
  kmpm_expr e, p;
  OK(kmpm_expr_compile("v*time", 0, 1, names, values, &e));
  OK(kmpm_expr_compile("v*time", 1, 1, names, values, &p));
  CHECK(e.n == 3 && e.op[2] == KMPM_OP_MUL, "program of v*time: %d ops", e.n);
  CHECK(p.n == 5 && p.op[4] == KMPM_OP_SUB, "program of v*time - dt: %d ops", p.n);
  kmpm_expr bad;
  CHECK(kmpm_expr_compile("v*tim", 0, 1, names, values, &bad) == KMPM_ERR_ARG, "unknown name");
  std::printf("host checks: %s\n", failures ? "FAILED" : "ok");
  return failures ? 1 : 0;
}

// USL::run / MUSL::run over the phase entry points.  The comments name the
// Modify hooks of the reference's loop at each place.
static int run_steps(kmpm_ctx *ctx, int scheme, int n) {
  for (int k = 0; k < n; k++) {
    int st = kmpm_compute_grid_weight_functions_and_gradients(ctx);
    // reset(); modify->initial_integrate()
    if (!st) st = kmpm_particles_to_grid(ctx);
    // modify->post_particles_to_grid(): body forces, inside the P2G phase
    if (!st) st = kmpm_update_grid_state(ctx);
    // modify->post_update_grid_state(): velocity fixes, inside the grid phase
    if (scheme == KMPM_SCHEME_USL) {
      // compute_rate_deformation_gradient(false), grid_to_points(),
      // post_grid_to_point, advance_particles(), post_advance_particles,
      // update_deformation_gradient(), update_stress(false): one fused phase
      if (!st) st = kmpm_grid_to_points_and_stress(ctx);
    } else {
      if (!st) st = kmpm_grid_to_points(ctx);          // grid_to_points + advance_particles
      // modify->post_advance_particles()
      if (!st) st = kmpm_velocities_to_grid(ctx);      // + post_velocities_to_grid
      if (!st) st = kmpm_update_stress(ctx);           // rate(true), F, update_stress(true)
    }
    if (!st) st = kmpm_adjust_dt(ctx);                 // update_time(), adjust_dt()
    // modify->final_integrate(): energies on output steps
    if (st) return st;
  }
  return KMPM_OK;
}

static int device_checks() {
  for (int scheme : {KMPM_SCHEME_USL, KMPM_SCHEME_MUSL}) {
    Setup s;
    setup(s, scheme);
    kmpm_ctx *ctx = nullptr;
    OK(kmpm_ctx_create(&s.cfg, &s.mat, 1, &ctx));
    if (!ctx) return 1;
    kmpm_particles all = s.p.view(s.np);
    OK(kmpm_upload_particles(ctx, &all));
    // a time-dependent fix on the nodes of the square's left column
    int64_t n[3];
    OK(kmpm_grid_size(&s.cfg, n));
    std::vector<uint32_t> mask((size_t)(n[0] * n[1]), 1u);
    for (int64_t j = 0; j < n[1]; j++) mask[(size_t)j] |= 2u;     // tags ny*i + j with i = 0
    OK(kmpm_set_node_mask(ctx, mask.data(), (int64_t)mask.size()));
    const char *names[1] = {"v"};
    const double values[1] = {0.05};
    kmpm_expr val[3], prev[3];
    std::memset(val, 0, sizeof(val));
    std::memset(prev, 0, sizeof(prev));
    OK(kmpm_expr_compile("v*time", 0, 1, names, values, &val[0]));
    OK(kmpm_expr_compile("v*time", 1, 1, names, values, &prev[0]));
    OK(kmpm_fix_velocity_nodes_expr(ctx, 2u, val, prev));
    double ek0 = 0, es0 = 0;
    OK(kmpm_reduce_energy(ctx, 1u, &ek0, &es0));
    OK(run_steps(ctx, scheme, 200));
    int64_t tag = -1;
    OK(kmpm_synchronize(ctx, &tag));
    double dt = 0, time = 0, ek = 0, es = 0, f[3];
    int64_t step = 0;
    OK(kmpm_get_time(ctx, &dt, &time, &step));
    CHECK(step == 200 && std::fabs(time - 200 * dt) <= 1e-12, "step %lld time %g",
          (long long)step, time);
    OK(kmpm_reduce_energy(ctx, 1u, &ek, &es));
    CHECK(std::isfinite(ek) && std::isfinite(es) && ek + es > 0.5 * ek0 && ek + es < 1.1 * ek0,
          "energy %g + %g against %g", ek, es, ek0);
    OK(kmpm_fix_velocity_nodes_force(ctx, 0, f));
    CHECK(std::isfinite(f[0]) && f[1] == 0 && f[2] == 0, "reaction %g %g %g", f[0], f[1], f[2]);
    Arrays out;
    out.resize(s.np);
    kmpm_particles pv = out.view(s.np);
    OK(kmpm_download_particles(ctx, &pv));
    bool ordered = true, finite = true;
    for (size_t i = 0; i < s.np; i++) {
      ordered = ordered && out.ptag[i] == s.p.ptag[i];
      for (int d = 0; d < 3; d++) finite = finite && std::isfinite(out.x[3 * i + d]);
    }
    CHECK(ordered && finite, "download: ptag order %d, finite %d", ordered, finite);
    std::printf("%s: 200 steps, Ek %.6e Es %.6e, reaction %.6e\n",
                scheme == KMPM_SCHEME_USL ? "usl" : "musl", ek, es, f[0]);
    OK(kmpm_ctx_destroy(ctx));
  }
  {
    // the error path: a particle thrown out of the box (solid.cpp:617-627)
    Setup s;
    setup(s, KMPM_SCHEME_USL);
    s.p.v[3 * 5] = -100.0;
    kmpm_ctx *ctx = nullptr;
    OK(kmpm_ctx_create(&s.cfg, &s.mat, 1, &ctx));
    if (!ctx) return 1;
    kmpm_particles all = s.p.view(s.np);
    OK(kmpm_upload_particles(ctx, &all));
    OK(kmpm_step(ctx, 20));
    int64_t tag = -1;
    const int st = kmpm_synchronize(ctx, &tag);
    CHECK(st == KMPM_ERR_PARTICLE_LEFT_BOX && tag == s.p.ptag[5], "error %d (%s), tag %lld",
          st, kmpm_status_string(st), (long long)tag);
    std::printf("error path: %s, particle %lld\n", kmpm_status_string(st), (long long)tag);
    OK(kmpm_ctx_destroy(ctx));
  }
  std::printf("device checks: %s\n", failures ? "FAILED" : "ok");
  return failures ? 1 : 0;
}

int main(int argc, char **argv) {
  if (argc > 1 && std::strcmp(argv[1], "device") == 0) return device_checks();
  return host_checks();
}

"""Problem descriptions for the explicit MPM step: the configurations the
input scripts of the reference would build (method/dimension/region/
material/solid/fix commands), expressed through the C ABI's POD types.

Numeric literals follow the reference's parser: plain decimals are read
with stof (float32) and widened (src/input.cpp:407,411); integer ratios
such as 1/20 are evaluated in double; `aeb` is a * 10^b (input.cpp:191).
"""
import dataclasses
from typing import List, Optional

import numpy as np

from . import native as N


def lit(x):
    """A decimal literal as the reference's parser stores it (stof, then double)."""
    return float(np.float32(x))


def lit_e(mant, exp):
    """`<mant>e<exp>` literal: mant * pow(10, exp) (input.cpp:191, powv)."""
    return float(np.float32(mant)) * float(10.0 ** exp)


@dataclasses.dataclass
class Fix:
    kind: str                 # "velocity_nodes" | "velocity_nodes_expr" | "body_force"
    bit: int
    set: tuple
    value: tuple              # velocity_nodes_expr: the script's texts (None: free)
    prev: Optional[tuple] = None
    variables: Optional[dict] = None    # velocity_nodes_expr: the script's constants


@dataclasses.dataclass
class Problem:
    name: str
    cfg: N.Config
    mats: List[N.Material]
    particles: dict
    node_mask: Optional[np.ndarray] = None
    fixes: List[Fix] = dataclasses.field(default_factory=list)
    steps: int = 0
    notes: str = ""

    @property
    def np(self):
        return len(self.particles["ptag"])


def make_config(dim, lo, hi, h, scheme="usl", shape="linear", flip=0.99, dt=1e-16,
                dt_constant=False, dt_factor=0.9, method="ulmpm", solid_region=None,
                sub_method="FLIP", np_per_cell=0, thermo=False, cpdi_style="R4"):
    """kmpm_config for dimension/method/scheme/shape commands; for TLMPM and
    TLCPDI the solid's region gives the limits of its grid (Region::limits).
    method is method()'s first word (ulmpm, tlmpm, ulcpdi, tlcpdi),
    sub_method its second (PIC, FLIP, APIC, AFLIP, ASFLIP, MLS); thermo its
    thermo-mechanical keyword; cpdi_style the CPDI methods' style (R4,
    Q4)."""
    c = N.Config()
    c.dimension = dim
    c.method = N.METHOD_KEYS[method]
    c.scheme = N.SCHEME_KEYS[scheme]
    c.shape = N.SHAPE_KEYS[shape]
    c.precision = N.PRECISION_FP64
    for d in range(3):
        c.boxlo[d] = lo[d] if d < dim else 0.0
        c.boxhi[d] = hi[d] if d < dim else 0.0
    c.cellsize = h
    c.pic_flip = flip
    c.dt = dt
    c.dt_constant = 1 if dt_constant else 0
    c.dt_factor = dt_factor
    c.sub_method = N.SUB_METHOD_KEYS[sub_method]
    c.np_per_cell = np_per_cell
    c.thermo = 1 if thermo else 0
    c.cpdi_style = N.CPDI_STYLE_KEYS[cpdi_style]
    if solid_region is not None:
        lo_, hi_ = N.region_limits(solid_region)
        for d in range(3):
            c.solidlo[d] = lo_[d] if d < dim else 0.0
            c.solidhi[d] = hi_[d] if d < dim else 0.0
    return c


def linear_material(rho, E, nu):
    m = N.Material()
    m.type = N.MAT_LINEAR
    m.lin_p[:] = [rho, E, nu]
    return N.material_init(m)


def eos_strength_material(eos, eos_p, strength, strength_p, damage=None, damage_p=()):
    m = N.Material()
    m.type = N.MAT_EOS_STRENGTH
    m.eos = N.EOS_KEYS[eos]
    m.strength = N.STRENGTH_KEYS[strength]
    m.damage = N.DAMAGE_KEYS[damage] if damage else N.DAMAGE_NONE
    m.eos_p[:len(eos_p)] = list(eos_p)
    m.strength_p[:len(strength_p)] = list(strength_p)
    m.damage_p[:len(damage_p)] = list(damage_p)
    return N.material_init(m)


def with_temperature(mat, chi, cp, kappa, alpha, T0, Tm, law="plastic_work"):
    """The material with a temperature law, the last argument of
    material(...): temperature(id, plastic_work, chi, cp, kappa, alpha, T0,
    Tm) (temperature_plastic_work.cpp:50-55)."""
    m = N.Material.from_buffer_copy(mat)
    m.temp = N.TEMP_KEYS[law]
    m.temp_p[:] = [chi, cp, kappa, alpha, T0, Tm]
    return m


def block(xlo, xhi, ylo, yhi, zlo=0.0, zhi=0.0):
    r = N.Region()
    r.kind = N.REGION_BLOCK
    r.interior = 1
    r.p[:] = [xlo, xhi, ylo, yhi, zlo, zhi]
    return r


def cylinder(c1, c2, R, lo=0.0, hi=0.0, axis=2):
    r = N.Region()
    r.kind = N.REGION_CYLINDER
    r.axis = axis
    r.interior = 1
    r.p[:] = [c1, c2, R, lo, hi, 0.0]
    return r


def node_coords(cfg):
    """Grid node positions in tag order (grid.cpp:227-231)."""
    n = N.grid_size(cfg)
    idx = [np.arange(n[d], dtype=np.float64) for d in range(3)]
    out = []
    for d in range(3):
        if d < cfg.dimension:
            out.append(cfg.boxlo[d] + idx[d] * cfg.cellsize)
        else:
            out.append(np.zeros(1))
    X, Y, Z = np.meshgrid(out[0], out[1], out[2], indexing="ij")
    return X.ravel(), Y.ravel(), Z.ravel()


def region_match_nodes(region, X, Y, Z):
    """Region::match on arrays (region_block.cpp:148-154, region_cylinder.cpp:147-168)."""
    p = list(region.p)
    if region.kind == N.REGION_BLOCK:
        inside = (X >= p[0]) & (X <= p[1]) & (Y >= p[2]) & (Y <= p[3]) & (Z >= p[4]) & (Z <= p[5])
    else:
        c1, c2, R, lo, hi = p[:5]
        a, b, al = {0: (Y, Z, X), 1: (X, Z, Y), 2: (X, Y, Z)}[region.axis]
        d2 = (a - c1) * (a - c1) + (b - c2) * (b - c2)
        inside = (al >= lo) & (al <= hi) & (d2 <= R * R)
    return inside if region.interior else ~inside


def populate_solids(cfg, specs, fn=None):
    """specs: list of (region, ppc1d, mat, mat_index) in solid order."""
    parts, off = [], 0
    for s, (region, ppc, mat, mi) in enumerate(specs):
        arrs = N.populate(cfg, region, ppc, mat, mi, s, off, fn=fn)
        off += len(arrs["ptag"])
        parts.append(arrs)
    return N.concat_particles(parts)


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def two_disks(steps=3000, ppc1d=2, flip=0.99, populate_fn=None):
    """Two elastic disks colliding (2D ULMPM USL, linear shape, FP64):
    unit square, cell 1/20, disks R=0.2 at (0.25,0.25) and (0.75,0.75),
    E=1000, nu=0.3, rho=1000, v0 = +-0.1 along the diagonal, dt = 1e-3,
    FLIP 0.99 (the reference's default PIC_FLIP, update.cpp:50; pure FLIP
    lets 1-ulp input differences reach J <= 0 within 3000 steps)."""
    h = 1.0 / 20.0
    cfg = make_config(2, (0.0, 0.0), (1.0, 1.0), h, scheme="usl", shape="linear", flip=flip,
                      dt=lit_e(1, -3), dt_constant=True)
    mat = linear_material(1000.0, 1000.0, lit(0.3))
    R = lit(0.2)
    specs = [(cylinder(0.25, 0.25, R), ppc1d, mat, 0), (cylinder(0.75, 0.75, R), ppc1d, mat, 0)]
    p = populate_solids(cfg, specs, fn=populate_fn)
    v = lit(0.1)
    s0 = p["solid"] == 0
    p["v"][s0, 0] = v
    p["v"][s0, 1] = v
    p["v"][~s0, 0] = -v
    p["v"][~s0, 1] = -v
    return Problem("two_disks", cfg, [mat], p, steps=steps,
                   notes="fix initial_velocity_particles applied at set-up (step-1 hook)")


def steel_jc(damage=True, d_params=(0.05, 3.44, -2.12, 0.002, 0.61)):
    """Steel with Johnson-Cook strength + shock EOS (+ JC damage), SI units.
    E = 200 GPa, nu = 0.3 -> K, G; JC A=792 MPa, B=510 MPa, n=0.26, C=0.014,
    epsdot0 = 1/s, m = 1, Tr = 0, Tm = 1800; shock rho0=7850, c0 = sqrt(K/rho0),
    S = 1.49, Gamma = 2, cv = 0, Q1 = Q2 = 0."""
    E, nu, rho = 200e9, lit(0.3), 7850.0
    K = E / (3 * (1 - 2 * nu))
    G = E / (2 * (1 + nu))
    c0 = float(np.sqrt(K / rho))
    eos_p = (rho, K, c0, lit(1.49), 2.0, 0.0, 0.0, 0.0, 0.0)
    jc_p = (G, 792e6, 510e6, lit(0.26), 1.0, lit(0.014), 1.0, 0.0, 1800.0)
    dmg_p = tuple(lit(d) for d in d_params) + (1.0, 0.0, 1800.0)
    if damage:
        return eos_strength_material("shock", eos_p, "johnson_cook", jc_p,
                                     "damage_johnson_cook", dmg_p)
    return eos_strength_material("shock", eos_p, "johnson_cook", jc_p)


def tensile_bar(h=6e-5, ppc1d=2, v_pull=10.0, shape="cubic-spline", scheme="usl",
                damage=True, d_params=(0.05, 3.44, -2.12, 0.002, 0.61), steps=0,
                populate_fn=None, length=0.02, width=0.0024, pull=None, material=None):
    """3D ULMPM tensile bar with JC strength, shock EOS and JC damage.
    Bar length x width x width centred at the origin in a box twice its
    size; nodes within 0.5 mm of each end are driven at -+v_pull along x
    (fix velocity_nodes).  h = 0.06 mm, 2x2x2 particles per cell gives
    4.27M particles.  pull: the right end's x velocity as an input-script
    expression of time with the variable v = v_pull (the left end takes
    "-(" pull ")"), instead of the constant v_pull.  material: another
    material than the steel."""
    L2, W2 = length / 2, width / 2
    cfg = make_config(3, (-2 * L2, -2 * W2, -2 * W2), (2 * L2, 2 * W2, 2 * W2), h,
                      scheme=scheme, shape=shape, flip=0.99)
    mat = material if material is not None else steel_jc(damage, d_params)
    p = populate_solids(cfg, [(block(-L2, L2, -W2, W2, -W2, W2), ppc1d, mat, 0)],
                        fn=populate_fn)
    X, Y, Z = node_coords(cfg)
    xbc = L2 - 0.0005
    left = region_match_nodes(block(-1e20, -xbc, -1e20, 1e20, -1e20, 1e20), X, Y, Z)
    right = region_match_nodes(block(xbc, 1e20, -1e20, 1e20, -1e20, 1e20), X, Y, Z)
    mask = np.ones(len(X), dtype=np.uint32)
    mask[left] |= 2
    mask[right] |= 4
    if pull is None:
        fixes = [Fix("velocity_nodes", 2, (1, 0, 0), (-v_pull, 0, 0), (-v_pull, 0, 0)),
                 Fix("velocity_nodes", 4, (1, 0, 0), (v_pull, 0, 0), (v_pull, 0, 0))]
    else:
        var = {"v": lit(v_pull)}
        fixes = [Fix("velocity_nodes_expr", 2, (1, 0, 0), ("-(%s)" % pull, None, None),
                     variables=var),
                 Fix("velocity_nodes_expr", 4, (1, 0, 0), (pull, None, None), variables=var)]
    return Problem("tensile_bar", cfg, [mat], p, node_mask=mask, fixes=fixes, steps=steps)


def dam_break(n_cells=32, ppc1d=2, steps=0, scheme="musl", shape="cubic-spline", jitter=0.0,
              seed=20261202, populate_fn=None, dt_factor=0.1):
    """3D ULMPM dam break: unit box, water column 0.4 x 0.4 x 0.8 in a corner,
    eos fluid (rho0=1000, K=2e6, Gamma=7) + strength fluid (mu=1e-3),
    gravity -9.81 along z (fix body_force on all nodes), slip walls on the six
    box faces (velocity_nodes with a zero normal component on nodes within
    h/4 of a face, as examples/Dam-break/water-dam.mpm does in 2D),
    dt_factor 0.1."""
    h = 1.0 / n_cells
    cfg = make_config(3, (0, 0, 0), (1, 1, 1), h, scheme=scheme, shape=shape, flip=0.99,
                      dt_factor=dt_factor)
    mat = eos_strength_material("fluid", (1000.0, 2e6, 7.0), "fluid", (lit_e(1, -3),))
    p = populate_solids(cfg, [(block(0, lit(0.4), 0, lit(0.4), 0, lit(0.8)), ppc1d, mat, 0)],
                        fn=populate_fn)
    if jitter:
        rng = np.random.default_rng(seed)
        p["x"] += rng.uniform(-jitter * h, jitter * h, size=p["x"].shape)
        p["x"] = np.clip(p["x"], 1e-9, 1 - 1e-9)
        p["x0"][:] = p["x"]
    X, Y, Z = node_coords(cfg)
    mask = np.ones(len(X), dtype=np.uint32)
    fixes = [Fix("body_force", 1, (0, 0, 1), (0, 0, lit(-9.81)))]
    big = 1e20
    for axis, C in enumerate((X, Y, Z)):
        for side in (0, 1):
            bit = 1 << (1 + 2 * axis + side)
            lo_ = [-big] * 3
            hi_ = [big] * 3
            if side == 0:
                hi_[axis] = h / 4
            else:
                lo_[axis] = 1.0 - h / 4
            r = block(lo_[0], hi_[0], lo_[1], hi_[1], lo_[2], hi_[2])
            mask[region_match_nodes(r, X, Y, Z)] |= bit
            sv = [0, 0, 0]
            sv[axis] = 1
            fixes.append(Fix("velocity_nodes", bit, tuple(sv), (0, 0, 0), (0, 0, 0)))
    return Problem("dam_break", cfg, [mat], p, node_mask=mask, fixes=fixes, steps=steps)


def copper_jc(q=(0.0, 0.0)):
    """Copper with Johnson-Cook strength and shock EOS, SI units (the Taylor
    bar of the reference's examples with the published JC constants):
    E = 115 GPa, nu = 0.31 -> G; JC A = 90 MPa, B = 292 MPa, n = 0.31,
    C = 0.025, epsdot0 = 1/s, m = 1.09, Tr = 0, Tm = 1356; shock rho0 = 8960,
    K = rho0 c0^2, c0 = 3940 m/s, S = 1.49, Gamma = 2.0, cv = 0, (Q1, Q2) = q
    (the artificial viscosity, eos_shock.cpp:53-61, off by default)."""
    E, nu, rho, c0 = 115e9, lit(0.31), 8960.0, 3940.0
    G = E / (2 * (1 + nu))
    K = rho * c0 * c0
    eos_p = (rho, K, c0, lit(1.49), 2.0, 0.0, 0.0, lit(q[0]), lit(q[1]))
    jc_p = (G, 90e6, 292e6, lit(0.31), 1.0, lit(0.025), lit(1.09), 0.0, 1356.0)
    return eos_strength_material("shock", eos_p, "johnson_cook", jc_p)


def taylor_bar(h=1e-4, ppc1d=1, v0=190.0, shape="cubic-spline", scheme="usl", radius=3.2e-3,
               length=32.4e-3, steps=0, populate_fn=None, dt_factor=0.5, material=None,
               method="tlmpm"):
    """3D TLMPM Taylor bar: copper cylinder along x (radius 3.2 mm, length
    32.4 mm, centred at the origin) moving at +v0 towards a rigid wall at
    its +x end, which is the velocity_nodes fix (zero x velocity) on the
    grid nodes of the end face, as in
    examples/Taylor-bar/cylindrical/TLMPM/cylindrical_taylor_bar.mpm:60-88.
    h = 0.1 mm with 1 particle per cell gives ~1.04M particles.  material:
    another material than the copper; method "ulmpm": the same impact on
    the box grid."""
    L2 = length / 2
    cyl = cylinder(0.0, 0.0, radius, -L2, L2, axis=0)
    margin = 2 * h
    cfg = make_config(3, (-L2 - margin, -radius - margin, -radius - margin),
                      (L2 + margin, radius + margin, radius + margin), h, scheme=scheme,
                      shape=shape, flip=0.99, dt_factor=dt_factor, method=method,
                      solid_region=cyl)
    mat = material if material is not None else copper_jc()
    p = populate_solids(cfg, [(cyl, ppc1d, mat, 0)], fn=populate_fn)
    p["v"][:, 0] = v0
    if method == "tlmpm":
        # TL grid nodes (reference positions of the solid's grid)
        n = N.grid_size(cfg)
        X = np.repeat(cfg.solidlo[0] + h * np.arange(n[0]), n[1] * n[2])
    else:
        X = node_coords(cfg)[0]
    mask = np.ones(len(X), dtype=np.uint32)
    wall = X >= L2 - 0.5 * h          # region(x_wall - h/2, INF, ...)
    mask[wall] |= 2
    fixes = [Fix("velocity_nodes", 2, (1, 0, 0), (0, 0, 0), (0, 0, 0))]
    return Problem("taylor_bar", cfg, [mat], p, node_mask=mask, fixes=fixes, steps=steps,
                   notes="fix initial_velocity_particles applied at set-up (step-1 hook)")


def context_for(problem):
    """Device context holding the problem (particles, node groups, fixes)."""
    ctx = N.Context(problem.cfg, problem.mats)
    ctx.upload(problem.particles)
    if problem.node_mask is not None:
        ctx.set_node_mask(problem.node_mask)
    for f in problem.fixes:
        if f.kind == "velocity_nodes":
            ctx.fix_velocity_nodes(f.bit, f.set, f.value, f.prev)
        elif f.kind == "velocity_nodes_expr":
            ctx.fix_velocity_nodes_expr(f.bit, f.value, f.variables)
        else:
            ctx.fix_body_force(f.bit, f.set, f.value)
    return ctx


# ---------------------------------------------------------------------------
# batched instances
# ---------------------------------------------------------------------------
INSTANCE_ALIGN = 16   # cells: instance ranges are multiples of every 2D tile extent


@dataclasses.dataclass
class Batch:
    """Independent instances packed side by side on one grid."""
    problem: Problem          # every instance's particles, materials, walls
    cell_lo: np.ndarray       # [ninst, 3] cell ranges of the instances
    cell_hi: np.ndarray
    params: list              # per instance: particles, v0, material, bar cells

    @property
    def ninst(self):
        return len(self.params)

    def instance(self, i):
        """Instance i alone on the same grid (same materials and walls)."""
        p = self.problem
        keep = p.particles["solid"] == i
        parts = {k: v[keep].copy() for k, v in p.particles.items()}
        return Problem("%s[%d]" % (p.name, i), p.cfg, p.mats, parts, node_mask=p.node_mask,
                       fixes=p.fixes, steps=p.steps)


def _jc_instance(base, rng, spread):
    """Copper or steel JC + shock material with A, B, n, C scaled by
    independent U(1 - spread, 1 + spread) factors; steel keeps JC damage."""
    m = copper_jc() if base == "copper" else steel_jc(damage=True)
    f = rng.uniform(1 - spread, 1 + spread, size=4)
    sp = list(m.strength_p)
    for k, idx in enumerate((1, 2, 3, 5)):        # A, B, n, C
        sp[idx] *= f[k]
    m.strength_p[:] = sp
    return N.material_init(m), f


def taylor_plates(n_inst=256, np_lo=10_000, np_hi=200_000, h=1e-4, ppc1d=2, seed=20261202,
                  aspect=5.0, v_lo=50.0, v_hi=300.0, spread=0.2, scheme="usl",
                  shape="cubic-spline", precision="fp64", steps=0, populate_fn=None,
                  dt_factor=0.5, align=INSTANCE_ALIGN):
    """Batched 2D plane-strain Taylor impacts (ULMPM): n_inst bars of
    log-uniform particle counts in [np_lo, np_hi] (length = aspect x width),
    each of copper or steel (chosen at random) with its Johnson-Cook A, B, n,
    C scaled by U(0.8, 1.2), moving at U(v_lo, v_hi) along +x into a rigid
    wall (velocity_nodes vx = 0 on the nodes at and beyond the bar's front
    face).  Each instance owns a tile-aligned cell range with room for the
    bar's flattening; the ranges are shelf-packed on one grid."""
    rng = np.random.default_rng(seed)
    spec = []
    for i in range(n_inst):
        n_target = float(np.exp(rng.uniform(np.log(np_lo), np.log(np_hi))))
        cells = n_target / (ppc1d * ppc1d)
        W = max(4, int(round(np.sqrt(cells / aspect))))
        L = max(4, int(round(cells / W)))
        base = "copper" if rng.uniform() < 0.5 else "steel"
        v0 = float(rng.uniform(v_lo, v_hi))
        mat, f = _jc_instance(base, rng, spread)
        side = int(np.ceil(0.75 * W)) + 6
        wx = -(-(6 + L + 8) // align) * align
        wy = -(-(W + 2 * side) // align) * align
        spec.append(dict(W=W, L=L, base=base, v0=v0, mat=mat, scale=f, wx=wx, wy=wy))
    # shelf packing, tallest first, rows of about the square root of the area
    order = sorted(range(n_inst), key=lambda i: -spec[i]["wy"])
    area = sum(s["wx"] * s["wy"] for s in spec)
    row_w = max(max(s["wx"] for s in spec), int(np.ceil(np.sqrt(area) * 1.05 / align)) * align)
    x = y = row_h = 0
    NX = 0
    for i in order:
        s = spec[i]
        if x + s["wx"] > row_w:
            y += row_h
            x = row_h = 0
        s["ox"], s["oy"] = x, y
        x += s["wx"]
        row_h = max(row_h, s["wy"])
        NX = max(NX, x)
    NY = y + row_h
    cfg = make_config(2, (0.0, 0.0), (NX * h, NY * h), h, scheme=scheme, shape=shape,
                      flip=0.99, dt_factor=dt_factor)
    if precision == "mixed":
        cfg.precision = N.PRECISION_MIXED
    mats = [s["mat"] for s in spec]
    nx, ny, _ = N.grid_size(cfg)
    mask = np.ones(nx * ny, dtype=np.uint32)
    parts, off = [], 0
    lo = np.zeros((n_inst, 3), dtype=np.int32)
    hi = np.zeros((n_inst, 3), dtype=np.int32)
    hi[:, 2] = 1
    params = []
    for i, s in enumerate(spec):
        ox, oy = s["ox"], s["oy"]
        lo[i, :2] = (ox, oy)
        hi[i, :2] = (ox + s["wx"], oy + s["wy"])
        bx0 = ox + 6
        by0 = oy + (s["wy"] - s["W"]) // 2
        bar = block(bx0 * h, (bx0 + s["L"]) * h, by0 * h, (by0 + s["W"]) * h)
        # populate on the instance's own cells (same lattice as the grid's)
        sub = make_config(2, (ox * h, oy * h), ((ox + s["wx"]) * h, (oy + s["wy"]) * h), h,
                          scheme=scheme, shape=shape)
        arrs = N.populate(sub, bar, ppc1d, s["mat"], i, i, off, fn=populate_fn)
        arrs["v"][:, 0] = s["v0"]
        off += len(arrs["ptag"])
        parts.append(arrs)
        # wall: nodes of the instance at and beyond the front face
        front = bx0 + s["L"]
        ii = np.arange(front, ox + s["wx"])
        jj = np.arange(oy, oy + s["wy"])
        ii = ii[ii < nx]
        jj = jj[jj < ny]
        mask[(ii[:, None] * ny + jj[None, :]).ravel()] |= 2
        params.append(dict(np=len(arrs["ptag"]), v0=s["v0"], base=s["base"],
                           scale=s["scale"].tolist(), bar_cells=(s["L"], s["W"])))
    p = N.concat_particles(parts)
    fixes = [Fix("velocity_nodes", 2, (1, 0, 0), (0, 0, 0), (0, 0, 0))]
    prob = Problem("taylor_plates", cfg, mats, p, node_mask=mask, fixes=fixes, steps=steps,
                   notes="fix initial_velocity_particles applied at set-up (step-1 hook)")
    return Batch(prob, lo, hi, params)


def context_for_batch(batch):
    ctx = context_for(batch.problem)
    ctx.set_instances(batch.cell_lo, batch.cell_hi)
    return ctx

"""GPU parity of the CPDI methods against the oracle, through the C ABI
(ULCPDI and TLCPDI, styles R4 and Q4; 2D, FP64).

- the collision of the two elastic disks (ulcpdi.cpp, solid.cpp:696-784,
  1188-1200, 2338-2352) with USL, MUSL and USF, linear and cubic shape
  functions: the weights of the particle domains, the node-by-node position
  update, the Q4 corners and their volume, the R4 vectors following F;
- a JC steel block on TLCPDI (tlcpdi.cpp: weights once from the reference
  domains, candidates from the box's corner) under an impact velocity;
- the domains download back in the reference's layouts.

Bar: 1e-10 relative L2 on every particle field and on the domains after N
steps (conditioned by the oracle's own 1-ulp sensitivity on the
collisions, as the other collision runs).
"""
import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import native as N
from karamelo_amd import problems as P
from test_gpu_parity import FIELDS, rel_l2, run_and_compare, say

pytestmark = pytest.mark.gpu

CPDI_FIELDS = FIELDS + ("vol", "J")


def cpdi_disks(style, scheme="usl", shape="linear", flip=0.99):
    """The two-disk collision on ULCPDI."""
    h = 1.0 / 20.0
    cfg = P.make_config(2, (0.0, 0.0), (1.0, 1.0), h, scheme=scheme, shape=shape, flip=flip,
                        dt=P.lit_e(1, -3), dt_constant=True, method="ulcpdi", cpdi_style=style)
    mat = P.linear_material(1000.0, 1000.0, P.lit(0.3))
    R = P.lit(0.2)
    specs = [(P.cylinder(0.25, 0.25, R), 2, mat, 0), (P.cylinder(0.75, 0.75, R), 2, mat, 0)]
    p = P.populate_solids(cfg, specs)
    v = P.lit(0.1)
    s0 = p["solid"] == 0
    p["v"][s0, :2] = v
    p["v"][~s0, :2] = -v
    return P.Problem("cpdi_disks", cfg, [mat], p)


def tl_block(style, scheme="usl", shape="cubic-spline"):
    """A JC steel block (TLCPDI) at the box's lower corner moving towards
    the box's far corner, its leading edge held by a velocity fix."""
    reg = P.block(0.0, 4e-3, 0.0, 2e-3)
    h = 2.5e-4
    cfg = P.make_config(2, (0.0, 0.0), (8e-3, 4e-3), h, scheme=scheme, shape=shape, flip=0.99,
                        method="tlcpdi", cpdi_style=style, solid_region=reg, dt_factor=0.5)
    mat = P.steel_jc(damage=False)
    p = P.populate_solids(cfg, [(reg, 2, mat, 0)])
    p["v"][:, 0] = 150.0
    n = N.grid_size(cfg)
    X = np.repeat(cfg.solidlo[0] + h * np.arange(n[0]), n[1] * n[2])
    mask = np.ones(len(X), dtype=np.uint32)
    mask[X >= 4e-3 - 0.5 * h] |= 2
    fixes = [P.Fix("velocity_nodes", 2, (1, 0, 0), (0, 0, 0), (0, 0, 0))]
    return P.Problem("cpdi_tl_block", cfg, [mat], p, node_mask=mask, fixes=fixes)


def compare_domains(prob, steps):
    """Both runs `steps` further; the current domains against each other."""
    ctx = P.context_for(prob)
    o = O.oracle_for(prob)
    ctx.step(steps)
    assert o.step(steps) == 0
    st, tag = ctx.synchronize()
    assert st == 0, (st, tag)
    pg, po = ctx.download(), o.download()
    key = "rp" if prob.cfg.cpdi_style == N.CPDI_R4 else "xpc"
    e = rel_l2(pg[key], po[key])
    say("  %s rel L2 %.2e" % (key, e))
    ref = "rp0" if key == "rp" else "xpc0"
    assert np.array_equal(pg[ref], po[ref])
    ctx.close()
    return e


UL_CASES = [(style, scheme, shape) for style in ("R4", "Q4") for scheme in ("usl", "musl", "usf")
            for shape in ("linear", "cubic-spline")]


@pytest.mark.parametrize("style,scheme,shape", UL_CASES)
def test_ul_two_disks(style, scheme, shape):
    say("ULCPDI %s two disks %s %s: 1400 steps" % (style, scheme, shape))
    prob = cpdi_disks(style, scheme, shape)
    run_and_compare(prob, 1400, 350, conditioned=True, fields=CPDI_FIELDS)
    assert compare_domains(prob, 300) <= 1e-10


@pytest.mark.parametrize("style,scheme", [("R4", "usl"), ("Q4", "usl"), ("R4", "musl"),
                                          ("Q4", "musl"), ("R4", "usf"), ("Q4", "usf")])
def test_tl_jc_block(style, scheme):
    say("TLCPDI %s JC block %s: 120 steps" % (style, scheme))
    prob = tl_block(style, scheme)
    run_and_compare(prob, 120, 40, conditioned=True, fields=CPDI_FIELDS)
    assert compare_domains(prob, 60) <= 1e-10


def test_domains_round_trip_and_rejections():
    say("CPDI: upload/download of the domains, rejected configurations")
    prob = cpdi_disks("Q4")
    ctx = P.context_for(prob)
    d = ctx.download()
    assert np.array_equal(d["xpc0"], prob.particles["xpc0"])
    assert np.array_equal(d["xpc"], prob.particles["xpc"])
    with pytest.raises(N.KmpmError):
        ctx.write_restart("/tmp/cpdi.restart")     # the reference's restart holds no domains
    ctx.close()
    for kw in ({"precision": N.PRECISION_MIXED}, {"sub_method": N.SUB_APIC},
               {"order": N.ORDER_REFERENCE}, {"thermo": 1}):
        cfg = N.Config.from_buffer_copy(prob.cfg)
        for k, v in kw.items():
            setattr(cfg, k, v)
        with pytest.raises(N.KmpmError):
            N.Context(cfg, prob.mats)

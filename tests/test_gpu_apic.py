"""GPU parity of the APIC family (APIC, AFLIP, ASFLIP, MLS) and of the USF
scheme against the oracle, through the C ABI.

- ULMPM cubic: the affine P2G momentum (solid.cpp:392-426), the affine rate
  of deformation (:1081-1153), the MLS force (:524-574) and ASFLIP's
  position update (:637-685, 791-796), with USL, MUSL and USF;
- TLMPM APIC (Fdot as the affine matrix, :1006-1079) with cubic and with
  linear shape functions (Di from the particles per cell, :1454-1462);
- USF (usf.cpp:34-94) with PIC/FLIP and APIC, its velocity-fix reaction;
- the mixed path of each against the FP64 oracle.

Bars: 1e-10 relative L2 on every particle field after N steps for FP64
(conditioned by the oracle's own 1-ulp sensitivity on the chaotic Taylor
impacts); the affine matrix within 1e-10 of the oracle's; mixed runs within
1e-3 of the FP64 oracle, as the other mixed-path tests.
"""
import dataclasses

import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import native as N
from karamelo_amd import problems as P
from test_gpu_parity import rel_l2, run_and_compare, say

pytestmark = pytest.mark.gpu

TOL_AFFINE = 1e-10
TOL_RUN_MIXED = 1e-3


def with_cfg(prob, **kw):
    """The same problem with config fields changed (sub_method given by name)."""
    cfg = N.Config.from_buffer_copy(prob.cfg)
    for k, v in kw.items():
        if k == "sub_method":
            v = N.SUB_METHOD_KEYS[v]
        elif k == "scheme":
            v = N.SCHEME_KEYS[v]
        setattr(cfg, k, v)
    return dataclasses.replace(prob, cfg=cfg,
                               particles={k: v.copy() for k, v in prob.particles.items()})


def cubic_disks():
    prob = P.two_disks()
    prob.cfg.shape = N.SHAPE_CUBIC_SPLINE
    return prob


def compare_affine(prob, steps):
    """Steps both, then the affine matrices against each other: relative L2
    with a floor of 1e-4 v_max / h per entry, below which L is the rounding
    noise of a rigid motion (USL and USF take the moments about the
    particle's own position, so a translating body's L is pure rounding)."""
    ctx = P.context_for(prob)
    o = O.oracle_for(prob)
    ctx.step(steps)
    assert o.step(steps) == 0
    st, tag = ctx.synchronize()
    assert st == 0, (st, tag)
    Cg, Co = ctx.affine(), o.affine()
    vmax = np.abs(o.download()["v"]).max()
    floor = 1e-4 * vmax / prob.cfg.cellsize * np.sqrt(Co.size)
    e = rel_l2(Cg, Co, floor)
    say("  affine rel L2 %.2e (|C| %.3e, floor %.3e)" % (e, np.linalg.norm(Co), floor))
    ctx.close()
    return e


UL_CASES = [(sub, scheme) for scheme in ("usl", "musl", "usf")
            for sub in ("APIC", "AFLIP", "ASFLIP", "MLS")]


@pytest.mark.parametrize("sub,scheme", UL_CASES)
def test_ul_two_disks(sub, scheme):
    """The elastic two-disk collision (conditioned, as its FLIP run)."""
    say("two disks cubic %s %s: 120 steps" % (sub, scheme))
    prob = with_cfg(cubic_disks(), sub_method=sub, scheme=scheme)
    run_and_compare(prob, 120, 40, conditioned=True)
    assert compare_affine(prob, 40) <= TOL_AFFINE


@pytest.mark.parametrize("sub,scheme", [("APIC", "usl"), ("MLS", "usl"), ("ASFLIP", "usl"),
                                        ("APIC", "usf"), ("MLS", "usf")])
def test_ul_tensile_bar_3d(sub, scheme):
    """3D cubic, JC steel with damage, velocity fixes at both ends."""
    say("tensile bar 3D %s %s: 60 steps" % (sub, scheme))
    prob = with_cfg(P.tensile_bar(h=6e-4, v_pull=40.0), sub_method=sub, scheme=scheme)
    run_and_compare(prob, 60, 30)
    assert compare_affine(prob, 30) <= TOL_AFFINE


@pytest.mark.parametrize("sub,scheme", [("APIC", "musl"), ("MLS", "musl"), ("AFLIP", "usf")])
def test_ul_dam_break(sub, scheme):
    """2D fluid column (cubic): the APIC family with MUSL and USF."""
    say("dam break %s %s: 100 steps" % (sub, scheme))
    prob = with_cfg(P.dam_break(n_cells=16), sub_method=sub, scheme=scheme)
    run_and_compare(prob, 100, 50)
    assert compare_affine(prob, 50) <= TOL_AFFINE


def test_usf_two_disks_flip_and_dam_break():
    say("USF FLIP: two disks cubic 120 steps, dam break 80 steps")
    run_and_compare(with_cfg(cubic_disks(), scheme="usf"), 120, 40)
    run_and_compare(with_cfg(P.dam_break(n_cells=16), scheme="usf"), 80, 40)


def test_usf_tensile_bar_fix_reaction():
    """The velocity fixes' reaction force of USF's second grid update (the
    grid velocity starts from the fixes' previous values)."""
    say("USF tensile bar: reaction force")
    prob = with_cfg(P.tensile_bar(h=6e-4, v_pull=40.0), scheme="usf")
    run_and_compare(prob, 40, 20)
    ctx = P.context_for(prob)
    o = O.oracle_for(prob)
    for _ in range(3):
        ctx.step(10)
        assert o.step(10) == 0
        for k in range(len(prob.fixes)):   # the tensile bar's fixes are all velocity fixes
            assert prob.fixes[k].kind.startswith("velocity_nodes")
            fg, fo = np.array(ctx.fix_force(k)), np.array(o.fix_force(k))
            assert np.abs(fg - fo).max() <= 1e-10 * max(np.abs(fo).max(), 1e-30), (k, fg, fo)
    ctx.close()


@pytest.mark.parametrize("scheme", ["usl", "musl", "usf"])
def test_tl_apic_taylor_cubic(scheme):
    say("TL APIC Taylor bar cubic %s" % scheme)
    prob = with_cfg(P.taylor_bar(h=8e-4, scheme=scheme), sub_method="APIC")
    run_and_compare(prob, 120, 60, conditioned=True)
    assert compare_affine(prob, 30) <= TOL_AFFINE


def test_tl_apic_taylor_linear_two_per_cell():
    say("TL APIC Taylor bar linear, 2 particles per cell")
    prob = with_cfg(P.taylor_bar(h=8e-4, shape="linear", ppc1d=2), sub_method="APIC",
                    np_per_cell=2)
    run_and_compare(prob, 100, 50, conditioned=True)


def test_tl_flip_usf():
    say("TL FLIP USF Taylor bar")
    run_and_compare(with_cfg(P.taylor_bar(h=8e-4), scheme="usf"), 120, 60, conditioned=True)


def _mixed(prob):
    return with_cfg(prob, precision=N.PRECISION_MIXED)


MIXED_CASES = [("tensile", "APIC", "usl"), ("tensile", "MLS", "usl"), ("tensile", "ASFLIP", "usl"),
               ("tensile", "FLIP", "usf"), ("tensile", "APIC", "usf"), ("dam", "MLS", "musl"),
               ("dam", "APIC", "musl")]


@pytest.mark.parametrize("name,sub,scheme", MIXED_CASES)
def test_mixed_against_fp64_oracle(name, sub, scheme):
    """200 steps.  (The pulled bar with the APIC family and MUSL loses a
    particle near step 180 in the oracle as well: MUSL takes the moments of
    the re-mapped velocities about the advanced positions with the step's
    first weights, so it runs on the dam break.)"""
    say("mixed %s %s %s: 200 steps" % (name, sub, scheme))
    base = P.tensile_bar(h=6e-4, v_pull=40.0) if name == "tensile" else P.dam_break(n_cells=16)
    prob = with_cfg(base, sub_method=sub, scheme=scheme)
    x0 = prob.particles["x"].copy()
    o = O.oracle_for(prob)
    ctx = P.context_for(_mixed(prob))
    G = max(m.G for m in prob.mats)
    worst = {"u": 0.0, "sigma": 0.0}
    for _ in range(2):
        ctx.step(100)
        assert o.step(100) == 0
        st, tag = ctx.synchronize()
        assert st == 0, (st, tag)
        pg, po = ctx.download(), o.download()
        worst["u"] = max(worst["u"], rel_l2(pg["x"] - x0, po["x"] - x0))
        worst["sigma"] = max(worst["sigma"],
                             rel_l2(pg["sigma"], po["sigma"], 1e-6 * G * np.sqrt(prob.np)))
    say("  worst rel L2 vs FP64: u=%.2e sigma=%.2e" % (worst["u"], worst["sigma"]))
    assert worst["u"] <= TOL_RUN_MIXED and worst["sigma"] <= TOL_RUN_MIXED
    ctx.close()


def test_mixed_usf_apic_graph_replay_is_bitwise():
    """Captured steps of the mixed USF + APIC step equal eager steps bit for
    bit (fixed-point node sums), over several replays."""
    say("mixed USF APIC: graph replays vs eager")
    prob = _mixed(with_cfg(P.tensile_bar(h=6e-4, v_pull=40.0), sub_method="APIC",
                           scheme="usf"))
    a, b = P.context_for(prob), P.context_for(prob)
    for _ in range(3):
        a.step(10)
        b.step_graph(10, 5)
        assert a.synchronize()[0] == 0 and b.synchronize()[0] == 0
        pa, pb = a.download(), b.download()
        for f in ("x", "v", "sigma", "F"):
            assert np.array_equal(pa[f], pb[f]), f
        assert np.array_equal(a.affine(), b.affine())
    a.close()
    b.close()


def test_rejected_configurations():
    say("APIC/USF: rejected configurations")
    lin = with_cfg(P.two_disks(), sub_method="APIC")          # ULMPM linear: no inertia tensor
    with pytest.raises(N.KmpmError):
        N.Context(lin.cfg, lin.mats)
    ref = with_cfg(cubic_disks(), sub_method="APIC", order=N.ORDER_REFERENCE)
    with pytest.raises(N.KmpmError):
        N.Context(ref.cfg, ref.mats)
    tl_lin = with_cfg(P.taylor_bar(h=8e-4, shape="linear"), sub_method="APIC", np_per_cell=3)
    with pytest.raises(N.KmpmError):
        N.Context(tl_lin.cfg, tl_lin.mats)
    prob = with_cfg(cubic_disks(), sub_method="APIC")
    ctx = P.context_for(prob)
    with pytest.raises(N.KmpmError):
        ctx.set_tile_variant(1)           # APIC kernels: tile variant 0 only
    plain = P.context_for(cubic_disks())
    with pytest.raises(N.KmpmError):
        plain.affine()                    # no affine matrix without APIC
    plain.close()
    ctx.close()

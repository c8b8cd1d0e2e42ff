"""The mixed-precision step (FP32 transfers and state, FP64 position and
deformation gradient, 64-bit fixed-point node sums) against the FP64 oracle.

Bars:
  * one P2G + grid update: node mass, v and v_update within 1e-5 relative
    max error of the FP64 oracle (FP32 transfer arithmetic), total node mass
    equal to the particle mass within 1e-6;
  * whole runs: displacement (x - x0) and stress within 1e-3 relative L2
    of the FP64 oracle (the accuracy the mixed path is specified to);
  * determinism: two runs from the same inputs agree bit for bit (the
    fixed-point node sums do not depend on the order of the adds).
"""
import dataclasses
import sys

import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import native as N
from karamelo_amd import problems as P

pytestmark = pytest.mark.gpu

TOL_GRID_MIXED = 1e-5
TOL_RUN_MIXED = 1e-3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def mixed(prob):
    """The same problem on the mixed-precision path."""
    cfg = N.Config.from_buffer_copy(prob.cfg)
    cfg.precision = N.PRECISION_MIXED
    return dataclasses.replace(prob, cfg=cfg)


def rel_l2(a, b, floor=0.0):
    a = np.asarray(a, float).ravel()
    b = np.asarray(b, float).ravel()
    nb = max(np.linalg.norm(b), floor)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _randomize(prob, seed=11, vscale=1.0, sscale=1e6):
    rng = np.random.default_rng(seed)
    p = prob.particles
    dim = prob.cfg.dimension
    p["v"][:, :dim] = rng.normal(scale=vscale, size=(len(p["v"]), dim))
    s = rng.normal(scale=sscale, size=(len(p["v"]), 3, 3))
    s = 0.5 * (s + s.transpose(0, 2, 1))
    if dim == 2:
        s[:, 2, :2] = 0
        s[:, :2, 2] = 0
    p["sigma"][:] = s.reshape(-1, 9)
    return prob


def small_problems():
    def cubic_disks():
        prob = P.two_disks()
        prob.cfg.shape = N.SHAPE_CUBIC_SPLINE
        return prob
    return {
        "two_disks": lambda: P.two_disks(),
        "two_disks_cubic": cubic_disks,
        "tensile_cubic": lambda: P.tensile_bar(h=6e-4),
        "tensile_linear": lambda: P.tensile_bar(h=6e-4, shape="linear"),
        "dam_cubic": lambda: P.dam_break(n_cells=16),
    }


@pytest.mark.parametrize("name", list(small_problems()))
def test_mixed_p2g_and_grid_update(name):
    say("mixed p2g/grid: %s" % name)
    prob = _randomize(small_problems()[name]())
    o = O.oracle_for(prob)
    ctx = P.context_for(mixed(prob))
    for ph in ("compute_grid_weight_functions_and_gradients", "particles_to_grid",
               "update_grid_state"):
        ctx.phase(ph)
    for ph in ("weights", "p2g", "grid_update"):
        o.phase(ph)
    mg, vg, vug = ctx.grid()
    mo, vo, vuo, fo = o.grid()
    st, _ = ctx.synchronize()
    assert st == 0
    assert np.abs(mg - mo).max() <= TOL_GRID_MIXED * np.abs(mo).max()
    assert np.abs(vg - vo).max() <= TOL_GRID_MIXED * np.abs(vo).max()
    assert np.abs(vug - vuo).max() <= TOL_GRID_MIXED * np.abs(vuo).max()
    assert abs(mg.sum() - prob.particles["mass"].sum()) <= 1e-6 * prob.particles["mass"].sum()
    ctx.close()


@pytest.mark.parametrize("name,variant", [(n, v) for n in ("two_disks_cubic", "tensile_cubic",
                                                             "tensile_linear")
                                          for v in (0, 1, 2)])
def test_mixed_p2g_and_grid_update_tile_variants(name, variant):
    """Every tile shape of the mixed grid update (64 to 512 nodes per tile,
    2D and 3D, both shape functions): one P2G + grid update against the
    FP64 oracle, to the bars of the default shape."""
    say("mixed p2g/grid: %s, tile variant %d" % (name, variant))
    prob = _randomize(small_problems()[name]())
    o = O.oracle_for(prob)
    ctx = P.context_for(mixed(prob))
    ctx.set_tile_variant(variant)
    say("  tile %s" % (ctx.tile_shape(),))
    for ph in ("compute_grid_weight_functions_and_gradients", "particles_to_grid",
               "update_grid_state"):
        ctx.phase(ph)
    for ph in ("weights", "p2g", "grid_update"):
        o.phase(ph)
    mg, vg, vug = ctx.grid()
    mo, vo, vuo, _ = o.grid()
    st, _ = ctx.synchronize()
    ctx.close()
    assert st == 0
    assert np.abs(mg - mo).max() <= TOL_GRID_MIXED * np.abs(mo).max()
    assert np.abs(vg - vo).max() <= TOL_GRID_MIXED * np.abs(vo).max()
    assert np.abs(vug - vuo).max() <= TOL_GRID_MIXED * np.abs(vuo).max()


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_mixed_musl_dam_break_tile_variants(variant):
    """MUSL (its second P2G re-maps momentum through the grid update's
    momentum-only form) on every tile shape: 60 mixed steps of the dam break
    within the mixed bar of the FP64 oracle."""
    say("mixed MUSL dam break, tile variant %d" % variant)
    w = run_mixed_vs_oracle(P.dam_break(n_cells=16), 60, 60, variant=variant)
    assert w["u"] <= TOL_RUN_MIXED and w["sigma"] <= TOL_RUN_MIXED


def test_mixed_p2g_crowded_cell_takes_the_exact_split():
    """300 particles next to one node: a cell run (segment) of up to 128
    particles sums to about 2^52 mass quanta, past the 2^51 range of the
    three-instruction fixed-point word, so those waves take the exact
    two-word split; the node sums still match the FP64 oracle."""
    say("mixed p2g: crowded cell")
    prob = _randomize(P.two_disks(ppc1d=4))
    p = prob.particles
    h = prob.cfg.cellsize
    idx = np.flatnonzero(p["solid"] == 0)[:300]
    assert len(idx) == 300
    k = np.arange(300)
    # node (5, 5) at (0.25, 0.25); offsets of 1e-3 h to 1e-2 h keep every
    # particle inside cell (5, 5) with a linear weight of about 0.99 to it
    p["x"][idx, 0] = 0.25 + h * 1e-3 * (1 + k % 10)
    p["x"][idx, 1] = 0.25 + h * 1e-3 * (1 + (k // 10) % 10)
    o = O.oracle_for(prob)
    ctx = P.context_for(mixed(prob))
    for ph in ("compute_grid_weight_functions_and_gradients", "particles_to_grid",
               "update_grid_state"):
        ctx.phase(ph)
    for ph in ("weights", "p2g", "grid_update"):
        o.phase(ph)
    mg, vg, _ = ctx.grid()
    mo, vo, _, _ = o.grid()
    st, _ = ctx.synchronize()
    ctx.close()
    assert st == 0
    node = int(np.argmax(mo))
    assert mo[node] > 250 * p["mass"][idx[0]]
    assert np.abs(mg - mo).max() <= TOL_GRID_MIXED * np.abs(mo).max()
    assert np.abs(vg - vo).max() <= TOL_GRID_MIXED * np.abs(vo).max()
    assert abs(mg.sum() - p["mass"].sum()) <= 1e-6 * p["mass"].sum()


def run_mixed_vs_oracle(prob, steps, chunk, variant=None):
    """Mixed GPU run (tile variant `variant`, or the default) against the
    FP64 oracle every `chunk` steps; returns the worst relative L2 error of
    displacement and stress."""
    x0 = prob.particles["x"].copy()
    o = O.oracle_for(prob)
    ctx = P.context_for(mixed(prob))
    if variant is not None:
        ctx.set_tile_variant(variant)
    G = max(m.G for m in prob.mats)
    worst = {"u": 0.0, "sigma": 0.0, "v": 0.0}
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        ctx.step(n)
        assert o.step(n) == 0
        done += n
        st, tag = ctx.synchronize()
        assert st == 0, (st, tag)
        pg, po = ctx.download(), o.download()
        assert np.array_equal(pg["ptag"], po["ptag"])
        floor_s = 1e-6 * G * np.sqrt(prob.np)
        e = {"u": rel_l2(pg["x"] - x0, po["x"] - x0),
             "sigma": rel_l2(pg["sigma"], po["sigma"], floor_s),
             "v": rel_l2(pg["v"], po["v"])}
        for k, val in e.items():
            worst[k] = max(worst[k], val)
    say("  worst rel L2 vs FP64: " + ", ".join("%s=%.2e" % kv for kv in worst.items()))
    ctx.close()
    return worst


def test_mixed_tensile_bar_usl_cubic_plastic():
    prob = P.tensile_bar(h=6e-4, v_pull=40.0)
    w = run_mixed_vs_oracle(prob, 300, 100)
    assert w["u"] <= TOL_RUN_MIXED and w["sigma"] <= TOL_RUN_MIXED


def test_mixed_tensile_bar_musl_linear():
    prob = P.tensile_bar(h=6e-4, v_pull=40.0, shape="linear", scheme="musl")
    w = run_mixed_vs_oracle(prob, 200, 100)
    assert w["u"] <= TOL_RUN_MIXED and w["sigma"] <= TOL_RUN_MIXED


def test_mixed_dam_break_musl_cubic():
    prob = P.dam_break(n_cells=16)
    w = run_mixed_vs_oracle(prob, 300, 100)
    assert w["u"] <= TOL_RUN_MIXED and w["sigma"] <= TOL_RUN_MIXED


def _jittered_disks(shape="linear", jitter=1e-3, seed=5):
    """Two disks with every particle moved by up to jitter x h.  On the
    exact lattice all 416 particles cross node lines at the same steps, and
    a particle within rounding of a node line leaves a node of vanishing
    mass whose linear-shape gradient is still +-1/h: FP32 stress noise over
    that mass is a huge node acceleration (the FP64 path's noise is 1e-9 of
    it).  Off the lattice no particle comes that close."""
    prob = P.two_disks(flip=1.0)
    if shape != "linear":
        prob.cfg.shape = N.SHAPE_CUBIC_SPLINE
    rng = np.random.default_rng(seed)
    h = prob.cfg.cellsize
    prob.particles["x"][:, :2] += rng.uniform(-jitter * h, jitter * h, size=(prob.np, 2))
    prob.particles["x0"][:] = prob.particles["x"]
    return prob


@pytest.mark.parametrize("shape", ["linear", "cubic-spline"])
def test_mixed_two_disks_energy(shape):
    prob = _jittered_disks(shape)
    o = O.oracle_for(prob)
    ctx = P.context_for(mixed(prob))
    for _ in range(3):
        ctx.step(500)
        assert o.step(500) == 0
        eg, eo = ctx.energy(), o.energy()
        tot = abs(eo[0]) + abs(eo[1])
        assert abs(eg[0] - eo[0]) <= 1e-3 * tot and abs(eg[1] - eo[1]) <= 1e-3 * tot, (eg, eo)
    assert ctx.synchronize()[0] == 0
    ctx.close()


def test_mixed_tl_p2g_and_grid_update():
    say("mixed TL p2g/grid")
    prob = _randomize(P.taylor_bar(h=8e-4), sscale=1e8)
    o = O.oracle_for(prob)
    ctx = P.context_for(mixed(prob))
    for ph in ("compute_grid_weight_functions_and_gradients", "particles_to_grid",
               "update_grid_state"):
        ctx.phase(ph)
    for ph in ("weights", "p2g", "grid_update"):
        o.phase(ph)
    mg, vg, vug = ctx.grid()
    mo, vo, vuo, fo = o.grid()
    assert ctx.synchronize()[0] == 0
    assert np.abs(mg - mo).max() <= TOL_GRID_MIXED * np.abs(mo).max()
    assert np.abs(vg - vo).max() <= TOL_GRID_MIXED * np.abs(vo).max()
    assert np.abs(vug - vuo).max() <= TOL_GRID_MIXED * np.abs(vuo).max()
    ctx.close()


@pytest.mark.parametrize("scheme,shape", [("usl", "cubic-spline"), ("musl", "linear")])
def test_mixed_tl_taylor_bar(scheme, shape):
    """TLMPM on the mixed path (FP32 state and transfers; F, its inverse,
    polar decomposition and J in FP64) against the FP64 oracle."""
    say("mixed TL Taylor bar: %s %s" % (scheme, shape))
    prob = P.taylor_bar(h=8e-4, scheme=scheme, shape=shape)
    w = run_mixed_vs_oracle(prob, 200, 100)
    assert w["u"] <= TOL_RUN_MIXED and w["sigma"] <= TOL_RUN_MIXED


def test_mixed_tl_run_is_bitwise_deterministic():
    prob = P.taylor_bar(h=8e-4)
    outs = []
    for _ in range(2):
        ctx = P.context_for(mixed(prob))
        ctx.step(60)
        assert ctx.synchronize()[0] == 0
        outs.append(ctx.download())
        ctx.close()
    for f in ("x", "v", "sigma", "F", "eff_plastic_strain"):
        assert np.array_equal(outs[0][f], outs[1][f]), f


@pytest.mark.parametrize("scheme", ["usl", "musl"])
def test_mixed_run_is_bitwise_deterministic(scheme):
    say("determinism: %s" % scheme)
    prob = _randomize(P.tensile_bar(h=3e-4, v_pull=40.0, scheme=scheme), sscale=1e7)
    outs = []
    for _ in range(2):
        ctx = P.context_for(mixed(prob))
        ctx.step(50)
        assert ctx.synchronize()[0] == 0
        outs.append((ctx.download(), ctx.grid()))
        ctx.close()
    (pa, ga), (pb, gb) = outs
    for f in ("x", "v", "sigma", "F", "strain_el", "eff_plastic_strain", "damage"):
        assert np.array_equal(pa[f], pb[f]), f
    for a, b in zip(ga, gb):
        assert np.array_equal(a, b)


def test_mixed_full_size_tensile_bar_conservation():
    """4.27M particles, mixed: one P2G conserves mass to the fixed-point
    resolution and momentum to FP32 rounding, with no fixes."""
    prob = P.tensile_bar()
    prob.fixes = []
    rng = np.random.default_rng(3)
    prob.particles["v"][:] = rng.normal(scale=5.0, size=prob.particles["v"].shape)
    m = prob.particles["mass"]
    p_tot = (m[:, None] * prob.particles["v"]).sum(axis=0)
    ctx = P.context_for(mixed(prob))
    for ph in ("compute_grid_weight_functions_and_gradients", "particles_to_grid",
               "update_grid_state"):
        ctx.phase(ph)
    mg, vg, vug = ctx.grid()
    assert ctx.synchronize()[0] == 0
    assert abs(mg.sum() - m.sum()) <= 1e-6 * m.sum()
    pg = (mg[:, None] * vg).sum(axis=0)
    assert np.all(np.abs(pg - p_tot) <= 1e-5 * np.abs(m[:, None] * prob.particles["v"]).sum(axis=0))
    ctx.close()

"""The workloads at their full sizes, through properties that do not depend
on the size (the oracle runs the small versions of the same problems in the
other test files):

* dam break, 256^3 cells, 17.2M fluid particles, MUSL cubic, mixed
  precision: one P2G conserves mass and momentum; two runs of 30 steps with
  gravity and the wall fixes agree bit for bit (deterministic P2G);
* Taylor bar, TLMPM cubic, 1.04M particles, FP64: mass and momentum
  conservation of the TL P2G; 100 steps of the impact without an error,
  the kinetic energy stays below its initial value and ends lower, and no
  particle passes the wall; the mixed path's 100 steps are within 1e-3
  relative L2 of the FP64 run (displacement and stress); the reference
  example's 20000 steps and on to 80 us of simulated time;
* 256 batched 2D instances, 15.1M particles, mixed: 20 steps as graph
  replays equal 20 eager steps bit for bit, each instance keeps its own
  time, and no particle leaves its instance.

The 256^3 dam break holds 205 x 205 x 410 particles at 2 x 2 x 2 per cell
(17.2M): the grid is the stated one, the lattice the reference's populate
rule at ppc1d = 2.
"""
import dataclasses
import sys

import numpy as np
import pytest

from karamelo_amd import problems as P
from test_gpu_mixed import mixed

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _p2g_totals(ctx):
    for ph in ("compute_grid_weight_functions_and_gradients", "particles_to_grid",
               "update_grid_state"):
        ctx.phase(ph)
    assert ctx.synchronize()[0] == 0
    mg, vg, _ = ctx.grid()
    return mg.sum(), (mg[:, None] * vg).sum(axis=0)


def test_dam_break_256_cubed_conservation_and_determinism():
    say("dam break 256^3: building the particles")
    prob = mixed(P.dam_break(n_cells=256, jitter=0.1))
    p = prob.particles
    say("  %d particles" % prob.np)
    assert prob.np > 17_000_000
    # one P2G without fixes: mass to the fixed-point resolution, momentum to
    # FP32 rounding
    free = dataclasses.replace(prob, particles=dict(p), node_mask=None, fixes=[])
    rng = np.random.default_rng(5)
    free.particles["v"] = rng.normal(scale=1.0, size=p["v"].shape)
    m = p["mass"]
    ctx = P.context_for(free)
    mt, pt = _p2g_totals(ctx)
    ctx.close()
    assert abs(mt - m.sum()) <= 1e-6 * m.sum()
    ref = (m[:, None] * free.particles["v"]).sum(axis=0)
    scale = np.abs(m[:, None] * free.particles["v"]).sum(axis=0)
    assert np.all(np.abs(pt - ref) <= 1e-5 * scale), (pt, ref)
    say("  conservation ok; two runs of 30 steps")
    outs = []
    for _ in range(2):
        ctx = P.context_for(prob)
        ctx.step(30)
        assert ctx.synchronize()[0] == 0
        q = ctx.download()
        outs.append(({f: q[f] for f in ("x", "v", "sigma")}, ctx.grid()[0], ctx.time()))
        ctx.close()
    (a, ga, ta), (b, gb, tb) = outs
    for f in a:
        assert np.array_equal(a[f], b[f]), f
    assert np.array_equal(ga, gb) and ta == tb
    # the column has started to fall, and stays in the box
    assert a["v"][:, 2].mean() < 0
    assert np.all((a["x"] >= 0) & (a["x"] <= 1))


def test_taylor_bar_full_size_impact():
    say("Taylor bar, 1.04M particles, TLMPM")
    prob = P.taylor_bar()
    assert prob.np > 1_000_000
    m = prob.particles["mass"]
    free = dataclasses.replace(prob, node_mask=None, fixes=[])
    ctx = P.context_for(free)
    mt, pt = _p2g_totals(ctx)
    ctx.close()
    assert abs(mt - m.sum()) <= 1e-10 * m.sum()
    assert abs(pt[0] - (m * prob.particles["v"][:, 0]).sum()) <= 1e-10 * abs(pt[0])
    ctx = P.context_for(prob)
    ek0 = ctx.energy()[0]
    ek = [ek0]
    for _ in range(5):
        ctx.step(20)
        assert ctx.synchronize()[0] == 0
        ek.append(ctx.energy()[0])
    say("  kinetic energy: " + ", ".join("%.4g" % e for e in ek))
    assert all(e <= ek0 * (1 + 1e-6) for e in ek)
    assert ek[-1] < ek0
    q = ctx.download()
    wall = 0.5 * 32.4e-3          # the bar's end face, where the wall nodes are fixed
    assert q["x"][:, 0].max() <= wall + prob.cfg.cellsize
    ctx.close()
    # the same 100 steps on the mixed path: within 1e-3 of the FP64 run
    ctx = P.context_for(mixed(prob))
    ctx.step(100)
    assert ctx.synchronize()[0] == 0
    r = ctx.download()
    ctx.close()
    x0 = prob.particles["x"]
    eu = np.linalg.norm((r["x"] - x0) - (q["x"] - x0)) / np.linalg.norm(q["x"] - x0)
    es = np.linalg.norm(r["sigma"] - q["sigma"]) / np.linalg.norm(q["sigma"])
    say("  mixed vs FP64 after 100 steps: displacement %.2e, stress %.2e" % (eu, es))
    assert eu <= 1e-3 and es <= 1e-3


def test_batched_256_instances_full_size_graph_equals_eager():
    say("256 batched instances: building")
    batch = P.taylor_plates(n_inst=256, precision="mixed")
    say("  %d particles" % batch.problem.np)
    assert batch.problem.np > 15_000_000
    eager = P.context_for_batch(batch)
    graph = P.context_for_batch(batch)
    eager.step(20)
    graph.step_graph(20, 10)
    assert eager.synchronize()[0] == 0 and graph.synchronize()[0] == 0
    pa, pb = eager.download(), graph.download()
    for f in ("x", "v", "sigma", "F", "damage"):
        assert np.array_equal(pa[f], pb[f], equal_nan=True), f
    dta, ta = eager.instance_time()
    dtb, tb = graph.instance_time()
    assert np.array_equal(dta, dtb) and np.array_equal(ta, tb)
    assert len(np.unique(ta)) > 1 and np.all(ta > 0)
    eager.close()
    graph.close()


def test_dam_break_256_cubed_2000_steps():
    """The 256^3 dam break (17.2M particles, mixed, MUSL cubic) for its 2000
    steps: no error, every particle inside the box, the column slumps, and
    the kinetic energy stays below the potential energy released."""
    say("dam break 256^3: 2000 steps")
    prob = mixed(P.dam_break(n_cells=256, jitter=0.1))
    m = prob.particles["mass"]
    z0 = prob.particles["x"][:, 2].copy()
    ctx = P.context_for(prob)
    for k in range(4):
        ctx.step(500)
        st, tag = ctx.synchronize()
        assert st == 0, (st, tag)
        say("  %d steps, t = %.4g s" % (500 * (k + 1), ctx.time()[1]))
    q = ctx.download()
    ek = ctx.energy()[0]
    ctx.close()
    assert np.all((q["x"] >= 0) & (q["x"] <= 1))
    dpe = 9.81 * float((m * (z0 - q["x"][:, 2])).sum())
    say("  kinetic energy %.4g J, potential energy released %.4g J" % (ek, dpe))
    assert dpe > 0 and q["x"][:, 2].mean() < z0.mean()
    assert ek <= 1.05 * dpe


def test_taylor_bar_full_size_80_us():
    """The 1.04M-particle TLMPM Taylor bar over the reference example's
    run(20000) (cylindrical_taylor_bar.mpm:87-88), then on with run_time to
    80 us of simulated time, the extent the workload is quoted at (about
    39500 steps: the CFL dt shrinks with the smallest stretch of F and
    settles near 1.7 ns).  At both points: no error, no particle past the
    wall, the kinetic energy spent in plastic work, the impact end
    mushroomed."""
    say("Taylor bar, 1.04M particles: 20000 steps, then to 80 us")
    prob = P.taylor_bar()
    r0 = np.hypot(prob.particles["x"][:, 1], prob.particles["x"][:, 2]).max()
    ctx = P.context_for(prob)
    ek0 = ctx.energy()[0]
    wall = 0.5 * 32.4e-3

    def check(q, ek, t_min):
        x = q["x"]
        assert x[:, 0].max() <= wall + prob.cfg.cellsize
        assert ek < 0.5 * ek0
        near = x[:, 0] > wall - 2e-3
        r_end = np.hypot(x[near, 1], x[near, 2]).max()
        say("  t = %.4g s: kinetic energy %.4g -> %.4g J, radius at the impact end %.4g mm "
            "(initial %.4g mm)" % (ctx.time()[1], ek0, ek, r_end * 1e3, r0 * 1e3))
        assert ctx.time()[1] >= t_min
        assert r_end > 1.1 * r0
        assert q["eff_plastic_strain"].max() > 0.1

    for k in range(4):
        ctx.step(5000)
        st, tag = ctx.synchronize()
        assert st == 0, (st, tag)
        say("  %d steps, t = %.4g s" % (5000 * (k + 1), ctx.time()[1]))
    check(ctx.download(), ctx.energy()[0], 30e-6)
    t = ctx.time()[1]
    n = ctx.run_time(80e-6 - t)
    st, tag = ctx.synchronize()
    assert st == 0, (st, tag)
    say("  run_time: %d more steps" % n)
    assert 10000 < n < 40000
    check(ctx.download(), ctx.energy()[0], 80e-6 - 1e-8)
    ctx.close()

"""fix velocity_nodes with time-dependent values and its reaction force on
the GPU, against the oracle, through the C ABI.

The values are the input script's expression texts (compiled once by
kmpm_expr_compile; the oracle re-parses them every step as the reference's
Vars do), with the previous value of the "time" -> "time - dt" rewrite; the
reaction force is sum m (value - v_update) / dt over the fix's nodes
(fix_velocity_nodes.cpp:130-225).  Bars: the fast FP64 path within 1e-10
(fields) and 1e-9 (forces, summed in another order); the reference-order
mode bit for bit; the mixed path within 1e-3 of the FP64 run.
"""
import dataclasses
import sys

import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import native as N
from karamelo_amd import problems as P

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "sigma", "F", "strain_el", "eff_plastic_strain")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b, floor=0.0):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    nb = max(np.linalg.norm(b), floor)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _with(prob, **cfg_changes):
    cfg = N.Config.from_buffer_copy(prob.cfg)
    for k, v in cfg_changes.items():
        setattr(cfg, k, v)
    return dataclasses.replace(prob, cfg=cfg)


def _compare(prob, steps, chunk, exact, tol=1e-10, ftol=1e-9):
    ctx = P.context_for(prob)
    o = O.oracle_for(prob)
    G = max(m.G for m in prob.mats)
    worst = {"force": 0.0}
    done = 0
    while done < steps:
        ctx.step(chunk)
        assert o.step(chunk) == 0
        done += chunk
        st, tag = ctx.synchronize()
        assert st == 0, (st, tag)
        pg, po = ctx.download(), o.download()
        for f in FIELDS:
            if exact:
                assert np.array_equal(pg[f], po[f]), (f, done)
            else:
                floor = 1e-6 * G * np.sqrt(prob.np) if f == "sigma" else 0.0
                e = _rel(pg[f], po[f], floor)
                worst[f] = max(worst.get(f, 0.0), e)
                assert e <= tol, (f, done, e)
        for k in range(len(prob.fixes)):
            fg, fo = ctx.fix_force(k), o.fix_force(k)
            assert np.linalg.norm(fo) > 0
            if exact:
                assert np.array_equal(fg, fo), (k, done, fg, fo)
            else:
                e = np.linalg.norm(fg - fo) / np.linalg.norm(fo)
                worst["force"] = max(worst["force"], e)
                assert e <= ftol, (k, done, fg, fo)
    say("  worst: " + ", ".join("%s=%.1e" % kv for kv in worst.items()))
    ctx.close()
    o.close()


def test_expression_pull_fast_path():
    """JC bar pulled by v (1 - exp(-t / 2 us)) at its right end and minus
    that at its left end (FP64, USL cubic)."""
    say("fixes: expression pull, fast FP64 path")
    prob = P.tensile_bar(h=6e-4, v_pull=40.0, pull="v*(1-exp(-time/2e-6))")
    _compare(prob, 150, 25, exact=False)


def test_expression_pull_musl():
    say("fixes: expression pull, MUSL linear")
    prob = P.tensile_bar(h=6e-4, v_pull=40.0, shape="linear", scheme="musl",
                         pull="v*time/1e-6*(time<1e-6)+v*(time>=1e-6)")
    _compare(prob, 120, 40, exact=False)


def test_expression_pull_reference_order_bit_exact():
    """A linear elastic bar pulled by the ramp v time / 1 us (+ - * / only):
    the reference-order mode equals the oracle bit for bit, fields and
    reaction forces (both sum the forces node by node in tag order)."""
    say("fixes: expression pull, reference order, bit for bit")
    lin = P.linear_material(7850.0, 200e9, P.lit(0.3))
    prob = P.tensile_bar(h=6e-4, v_pull=40.0, material=lin, pull="v*time/1e-6")
    _compare(_with(prob, order=N.ORDER_REFERENCE), 100, 25, exact=True)


def test_constant_fix_reaction_force():
    """The constant end velocities of the benchmark bar: reaction forces
    against the oracle's."""
    say("fixes: constant pull, reaction force")
    _compare(P.tensile_bar(h=6e-4, v_pull=40.0), 100, 50, exact=False)


def test_mixed_reaction_force_close_to_fp64():
    say("fixes: mixed path reaction force vs FP64")
    prob = P.tensile_bar(h=6e-4, v_pull=40.0, pull="v*(1-exp(-time/2e-6))")
    out = {}
    for prec in (N.PRECISION_FP64, N.PRECISION_MIXED):
        ctx = P.context_for(_with(prob, precision=prec))
        ctx.step(200)
        assert ctx.synchronize()[0] == 0
        out[prec] = np.array([ctx.fix_force(k) for k in range(2)])
        ctx.close()
    a, b = out[N.PRECISION_MIXED], out[N.PRECISION_FP64]
    e = np.linalg.norm(a - b) / np.linalg.norm(b)
    say("  mixed vs FP64 reaction force: %.2e" % e)
    assert e <= 1e-3


def test_expression_fix_graph_replays_equal_eager_steps():
    """Time-dependent values are evaluated on the device from the device's
    time, so graph replays equal eager steps (mixed: the fields bit for
    bit)."""
    say("fixes: graph replays with expression values")
    prob = _with(P.tensile_bar(h=6e-4, v_pull=40.0, pull="v*(1-exp(-time/2e-6))"),
                 precision=N.PRECISION_MIXED)
    a, b = P.context_for(prob), P.context_for(prob)
    a.step(1)
    b.step_graph(1, 5)
    for rep in range(3):
        a.step(5)
        b.step_graph(5, 5)
        pa, pb = a.download(), b.download()
        for f in ("x", "v", "sigma"):
            assert np.array_equal(pa[f], pb[f]), (rep, f)
        # the force is a sum of FP64 block partials added in arrival order
        fa, fb = a.fix_force(1), b.fix_force(1)
        assert np.linalg.norm(fa - fb) <= 1e-12 * np.linalg.norm(fa)
    a.close()
    b.close()

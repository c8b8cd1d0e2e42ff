"""Batched instances and captured steps on the GPU, through the C ABI.

* Every instance of a batched FP64 context equals the oracle run of that
  instance alone on the same grid: fields within 1e-10 relative L2, its own
  dt and time.
* A captured graph of G steps, replayed with eager work (downloads,
  synchronisations, energy reductions) in between, equals the eager step
  after every replay, bit for bit.
* The mixed-precision batched run stays within 1e-3 of the FP64 oracle per
  instance.
* A particle whose stencil leaves its instance's cells is reported.
"""
import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import native as N
from karamelo_amd import problems as P
from test_gpu_parity import FIELDS, TOL_STEP, _perturbed, field_floors, rel_l2, say

pytestmark = pytest.mark.gpu


def small_batch(**kw):
    args = dict(n_inst=4, np_lo=400, np_hi=3000, h=1e-4, seed=13)
    args.update(kw)
    return P.taylor_plates(**args)


def _instance_fields(p, i):
    keep = p["solid"] == i
    return {k: v[keep] for k, v in p.items()}


def compare_instances(batch, steps, chunk, tol=TOL_STEP, conditioned=False, tile_variant=0):
    """Each instance against its oracle run alone.  With conditioned=True the
    bar per field is max(tol, 10 x e_self), e_self being the largest
    relative difference between the instance's oracle and its oracle runs
    with every position (+-1 ulp) or velocity (+1 ulp) perturbed."""
    ctx = P.context_for_batch(batch)
    ctx.set_tile_variant(tile_variant)
    oracles = [O.oracle_for(batch.instance(i)) for i in range(batch.ninst)]
    others = [[O.oracle_for(_perturbed(batch.instance(i), k)) for k in ("x+", "x-", "v+")]
              if conditioned else [] for i in range(batch.ninst)]
    floors = field_floors(batch.problem)
    worst = {}
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        ctx.step(n)
        done += n
        st, tag = ctx.synchronize()
        assert st == 0, (st, tag)
        dtg, tg = ctx.instance_time()
        pg = ctx.download()
        for i, o in enumerate(oracles):
            assert o.step(n) == 0
            dto, to, so = o.time()
            assert so == done
            assert abs(dtg[i] - dto) <= tol * abs(dto), (i, dtg[i], dto)
            assert abs(tg[i] - to) <= tol * abs(to), (i, tg[i], to)
            po = o.download()
            for q in others[i]:
                assert q.step(n) == 0
            pq = [q.download() for q in others[i]]
            gi = _instance_fields(pg, i)
            assert np.array_equal(gi["ptag"], po["ptag"])
            for f in FIELDS:
                bar = max([tol] + [10 * rel_l2(p2[f], po[f], floors[f]) for p2 in pq])
                e = rel_l2(gi[f], po[f], floors[f])
                worst[f] = max(worst.get(f, 0.0), e)
                assert e <= bar, (i, f, done, e, bar)
    say("  worst rel L2 over instances: " + ", ".join("%s=%.2e" % kv for kv in worst.items()))
    # instances advance with their own dt: the batch holds several
    assert len(set(np.round(ctx.instance_time()[0], 20))) > 1
    ctx.close()
    return oracles


def test_batched_instances_match_the_oracle_alone_usl_cubic():
    say("batched instances vs oracle alone: usl cubic")
    compare_instances(small_batch(), 150, 50)


def test_batched_instances_match_the_oracle_alone_musl_linear():
    """Linear shapes: plastic flow at the wall amplifies 1-ulp differences
    (as in the two-disk collision), so the bar is the oracle's own
    sensitivity when that exceeds 1e-10."""
    say("batched instances vs oracle alone: musl linear, conditioned")
    compare_instances(small_batch(scheme="musl", shape="linear"), 150, 50, conditioned=True)


def _snapshot(ctx):
    p = ctx.download()
    dt, t = ctx.instance_time()
    return p, dt, t


def _assert_same(a, b, exact):
    """Bit for bit (mixed path: fixed-point node sums make every step
    deterministic) or, FP64, within the parity bar: the FP64 P2G adds the
    waves' contributions to a node in LDS in arrival order."""
    pa, dta, ta = a
    pb, dtb, tb = b
    assert np.array_equal(pa["ptag"], pb["ptag"])
    for f in FIELDS:
        if exact:
            assert np.array_equal(pa[f], pb[f], equal_nan=True), f
        else:
            assert rel_l2(pa[f], pb[f]) <= TOL_STEP, f
    if exact:
        assert np.array_equal(dta, dtb) and np.array_equal(ta, tb)
    else:
        assert np.allclose(dta, dtb, rtol=TOL_STEP, atol=0)
        assert np.allclose(ta, tb, rtol=TOL_STEP, atol=0)


@pytest.mark.parametrize("precision,graph_steps", [("fp64", 3), ("mixed", 4)])
def test_graph_replays_equal_eager_steps(precision, graph_steps):
    """Four replays of a captured graph (an odd step count alternates two
    graphs over the two buffer orders), each checked against the eager run,
    with downloads and energy reductions between the replays."""
    say("graph replays: %s, %d steps per graph" % (precision, graph_steps))
    batch = small_batch(precision=precision)
    eager = P.context_for_batch(batch)
    graph = P.context_for_batch(batch)
    exact = precision == "mixed"
    eager.step(1)
    graph.step_graph(1, graph_steps)          # first step (cell keys) runs eagerly
    for rep in range(4):
        eager.step(graph_steps)
        graph.step_graph(graph_steps, graph_steps)
        assert eager.synchronize()[0] == 0 and graph.synchronize()[0] == 0
        _assert_same(_snapshot(eager), _snapshot(graph), exact)
        ea, eb = eager.energy(), graph.energy()
        if exact:
            assert ea == eb or (np.isnan(ea).any() and np.array_equal(ea, eb, equal_nan=True))
    # a remainder that runs eagerly after the replays
    eager.step(2 * graph_steps + 1)
    graph.step_graph(2 * graph_steps + 1, graph_steps)
    _assert_same(_snapshot(eager), _snapshot(graph), exact)
    eager.close()
    graph.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_steps_of_a_plain_context(precision):
    """Captured steps of an unbatched 3D context (context dt and time)."""
    prob = P.tensile_bar(h=6e-4)
    if precision == "mixed":
        prob.cfg.precision = N.PRECISION_MIXED
    eager = P.context_for(prob)
    graph = P.context_for(prob)
    eager.step(10)
    graph.step_graph(10, 3)
    assert eager.synchronize()[0] == 0 and graph.synchronize()[0] == 0
    exact = precision == "mixed"
    ta, tb = eager.time(), graph.time()
    _assert_same((eager.download(), np.array(ta[:2]), np.array(ta[2:])),
                 (graph.download(), np.array(tb[:2]), np.array(tb[2:])), exact)
    eager.close()
    graph.close()


def test_mixed_batch_within_1e3_of_the_fp64_oracle():
    say("mixed batch vs FP64 oracle")
    batch = small_batch(precision="mixed")
    ctx = P.context_for_batch(batch)
    ctx.step(120)
    assert ctx.synchronize()[0] == 0
    pg = ctx.download()
    dtg, tg = ctx.instance_time()
    floors = field_floors(batch.problem)
    for i in range(batch.ninst):
        inst = batch.instance(i)
        inst.cfg = N.Config.from_buffer_copy(inst.cfg)
        inst.cfg.precision = N.PRECISION_FP64
        o = O.oracle_for(inst)
        assert o.step(120) == 0
        po = o.download()
        gi = _instance_fields(pg, i)
        for f in ("x", "v", "sigma"):
            e = rel_l2(gi[f] - (gi["x0"] if f == "x" else 0), po[f] - (po["x0"] if f == "x" else 0),
                       floors[f])
            assert e <= 1e-3, (i, f, e)
        assert abs(tg[i] - o.time()[1]) <= 1e-5 * o.time()[1]
    ctx.close()


def test_particle_outside_its_instance_is_reported():
    batch = small_batch()
    p = batch.problem.particles
    # move instance 0's bar against the upper edge of its cells
    i0 = p["solid"] == 0
    top = batch.cell_hi[0, 1] * batch.problem.cfg.cellsize
    p["x"][i0, 1] += top - p["x"][i0, 1].max() - 0.1 * batch.problem.cfg.cellsize
    p["x0"][i0] = p["x"][i0]
    ctx = P.context_for_batch(batch)
    ctx.step(1)
    st, tag = ctx.synchronize()
    assert st == N.ERR_INSTANCE
    assert tag in set(p["ptag"][i0].tolist())
    ctx.close()


def test_set_instances_rejects_bad_ranges():
    batch = small_batch()
    ctx = P.context_for(batch.problem)
    T = ctx.tile_shape()
    lo, hi = batch.cell_lo.copy(), batch.cell_hi.copy()
    bad = hi.copy()
    bad[0, 0] += T[0] // 2                       # not tile-aligned
    with pytest.raises(N.KmpmError):
        ctx.set_instances(lo, bad)
    over = lo.copy()
    over[1] = lo[0]                              # overlaps instance 0
    over_hi = hi.copy()
    over_hi[1] = hi[0]
    with pytest.raises(N.KmpmError):
        ctx.set_instances(over, over_hi)
    z = hi.copy()
    z[0, 2] = 2                                  # 2D: z range must be [0, 1)
    with pytest.raises(N.KmpmError):
        ctx.set_instances(lo, z)
    ctx.set_instances(lo, hi)
    ctx.set_instances(np.zeros((0, 3)), np.zeros((0, 3)))   # back to one simulation
    ctx.step(2)
    assert ctx.synchronize()[0] == 0
    ctx.close()

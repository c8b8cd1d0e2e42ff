"""GPU parity of the HIP step against the oracle, through the C ABI.

Bar: particle-cell indices and the sort permutation bit-exact; shape
function weights and gradients bit-exact; grid and particle fields within
1e-12 (one P2G) / 1e-10 relative L2 (after N steps); energies within 1e-10.
"""
import dataclasses
import sys

import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import native as N
from karamelo_amd import problems as P

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-10      # relative L2 of particle fields after N steps
TOL_GRID = 1e-12      # relative max error of grid fields after one P2G


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b, floor=0.0):
    """Relative L2 error over the finite entries, relative to
    max(||b||, floor); non-finite entries (the reference's 0/0 strain of
    fully damaged particles) must coincide."""
    a = np.asarray(a, float).ravel()
    b = np.asarray(b, float).ravel()
    fa, fb = np.isfinite(a), np.isfinite(b)
    if not np.array_equal(fa, fb):
        return np.inf
    if not np.array_equal(np.isnan(a), np.isnan(b)) or not np.array_equal(a[~fa & ~np.isnan(a)],
                                                                      b[~fb & ~np.isnan(b)]):
        return np.inf
    a, b = a[fa], b[fb]
    nb = max(np.linalg.norm(b), floor)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def field_floors(prob):
    """Noise floors of the relative errors: a stress of 1e-6 G and a strain
    of 1e-6 per particle (below them the fields are rounding noise, e.g. the
    stress of a rigidly translating body)."""
    n = np.sqrt(prob.np)
    G = max(m.G for m in prob.mats)
    return {"sigma": 1e-6 * G * n, "strain_el": 1e-6 * n, "eff_plastic_strain": 1e-6 * n,
            "damage_init": 1e-6 * n, "damage": 1e-6 * n, "F": 0.0, "x": 0.0, "v": 0.0}


def _randomize(prob, seed=7, vscale=1.0, sscale=1e6):
    """Random velocities and symmetric stresses so every P2G term is exercised."""
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
    return {
        "two_disks": lambda: P.two_disks(),
        "two_disks_cubic": lambda: _cubic_disks(),
        "tensile_cubic": lambda: P.tensile_bar(h=6e-4),
        "tensile_linear": lambda: P.tensile_bar(h=6e-4, shape="linear"),
        "dam_cubic": lambda: P.dam_break(n_cells=16),
    }


def _cubic_disks():
    prob = P.two_disks()
    prob.cfg.shape = N.SHAPE_CUBIC_SPLINE
    return prob


@pytest.mark.parametrize("name", list(small_problems()))
def test_weights_cells_and_sort_bit_exact(name):
    say("weights/cells/sort: %s" % name)
    prob = small_problems()[name]()
    ctx = P.context_for(prob)
    o = O.oracle_for(prob)
    ctx.phase("compute_grid_weight_functions_and_gradients")
    o.phase("weights")
    ijk_g, sorted_g = ctx.cells()
    assert np.array_equal(ijk_g, o.cells())
    assert np.array_equal(sorted_g, o.sort_permutation(ctx.tile_shape()))
    tag_g, w_g, wd_g = ctx.debug_weights()
    nn, tag_o, w_o, wd_o = o.neighbours()
    S = tag_g.shape[1]
    for p in range(prob.np):
        keep = tag_g[p] >= 0
        tg, wg, dg = tag_g[p][keep], w_g[p][keep], wd_g[p][keep]
        og = np.argsort(tg)
        k = nn[p]
        to, wo, do = tag_o[p, :k], w_o[p, :k], wd_o[p, :k]
        oo = np.argsort(to)
        assert np.array_equal(tg[og], to[oo]), p
        assert np.array_equal(wg[og], wo[oo]), p
        assert np.array_equal(dg[og], do[oo]), p
    ctx.close()


@pytest.mark.parametrize("name", list(small_problems()))
def test_p2g_and_grid_update(name):
    say("p2g/grid: %s" % name)
    prob = _randomize(small_problems()[name]())
    ctx = P.context_for(prob)
    o = O.oracle_for(prob)
    for ph in ("compute_grid_weight_functions_and_gradients", "particles_to_grid",
               "update_grid_state"):
        ctx.phase(ph)
    for ph in ("weights", "p2g", "grid_update"):
        o.phase(ph)
    mg, vg, vug = ctx.grid()
    mo, vo, vuo, fo = o.grid()
    assert np.abs(mg - mo).max() <= TOL_GRID * np.abs(mo).max()
    assert np.abs(vg - vo).max() <= TOL_GRID * np.abs(vo).max()
    assert np.abs(vug - vuo).max() <= TOL_GRID * np.abs(vuo).max()
    st, _ = ctx.synchronize()
    assert st == 0
    ctx.close()


FIELDS = ("x", "v", "sigma", "F", "strain_el", "eff_plastic_strain", "damage_init", "damage")


def _perturbed(prob, kind):
    """Same problem with every active position (kind 'x+', 'x-') or velocity
    ('v+') component moved by one ulp."""
    q = dataclasses.replace(prob, particles={k: v.copy() for k, v in prob.particles.items()})
    d = q.cfg.dimension
    f, direction = ("x" if kind[0] == "x" else "v"), (np.inf if kind[1] == "+" else -np.inf)
    q.particles[f][:, :d] = np.nextafter(q.particles[f][:, :d], direction)
    return q


def run_and_compare(prob, steps, chunk, tol=TOL_STEP, check_energy=True, conditioned=False,
                    tile_variant=0, fields=FIELDS):
    """GPU vs oracle every `chunk` steps.  With conditioned=True the bar per
    field is max(tol, 10 x e_self), e_self being the largest relative
    difference between the oracle and three oracle runs whose input
    positions (+-1 ulp) or velocities (+1 ulp) are perturbed: the GPU must be
    as close to the oracle as the oracle is to itself."""
    floors = field_floors(prob)
    ctx = P.context_for(prob)
    ctx.set_tile_variant(tile_variant)
    o = O.oracle_for(prob)
    others = [O.oracle_for(_perturbed(prob, k)) for k in ("x+", "x-", "v+")] if conditioned else []
    worst = {}
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        ctx.step(n)
        assert o.step(n) == 0
        for q in others:
            assert q.step(n) == 0
        done += n
        st, tag = ctx.synchronize()
        assert st == 0, (st, tag)
        dtg, tg, sg = ctx.time()
        dto, to, so = o.time()
        assert sg == so == done
        pg, po = ctx.download(), o.download()
        pq = [q.download() for q in others]
        bar = {f: max([tol] + [10 * rel_l2(p2[f], po[f], floors.get(f, 0.0)) for p2 in pq])
               for f in fields}
        tbar = max([tol] + [10 * abs(q.time()[0] - dto) / abs(dto) for q in others])
        assert abs(dtg - dto) <= tbar * abs(dto)
        assert abs(tg - to) <= tbar * abs(to)
        if check_energy:
            eg, eo = ctx.energy(), o.energy()
            if all(np.isfinite(eo)):
                scale = abs(eo[0]) + abs(eo[1])
                ebar = tol * scale
                for q in others:
                    e2 = q.energy()
                    ebar = max(ebar, 10 * max(abs(e2[0] - eo[0]), abs(e2[1] - eo[1])))
                for a, b in zip(eg, eo):
                    assert abs(a - b) <= ebar, (done, eg, eo, ebar)
            else:   # NaN strain of fully damaged particles (0/0, solid.cpp:1358-1366)
                assert [np.isfinite(v) for v in eg] == [np.isfinite(v) for v in eo]
        assert np.array_equal(pg["ptag"], po["ptag"])
        for f in fields:
            e = rel_l2(pg[f], po[f], floors.get(f, 0.0))
            worst[f] = max(worst.get(f, 0.0), e)
            assert e <= bar[f], (f, done, e, bar[f])
    say("  worst rel L2: " + ", ".join("%s=%.2e" % kv for kv in worst.items()))
    ctx.close()
    return worst, po


def test_dense_tiles_sort_and_steps():
    """125 particles per cell: the bar's full tiles hold more particles than
    the per-tile sort orders in local memory (1024), its edge tiles fewer, so
    both orderings run.  Sort permutation bit-exact, then 10 steps vs the
    oracle."""
    say("dense tiles: sort and 10 steps")
    prob = P.tensile_bar(h=6e-4, ppc1d=5)
    ctx = P.context_for(prob)
    o = O.oracle_for(prob)
    ctx.phase("compute_grid_weight_functions_and_gradients")
    o.phase("weights")
    ijk_g, sorted_g = ctx.cells()
    assert np.array_equal(ijk_g, o.cells())
    T = ctx.tile_shape()
    assert np.array_equal(sorted_g, o.sort_permutation(T))
    n = N.grid_size(prob.cfg)
    lo = np.array(prob.cfg.boxlo)
    c = np.minimum(((prob.particles["x"] - lo) / prob.cfg.cellsize).astype(np.int64),
                   np.array(n) - 1)
    nt = [(n[d] + T[d] - 1) // T[d] for d in range(3)]
    tile = ((c[:, 0] // T[0]) * nt[1] + c[:, 1] // T[1]) * nt[2] + c[:, 2] // T[2]
    per_tile = np.bincount(tile)
    per_tile = per_tile[per_tile > 0]
    assert per_tile.max() > 1024 and per_tile.min() <= 1024, (per_tile.min(), per_tile.max())
    ctx.close()
    run_and_compare(prob, 10, 5)


def _sort_keys(prob, T, x):
    """Tile-major cell keys of positions x (cell index clamped to the grid)."""
    n = N.grid_size(prob.cfg)
    lo = np.array(prob.cfg.boxlo)
    u = (x - lo) * (1.0 / prob.cfg.cellsize)
    c = np.where(u >= 0.0, np.minimum(u, np.array(n) - 1.0), 0.0).astype(np.int64)
    c[:, 2] = np.where(prob.cfg.dimension == 3, c[:, 2], 0)
    nt = [(n[d] + T[d] - 1) // T[d] for d in range(3)]
    tile = ((c[:, 0] // T[0]) * nt[1] + c[:, 1] // T[1]) * nt[2] + c[:, 2] // T[2]
    loc = ((c[:, 0] % T[0]) * T[1] + c[:, 1] % T[1]) * T[2] + c[:, 2] % T[2]
    return tile * (T[0] * T[1] * T[2]) + loc


def test_sort_is_stable_across_steps():
    """Ties in a cell keep the previous storage order, which is the previous
    sort's order: each step's permutation is the stable sort, by cell key, of
    the previous one (the first is by key, then ascending tag)."""
    say("stable sort across 60 steps")
    prob = P.two_disks()
    ctx = P.context_for(prob)
    T = ctx.tile_shape()
    prev = prob.particles["ptag"].copy()
    moved = 0
    for step in range(60):
        ctx.phase("compute_grid_weight_functions_and_gradients")
        _, got = ctx.cells()
        p = ctx.download()
        keys = _sort_keys(prob, T, p["x"])[np.searchsorted(p["ptag"], prev)]
        want = prev[np.argsort(keys, kind="stable")]
        assert np.array_equal(got, want), step
        moved += int((got != prev).sum())
        prev = got
        for ph in ("particles_to_grid", "update_grid_state", "grid_to_points_and_stress",
                   "adjust_dt"):
            ctx.phase(ph)
        assert ctx.synchronize()[0] == 0
    say("  positions changed over the steps: %d" % moved)
    assert moved > 0
    ctx.close()


def _run_time_rule(t0, duration, times):
    """Steps the reference's run_time(duration) takes, given the time after
    each step: the loop runs while time < stof(to_string(t0 + duration))
    and breaks after a step whose time exceeds t0 + duration."""
    t_end = duration + t0
    cond = float(np.float32("%f" % t_end))
    t, n = t0, 0
    while t < cond:
        t = times(n)
        n += 1
        if t > t_end:
            break
    return n


def test_run_time_follows_the_reference_loop():
    """run_time(duration) against the reference's loop rule: with set_dt
    (two disks) the step count follows from the time sums alone; with the
    CFL dt (tensile bar) from the same rule applied step by step on a
    second context.  The states then match the oracle at that count."""
    say("run_time")
    prob = P.two_disks()
    dt = prob.cfg.dt
    acc = [0.0]

    def sums(k):
        acc[0] += dt
        return acc[0]
    duration = float(np.float32(0.05))            # the script literal 0.05
    want = _run_time_rule(0.0, duration, sums)
    ctx = P.context_for(prob)
    got = ctx.run_time(duration)
    assert ctx.synchronize()[0] == 0
    assert got == want, (got, want)
    _, t, steps = ctx.time()
    assert steps == got and abs(t - acc[0]) <= 1e-15
    o = O.oracle_for(prob)
    assert o.step(got) == 0
    pg, po = ctx.download(), o.download()
    floors = field_floors(prob)
    for f in FIELDS:
        assert rel_l2(pg[f], po[f], floors[f]) <= TOL_STEP, f
    ctx.close()
    # CFL dt, after a few steps (run_time starts from the current time)
    prob = P.tensile_bar(h=6e-4)
    a, b = P.context_for(prob), P.context_for(prob)
    a.step(5)
    b.step(5)
    _, t0, _ = b.time()
    duration = 40 * b.time()[0]

    def by_hand(k):
        b.step(1)
        return b.time()[1]
    want = _run_time_rule(t0, duration, by_hand)
    got = a.run_time(duration)
    assert got == want and got > 30, (got, want)
    (_, ta, sa), (_, tb, sb) = a.time(), b.time()
    assert sa == sb and abs(ta - tb) <= 1e-10 * tb
    pa, pb = a.download(), b.download()
    floors = field_floors(prob)
    for f in FIELDS:
        assert rel_l2(pa[f], pb[f], floors[f]) <= TOL_STEP, f
    assert a.run_time(duration, max_steps=3) == 3
    a.close()
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_empty_and_single_particle_contexts(precision):
    """No particles: steps run, the time advances and downloads are empty.
    One particle: 20 steps against the oracle (FP64: every field; mixed:
    position and velocity within 1e-3).  An isolated particle is the mixed
    path's worst case for stress: every node of its stencil has a small
    mass, and the fixed-point momentum quantum (set by the material's wave
    speed) is then a large relative error of the node velocity, which the
    velocity gradient turns into stress noise."""
    say("empty and single-particle contexts: %s" % precision)
    base = P.tensile_bar(h=6e-4)
    if precision == "mixed":
        cfg = N.Config.from_buffer_copy(base.cfg)
        cfg.precision = N.PRECISION_MIXED
        base = dataclasses.replace(base, cfg=cfg)
    empty = dataclasses.replace(base, particles={k: v[:0].copy() for k, v in base.particles.items()})
    ctx = P.context_for(empty)
    ctx.step(5)
    assert ctx.synchronize()[0] == 0
    dt, t, steps = ctx.time()
    assert steps == 5 and t > 0
    assert len(ctx.download()["x"]) == 0
    assert ctx.energy() == (0.0, 0.0)
    ctx.close()
    mid = len(base.particles["x"]) // 2
    one = dataclasses.replace(base, particles={k: v[mid:mid + 1].copy()
                                               for k, v in base.particles.items()})
    one.particles["v"][:] = (3.0, -2.0, 1.0)
    ctx = P.context_for(one)
    ref = dataclasses.replace(one, cfg=N.Config.from_buffer_copy(one.cfg))
    ref.cfg.precision = N.PRECISION_FP64
    o = O.oracle_for(ref)
    ctx.step(20)
    assert o.step(20) == 0
    assert ctx.synchronize()[0] == 0
    pg, po = ctx.download(), o.download()
    tol = TOL_STEP if precision == "fp64" else 1e-3
    floors = field_floors(one)
    for f in ("x", "v", "sigma", "F") if precision == "fp64" else ("x", "v"):
        assert rel_l2(pg[f], po[f], floors[f]) <= tol, f
    ctx.close()


def test_two_disks_3000_steps_conditioned():
    """The full collision.  Linear shapes with lattice-aligned uniform motion
    amplify 1-ulp differences (derivative_linear(0) == 0 at node crossings),
    so the bar is set by the oracle's own 1-ulp sensitivity."""
    say("two disks, 3000 steps, conditioned")
    run_and_compare(P.two_disks(), 3000, 250, conditioned=True)


def test_two_disks_cubic_1000_steps():
    say("two disks cubic, 1000 steps")
    run_and_compare(_cubic_disks(), 1000, 250)


def test_tensile_bar_usl_cubic_plastic():
    say("tensile bar USL cubic, 300 steps")
    prob = P.tensile_bar(h=6e-4, v_pull=40.0)
    w, po = run_and_compare(prob, 300, 100)
    assert po["eff_plastic_strain"].max() > 0, "the bar must yield in this run"


def test_tensile_bar_musl_linear():
    say("tensile bar MUSL linear, 200 steps")
    run_and_compare(P.tensile_bar(h=6e-4, shape="linear", scheme="musl", v_pull=40.0), 200, 100)


def test_tensile_bar_damage_onset():
    """Low failure strain so damage initiates and grows within the run."""
    say("tensile bar with damage onset")
    prob = P.tensile_bar(h=6e-4, v_pull=80.0, d_params=(0.002, 0.0, 0.0, 0.0, 0.0))
    w, po = run_and_compare(prob, 300, 50, conditioned=True)
    assert po["damage"].max() > 0, "damage must initiate in this configuration"


def test_dam_break_musl_cubic():
    say("dam break MUSL cubic, 150 steps")
    run_and_compare(P.dam_break(n_cells=16), 150, 50)


def test_dam_break_usl_linear_jittered():
    say("dam break USL linear jittered, 150 steps")
    run_and_compare(P.dam_break(n_cells=16, scheme="usl", shape="linear", jitter=0.1), 150, 50)


def test_particle_leaving_box_is_reported():
    say("error: particle left the box")
    prob = P.two_disks()
    prob.particles["v"][5, 0] = -100.0
    ctx = P.context_for(prob)
    o = O.oracle_for(prob)
    ctx.step(20)
    st, tag = ctx.synchronize()
    assert o.step(20) == N.ERR_PARTICLE_LEFT_BOX
    assert st == N.ERR_PARTICLE_LEFT_BOX
    assert tag == O.load().orc_error_particle(o.h) == prob.particles["ptag"][5]
    ctx.close()


def test_negative_jacobian_is_reported():
    say("error: J <= 0")
    prob = P.two_disks()
    prob.particles["F"][7] = [1, 0, 0, 0, 1, 0, 0, 0, -1]
    ctx = P.context_for(prob)
    ctx.step(1)
    st, tag = ctx.synchronize()
    assert st == N.ERR_JACOBIAN and tag == prob.particles["ptag"][7]
    ctx.close()


@pytest.mark.parametrize("graph", [False, True])
def test_steps_after_an_error_change_nothing(graph):
    """The reference aborts on a particle leaving the box (solid.cpp:617-627):
    after the step that raised the error, further steps (eager or graph
    replays) leave the state, time and step count as they are."""
    say("error freezes the context (graph=%s)" % graph)
    prob = P.two_disks()
    prob.particles["v"][5, 0] = -100.0
    ctx = P.context_for(prob)
    ctx.step(20)
    st, tag = ctx.synchronize()
    assert st == N.ERR_PARTICLE_LEFT_BOX
    before, t_before = ctx.download(), ctx.time()
    if graph:
        ctx.step_graph(12, 4)
    else:
        ctx.step(7)
    st2, tag2 = ctx.synchronize()
    assert (st2, tag2) == (st, tag)
    after, t_after = ctx.download(), ctx.time()
    assert t_after == t_before
    for f in ("x", "v", "sigma", "F", "strain_el", "ptag"):
        assert np.array_equal(after[f], before[f]), f
    ctx.close()


def test_rejected_upload_keeps_the_previous_particles():
    """An upload rejected by its checks (here a particle whose F has no
    polar decomposition on a TLMPM context) changes nothing: the context
    keeps its particles and steps on as if the call had not been made."""
    say("rejected upload keeps the previous particles")
    prob = P.taylor_bar(h=8e-4)
    a = P.context_for(prob)
    b = P.context_for(prob)
    a.step(3)
    b.step(3)
    bad = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in prob.particles.items()}
    bad["F"][4] = [1, 0, 0, 0, 1, 0, 0, 0, -1]
    with pytest.raises(N.KmpmError) as e:
        a.upload(bad)
    assert e.value.code == N.ERR_JACOBIAN
    a.step(5)
    b.step(5)
    assert a.synchronize()[0] == 0 and b.synchronize()[0] == 0
    pa, pb = a.download(), b.download()
    # the FP64 P2G adds node sums through LDS atomics, whose order follows
    # wave timing, so two identical runs agree to round-off, not bit for
    # bit; a partly applied upload (F = diag(1, 1, -1) at one particle)
    # would differ at O(1)
    for f in ("x", "v", "sigma", "F"):
        assert rel_l2(pa[f], pb[f]) <= 1e-12, f
    assert np.array_equal(pa["ptag"], pb["ptag"])
    a.close()
    b.close()


def test_full_size_tensile_bar_properties():
    """4.27M particles: cell indices of the reference rule, a complete sort
    permutation ordered by (tile, cell, ptag), and mass and momentum
    conservation of one P2G."""
    say("full-size tensile bar properties")
    prob = _randomize(P.tensile_bar(), vscale=1.0)
    p = prob.particles
    ctx = P.context_for(prob)
    ctx.phase("compute_grid_weight_functions_and_gradients")
    ijk, sp = ctx.cells()
    lo = np.array(prob.cfg.boxlo)
    inv_h = 1.0 / prob.cfg.cellsize
    ref = ((p["x"] - lo) * inv_h - 1).astype(np.int64)   # ulmpm.cpp:206-208 (truncation)
    assert np.array_equal(ijk, ref)
    assert np.array_equal(np.sort(sp), np.sort(p["ptag"]))
    T = ctx.tile_shape()
    n = N.grid_size(prob.cfg)
    c = np.minimum(((p["x"] - lo) * inv_h).astype(np.int64), np.array(n) - 1)
    nt = [(n[d] + T[d] - 1) // T[d] for d in range(3)]
    tile = ((c[:, 0] // T[0]) * nt[1] + c[:, 1] // T[1]) * nt[2] + c[:, 2] // T[2]
    loc = ((c[:, 0] % T[0]) * T[1] + c[:, 1] % T[1]) * T[2] + c[:, 2] % T[2]
    key = tile * (T[0] * T[1] * T[2]) + loc
    order = np.lexsort((p["ptag"], key))
    assert np.array_equal(sp, p["ptag"][order])
    assert ctx.synchronize()[0] == 0
    ctx.close()
    # conservation of one P2G, without the velocity fixes that overwrite nodes
    free = dataclasses.replace(prob, fixes=[], node_mask=None)
    ctx = P.context_for(free)
    for ph in ("compute_grid_weight_functions_and_gradients", "particles_to_grid",
               "update_grid_state"):
        ctx.phase(ph)
    m, v, vu = ctx.grid()
    assert abs(m.sum() - p["mass"].sum()) <= 1e-12 * p["mass"].sum()
    mom = (m[:, None] * v).sum(axis=0)
    mp = (p["mass"][:, None] * p["v"]).sum(axis=0)
    assert np.abs(mom - mp).max() <= 1e-12 * np.abs(p["mass"][:, None] * p["v"]).sum()
    assert ctx.synchronize()[0] == 0
    ctx.close()

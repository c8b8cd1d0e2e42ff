"""TLMPM on the GPU against the oracle, through the C ABI (FP64).

Bars as for ULMPM: particle-cell indices of x0 on the solid's grid, the
sort permutation and every weight/gradient bit-exact; one P2G + grid update
within 1e-12; whole Taylor-bar runs within 1e-10 relative L2 (conditioned
by the oracle's own 1-ulp sensitivity where the run is chaotic).
"""
import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import native as N
from karamelo_amd import problems as P
from test_gpu_parity import TOL_GRID, _randomize, run_and_compare, say

pytestmark = pytest.mark.gpu


def tl_problems():
    return {
        "taylor_cubic": lambda: P.taylor_bar(h=8e-4),
        "taylor_linear": lambda: P.taylor_bar(h=8e-4, shape="linear"),
        "taylor_aligned_cubic": lambda: P.taylor_bar(h=8e-4, length=0.032),
    }


@pytest.mark.parametrize("name", list(tl_problems()))
def test_tl_weights_cells_and_sort_bit_exact(name):
    say("TL weights/cells/sort: %s" % name)
    prob = tl_problems()[name]()
    ctx = P.context_for(prob)
    o = O.oracle_for(prob)
    ctx.phase("compute_grid_weight_functions_and_gradients")
    o.phase("weights")
    ijk_g, sorted_g = ctx.cells()
    assert np.array_equal(ijk_g, o.cells())
    assert np.array_equal(sorted_g, o.sort_permutation(ctx.tile_shape()))
    tag_g, w_g, wd_g = ctx.debug_weights()
    nn, tag_o, w_o, wd_o = o.neighbours()
    for p in range(prob.np):
        keep = tag_g[p] >= 0
        tg, wg, dg = tag_g[p][keep], w_g[p][keep], wd_g[p][keep]
        og = np.argsort(tg)
        k = nn[p]
        oo = np.argsort(tag_o[p, :k])
        assert np.array_equal(tg[og], tag_o[p, :k][oo]), p
        assert np.array_equal(wg[og], w_o[p, :k][oo]), p
        assert np.array_equal(dg[og], wd_o[p, :k][oo]), p
    ctx.close()


@pytest.mark.parametrize("name", list(tl_problems()))
def test_tl_p2g_and_grid_update(name):
    """Random velocities and stresses: the force is -vol0 PK1 . grad w0 with
    PK1 = vol0 sigma at F = I."""
    say("TL p2g/grid: %s" % name)
    prob = _randomize(tl_problems()[name]())
    ctx = P.context_for(prob)
    o = O.oracle_for(prob)
    for ph in ("compute_grid_weight_functions_and_gradients", "particles_to_grid",
               "update_grid_state"):
        ctx.phase(ph)
    for ph in ("weights", "p2g", "grid_update"):
        o.phase(ph)
    mg, vg, vug = ctx.grid()
    mo, vo, vuo, fo = o.grid()
    assert ctx.synchronize()[0] == 0
    assert np.abs(mg - mo).max() <= TOL_GRID * np.abs(mo).max()
    assert np.abs(vg - vo).max() <= TOL_GRID * np.abs(vo).max()
    assert np.abs(vug - vuo).max() <= TOL_GRID * np.abs(vuo).max()
    ctx.close()


def test_tl_taylor_bar_usl_cubic():
    """The impact: plastic flow at the wall, F far from I, dt set by the
    smallest |Re eig(F)|."""
    run_and_compare(P.taylor_bar(h=8e-4), 240, 80, conditioned=True)


def test_tl_taylor_bar_musl_linear():
    run_and_compare(P.taylor_bar(h=8e-4, shape="linear", scheme="musl"), 160, 80,
                    conditioned=True)


def test_tl_taylor_bar_aligned_usl_cubic():
    run_and_compare(P.taylor_bar(h=8e-4, length=0.032), 200, 100, conditioned=True)


def test_tl_sort_is_kept_after_the_first_step():
    """The TL state stays in its first sorted order: later steps reuse it."""
    prob = P.taylor_bar(h=8e-4)
    ctx = P.context_for(prob)
    ctx.step(3)
    assert ctx.synchronize()[0] == 0
    ctx.phase("compute_grid_weight_functions_and_gradients")
    _, sorted_g = ctx.cells()
    o = O.oracle_for(prob)
    assert np.array_equal(sorted_g, o.sort_permutation(ctx.tile_shape()))
    ctx.close()


def test_tl_rejects_several_solids():
    prob = P.taylor_bar(h=8e-4)
    ctx = N.Context(prob.cfg, prob.mats)
    two = {k: v.copy() for k, v in prob.particles.items()}
    two["solid"][len(two["solid"]) // 2:] = 1
    with pytest.raises(N.KmpmError):
        ctx.upload(two)
    ctx.close()

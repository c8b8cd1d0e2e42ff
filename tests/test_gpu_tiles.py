"""Tile variants and the autotuner on the GPU, through the C ABI.

* Every tile variant gives the oracle's cell indices, the oracle's sort
  permutation for that tile shape and the oracle's weights bit for bit, and
  the oracle's fields after N steps within 1e-10 relative L2 (FP64), for
  each dimension and shape function, ULMPM USL/MUSL and TLMPM.
* Batched instances keep matching the oracle run alone with the widest
  2D tiles.
* kmpm_autotune leaves the state, dt, time and step count as they were and
  keeps the variant it reports: the run that follows equals, bit for bit
  (mixed path), a run that was set to that variant by hand.
"""
import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import native as N
from karamelo_amd import problems as P
from test_gpu_batch import compare_instances, small_batch
from test_gpu_parity import FIELDS, _cubic_disks, run_and_compare, say

pytestmark = pytest.mark.gpu


def problems():
    return {
        "disks_linear_2d": (lambda: P.two_disks(), 300, True),
        "disks_cubic_2d": (_cubic_disks, 300, True),
        "tensile_cubic_3d": (lambda: P.tensile_bar(h=6e-4), 100, False),
        # linear shapes with lattice-aligned motion: node crossings amplify
        # 1-ulp differences, so the bar is the oracle's own sensitivity
        "tensile_linear_3d": (lambda: P.tensile_bar(h=6e-4, shape="linear"), 100, True),
        "dam_musl_cubic_3d": (lambda: P.dam_break(n_cells=16), 100, False),
        "taylor_tl_cubic_3d": (lambda: P.taylor_bar(h=8e-4), 120, True),
    }


CASES = [(name, v) for name in problems() for v in (1, 2)]


@pytest.mark.parametrize("name,variant", CASES)
def test_variant_cells_sort_and_weights_bit_exact(name, variant):
    say("tile variant %d: cells/sort/weights of %s" % (variant, name))
    prob = problems()[name][0]()
    ctx = P.context_for(prob)
    ctx.set_tile_variant(variant)
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


@pytest.mark.parametrize("name,variant", CASES)
def test_variant_steps_match_the_oracle(name, variant):
    make, steps, conditioned = problems()[name]
    say("tile variant %d: %d steps of %s" % (variant, steps, name))
    run_and_compare(make(), steps, steps // 2, conditioned=conditioned, tile_variant=variant)


def test_tile_shapes_of_the_variants():
    shapes = {}
    for name in ("disks_cubic_2d", "tensile_cubic_3d"):
        ctx = P.context_for(problems()[name][0]())
        shapes[name] = []
        for v in range(N.NTILE_VARIANTS):
            ctx.set_tile_variant(v)
            shapes[name].append(ctx.tile_shape())
        ctx.close()
    assert shapes["disks_cubic_2d"] == [(8, 8, 1), (8, 16, 1), (16, 16, 1)]
    assert shapes["tensile_cubic_3d"] == [(4, 4, 4), (4, 4, 8), (4, 8, 8)]


def test_batched_instances_with_the_widest_2d_tiles():
    say("batched instances, 16x16 tiles")
    compare_instances(small_batch(), 100, 50, tile_variant=2)


def _state(ctx):
    return ctx.download(), ctx.time()


def test_autotune_keeps_the_state_and_the_chosen_variant():
    prob = P.tensile_bar(h=3e-4, v_pull=40.0)
    prob.cfg.precision = N.PRECISION_MIXED
    ctx = P.context_for(prob)
    ctx.step(5)
    before = _state(ctx)
    variant, ms = ctx.autotune(4)
    say("  autotune: variant %d, ms per step %s" % (variant, np.round(ms, 4)))
    assert 0 <= variant < N.NTILE_VARIANTS and np.all(ms > 0)
    assert ms[variant] == ms.min()
    after = _state(ctx)
    for f in FIELDS + ("ptag",):
        assert np.array_equal(before[0][f], after[0][f], equal_nan=True), f
    assert before[1] == after[1]
    ref = P.context_for(prob)
    ref.step(5)
    ref.set_tile_variant(variant)
    assert ref.tile_shape() == ctx.tile_shape()
    ctx.step(20)
    ref.step(20)
    assert ctx.synchronize()[0] == 0 and ref.synchronize()[0] == 0
    a, b = _state(ctx), _state(ref)
    for f in FIELDS:
        assert np.array_equal(a[0][f], b[0][f], equal_nan=True), f
    assert a[1] == b[1]
    ctx.close()
    ref.close()


def test_autotune_skips_variants_the_instances_do_not_align_with():
    """Instance ranges packed to multiples of 8 cells fit the 8x8 tiles only."""
    batch = small_batch(align=8)
    assert (batch.cell_lo[:, 1] % 16).any() or (batch.cell_hi[:, 1] % 16).any()
    ctx = P.context_for_batch(batch)
    with pytest.raises(N.KmpmError):
        ctx.set_tile_variant(2)
    assert ctx.tile_shape() == (8, 8, 1)
    variant, ms = ctx.autotune(2)
    assert ms[1] == -1 and ms[2] == -1 and variant == 0
    ctx.close()

"""Pins of the oracle's CPDI methods (ULCPDI, TLCPDI; CPU).

Checks that follow from the geometry of the particle domains, not from the
restated loops (ulcpdi.cpp:111-393, tlcpdi.cpp:96-329, solid.cpp:696-764,
1188-1200, 2188-2222, 2338-2352):
  * R4 weights are a partition of unity that reproduces linear fields and
    whose gradients reproduce their gradient, Σ ∇w x_Iᵀ = I, for square and
    sheared domains (a wrong sign or a swapped component of the domain
    vectors breaks the last one);
  * Q4 weights sum to 24 x 0.0417 = 1.0008 (the reference's rounded 1/24)
    for any quadrilateral, and their gradients give Σ ∇w x_Iᵀ = (A/V) I
    with A the corners' polygon area;
  * a uniformly moving R4 block keeps its velocity, F = I and rp = rp0
    under PIC, while a Q4 block's velocity grows by 1.0008 per step;
  * R4 domain vectors follow F (rp = F rp0, not Fᵀ rp0); a Q4 particle's
    volume is its corners' polygon area and J = vol / vol0;
  * TLCPDI takes its candidate nodes from the box's corner, not the
    solid's: a solid away from the box's lower corner finds no nodes.
"""
import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import native as N
from karamelo_amd import problems as P

H = 0.05


def _block(method="ulcpdi", style="R4", shape="linear", reg=(0.3, 0.7, 0.3, 0.7), flip=0.99,
           ppc=2, dt=1e-4):
    region = P.block(*reg)
    cfg = P.make_config(2, (0, 0), (1, 1), H, shape=shape, method=method, cpdi_style=style,
                        dt=dt, dt_constant=True, flip=flip,
                        solid_region=region if method == "tlcpdi" else None)
    mat = P.linear_material(1000.0, 1e6, 0.3)
    p = P.populate_solids(cfg, [(region, ppc, mat, 0)], fn=O.oracle_populate_fn())
    return cfg, mat, p


def _weights(cfg, mat, p):
    o = O.Oracle(cfg, [mat])
    o.upload(p)
    assert O.load().orc_phase_weights(o.h) == 0
    nn, tags, w, wd = o.neighbours()
    o.close()
    ny = _ny(cfg)
    lo = cfg.solidlo if cfg.method == N.METHOD_TLCPDI else cfg.boxlo
    valid = tags >= 0
    xi = np.where(valid, lo[0] + H * (tags // ny), 0.0)
    yi = np.where(valid, lo[1] + H * (tags % ny), 0.0)
    w = np.where(valid, w, 0.0)
    wd = np.where(valid[..., None], wd, 0.0)
    return nn, w, wd, xi, yi


def _ny(cfg):
    n = O.N.I64_3()
    assert O.load().orc_grid_size(O.C.byref(cfg), n) == 0
    return n[1]


def _interior(p, reg=(0.3, 0.7, 0.3, 0.7), margin=3):
    x = p["x"]
    return np.all((x[:, :2] > np.array(reg[0::2]) + margin * H) &
                  (x[:, :2] < np.array(reg[1::2]) - margin * H), axis=1)


def _gradient_moments(w, wd, xi, yi):
    """Σ w, Σ w x_I, Σ w y_I and G_ab = Σ ∂_b w x_I,a."""
    G = np.stack([np.stack([(wd[..., 0] * xi).sum(1), (wd[..., 1] * xi).sum(1)], -1),
                  np.stack([(wd[..., 0] * yi).sum(1), (wd[..., 1] * yi).sum(1)], -1)], 1)
    return w.sum(1), (w * xi).sum(1), (w * yi).sum(1), G


@pytest.mark.parametrize("shape", ["linear", "cubic-spline"])
@pytest.mark.parametrize("shear", [False, True])
def test_r4_weights_reproduce_linear_fields(shape, shear):
    cfg, mat, p = _block(shape=shape)
    if shear:     # a deformed domain: rp = F rp0
        F = np.array([[1.1, 0.3], [-0.2, 0.9]])
        r = p["rp0"].reshape(-1, 2, 3)
        rr = r.copy()
        rr[:, :, :2] = np.einsum("ij,pkj->pki", F, r[:, :, :2])
        p["rp"][:] = rr.reshape(-1, 6)
        p["vol"][:] = p["vol0"] * np.linalg.det(F)
    nn, w, wd, xi, yi = _weights(cfg, mat, p)
    inside = _interior(p)
    sw, sx, sy, G = _gradient_moments(w, wd, xi, yi)
    x = p["x"]
    np.testing.assert_allclose(sw[inside], 1.0, atol=1e-13)
    np.testing.assert_allclose(sx[inside], x[inside, 0], atol=1e-13)
    np.testing.assert_allclose(sy[inside], x[inside, 1], atol=1e-13)
    np.testing.assert_allclose(G[inside], np.broadcast_to(np.eye(2), G[inside].shape), atol=1e-12)
    np.testing.assert_allclose(wd[inside].sum(1), 0.0, atol=1e-11)


@pytest.mark.parametrize("shape", ["linear", "cubic-spline"])
def test_q4_weights_sum_to_the_rounded_constant(shape):
    cfg, mat, p = _block(style="Q4", shape=shape)
    rng = np.random.default_rng(5)
    c = p["xpc"].reshape(-1, 4, 3)
    c[:, :, :2] += rng.uniform(-0.15, 0.15, size=c[:, :, :2].shape) * 0.25 * H   # general quads
    p["xpc"][:] = c.reshape(-1, 12)
    nn, w, wd, xi, yi = _weights(cfg, mat, p)
    inside = _interior(p)
    sw, sx, sy, G = _gradient_moments(w, wd, xi, yi)
    np.testing.assert_allclose(sw[inside], 24 * 0.0417, atol=1e-13)
    X, Y = c[:, :, 0], c[:, :, 1]
    area = 0.5 * (X * np.roll(Y, -1, 1) - np.roll(X, -1, 1) * Y).sum(1)
    ratio = area / p["vol"]
    assert np.abs(ratio - 1).max() > 1e-3          # the corners moved off the square
    want = ratio[:, None, None] * np.eye(2)
    # nodes whose weight is below the 1e-12 cut drop their gradient too
    # (wfd ~ wfc / h): 1e-10
    np.testing.assert_allclose(G[inside], want[inside], atol=1e-10)


def test_uniform_motion_r4_keeps_its_state_q4_accelerates():
    v0 = 0.5
    for style in ("R4", "Q4"):
        cfg, mat, p = _block(style=style, flip=0.0)
        p["v"][:, 0] = v0
        o = O.Oracle(cfg, [mat])
        o.upload(p)
        assert o.step(10) == 0
        d = o.download()
        o.close()
        if style == "R4":
            np.testing.assert_allclose(d["v"][:, 0], v0, rtol=1e-13)
            np.testing.assert_allclose(d["F"], p["F"], atol=1e-13)
            np.testing.assert_allclose(d["rp"], d["rp0"], atol=1e-15)
        else:
            # PIC: every step's grid velocity is the last one x 1.0008
            np.testing.assert_allclose(d["v"][:, 0], v0 * (24 * 0.0417) ** 10, rtol=1e-12)
            np.testing.assert_allclose(d["J"], 1.0, atol=1e-12)


def _sheared(style):
    """30 steps of a block under a non-uniform shear flow."""
    cfg, mat, p = _block(style=style, shape="cubic-spline")
    x = p["x"]
    p["v"][:, 0] = 0.4 * np.sin(2 * np.pi * x[:, 1])
    p["v"][:, 1] = -0.1 * np.sin(2 * np.pi * x[:, 0])
    if style == "Q4":    # corners 1 % out from the centre: area 1.0201 vol0 at F = I
        c = p["xpc"].reshape(-1, 4, 3)
        c[:] = x[:, None, :] + 1.01 * (c - x[:, None, :])
        p["xpc"][:] = c.reshape(-1, 12)
    o = O.Oracle(cfg, [mat])
    o.upload(p)
    assert o.step(30) == 0
    d = o.download()
    o.close()
    return d


def test_r4_domain_vectors_follow_F():
    d = _sheared("R4")
    F = d["F"].reshape(-1, 3, 3)
    r0 = d["rp0"].reshape(-1, 2, 3)
    r = d["rp"].reshape(-1, 2, 3)
    want = np.einsum("pij,pkj->pki", F, r0)
    assert np.abs(F[:, 0, 1] - F[:, 1, 0]).max() > 1e-4     # F is not symmetric
    np.testing.assert_allclose(r, want, atol=1e-15)
    assert np.abs(r - np.einsum("pji,pkj->pki", F, r0)).max() > 1e-7


def test_q4_volume_is_the_corner_polygon():
    d = _sheared("Q4")
    c = d["xpc"].reshape(-1, 4, 3)
    X, Y = c[:, :, 0], c[:, :, 1]
    area = 0.5 * (X * np.roll(Y, -1, 1) - np.roll(X, -1, 1) * Y).sum(1)
    np.testing.assert_allclose(d["vol"], area, rtol=1e-13)
    np.testing.assert_allclose(d["J"], d["vol"] / d["vol0"], rtol=1e-15)
    detF = np.linalg.det(d["F"].reshape(-1, 3, 3))
    assert np.abs(d["J"] / detF - 1.0201).max() < 1e-3     # not det F


def test_tlcpdi_candidates_are_taken_from_the_box_corner():
    cfg, mat, p = _block(method="tlcpdi", reg=(0.3, 0.7, 0.3, 0.7))
    nn, w, wd, xi, yi = _weights(cfg, mat, p)
    assert nn.max() == 0
    cfg, mat, p = _block(method="tlcpdi", reg=(0.0, 0.4, 0.0, 0.4))
    nn, w, wd, xi, yi = _weights(cfg, mat, p)
    assert nn.min() == 4
    np.testing.assert_allclose(w.sum(1), 1.0, atol=1e-13)

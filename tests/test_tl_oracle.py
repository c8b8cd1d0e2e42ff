"""Pins of the oracle's TLMPM restatement (CPU):
  * the polar rotation equals U V^T of LAPACK's SVD (numpy) and is a proper
    rotation with R^T F symmetric positive definite;
  * the eigenvalue real parts equal numpy's on well-separated spectra,
    rotations (cos theta) and triangular matrices;
  * Region::limits, the TL grid (nodes over the solid's limits) and the TL
    lattice (cells from the region's lower corner);
  * a rigidly translating TL solid keeps F = I and zero stress; without
    fixes TLMPM conserves momentum; the Taylor bar decelerates and yields
    most at its impact end.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import native as N
from karamelo_amd import problems as P


def _polar(F):
    R = np.zeros(9)
    ok = O.load().orc_polar_rotation(O._dp(np.ascontiguousarray(F, float).ravel().copy()),
                                     O._dp(R))
    return ok, R.reshape(3, 3)


def _eig_re(F):
    re = np.zeros(3)
    O.load().orc_eig_real_parts(O._dp(np.ascontiguousarray(F, float).ravel().copy()), O._dp(re))
    return np.sort(re)


def test_polar_rotation_matches_svd():
    rng = np.random.default_rng(1)
    n = 0
    for _ in range(500):
        F = np.eye(3) + rng.normal(scale=0.4, size=(3, 3))
        if np.linalg.det(F) <= 0.05:
            continue
        ok, R = _polar(F)
        assert ok
        U, S, Vt = np.linalg.svd(F)
        assert np.abs(R - U @ Vt).max() <= 1e-13
        assert np.abs(R.T @ R - np.eye(3)).max() <= 1e-14
        assert np.linalg.det(R) > 0
        Us = R.T @ F
        assert np.abs(Us - Us.T).max() <= 1e-13 * np.abs(Us).max()
        assert np.all(np.linalg.eigvalsh(0.5 * (Us + Us.T)) > 0)
        n += 1
    assert n > 300
    ok, _ = _polar(np.diag([1.0, 1.0, -1.0]))      # improper: rejected
    assert not ok


def test_eigenvalue_real_parts():
    rng = np.random.default_rng(2)
    checked = 0
    for _ in range(500):
        F = np.eye(3) + rng.normal(scale=0.3, size=(3, 3))
        ev = np.linalg.eigvals(F)
        gaps = [abs(ev[i] - ev[j]) for i in range(3) for j in range(i)]
        if min(gaps) < 1e-2:
            continue
        assert np.abs(_eig_re(F) - np.sort(ev.real)).max() <= 1e-12
        checked += 1
    assert checked > 300
    for th in (0.1, 0.7, 2.0):                      # rotation about z: cos th +- i sin th, 1
        c, s = np.cos(th), np.sin(th)
        R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
        assert np.allclose(_eig_re(R), np.sort([c, c, 1.0]), atol=1e-14)
    T = np.array([[0.7, 0.3, -0.2], [0, 1.2, 0.5], [0, 0, 0.9]])
    assert np.allclose(_eig_re(T), [0.7, 0.9, 1.2], atol=1e-14)
    assert np.allclose(_eig_re(2.5 * np.eye(3)), [2.5] * 3)


def test_region_limits_and_tl_grid():
    cyl = P.cylinder(0.1, -0.2, 0.5, -1.0, 2.0, axis=0)
    lo, hi = N.region_limits(cyl)
    assert lo == (-1.0, 0.1 - 0.5, -0.2 - 0.5) and hi == (2.0, 0.1 + 0.5, -0.2 + 0.5)
    lo2, hi2 = C.c_double * 3, C.c_double * 3
    a, b = lo2(), hi2()
    O.load().orc_region_limits(C.byref(cyl), a, b)
    assert tuple(a) == lo and tuple(b) == hi
    for axis in (1, 2):
        cz = P.cylinder(0.3, 0.4, 0.25, -0.5, 0.5, axis=axis)
        l, h = N.region_limits(cz)
        assert l[axis] == -0.5 and h[axis] == 0.5
    prob = P.taylor_bar(h=8e-4)
    n = N.grid_size(prob.cfg)
    L = np.array(prob.cfg.solidhi) - np.array(prob.cfg.solidlo)
    h = prob.cfg.cellsize
    # nodes span the solid: the last node is the first at or beyond hi - h/2
    for d in range(3):
        assert (n[d] - 1) * h >= L[d] - 0.5 * h - 1e-12 and (n[d] - 2) * h < L[d] - 0.5 * h + 1e-12
    assert O.load().orc_grid_size(C.byref(prob.cfg), N.I64_3()) == 0


def test_tl_lattice_starts_at_region_limits():
    prob = P.taylor_bar(h=8e-4)
    x = prob.particles["x"]
    lo = np.array(prob.cfg.solidlo)
    h = prob.cfg.cellsize
    # one particle per cell, at the cell centres of a lattice anchored at lo
    k = (x - lo) / h - 0.5
    assert np.abs(k - np.round(k)).max() <= 1e-9
    # same particles from the oracle's populate
    q = P.taylor_bar(h=8e-4, populate_fn=O.oracle_populate_fn())
    assert np.array_equal(prob.particles["x"], q.particles["x"])


def test_tl_translation_keeps_F_identity():
    # length and diameter multiples of h: every particle lies inside the
    # span of the solid's grid, where the shape functions are a partition of
    # unity (32.4 mm is not, and its end-face particles see a cut stencil, as
    # in the reference)
    prob = P.taylor_bar(h=8e-4, v0=20.0, length=0.032)
    prob.fixes = []
    o = O.oracle_for(prob)
    assert o.step(30) == 0
    p = o.download()
    F = p["F"].reshape(-1, 3, 3)
    assert np.abs(F - np.eye(3)).max() <= 1e-12
    assert np.abs(p["sigma"]).max() <= 1e-6 * prob.mats[0].G * 1e-6
    assert np.allclose(p["v"][:, 0], 20.0, rtol=0, atol=1e-9)


def test_tl_momentum_conservation_without_fixes():
    prob = P.taylor_bar(h=8e-4, v0=100.0, length=0.032)
    prob.fixes = []
    rng = np.random.default_rng(4)
    prob.particles["v"][:] += rng.normal(scale=5.0, size=prob.particles["v"].shape)
    m = prob.particles["mass"]
    p0 = (m[:, None] * prob.particles["v"]).sum(axis=0)
    o = O.oracle_for(prob)
    assert o.step(50) == 0
    p = o.download()
    p1 = (m[:, None] * p["v"]).sum(axis=0)
    scale = (m[:, None] * np.abs(prob.particles["v"])).sum()
    assert np.abs(p1 - p0).max() <= 1e-10 * scale


def test_tl_taylor_bar_decelerates_and_yields_at_the_wall():
    prob = P.taylor_bar(h=8e-4)
    o = O.oracle_for(prob)
    assert o.step(150) == 0
    p = o.download()
    assert p["v"][:, 0].mean() < 0.95 * 190.0
    ep = p["eff_plastic_strain"]
    x0 = p["x0"][:, 0]
    L2 = 0.0162
    assert ep[x0 > L2 - 0.002].mean() > 10 * max(ep[x0 < -L2 + 0.01].mean(), 1e-12)

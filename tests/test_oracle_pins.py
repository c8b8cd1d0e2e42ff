"""Pins of the oracle (CPU): the reference's own compiled shape functions,
closed forms, invariants and known-answer cases chosen so that a dropped
term, a wrong sign or a transposed operand fails."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import native as N
from karamelo_amd import problems as P


def _rs():
    # exact breakpoints, their neighbours and a dense sweep
    pts = [-2.5, -2.0, -1.5, -1.0, -0.5, 0.0, 0.5, 1.0, 1.5, 2.0, 2.5]
    eps = [math.nextafter(p, -math.inf) for p in pts] + [math.nextafter(p, math.inf) for p in pts]
    sweep = np.linspace(-2.6, 2.6, 2001).tolist()
    return pts + eps + sweep


def test_basis_matches_reference_header():
    """orc_basis == src/basis_functions.h compiled from the reference tree, bit for bit."""
    ref = O.load_ref_basis()
    if ref is None:
        pytest.skip("reference tree not present: basis pinned by the closed-form tests below")
    L = O.load()
    inv_h = 1.0 / 0.05
    for shape in (0, 1):
        for der in (0, 1):
            for nt in (-2, -1, 0, 1, 2):
                for r in _rs():
                    a = L.orc_basis(shape, der, r, nt, inv_h)
                    b = ref.ref_basis(shape, der, r, nt, inv_h)
                    assert a == b or (math.isnan(a) and math.isnan(b)), (shape, der, nt, r, a, b)


def _weights(o):
    nn, tags, w, wd = o.neighbours()
    return nn, tags, w, wd


def _grid_xyz(cfg):
    X, Y, Z = P.node_coords(cfg)
    return np.stack([X, Y, Z], axis=1)


@pytest.mark.parametrize("shape", ["linear", "cubic-spline"])
@pytest.mark.parametrize("dim", [2, 3])
def test_weights_partition_of_unity_and_reproduction(shape, dim):
    """sum w = 1, sum grad w = 0, sum w x_I = x_p, sum x_I (x) grad w = I for
    particles anywhere in the box, including the boundary-modified splines."""
    rng = np.random.default_rng(1)
    h = 0.1
    lo, hi = (0.0, 0.0, 0.0), (1.0, 0.8, 0.6)
    cfg = P.make_config(dim, lo, hi, h, shape=shape)
    mat = P.linear_material(1000.0, 1e3, 0.3)
    n = 400
    arrs = N.empty_particles(n)
    for d in range(dim):
        arrs["x"][:, d] = rng.uniform(lo[d], hi[d], n)
    arrs["ptag"][:] = np.arange(1, n + 1)
    arrs["F"][:, [0, 4, 8]] = 1
    arrs["mass"][:] = 1
    arrs["vol0"][:] = 1
    o = O.Oracle(cfg, [mat])
    o.upload(arrs)
    o.phase("weights")
    nn, tags, w, wd = _weights(o)
    xg = _grid_xyz(cfg)
    for p in range(n):
        k = nn[p]
        t, ww, gg = tags[p, :k], w[p, :k], wd[p, :k]
        assert abs(ww.sum() - 1) < 1e-12
        assert np.abs(gg.sum(axis=0)).max() < 1e-9 / h
        rep = (ww[:, None] * xg[t]).sum(axis=0)
        assert np.abs(rep[:dim] - arrs["x"][p, :dim]).max() < 1e-12
        Jm = (xg[t][:, :dim, None] * gg[:, None, :dim]).sum(axis=0)
        assert np.abs(Jm - np.eye(dim)).max() < 1e-10


def test_cubic_weights_interior_closed_form():
    """Interior cubic B-spline weights equal the textbook polynomials of the
    fractional position t: (1-t)^3/6, (3t^3-6t^2+4)/6, (-3t^3+3t^2+3t+1)/6, t^3/6."""
    h = 0.1
    cfg = P.make_config(2, (0, 0), (2, 2), h, shape="cubic-spline")
    mat = P.linear_material(1000.0, 1e3, 0.3)
    ts = np.linspace(0.01, 0.99, 37)
    n = len(ts)
    arrs = N.empty_particles(n)
    arrs["x"][:, 0] = 1.0 + ts * h
    arrs["x"][:, 1] = 1.0 + 0.5 * h
    arrs["ptag"][:] = np.arange(1, n + 1)
    arrs["F"][:, [0, 4, 8]] = 1
    arrs["mass"][:] = 1
    arrs["vol0"][:] = 1
    o = O.Oracle(cfg, [mat])
    o.upload(arrs)
    o.phase("weights")
    nn, tags, w, wd = _weights(o)
    ny = N.grid_size(cfg)[1]
    for p, t in enumerate(ts):
        wx = {}
        for k in range(nn[p]):
            i, j = divmod(int(tags[p, k]), ny)
            wx.setdefault(i, 0.0)
            wx[i] += w[p, k]
        expect = [(1 - t) ** 3 / 6, (3 * t**3 - 6 * t**2 + 4) / 6,
                  (-3 * t**3 + 3 * t**2 + 3 * t + 1) / 6, t**3 / 6]
        got = [wx[i] for i in sorted(wx)]
        assert np.allclose(got, expect, atol=1e-13), (t, got, expect)


def test_linear_derivative_quirk_and_pair_filter():
    """derivative_linear(0) == 0 and pairs with zero weight are dropped
    (basis_functions.h:30-38, ulmpm.cpp:252-265)."""
    L = O.load()
    assert L.orc_basis(0, 1, 0.0, 0, 20.0) == 0.0
    assert L.orc_basis(0, 1, 0.5, 0, 20.0) == -20.0
    assert L.orc_basis(0, 1, -0.5, 0, 20.0) == 20.0
    # cubic ntype==1 at r in [1,2): value 0, derivative -1/h
    assert L.orc_basis(1, 0, 1.5, 1, 20.0) == 0.0
    assert L.orc_basis(1, 1, 1.5, 1, 20.0) == -20.0


def test_grid_size_closed_form():
    for L_, h, n in [(1.0, 0.05, 21), (1.0, 0.1, 11), (0.8, 0.1, 9), (2.0, 0.3, 8)]:
        cfg = P.make_config(2, (0, 0), (L_, L_), h)
        nx = O.load().orc_grid_size
        out = N.I64_3()
        assert nx(C.byref(cfg), out) == 0
        # smallest n with n*h > L + h/2 (grid.cpp:151-171)
        assert out[0] == out[1] == n == int(math.floor((L_ + 0.5 * h) / h)) + 1


def _mat(kind="jc"):
    return P.steel_jc(True)


def _sym(a):
    a = np.asarray(a, float).reshape(3, 3)
    return 0.5 * (a + a.T)


def _jc(m, sigma, D, dt, ep, epdot, dmg):
    L = O.load()
    out = np.zeros(9)
    dps = C.c_double(0.0)
    s = np.ascontiguousarray(sigma, float).ravel()
    d = np.ascontiguousarray(D, float).ravel()
    L.orc_jc_strength(C.byref(m), O._dp(s), O._dp(d), dt, ep, epdot, dmg, O._dp(out), C.byref(dps))
    return out.reshape(3, 3), dps.value


def test_jc_elastic_trial():
    m = _mat()
    G = m.strength_p[0]
    D = _sym([[1, 2, 0], [2, -3, 1], [0, 1, 0.5]]) * 1e-3
    sig = np.diag([1e6, -2e6, 5e5])
    dt = 1e-8
    s, dps = _jc(m, sig, D, dt, 0.0, 0.0, 0.0)
    trial = sig + 2 * G * dt * D
    dev = trial - np.trace(trial) / 3 * np.eye(3)
    assert dps == 0.0
    assert np.allclose(s, dev, rtol=1e-14, atol=1e-6)


def test_jc_radial_return_on_yield_surface():
    """Plastic step: 1.224744871*|s| == sigma_y (A + B ep^n)(1+max(epsdot/epsdot0,1))^C,
    dps = (J2_trial - sigma_y)/(3G), s parallel to the trial deviator."""
    m = _mat()
    G, A, B, n, e0, Cc = (m.strength_p[i] for i in range(6))
    D = _sym([[1, 0.2, 0], [0.2, -0.5, 0], [0, 0, -0.5]])
    sig = np.zeros((3, 3))
    dt, ep, epdot = 1e-2, 0.05, 5.0
    s, dps = _jc(m, sig, D, dt, ep, epdot, 0.0)
    ys = (A + B * ep**n) * (1 + max(epdot / e0, 1.0)) ** Cc
    assert dps > 0
    assert abs(1.224744871 * np.linalg.norm(s) - ys) < 1e-9 * ys
    trial = 2 * G * dt * D
    dev = trial - np.trace(trial) / 3 * np.eye(3)
    j2 = 1.224744871 * np.linalg.norm(dev)
    assert abs(dps - (j2 - ys) / (3 * G)) < 1e-12 * max(1.0, dps)
    assert np.allclose(s / np.linalg.norm(s), dev / np.linalg.norm(dev), atol=1e-14)


def test_jc_initial_yield_and_damage_scaling():
    m = _mat()
    G, A = m.strength_p[0], m.strength_p[1]
    Cc = m.strength_p[5]
    D = _sym([[1, 0, 0], [0, -0.5, 0], [0, 0, -0.5]])
    # ep < 1e-10 -> A; epsdot = 0 -> rate factor 2^C
    s, _ = _jc(m, np.zeros((3, 3)), D, 1e-2, 0.0, 0.0, 0.0)
    assert abs(1.224744871 * np.linalg.norm(s) - A * 2**Cc) < 1e-9 * A
    # damage 0.5 scales G and the yield stress
    s2, dps2 = _jc(m, np.zeros((3, 3)), D, 1e-2, 0.0, 0.0, 0.5)
    assert abs(1.224744871 * np.linalg.norm(s2) - 0.5 * A * 2**Cc) < 1e-9 * A
    trial = 2 * G * 0.5 * 1e-2 * D
    dev = trial - np.trace(trial) / 3 * np.eye(3)
    assert abs(dps2 - (1.224744871 * np.linalg.norm(dev) - 0.5 * A * 2**Cc) / (3 * G * 0.5)) < 1e-12
    # damage >= 1 -> zero deviator
    s3, dps3 = _jc(m, np.eye(3) * 1e8, D, 1e-2, 0.1, 0.0, 1.0)
    assert np.all(s3 == 0) and dps3 == 0.0


def test_shock_eos_known_answers():
    m = _mat()
    L = O.load()
    rho0, K, c0, S, Gam = (m.eos_p[i] for i in range(5))
    D = np.zeros(9)
    # mu = 0 -> p = 0
    assert L.orc_shock_pressure(C.byref(m), 1.0, rho0, 0.0, O._dp(D), 1e-4) == 0.0
    # Hugoniot closed form at mu = 0.05
    rho = rho0 * 1.05
    mu = rho / rho0 - 1
    pH = rho0 * c0**2 * mu * (1 + mu) / (1 - (S - 1) * mu) ** 2
    p = L.orc_shock_pressure(C.byref(m), 1 / 1.05, rho, 0.0, O._dp(D), 1e-4)
    assert abs(p - pH) < 1e-12 * abs(pH)
    # tension with damage: p(1-D); D >= 1 -> 0
    rho = rho0 * 0.99
    mu = rho / rho0 - 1
    pT = rho0 * c0**2 * mu * (1 + mu) / (1 - (S - 1) * mu) ** 2
    assert pT < 0
    p = L.orc_shock_pressure(C.byref(m), 1 / 0.99, rho, 0.3, O._dp(D), 1e-4)
    assert abs(p - 0.7 * pT) < 1e-12 * abs(pT)
    assert L.orc_shock_pressure(C.byref(m), 1 / 0.99, rho, 1.0, O._dp(D), 1e-4) == 0.0


def _dmg(m, dinit, dmg, pH, sdev, epdot, dps):
    L = O.load()
    a, b = C.c_double(dinit), C.c_double(dmg)
    s = np.ascontiguousarray(sdev, float).ravel()
    L.orc_jc_damage(C.byref(m), C.byref(a), C.byref(b), pH, O._dp(s), epdot, dps)
    return a.value, b.value


def test_jc_damage_known_answers():
    m = _mat()
    d1, d2, d3, d4 = (m.damage_p[i] for i in range(4))
    e0 = m.damage_p[5]
    s = np.diag([2.0, -1.0, -1.0]) * 1e8       # von Mises = 1.224744871*|s|
    vm = 1.224744871 * np.linalg.norm(s)
    pH = -vm / 3 * (1 - 1e-12)                    # triaxiality ~ 1/3 (with the 0.001|p| softening)
    triax = -pH / (vm + 0.001 * abs(pH))
    epdot = 100.0
    ef = (d1 + d2 * math.exp(d3 * triax)) * (1 + d4 * math.log(epdot / e0))
    a, b = _dmg(m, 0.0, 0.0, pH, s, epdot, 0.01)
    assert abs(a - 0.01 / ef) < 1e-15 and b == 0.0
    # onset: damage = min(10*(Dinit-1), 1)
    a, b = _dmg(m, 1.0 - 0.5 * 0.01 / ef, 0.0, pH, s, epdot, 0.01)
    assert abs(b - 10 * (a - 1)) < 1e-12 and 0 < b < 1
    # triaxiality <= -3 resets the initiation variable
    a, b = _dmg(m, 0.7, 0.0, 1e10, s, epdot, 0.01)
    assert a == 0.0 and b == 0.0
    # fully damaged and no plastic increment: untouched
    a, b = _dmg(m, 1.5, 1.0, pH, s, epdot, 0.0)
    assert a == 1.5 and b == 1.0


def test_two_disks_conservation_and_symmetry():
    """Whole USL step on the two-disk collision: total momentum stays 0 by
    symmetry, mass is conserved, Ek + Es stays near its initial value."""
    prob = P.two_disks(flip=1.0, populate_fn=O.oracle_populate_fn())
    assert prob.np == 416
    o = O.oracle_for(prob)
    p0 = o.download()
    ek0, es0 = o.energy()
    assert es0 == 0.0 and abs(ek0 - 0.5 * p0["mass"].sum() * 2 * P.lit(0.1) ** 2) < 1e-12
    etot = []
    for k in range(15):
        assert o.step(100) == 0
        ek, es = o.energy()
        etot.append(ek + es)
        p = o.download()
        mom = (p["mass"][:, None] * p["v"]).sum(axis=0)
        assert np.abs(mom).max() < 1e-12
    # linear-shape MPM loses a few percent of energy during contact
    assert max(abs(e - ek0) for e in etot) < 0.10 * ek0
    eks = []
    o2 = O.oracle_for(prob)
    for k in range(30):
        o2.step(100)
        eks.append(o2.energy()[0])
    # elastic rebound: kinetic energy nearly vanishes at peak compression,
    # then most of it comes back
    assert min(eks) < 0.1 * ek0 and eks[-1] > 0.8 * ek0


def test_p2g_conservation_oracle():
    """sum_I m_I = sum_p m_p and sum_I m_I v_I = sum_p m_p v_p after P2G."""
    prob = P.tensile_bar(h=4e-4, populate_fn=O.oracle_populate_fn())
    rng = np.random.default_rng(3)
    prob.particles["v"][:] = rng.normal(size=prob.particles["v"].shape)
    o = O.oracle_for(prob)
    o.phase("weights")
    o.phase("p2g")
    m, v, vu, f = o.grid()
    p = prob.particles
    assert abs(m.sum() - p["mass"].sum()) < 1e-12 * p["mass"].sum()
    mom = (m[:, None] * v).sum(axis=0)
    mp = (p["mass"][:, None] * p["v"]).sum(axis=0)
    assert np.abs(mom - mp).max() < 1e-12 * np.abs(p["mass"][:, None] * p["v"]).sum()
    # internal force of a zero-stress state is zero
    assert np.abs(f).max() == 0.0


def test_sort_permutation_is_sorted_and_complete():
    prob = P.tensile_bar(h=4e-4, populate_fn=O.oracle_populate_fn())
    o = O.oracle_for(prob)
    perm = o.sort_permutation((4, 4, 4))
    assert sorted(perm.tolist()) == sorted(prob.particles["ptag"].tolist())

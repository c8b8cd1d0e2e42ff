"""GPU parity of the thermo-mechanical coupling against the oracle, through
the C ABI (FP64).

- the grid temperature, the thermal driving forces and T_update
  (solid.cpp:2743-2797, grid.cpp:1354-1362), the particles' temperature
  (solid.cpp:2799-2808) and heat flux (:2810-2839) after every scheme's
  passes: USL, MUSL (the re-mapped grid temperature, ulmpm.cpp:491-493) and
  USF (usf.cpp:53-64);
- the temperature-aware constitutive update: JC thermal softening, the
  shock EOS's and the law's temperature terms, JC damage's temperature
  factor and the heat source of plastic work (solid.cpp:1292-1351,
  temperature_plastic_work.cpp:65-76);
- ULMPM (a pulled JC bar) and TLMPM (a Taylor impact), and conduction
  through a resting 2D block;
- restarts carry the temperature (solid.cpp:2881-2883, 2950-2952).

Bar: 1e-10 relative L2 on every particle field, T and gamma included
(conditioned by the oracle's own 1-ulp sensitivity on the Taylor impacts).
"""
import os

import numpy as np
import pytest

from karamelo_amd import native as N
from karamelo_amd import problems as P
from test_gpu_parity import FIELDS, run_and_compare, say
from test_gpu_apic import with_cfg

pytestmark = pytest.mark.gpu

THERMAL_FIELDS = FIELDS + ("T", "gamma")
CHI, CP, KAPPA, ALPHA, T0, TM = 0.9, 477.0, 50.0, 3.0e6, 293.0, 1800.0


def _hot_steel():
    """The tensile bar's steel with a shock EOS heat capacity, JC softening
    and damage from 293 K, and the plastic-work temperature law."""
    E, nu, rho = 200e9, P.lit(0.3), 7850.0
    K = E / (3 * (1 - 2 * nu))
    G = E / (2 * (1 + nu))
    c0 = float(np.sqrt(K / rho))
    mat = P.eos_strength_material(
        "shock", (rho, K, c0, P.lit(1.49), 2.0, 477.0, T0, 0.0, 0.0),
        "johnson_cook", (G, 792e6, 510e6, P.lit(0.26), 1.0, P.lit(0.014), 1.0, T0, TM),
        "damage_johnson_cook", (0.05, 3.44, -2.12, 0.002, 0.61, 1.0, T0, TM))
    return P.with_temperature(mat, CHI, CP, KAPPA, ALPHA, T0, TM)


def _thermo(prob, T=None):
    """The problem as a thermo-mechanical run, particles at T (default: a
    gradient along x from T0 to T0 + 200 K, so that heat flows)."""
    q = with_cfg(prob, thermo=1)
    x = q.particles["x"][:, 0]
    q.particles["T"][:] = T if T is not None else T0 + 200.0 * (x - x.min()) / np.ptp(x)
    return q


def tensile_bar(scheme):
    prob = P.tensile_bar(h=6e-4, v_pull=40.0, scheme=scheme, material=_hot_steel())
    return _thermo(prob)


def taylor_bar(scheme):
    prob = P.taylor_bar(h=8e-4, scheme=scheme)
    mats = [P.with_temperature(m, CHI, 385.0, 400.0, 0.0, T0, 1356.0) for m in prob.mats]
    prob.mats[:] = mats
    return _thermo(prob)


@pytest.mark.parametrize("scheme", ["usl", "musl", "usf"])
def test_ul_tensile_bar(scheme):
    say("thermo tensile bar %s: 120 steps" % scheme)
    worst, po = run_and_compare(tensile_bar(scheme), 120, 40, fields=THERMAL_FIELDS)
    assert np.abs(po["gamma"]).max() > 0          # plastic work heats
    assert worst["T"] <= 1e-10


@pytest.mark.parametrize("scheme", ["usl", "musl", "usf"])
def test_tl_taylor_bar(scheme):
    say("thermo TL Taylor bar %s: 100 steps" % scheme)
    worst, po = run_and_compare(taylor_bar(scheme), 100, 50, conditioned=True,
                                fields=THERMAL_FIELDS)
    assert np.abs(po["gamma"]).max() > 0


def conduction_block(scheme, shape):
    """A resting 2D block (linear EOS and strength) with T = T0 + a r^2."""
    cfg = P.make_config(2, (0, 0), (1, 1), 0.025, scheme=scheme, shape=shape, dt=1e-3,
                        dt_constant=True, thermo=True)
    mat = P.eos_strength_material("linear", (1000.0, 1e6), "linear", (5e5,))
    mat = P.with_temperature(mat, CHI, CP, KAPPA, 0.0, T0, TM)
    p = P.populate_solids(cfg, [(P.block(0.25, 0.75, 0.25, 0.75), 2, mat, 0)])
    x = p["x"]
    p["T"][:] = T0 + 500.0 * ((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2)
    return P.Problem("conduction", cfg, [mat], p)


@pytest.mark.parametrize("scheme,shape", [("usl", "cubic-spline"), ("musl", "cubic-spline"),
                                          ("usf", "cubic-spline"), ("usl", "linear")])
def test_conduction_block(scheme, shape):
    say("thermo conduction block %s %s: 60 steps" % (scheme, shape))
    prob = conduction_block(scheme, shape)
    T_start = prob.particles["T"].copy()
    worst, po = run_and_compare(prob, 60, 20, fields=THERMAL_FIELDS, check_energy=False)
    assert np.abs(po["T"] - T_start).max() > 1e-3   # the field has conducted


def test_restart_carries_the_temperature(tmp_path):
    say("thermo restart: T written and read back")
    prob = tensile_bar("usl")
    a = P.context_for(prob)
    a.step(20)
    assert a.synchronize()[0] == 0
    f = os.path.join(str(tmp_path), "thermo.restart")
    a.write_restart(f)
    b = P.context_for(prob)
    b.read_restart(f)
    pa, pb = a.download(), b.download()
    assert np.array_equal(pa["T"], pb["T"])
    assert np.array_equal(pa["x"], pb["x"])
    a.close()
    b.close()
    # a mechanical context refuses a thermo-mechanical restart
    c = P.context_for(with_cfg(prob, thermo=0))
    with pytest.raises(N.KmpmError):
        c.read_restart(f)
    c.close()


def test_rejected_configurations():
    say("thermo: rejected configurations")
    prob = tensile_bar("usl")
    for kw in ({"precision": N.PRECISION_MIXED}, {"sub_method": "APIC"},
               {"order": N.ORDER_REFERENCE}):
        bad = with_cfg(prob, **kw)
        with pytest.raises(N.KmpmError):
            N.Context(bad.cfg, bad.mats)
    ctx = P.context_for(prob)
    with pytest.raises(N.KmpmError):
        ctx.set_tile_variant(1)           # the thermal kernels: tile variant 0 only
    ctx.close()


def test_dumps_carry_the_temperatures(tmp_path):
    """A uniform T stays uniform on the grid: the grid dump's T column holds
    it on every node with mass and 0 elsewhere; the particle dump's T and
    gamma columns hold the particles' values."""
    say("thermo dumps: grid and particle temperatures")
    prob = conduction_block("usl", "cubic-spline")
    prob.particles["T"][:] = 300.0
    ctx = P.context_for(prob)
    ctx.step(1)
    ctx.dump_grid(str(tmp_path / "grid.*.LAMMPS"), ["mass", "T"])
    with open(tmp_path / "grid.1.LAMMPS") as f:
        lines = f.read().split("\n")
    nn = int(lines[3])
    rows = np.array([[float(x) for x in ln.split()] for ln in lines[9:9 + nn]])
    mass, T = rows[:, 3], rows[:, 4]
    assert (mass > 0).sum() > 0
    assert np.all(T[mass > 0] == 300.0) and np.all(T[mass == 0] == 0.0)
    ctx.dump_particles(str(tmp_path / "p.*.LAMMPS"), ["T", "gamma"])
    with open(tmp_path / "p.1.LAMMPS") as f:
        lines = f.read().split("\n")
    rows = np.array([[float(x) for x in ln.split()] for ln in lines[9:9 + prob.np]])
    p = ctx.download()
    assert np.all(np.abs(rows[:, 3] - p["T"]) <= 5e-6 * np.abs(p["T"]))
    assert np.all(rows[:, 4] == 0.0)      # a linear material does no plastic work
    ctx.close()

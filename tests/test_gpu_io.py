"""The reference's on-disk formats written from a context (kmpm_dump_particles,
kmpm_dump_grid, kmpm_write_restart / kmpm_read_restart), through the C ABI.

Bars:
  * particle dump: the header lines of DumpParticle::write
    (dump_particle.cpp:76-93) verbatim, the file name with the step in place
    of '*', one line per particle in solid then upload order, every value
    the downloaded field printed with %g (6 significant digits), seq the von
    Mises stress, TLMPM stresses rotated by the polar R of F;
  * grid dump: one line per node in tag order; ntype equal to the closed form
    min(2, i) - min(n - 1 - i, 2) (grid.cpp:236-238) evaluated per node,
    the body force m g on the group's nodes with mass;
  * restart: the update and solid blocks parsed back field by field from the
    bytes (Eigen column-major matrices), and a run resumed from the file
    equal to the uninterrupted run (reference-order FP64: 1e-13, the mass of
    a restarted particle being rho0 vol0 as in Solid::read_restart).
"""
import dataclasses
import os
import struct
import sys

import numpy as np
import pytest

from karamelo_amd import native as N
from karamelo_amd import problems as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _g(v):
    return float("%g" % v)


def _read_dump(path):
    with open(path) as f:
        lines = f.read().split("\n")
    return lines


def test_particle_dump_format(tmp_path, monkeypatch):
    say("particle dump: 3D tensile bar after 7 steps")
    prob = P.tensile_bar(h=6e-4, v_pull=40.0)
    ctx = P.context_for(prob)
    ctx.step(7)
    names = ["x", "y", "z", "vx", "s11", "s12", "s23", "seq", "e22", "volume", "mass", "damage",
             "ep", "epdot", "bx", "T", "gamma", "ienergy"]
    ctx.dump_particles(str(tmp_path / "dump_p.*.LAMMPS"), names)
    path = tmp_path / "dump_p.7.LAMMPS"
    assert path.exists()
    lines = _read_dump(path)
    p = ctx.download()
    n = prob.np
    c = prob.cfg
    head = ["ITEM: TIMESTEP", "0", "ITEM: NUMBER OF ATOMS", str(n), "ITEM: BOX BOUNDS sm sm sm",
            "%g %g" % (c.boxlo[0], c.boxhi[0]), "%g %g" % (c.boxlo[1], c.boxhi[1]),
            "%g %g" % (c.boxlo[2], c.boxhi[2]), "ITEM: ATOMS id type tag " + " ".join(names) + " "]
    assert lines[:9] == head
    rows = [ln.split() for ln in lines[9:9 + n]]
    assert len(rows) == n and lines[9 + n] == ""
    s = p["sigma"].reshape(n, 3, 3)
    dev = s - np.trace(s, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
    seq = np.sqrt(1.5) * np.sqrt(np.sum(dev * dev, axis=(1, 2)))
    e = p["strain_el"].reshape(n, 3, 3)
    want = np.stack([p["x"][:, 0], p["x"][:, 1], p["x"][:, 2], p["v"][:, 0], s[:, 0, 0],
                     s[:, 0, 1], s[:, 1, 2], seq, e[:, 1, 1], p["vol"], p["mass"], p["damage"],
                     p["eff_plastic_strain"], p["eff_plastic_strain_rate"], np.zeros(n),
                     np.zeros(n), np.zeros(n), np.zeros(n)], axis=1)
    for i in (0, 1, n // 2, n - 1):
        r = rows[i]
        assert r[0] == str(p["ptag"][i]) and r[1] == "1" and r[2] == str(p["ptag"][i])
        got = [float(x) for x in r[3:]]
        exp = [_g(v) for v in want[i]]
        assert np.allclose(got, exp, rtol=1e-5, atol=0), (i, got, exp)
    # every value: within the 6 significant digits of %g
    got = np.array([[float(x) for x in r[3:]] for r in rows])
    assert np.all(np.abs(got - want) <= 5e-6 * np.abs(want) + 1e-300)
    # an unknown variable is rejected, as the reference's constructor does
    with pytest.raises(N.KmpmError):
        ctx.dump_particles(str(tmp_path / "bad.*"), ["x", "pressure"])
    # a name without '*' becomes <name><step><name> (the reference's wrapped test)
    monkeypatch.chdir(tmp_path)
    ctx.dump_particles("plain.txt", ["x"])
    assert os.path.exists("plain.txt7plain.txt")
    ctx.close()


def test_tl_particle_dump_rotates_the_stress(tmp_path):
    say("particle dump: TLMPM Taylor bar, rotated stress")
    prob = P.taylor_bar(h=8e-4)
    ctx = P.context_for(prob)
    ctx.step(20)
    ctx.dump_particles(str(tmp_path / "tl.*"), ["s11", "s22", "s12"])
    lines = _read_dump(tmp_path / "tl.20")
    p = ctx.download()
    n = prob.np
    F = p["F"].reshape(n, 3, 3)
    U, _, Vt = np.linalg.svd(F)
    R = U @ Vt
    s = R @ p["sigma"].reshape(n, 3, 3) @ np.transpose(R, (0, 2, 1))
    got = np.array([[float(x) for x in ln.split()[3:]] for ln in lines[9:9 + n]])
    want = np.stack([s[:, 0, 0], s[:, 1, 1], s[:, 0, 1]], axis=1)
    scale = np.abs(want).max()
    assert np.all(np.abs(got - want) <= 5e-6 * np.abs(want) + 1e-9 * scale)
    ctx.close()


def test_grid_dump_format(tmp_path):
    say("grid dump: 3D dam break (cubic), body force group")
    prob = P.dam_break(n_cells=16, scheme="usl")
    ctx = P.context_for(prob)
    ctx.step(3)
    names = ["x", "y", "vx", "vy", "mass", "mask", "ntypex", "ntypey", "bz", "rigid"]
    ctx.dump_grid(str(tmp_path / "grid.*.LAMMPS"), names)
    lines = _read_dump(tmp_path / "grid.3.LAMMPS")
    nn = int(lines[3])
    n = N.grid_size(prob.cfg)
    assert nn == int(np.prod(n))
    rows = np.array([[float(x) for x in ln.split()] for ln in lines[9:9 + nn]])
    mass, v, _ = ctx.grid()
    tags = np.arange(nn)
    assert np.array_equal(rows[:, 0], tags) and np.all(rows[:, 1] == 1)
    i = tags // (n[1] * n[2])
    j = (tags // n[2]) % n[1]
    X, Y, Z = P.node_coords(prob.cfg)
    for k, (got, want) in enumerate(((rows[:, 3], X), (rows[:, 4], Y), (rows[:, 5], v[:, 0]),
                                     (rows[:, 6], v[:, 1]), (rows[:, 7], mass))):
        assert np.all(np.abs(got - want) <= 5e-6 * np.abs(want) + 1e-300), k
    assert np.array_equal(rows[:, 8].astype(np.int64), prob.node_mask.astype(np.int64))
    # ntype: the closed form per node (grid.cpp:236-238), written out per tag
    for col, idx, nd in ((9, i, n[0]), (10, j, n[1])):
        want = np.array([min(2, a) - min(nd - 1 - a, 2) for a in idx])
        assert np.array_equal(rows[:, col].astype(np.int64), want)
    # the body force: m g on nodes of the fix's group with mass
    bf = [f for f in prob.fixes if f.kind == "body_force"][0]
    g = bf.value[2]
    want = np.where((mass > 0) & ((prob.node_mask & bf.bit) != 0), mass * g, 0.0)
    assert np.all(np.abs(rows[:, 11] - want) <= 5e-6 * np.abs(want) + 1e-300)
    assert np.all(rows[:, 12] == 0)
    ctx.close()


def _ref_order(prob):
    cfg = N.Config.from_buffer_copy(prob.cfg)
    cfg.order = N.ORDER_REFERENCE
    return dataclasses.replace(prob, cfg=cfg)


def test_restart_blocks_and_resumed_run(tmp_path):
    say("restart: tensile bar, write at step 15, resume 15 more")
    prob = _ref_order(P.tensile_bar(h=6e-4, v_pull=40.0))
    a = P.context_for(prob)
    a.step(15)
    f = str(tmp_path / "restart.bin")
    a.write_restart(f)
    pa = a.download()
    dt, t, st = a.time()
    raw = open(f, "rb").read()
    # ---- parse the blocks back from the bytes ----
    o = 0

    def take(fmt):
        nonlocal o
        v = struct.unpack_from("<" + fmt, raw, o)
        o += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    def string():
        n = take("Q")
        nonlocal o
        s = raw[o:o + n].decode()
        o += n
        return s

    assert raw[:8] == b"KMPMRST1"
    o = 8
    assert take("i") == 3
    assert string() == "ulmpm" and take("?") is False and string() == "usl"
    assert take("i") == 1 and take("d") == prob.cfg.pic_flip and take("i") == 1
    assert take("Q") == 0
    assert take("d") == t and take("q") == st and take("d") == dt
    assert take("d") == prob.cfg.dt_factor and take("?") == bool(prob.cfg.dt_constant)
    box = take("6d")
    assert list(box[:3]) == list(prob.cfg.boxlo) and list(box[3:]) == list(prob.cfg.boxhi)
    assert take("i") == 1
    assert string() == "solid1"
    take("12d")
    np_ = take("q")
    assert np_ == prob.np and take("i") == prob.np and take("i") == 8 and take("i") == 0
    assert take("d") == prob.cfg.cellsize
    for i in range(3):
        assert take("q") == pa["ptag"][i]
        assert list(take("3d")) == list(pa["x0"][i])
        assert list(take("3d")) == list(pa["x"][i])
        assert list(take("3d")) == list(pa["v"][i])
        sig = np.array(take("9d")).reshape(3, 3).T          # column-major
        assert np.array_equal(sig, pa["sigma"][i].reshape(3, 3))
        eel = np.array(take("9d")).reshape(3, 3).T
        assert np.array_equal(eel, pa["strain_el"][i].reshape(3, 3))
        Fm = np.array(take("9d")).reshape(3, 3).T
        assert np.array_equal(Fm, pa["F"][i].reshape(3, 3))
        J, vol0, rho0, ep, epd, dmg, dinit, ien = take("8d")
        assert (J, vol0, ep, epd, dmg, dinit) == (pa["J"][i], pa["vol0"][i],
                                                  pa["eff_plastic_strain"][i],
                                                  pa["eff_plastic_strain_rate"][i],
                                                  pa["damage"][i], pa["damage_init"][i])
        assert rho0 == prob.mats[0].rho0 and ien == 0.0
        assert take("i") == pa["mask"][i]
    # ---- resume: a fresh context from the file, 15 more steps ----
    b = P.context_for(prob)
    b.read_restart(f)
    assert b.time() == (dt, t, st)
    a.step(15)
    b.step(15)
    assert a.synchronize()[0] == 0 and b.synchronize()[0] == 0
    ra, rb = a.download(), b.download()
    assert a.time()[2] == b.time()[2] == 30
    for k in ("x", "v", "sigma", "F", "eff_plastic_strain"):
        e = np.linalg.norm(ra[k] - rb[k]) / max(np.linalg.norm(ra[k]), 1e-300)
        say("  %s: resumed vs uninterrupted rel L2 %.1e" % (k, e))
        assert e <= 1e-13, k
    # a file for another scheme is refused
    other = P.context_for(_ref_order(P.tensile_bar(h=6e-4, v_pull=40.0, scheme="musl")))
    with pytest.raises(N.KmpmError):
        other.read_restart(f)
    for c in (a, b, other):
        c.close()


def test_tl_restart_resumed_run(tmp_path):
    say("restart: TLMPM Taylor bar, write at step 10, resume 10 more")
    prob = P.taylor_bar(h=8e-4)
    a = P.context_for(prob)
    a.step(10)
    f = str(tmp_path / "tl.bin")
    a.write_restart(f)
    b = P.context_for(prob)
    b.read_restart(f)
    a.step(10)
    b.step(10)
    ra, rb = a.download(), b.download()
    for k in ("x", "v", "sigma", "F"):
        e = np.linalg.norm(ra[k] - rb[k]) / max(np.linalg.norm(ra[k]), 1e-300)
        say("  %s: resumed vs uninterrupted rel L2 %.1e" % (k, e))
        # FP64 P2G sums in LDS-atomic order, and the PK1 is recomputed from
        # F and sigma at upload: round-off only
        assert e <= 1e-11, k
    a.close()
    b.close()

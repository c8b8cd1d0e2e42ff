"""ctypes binding of the oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg load it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from karamelo_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
REF_BASIS = os.path.join(ORACLE_DIR, "_ref", "libref_basis.so")

_LIBS = {}


def _build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "all"], check=True)


def load(variant="liboracle.so"):
    if variant not in _LIBS:
        path = os.path.join(ORACLE_DIR, variant)
        if not os.path.exists(path):
            _build()
        lib = C.CDLL(path)
        P = C.POINTER
        vp = C.c_void_p
        D = P(C.c_double)
        sigs = {
            "orc_material_init": (C.c_int, [P(N.Material)]),
            "orc_grid_size": (C.c_int, [P(N.Config), N.I64_3]),
            "orc_region_match": (C.c_int, [P(N.Region), C.c_int, C.c_double, C.c_double, C.c_double]),
            "orc_region_limits": (C.c_int, [P(N.Region), N.D3, N.D3]),
            "orc_polar_rotation": (C.c_int, [D, D]),
            "orc_eig_real_parts": (None, [D, D]),
            "orc_populate": (C.c_int, [P(N.Config), P(N.Region), C.c_int, P(N.Material), C.c_int,
                                       C.c_int, C.c_int64, P(N.Particles), C.c_int64, P(C.c_int64)]),
            "orc_create": (vp, [P(N.Config), P(N.Material), C.c_int]),
            "orc_destroy": (None, [vp]),
            "orc_set_particles": (C.c_int, [vp, P(N.Particles)]),
            "orc_get_particles": (C.c_int, [vp, P(N.Particles)]),
            "orc_set_node_mask": (C.c_int, [vp, P(C.c_uint32), C.c_int64]),
            "orc_fix_velocity_nodes": (C.c_int, [vp, C.c_uint32, N.I3, N.D3, N.D3]),
            "orc_fix_body_force": (C.c_int, [vp, C.c_uint32, N.I3, N.D3]),
            "orc_step": (C.c_int, [vp, C.c_int64]),
            "orc_get_affine": (C.c_int, [vp, D, C.c_int64]),
            "orc_get_heat_flux": (C.c_int, [vp, D, C.c_int64]),
            "orc_update_stress_point_T": (None, [P(N.Material), C.c_double, C.c_double,
                                                 C.c_double, C.c_double, D, D, D, D, D, D, D,
                                                 C.c_double, C.c_int, C.c_double, D]),
            "orc_set_affine": (C.c_int, [vp, D, C.c_int64]),
            "orc_get_time": (C.c_int, [vp, D, D, P(C.c_int64)]),
            "orc_energy": (C.c_int, [vp, C.c_uint32, D, D]),
            "orc_phase_weights": (C.c_int, [vp]),
            "orc_phase_p2g": (C.c_int, [vp]),
            "orc_phase_grid_update": (C.c_int, [vp]),
            "orc_get_grid": (C.c_int, [vp, D, D, D, D, C.c_int64]),
            "orc_get_neighbours": (C.c_int, [vp, P(C.c_int32), P(C.c_int64), D, D, C.c_int64, C.c_int]),
            "orc_cell_index": (C.c_int, [vp, P(C.c_int32), C.c_int64]),
            "orc_sort_permutation": (C.c_int, [vp, N.I3, P(C.c_int64), C.c_int64]),
            "orc_error_particle": (C.c_int64, [vp]),
            "orc_jc_strength": (None, [P(N.Material), D, D, C.c_double, C.c_double, C.c_double,
                                       C.c_double, D, D]),
            "orc_shock_pressure": (C.c_double, [P(N.Material), C.c_double, C.c_double, C.c_double,
                                                D, C.c_double]),
            "orc_jc_damage": (None, [P(N.Material), D, D, C.c_double, D, C.c_double, C.c_double]),
            "orc_basis": (C.c_double, [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double]),
            "orc_fix_velocity_nodes_expr": (C.c_int, [vp, C.c_uint32, P(C.c_char_p) , C.c_int,
                                                      P(C.c_char_p), D]),
            "orc_fix_velocity_force": (C.c_int, [vp, C.c_int, D]),
            "orc_var_eval": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_int,
                                       P(C.c_char_p), D, D]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _LIBS[variant] = lib
    return _LIBS[variant]


def load_ref_basis():
    """The reference's own basis_functions.h, compiled by `make -C oracle ref`
    where the reference tree exists; None otherwise."""
    if not os.path.exists(REF_BASIS):
        if os.path.isdir("/root/reference/src"):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR, "ref"], check=False)
    if not os.path.exists(REF_BASIS):
        return None
    lib = C.CDLL(REF_BASIS)
    lib.ref_basis.restype = C.c_double
    lib.ref_basis.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double]
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """Oracle simulation with the same surface as karamelo_amd.native.Context."""

    def __init__(self, cfg, mats, variant="liboracle.so"):
        self.L = load(variant)
        self.cfg = cfg
        self._mats = (N.Material * len(mats))(*mats)
        self.h = self.L.orc_create(C.byref(cfg), self._mats, len(mats))
        if not self.h:
            raise RuntimeError("orc_create failed")
        n = N.I64_3()
        assert self.L.orc_grid_size(C.byref(cfg), n) == 0
        self.nn = int(n[0] * n[1] * n[2])
        self.np = 0

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, arrs):
        self._keep = arrs
        ps = N.particles_struct(arrs)
        assert self.L.orc_set_particles(self.h, C.byref(ps)) == 0
        self.np = ps.np

    def download(self):
        arrs = N.empty_particles(self.np, cpdi=N.is_cpdi(self.cfg))
        ps = N.particles_struct(arrs, self.np)
        assert self.L.orc_get_particles(self.h, C.byref(ps)) == 0
        return arrs

    def set_node_mask(self, mask):
        mask = np.ascontiguousarray(mask, dtype=np.uint32)
        assert self.L.orc_set_node_mask(self.h, mask.ctypes.data_as(C.POINTER(C.c_uint32)),
                                        len(mask)) == 0

    def fix_velocity_nodes(self, bit, setv, val, prev):
        assert self.L.orc_fix_velocity_nodes(self.h, bit, N.I3(*setv), N.D3(*val), N.D3(*prev)) == 0

    def fix_body_force(self, bit, setv, val):
        assert self.L.orc_fix_body_force(self.h, bit, N.I3(*setv), N.D3(*val)) == 0

    def fix_velocity_nodes_expr(self, bit, texts, variables=None):
        t = (C.c_char_p * 3)(*[None if x is None else x.encode() for x in texts])
        names, vals = _vars(variables)
        assert self.L.orc_fix_velocity_nodes_expr(self.h, bit, t, len(variables or {}), names,
                                                  _dp(vals)) == 0

    def fix_force(self, fix):
        f = np.zeros(3)
        assert self.L.orc_fix_velocity_force(self.h, fix, _dp(f)) == 0
        return f

    def step(self, n):
        return self.L.orc_step(self.h, int(n))

    def affine(self):
        """L (ULMPM) or Fdot (TLMPM) per particle, (np, 3, 3)."""
        out = np.zeros((self.np, 3, 3))
        assert self.L.orc_get_affine(self.h, _dp(out), self.np) == 0
        return out

    def set_affine(self, Cm):
        Cm = np.ascontiguousarray(Cm, dtype=np.float64).reshape(self.np, 3, 3)
        assert self.L.orc_set_affine(self.h, _dp(Cm), self.np) == 0

    def phase(self, name):
        assert getattr(self.L, "orc_phase_" + name)(self.h) == 0

    def time(self):
        dt, t, s = C.c_double(), C.c_double(), C.c_int64()
        self.L.orc_get_time(self.h, C.byref(dt), C.byref(t), C.byref(s))
        return dt.value, t.value, s.value

    def energy(self, bit=1):
        ek, es = C.c_double(), C.c_double()
        self.L.orc_energy(self.h, bit, C.byref(ek), C.byref(es))
        return ek.value, es.value

    def grid(self):
        m = np.zeros(self.nn)
        v = np.zeros((self.nn, 3))
        vu = np.zeros((self.nn, 3))
        f = np.zeros((self.nn, 3))
        assert self.L.orc_get_grid(self.h, _dp(m), _dp(v), _dp(vu), _dp(f), self.nn) == 0
        return m, v, vu, f

    def neighbours(self):
        S1 = 2 if self.cfg.shape == N.SHAPE_LINEAR else 4
        S = N.CPDI_MAX_SPAN ** 2 if N.is_cpdi(self.cfg) else S1 ** self.cfg.dimension
        nn = np.zeros(self.np, dtype=np.int32)
        tags = np.zeros((self.np, S), dtype=np.int64)
        w = np.zeros((self.np, S))
        wd = np.zeros((self.np, S, 3))
        assert self.L.orc_get_neighbours(self.h, nn.ctypes.data_as(C.POINTER(C.c_int32)),
                                         tags.ctypes.data_as(C.POINTER(C.c_int64)), _dp(w),
                                         _dp(wd), self.np, S) == 0
        return nn, tags, w, wd

    def cells(self):
        ijk = np.zeros((self.np, 3), dtype=np.int32)
        assert self.L.orc_cell_index(self.h, ijk.ctypes.data_as(C.POINTER(C.c_int32)), self.np) == 0
        return ijk

    def sort_permutation(self, tile):
        out = np.zeros(self.np, dtype=np.int64)
        assert self.L.orc_sort_permutation(self.h, N.I3(*tile),
                                           out.ctypes.data_as(C.POINTER(C.c_int64)), self.np) == 0
        return out


def _vars(variables):
    variables = variables or {}
    names = (C.c_char_p * max(1, len(variables)))(*[k.encode() for k in variables])
    vals = np.array(list(variables.values()) or [0.0], dtype=float)
    return names, vals


def var_eval(text, previous, time, dt, variables=None):
    """The oracle's Var of a script expression at (time, dt)."""
    names, vals = _vars(variables)
    out = C.c_double()
    st = load().orc_var_eval(text.encode(), 1 if previous else 0, time, dt,
                             len(variables or {}), names, _dp(vals), C.byref(out))
    return out.value if st == 0 else None


def oracle_for(problem, variant="liboracle.so"):
    o = Oracle(problem.cfg, problem.mats, variant)
    o.upload(problem.particles)
    if problem.node_mask is not None:
        o.set_node_mask(problem.node_mask)
    for f in problem.fixes:
        if f.kind == "velocity_nodes":
            o.fix_velocity_nodes(f.bit, f.set, f.value, f.prev)
        elif f.kind == "velocity_nodes_expr":
            o.fix_velocity_nodes_expr(f.bit, f.value, f.variables)
        else:
            o.fix_body_force(f.bit, f.set, f.value)
    return o


def oracle_populate_fn():
    return load().orc_populate

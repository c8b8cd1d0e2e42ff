"""ctypes binding of libkmpm.so (include/kmpm.h).

The product path is the HIP library: loading fails loudly when libkmpm.so is
missing, and there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkmpm.so")

# ---- constants (include/kmpm.h) ------------------------------------------
OK = 0
ERR_ARG, ERR_HIP, ERR_NO_DEVICE, ERR_CAPACITY = 1, 2, 3, 4
ERR_PARTICLE_LEFT_BOX, ERR_JACOBIAN, ERR_DTCFL, ERR_WAVE_SPEED, ERR_GRID = 10, 11, 12, 13, 14
ERR_FIXED_POINT, ERR_INTERNAL, ERR_INSTANCE, ERR_DOMAIN = 15, 16, 17, 18
NTILE_VARIANTS = 3
SURROGATE_NPARAM = 1185
METHOD_ULMPM, METHOD_TLMPM, METHOD_ULCPDI, METHOD_TLCPDI = 0, 1, 2, 3
CPDI_R4, CPDI_Q4 = 0, 1
CPDI_MAX_SPAN = 8
SCHEME_USL, SCHEME_MUSL, SCHEME_USF = 0, 1, 2
SUB_PIC, SUB_FLIP, SUB_APIC, SUB_AFLIP, SUB_ASFLIP, SUB_MLS = 0, 1, 2, 3, 4, 5
SHAPE_LINEAR, SHAPE_CUBIC_SPLINE = 0, 1
PRECISION_FP64, PRECISION_MIXED = 0, 1
ORDER_FAST, ORDER_REFERENCE = 0, 1
MAT_LINEAR, MAT_EOS_STRENGTH = 0, 1
EOS_NONE, EOS_LINEAR, EOS_SHOCK, EOS_FLUID = 0, 1, 2, 3
STRENGTH_NONE, STRENGTH_LINEAR, STRENGTH_JOHNSON_COOK, STRENGTH_FLUID = 0, 1, 2, 3
DAMAGE_NONE, DAMAGE_JOHNSON_COOK = 0, 1
REGION_BLOCK, REGION_CYLINDER = 0, 1
PHASES = ["sort", "p2g", "grid", "g2p", "g2p_a", "p2g_b", "grid_b", "stress_b", "dt",
          "learned"]
NPHASE = len(PHASES)

# style keys of the reference's plugin registries -> enum values
EOS_KEYS = {"linear": EOS_LINEAR, "shock": EOS_SHOCK, "fluid": EOS_FLUID}
STRENGTH_KEYS = {"linear": STRENGTH_LINEAR, "johnson_cook": STRENGTH_JOHNSON_COOK,
                 "fluid": STRENGTH_FLUID}
DAMAGE_KEYS = {"damage_johnson_cook": DAMAGE_JOHNSON_COOK}
TEMP_NONE, TEMP_PLASTIC_WORK = 0, 1
TEMP_KEYS = {"plastic_work": TEMP_PLASTIC_WORK}
SCHEME_KEYS = {"usl": SCHEME_USL, "musl": SCHEME_MUSL, "usf": SCHEME_USF}
# method(<style>, <sub-method>, ...) words (update.h:78-83)
SUB_METHOD_KEYS = {"PIC": SUB_PIC, "FLIP": SUB_FLIP, "APIC": SUB_APIC, "AFLIP": SUB_AFLIP,
                   "ASFLIP": SUB_ASFLIP, "MLS": SUB_MLS}
METHOD_KEYS = {"ulmpm": METHOD_ULMPM, "tlmpm": METHOD_TLMPM, "ulcpdi": METHOD_ULCPDI,
               "tlcpdi": METHOD_TLCPDI}
CPDI_STYLE_KEYS = {"R4": CPDI_R4, "Q4": CPDI_Q4}   # ulcpdi.h:59
SHAPE_KEYS = {"linear": SHAPE_LINEAR, "cubic-spline": SHAPE_CUBIC_SPLINE}

D3 = C.c_double * 3
I3 = C.c_int * 3
I64_3 = C.c_int64 * 3


class Config(C.Structure):
    _fields_ = [("dimension", C.c_int), ("method", C.c_int), ("scheme", C.c_int),
                ("shape", C.c_int), ("precision", C.c_int), ("order", C.c_int),
                ("boxlo", D3), ("boxhi", D3), ("cellsize", C.c_double),
                ("pic_flip", C.c_double), ("dt", C.c_double), ("dt_constant", C.c_int),
                ("sub_method", C.c_int), ("dt_factor", C.c_double),
                ("solidlo", D3), ("solidhi", D3), ("np_per_cell", C.c_int),
                ("thermo", C.c_int), ("cpdi_style", C.c_int)]


class Material(C.Structure):
    _fields_ = [("type", C.c_int), ("eos", C.c_int), ("strength", C.c_int), ("damage", C.c_int),
                ("lin_p", C.c_double * 3), ("eos_p", C.c_double * 9),
                ("strength_p", C.c_double * 9), ("damage_p", C.c_double * 8),
                ("rho0", C.c_double), ("K", C.c_double), ("G", C.c_double), ("E", C.c_double),
                ("nu", C.c_double), ("lambda_", C.c_double), ("signal_velocity", C.c_double),
                ("temp", C.c_int), ("reserved_t", C.c_int), ("temp_p", C.c_double * 6)]


class Region(C.Structure):
    _fields_ = [("kind", C.c_int), ("axis", C.c_int), ("interior", C.c_int),
                ("reserved", C.c_int), ("p", C.c_double * 6)]


_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)


class SurrogateConfig(C.Structure):
    _fields_ = [("eps_p_max", C.c_double), ("rate_max", C.c_double), ("triax_max", C.c_double),
                ("samples", C.c_int64), ("iterations", C.c_int64), ("batch", C.c_int32),
                ("reserved", C.c_int32), ("learning_rate", C.c_double), ("seed", C.c_uint64)]


class CutNetConfig(C.Structure):
    _fields_ = [("eps_p_max", C.c_double), ("rate_min", C.c_double), ("rate_max", C.c_double),
                ("damage_max", C.c_double), ("damage_init_max", C.c_double),
                ("pressure_max", C.c_double), ("increment_min", C.c_double),
                ("increment_max", C.c_double), ("dt", C.c_double), ("cellsize", C.c_double),
                ("samples", C.c_int64), ("iterations", C.c_int64), ("batch", C.c_int32),
                ("judge_stride", C.c_int32), ("learning_rate", C.c_double), ("seed", C.c_uint64)]


CUTNET_NPARAM, CUTNET_NI, CUTNET_NO = 18441, 18, 9


class Particles(C.Structure):
    _fields_ = [("np", C.c_int64), ("ptag", C.POINTER(C.c_int64)),
                ("solid", C.POINTER(C.c_int32)), ("mat", C.POINTER(C.c_int32)),
                ("mask", C.POINTER(C.c_uint32)), ("x0", _DP), ("x", _DP), ("v", _DP),
                ("F", _DP), ("sigma", _DP), ("strain_el", _DP), ("mass", _DP), ("vol0", _DP),
                ("vol", _DP), ("rho0", _DP), ("rho", _DP), ("J", _DP),
                ("eff_plastic_strain", _DP), ("eff_plastic_strain_rate", _DP),
                ("damage", _DP), ("damage_init", _DP), ("T", _DP), ("gamma", _DP),
                ("rp0", _DP), ("rp", _DP), ("xpc0", _DP), ("xpc", _DP)]


# particle arrays: name -> (dtype, components)
PARTICLE_FIELDS = {
    "ptag": (np.int64, 1), "solid": (np.int32, 1), "mat": (np.int32, 1), "mask": (np.uint32, 1),
    "x0": (np.float64, 3), "x": (np.float64, 3), "v": (np.float64, 3), "F": (np.float64, 9),
    "sigma": (np.float64, 9), "strain_el": (np.float64, 9), "mass": (np.float64, 1),
    "vol0": (np.float64, 1), "vol": (np.float64, 1), "rho0": (np.float64, 1),
    "rho": (np.float64, 1), "J": (np.float64, 1), "eff_plastic_strain": (np.float64, 1),
    "eff_plastic_strain_rate": (np.float64, 1), "damage": (np.float64, 1),
    "damage_init": (np.float64, 1), "T": (np.float64, 1), "gamma": (np.float64, 1),
}
# CPDI particle domains (solid.h): R4 vectors [np][2][3], Q4 corners [np][4][3];
# present in a particle dict only for CPDI problems
CPDI_FIELDS = {"rp0": (np.float64, 6), "rp": (np.float64, 6), "xpc0": (np.float64, 12),
               "xpc": (np.float64, 12)}


def is_cpdi(cfg):
    return cfg.method in (METHOD_ULCPDI, METHOD_TLCPDI)


def empty_particles(n, cpdi=False):
    """Dict of zeroed particle arrays in the reference layouts (with the CPDI
    domain arrays when cpdi)."""
    out = {}
    fields = dict(PARTICLE_FIELDS, **CPDI_FIELDS) if cpdi else PARTICLE_FIELDS
    for k, (dt, c) in fields.items():
        out[k] = np.zeros((n, c) if c > 1 else (n,), dtype=dt)
    return out


def particles_struct(arrs, n=None):
    """Build a Particles struct pointing at the numpy arrays in `arrs`
    (the arrays must stay alive while the struct is used)."""
    p = Particles()
    p.np = int(n if n is not None else len(arrs["ptag"]))
    ctmap = {np.int64: C.c_int64, np.int32: C.c_int32, np.uint32: C.c_uint32,
             np.float64: C.c_double}
    for k, (dt, _) in dict(PARTICLE_FIELDS, **CPDI_FIELDS).items():
        a = arrs.get(k)
        if a is None:
            setattr(p, k, None)
            continue
        assert a.dtype == dt and a.flags["C_CONTIGUOUS"], k
        setattr(p, k, a.ctypes.data_as(C.POINTER(ctmap[dt])))
    return p


def concat_particles(parts):
    keys = [k for k in dict(PARTICLE_FIELDS, **CPDI_FIELDS) if all(k in p for p in parts)]
    return {k: np.concatenate([p[k] for p in parts]) for k in keys}


EXPR_MAX = 32


class Expr(C.Structure):
    """kmpm_expr: a fix value program of time and dt."""
    _fields_ = [("n", C.c_int32), ("op", C.c_int32 * EXPR_MAX), ("c", C.c_double * EXPR_MAX)]


Expr3 = Expr * 3


def _declare(lib):
    vp = C.c_void_p
    P = C.POINTER
    sigs = {
        "kmpm_status_string": (C.c_char_p, [C.c_int]),
        "kmpm_material_init": (C.c_int, [P(Material)]),
        "kmpm_grid_size": (C.c_int, [P(Config), I64_3]),
        "kmpm_populate": (C.c_int, [P(Config), P(Region), C.c_int, P(Material), C.c_int, C.c_int,
                                    C.c_int64, P(Particles), C.c_int64, P(C.c_int64)]),
        "kmpm_region_match": (C.c_int, [P(Region), C.c_int, C.c_double, C.c_double, C.c_double]),
        "kmpm_region_limits": (C.c_int, [P(Region), D3, D3]),
        "kmpm_ctx_create": (C.c_int, [P(Config), P(Material), C.c_int, P(vp)]),
        "kmpm_ctx_destroy": (C.c_int, [vp]),
        "kmpm_upload_particles": (C.c_int, [vp, P(Particles)]),
        "kmpm_download_particles": (C.c_int, [vp, P(Particles)]),
        "kmpm_set_node_mask": (C.c_int, [vp, P(C.c_uint32), C.c_int64]),
        "kmpm_fix_velocity_nodes": (C.c_int, [vp, C.c_uint32, I3, D3, D3]),
        "kmpm_fix_body_force": (C.c_int, [vp, C.c_uint32, I3, D3]),
        "kmpm_expr_compile": (C.c_int, [C.c_char_p, C.c_int, C.c_int, P(C.c_char_p), _DP,
                                        P(Expr)]),
        "kmpm_fix_velocity_nodes_expr": (C.c_int, [vp, C.c_uint32, Expr3, Expr3]),
        "kmpm_fix_velocity_nodes_force": (C.c_int, [vp, C.c_int, D3]),
        "kmpm_compute_grid_weight_functions_and_gradients": (C.c_int, [vp]),
        "kmpm_particles_to_grid": (C.c_int, [vp]),
        "kmpm_update_grid_state": (C.c_int, [vp]),
        "kmpm_grid_to_points_and_stress": (C.c_int, [vp]),
        "kmpm_grid_to_points": (C.c_int, [vp]),
        "kmpm_velocities_to_grid": (C.c_int, [vp]),
        "kmpm_update_stress": (C.c_int, [vp]),
        "kmpm_particles_to_grid_usf_1": (C.c_int, [vp]),
        "kmpm_particles_to_grid_usf_2": (C.c_int, [vp]),
        "kmpm_download_affine": (C.c_int, [vp, C.POINTER(C.c_double), C.c_int64]),
        "kmpm_adjust_dt": (C.c_int, [vp]),
        "kmpm_step": (C.c_int, [vp, C.c_int64]),
        "kmpm_run_time": (C.c_int, [vp, C.c_double, C.c_int64, P(C.c_int64)]),
        "kmpm_synchronize": (C.c_int, [vp, P(C.c_int64)]),
        "kmpm_set_instances": (C.c_int, [vp, C.c_int, P(C.c_int32), P(C.c_int32)]),
        "kmpm_instance_time": (C.c_int, [vp, C.c_int, _DP, _DP]),
        "kmpm_step_graph": (C.c_int, [vp, C.c_int64, C.c_int64]),
        "kmpm_set_tile_variant": (C.c_int, [vp, C.c_int]),
        "kmpm_autotune": (C.c_int, [vp, C.c_int, P(C.c_int), _DP]),
        "kmpm_get_time": (C.c_int, [vp, P(C.c_double), P(C.c_double), P(C.c_int64)]),
        "kmpm_reduce_energy": (C.c_int, [vp, C.c_uint32, P(C.c_double), P(C.c_double)]),
        "kmpm_download_grid": (C.c_int, [vp, _DP, _DP, _DP, C.c_int64]),
        "kmpm_download_cells": (C.c_int, [vp, P(C.c_int32), P(C.c_int64), C.c_int64]),
        "kmpm_debug_weights": (C.c_int, [vp, P(C.c_int64), _DP, _DP, C.c_int64]),
        "kmpm_get_tile_shape": (C.c_int, [vp, I3]),
        "kmpm_timing_enable": (C.c_int, [vp, C.c_int]),
        "kmpm_timing_read": (C.c_int, [vp, P(C.c_double), P(C.c_int64)]),
        "kmpm_copy_probe": (C.c_int, [C.c_int64, C.c_int, P(C.c_double)]),
        "kmpm_surrogate_create": (C.c_int, [P(Material), P(SurrogateConfig), P(vp)]),
        "kmpm_surrogate_destroy": (C.c_int, [vp]),
        "kmpm_surrogate_train": (C.c_int, [vp, _DP]),
        "kmpm_surrogate_tuples": (C.c_int, [vp, C.c_int, C.c_int64, C.c_uint64, _FP, _FP]),
        "kmpm_surrogate_tuples_from_ctx": (C.c_int, [vp, vp, C.c_int, C.c_int64, _FP, _FP,
                                                      C.c_int64]),
        "kmpm_surrogate_set_tuples": (C.c_int, [vp, C.c_int, C.c_int64, _FP, _FP, C.c_int64]),
        "kmpm_surrogate_eval": (C.c_int, [vp, C.c_int, _FP, C.c_int64, _FP]),
        "kmpm_surrogate_gradient": (C.c_int, [vp, C.c_int, _FP, _FP, C.c_int64, _FP, _DP]),
        "kmpm_surrogate_params": (C.c_int, [vp, C.c_int, _FP, C.c_int, C.c_int]),
        "kmpm_surrogate_normalization": (C.c_int, [vp, C.c_int, _DP]),
        "kmpm_surrogate_envelope": (C.c_int, [vp, _DP, _DP]),
        "kmpm_ctx_set_surrogate": (C.c_int, [vp, vp]),
        "kmpm_surrogate_stats": (C.c_int, [vp, P(C.c_int64), P(C.c_int64)]),
        "kmpm_dump_particles": (C.c_int, [vp, C.c_char_p, P(C.c_char_p), C.c_int]),
        "kmpm_dump_grid": (C.c_int, [vp, C.c_char_p, P(C.c_char_p), C.c_int]),
        "kmpm_write_restart": (C.c_int, [vp, C.c_char_p]),
        "kmpm_read_restart": (C.c_int, [vp, C.c_char_p]),
        "kmpm_set_time": (C.c_int, [vp, C.c_double, C.c_double, C.c_int64]),
        "kmpm_cutnet_create": (C.c_int, [P(Material), P(CutNetConfig), P(vp)]),
        "kmpm_cutnet_destroy": (C.c_int, [vp]),
        "kmpm_cutnet_train": (C.c_int, [vp, C.c_int64, _DP, _DP]),
        "kmpm_cutnet_tuples": (C.c_int, [vp, C.c_int64, C.c_uint64, _FP, _FP]),
        "kmpm_cutnet_eval": (C.c_int, [vp, _FP, C.c_int64, _FP]),
        "kmpm_cutnet_gradient": (C.c_int, [vp, _FP, _FP, C.c_int64, _FP, _DP]),
        "kmpm_cutnet_gradient_noalloc": (C.c_int, [vp, vp, vp, C.c_int64, vp, vp]),
        "kmpm_cutnet_params": (C.c_int, [vp, _FP, C.c_int, C.c_int]),
        "kmpm_cutnet_maps": (C.c_int, [vp, _FP, _FP, _FP, _FP, C.c_int]),
        "kmpm_cutnet_envelope": (C.c_int, [vp, _DP]),
        "kmpm_ctx_set_cutnet": (C.c_int, [vp, vp]),
        "kmpm_cutnet_stats": (C.c_int, [vp, P(C.c_int64), P(C.c_int64), P(C.c_int64), _DP, _DP]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_LIB = None


def lib():
    """Load libkmpm.so (built by karamelo_amd.build); raises if it is missing."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libkmpm.so is not built (run karamelo_amd.build.build()); "
                               "there is no CPU fallback for the MPM step")
        _LIB = _declare(C.CDLL(LIB_PATH))
    return _LIB


class KmpmError(RuntimeError):
    def __init__(self, code, what=""):
        msg = lib().kmpm_status_string(code).decode()
        super().__init__("%s: status %d (%s)" % (what, code, msg))
        self.code = code


def check(code, what=""):
    if code != OK:
        raise KmpmError(code, what)
    return code


def material_init(m):
    check(lib().kmpm_material_init(C.byref(m)), "kmpm_material_init")
    return m


def region_limits(region):
    """Region::limits: (lo[3], hi[3]) of a region's bounding box."""
    lo, hi = D3(), D3()
    check(lib().kmpm_region_limits(C.byref(region), lo, hi), "kmpm_region_limits")
    return tuple(lo), tuple(hi)


def grid_size(cfg):
    n = I64_3()
    check(lib().kmpm_grid_size(C.byref(cfg), n), "kmpm_grid_size")
    return tuple(int(v) for v in n)


def populate(cfg, region, ppc1d, mat, mat_index, solid_index, ptag_offset, fn=None):
    """Solid::populate through the C ABI (or an equivalent `fn` with the same
    signature, e.g. the oracle's); returns a dict of particle arrays."""
    fn = fn or lib().kmpm_populate
    npo = C.c_int64(0)
    empty = particles_struct(empty_particles(0), 0)
    st = fn(C.byref(cfg), C.byref(region), ppc1d, C.byref(mat), mat_index, solid_index,
            ptag_offset, C.byref(empty), 0, C.byref(npo))
    if st not in (OK, ERR_CAPACITY):
        raise KmpmError(st, "populate (count)")
    n = npo.value
    arrs = empty_particles(n, cpdi=is_cpdi(cfg))
    ps = particles_struct(arrs, n)
    check(fn(C.byref(cfg), C.byref(region), ppc1d, C.byref(mat), mat_index, solid_index,
             ptag_offset, C.byref(ps), n, C.byref(npo)), "populate")
    return arrs


class Context:
    """One device context (kmpm_ctx) on the current HIP device."""

    def __init__(self, cfg, mats):
        self.L = lib()
        self.cfg = cfg
        arr = (Material * len(mats))(*mats)
        self._mats = arr
        h = C.c_void_p()
        check(self.L.kmpm_ctx_create(C.byref(cfg), arr, len(mats), C.byref(h)), "kmpm_ctx_create")
        self.h = h
        self.nn = int(np.prod(grid_size(cfg)))
        self.np = 0

    def close(self):
        if self.h:
            self.L.kmpm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, arrs):
        self._upload_keep = arrs
        ps = particles_struct(arrs)
        check(self.L.kmpm_upload_particles(self.h, C.byref(ps)), "upload")
        self.np = ps.np

    def download(self):
        arrs = empty_particles(self.np, cpdi=is_cpdi(self.cfg))
        ps = particles_struct(arrs, self.np)
        check(self.L.kmpm_download_particles(self.h, C.byref(ps)), "download")
        return arrs

    def set_node_mask(self, mask):
        mask = np.ascontiguousarray(mask, dtype=np.uint32)
        check(self.L.kmpm_set_node_mask(self.h, mask.ctypes.data_as(C.POINTER(C.c_uint32)),
                                        len(mask)), "set_node_mask")

    def fix_velocity_nodes(self, bit, setv, val, prev):
        check(self.L.kmpm_fix_velocity_nodes(self.h, bit, I3(*setv), D3(*val), D3(*prev)),
              "fix_velocity_nodes")

    def fix_body_force(self, bit, setv, val):
        check(self.L.kmpm_fix_body_force(self.h, bit, I3(*setv), D3(*val)), "fix_body_force")

    def fix_velocity_nodes_expr(self, bit, texts, variables=None):
        """fix(id, velocity_nodes, group, vx, vy, vz) with the script's
        expression texts (None or "NULL": free component) and its constant
        user variables {name: value}."""
        val, prev = compile_fix_values(texts, variables)
        check(self.L.kmpm_fix_velocity_nodes_expr(self.h, bit, val, prev),
              "fix_velocity_nodes_expr")

    def fix_force(self, fix):
        """Reaction force of velocity fix number `fix` in the last step."""
        f = D3()
        check(self.L.kmpm_fix_velocity_nodes_force(self.h, int(fix), f), "fix_velocity_nodes_force")
        return np.array(f[:])

    def phase(self, name):
        check(getattr(self.L, "kmpm_" + name)(self.h), name)

    def affine(self):
        """The APIC affine matrix per particle (L, or Fdot on TLMPM), (np, 3, 3)."""
        out = np.zeros((self.np, 3, 3))
        check(self.L.kmpm_download_affine(self.h, out.ctypes.data_as(C.POINTER(C.c_double)),
                                          self.np), "download_affine")
        return out

    def step(self, n):
        check(self.L.kmpm_step(self.h, int(n)), "step")

    def run_time(self, duration, max_steps=1 << 40):
        """run_time(duration) of an input script; returns the steps run."""
        n = C.c_int64()
        check(self.L.kmpm_run_time(self.h, float(duration), int(max_steps), C.byref(n)), "run_time")
        return n.value

    def step_graph(self, n, graph_steps):
        """n steps as replays of a captured graph of `graph_steps` steps."""
        check(self.L.kmpm_step_graph(self.h, int(n), int(graph_steps)), "step_graph")

    def set_instances(self, cell_lo, cell_hi):
        """Batched instances: instance i owns cells [cell_lo[i], cell_hi[i])
        (ninst x 3, tile-aligned)."""
        lo = np.ascontiguousarray(cell_lo, dtype=np.int32).reshape(-1, 3)
        hi = np.ascontiguousarray(cell_hi, dtype=np.int32).reshape(-1, 3)
        assert lo.shape == hi.shape
        ip = C.POINTER(C.c_int32)
        check(self.L.kmpm_set_instances(self.h, len(lo), lo.ctypes.data_as(ip),
                                        hi.ctypes.data_as(ip)), "set_instances")
        self.ninst = len(lo)

    def instance_time(self):
        n = getattr(self, "ninst", 0)
        dt, t = np.zeros(n), np.zeros(n)
        check(self.L.kmpm_instance_time(self.h, n, dt.ctypes.data_as(_DP), t.ctypes.data_as(_DP)),
              "instance_time")
        return dt, t

    def synchronize(self):
        t = C.c_int64(-1)
        st = self.L.kmpm_synchronize(self.h, C.byref(t))
        return st, t.value

    def time(self):
        dt, t, s = C.c_double(), C.c_double(), C.c_int64()
        check(self.L.kmpm_get_time(self.h, C.byref(dt), C.byref(t), C.byref(s)), "get_time")
        return dt.value, t.value, s.value

    def energy(self, bit=1):
        ek, es = C.c_double(), C.c_double()
        check(self.L.kmpm_reduce_energy(self.h, bit, C.byref(ek), C.byref(es)), "energy")
        return ek.value, es.value

    def grid(self):
        m = np.zeros(self.nn)
        v = np.zeros((self.nn, 3))
        vu = np.zeros((self.nn, 3))
        check(self.L.kmpm_download_grid(self.h, m.ctypes.data_as(_DP), v.ctypes.data_as(_DP),
                                        vu.ctypes.data_as(_DP), self.nn), "download_grid")
        return m, v, vu

    def cells(self, sorted_ptags=True):
        ijk = np.zeros((self.np, 3), dtype=np.int32)
        sp = np.zeros(self.np, dtype=np.int64)
        check(self.L.kmpm_download_cells(
            self.h, ijk.ctypes.data_as(C.POINTER(C.c_int32)),
            sp.ctypes.data_as(C.POINTER(C.c_int64)) if sorted_ptags else None, self.np),
            "download_cells")
        return ijk, (sp if sorted_ptags else None)

    def timing_enable(self, on=True):
        check(self.L.kmpm_timing_enable(self.h, 1 if on else 0), "timing_enable")

    def timing_read(self):
        ms = (C.c_double * NPHASE)()
        cnt = (C.c_int64 * NPHASE)()
        check(self.L.kmpm_timing_read(self.h, ms, cnt), "timing_read")
        return {PHASES[k]: (ms[k], cnt[k]) for k in range(NPHASE)}

    def tile_shape(self):
        t = I3()
        check(self.L.kmpm_get_tile_shape(self.h, t), "get_tile_shape")
        return tuple(t)

    def set_tile_variant(self, variant):
        check(self.L.kmpm_set_tile_variant(self.h, int(variant)), "set_tile_variant")

    def autotune(self, trial_steps=10):
        """Times every tile variant from the current state (which is kept)
        and keeps the fastest; returns (variant, ms per step per variant)."""
        v = C.c_int(-1)
        ms = np.zeros(NTILE_VARIANTS)
        check(self.L.kmpm_autotune(self.h, int(trial_steps), C.byref(v), ms.ctypes.data_as(_DP)),
              "autotune")
        return v.value, ms

    def set_surrogate(self, surrogate):
        check(self.L.kmpm_ctx_set_surrogate(self.h, surrogate.h if surrogate else None),
              "set_surrogate")

    def dump_particles(self, filename, variables):
        v = (C.c_char_p * len(variables))(*[x.encode() for x in variables])
        check(self.L.kmpm_dump_particles(self.h, filename.encode(), v, len(variables)),
              "dump_particles")

    def dump_grid(self, filename, variables):
        v = (C.c_char_p * len(variables))(*[x.encode() for x in variables])
        check(self.L.kmpm_dump_grid(self.h, filename.encode(), v, len(variables)), "dump_grid")

    def write_restart(self, filename):
        check(self.L.kmpm_write_restart(self.h, filename.encode()), "write_restart")

    def read_restart(self, filename):
        check(self.L.kmpm_read_restart(self.h, filename.encode()), "read_restart")

    def set_time(self, dt, time, step):
        check(self.L.kmpm_set_time(self.h, dt, time, step), "set_time")

    def set_cutnet(self, net):
        check(self.L.kmpm_ctx_set_cutnet(self.h, net.h if net else None), "set_cutnet")

    def cutnet_stats(self):
        """(network updates, fallbacks, judged updates, judged relative L2
        error of the deviatoric stress) since the last call."""
        ev, fb, nj = C.c_int64(), C.c_int64(), C.c_int64()
        e2, r2 = C.c_double(), C.c_double()
        check(self.L.kmpm_cutnet_stats(self.h, C.byref(ev), C.byref(fb), C.byref(nj),
                                       C.byref(e2), C.byref(r2)), "cutnet_stats")
        rel = (e2.value / r2.value) ** 0.5 if r2.value > 0 else 0.0
        return ev.value, fb.value, nj.value, rel

    def surrogate_stats(self):
        ev, fb = C.c_int64(), C.c_int64()
        check(self.L.kmpm_surrogate_stats(self.h, C.byref(ev), C.byref(fb)), "surrogate_stats")
        return ev.value, fb.value

    def debug_weights(self):
        d = self.cfg.dimension
        S1 = 2 if self.cfg.shape == SHAPE_LINEAR else 4
        S = S1 ** d
        tag = np.zeros((self.np, S), dtype=np.int64)
        w = np.zeros((self.np, S))
        wd = np.zeros((self.np, S, 3))
        check(self.L.kmpm_debug_weights(self.h, tag.ctypes.data_as(C.POINTER(C.c_int64)),
                                        w.ctypes.data_as(_DP), wd.ctypes.data_as(_DP), self.np),
              "debug_weights")
        return tag, w, wd


def expr_compile(text, previous=False, variables=None):
    """kmpm_expr_compile: the program of an input-script expression."""
    variables = variables or {}
    names = (C.c_char_p * max(1, len(variables)))(*[k.encode() for k in variables])
    vals = (C.c_double * max(1, len(variables)))(*[float(v) for v in variables.values()])
    e = Expr()
    check(lib().kmpm_expr_compile(text.encode(), 1 if previous else 0, len(variables), names,
                                  vals, C.byref(e)), "expr_compile(%r)" % text)
    return e


def compile_fix_values(texts, variables=None):
    """Value and previous-value programs of the three components of a
    velocity_nodes fix (None or "NULL": the component is free)."""
    val, prev = Expr3(), Expr3()
    for d, t in enumerate(texts):
        if t is None or t == "NULL":
            continue
        val[d] = expr_compile(t, False, variables)
        prev[d] = expr_compile(t, True, variables)
    return val, prev


def copy_probe(nbytes, reps=20):
    """GB/s of a streaming device copy (read + write bytes)."""
    g = C.c_double()
    check(lib().kmpm_copy_probe(int(nbytes), int(reps), C.byref(g)), "copy_probe")
    return g.value


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Surrogate:
    """Learned JC flow-stress / failure-strain networks (kmpm_surrogate)."""

    def __init__(self, mat, eps_p_max=2.0, rate_max=1e6, triax_max=2.0, samples=1 << 20,
                 iterations=4000, batch=4096, learning_rate=3e-3, seed=20261202):
        self.L = lib()
        c = SurrogateConfig()
        c.eps_p_max, c.rate_max, c.triax_max = eps_p_max, rate_max, triax_max
        c.samples, c.iterations, c.batch = samples, iterations, batch
        c.learning_rate, c.seed = learning_rate, seed
        self.cfg = c
        self._mat = mat
        h = C.c_void_p()
        check(self.L.kmpm_surrogate_create(C.byref(mat), C.byref(c), C.byref(h)),
              "kmpm_surrogate_create")
        self.h = h

    def close(self):
        if self.h:
            self.L.kmpm_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train(self):
        loss = (C.c_double * 2)()
        check(self.L.kmpm_surrogate_train(self.h, loss), "surrogate_train")
        return loss[0], loss[1]

    def tuples(self, net, n, seed):
        X = np.zeros((n, 2), np.float32)
        Y = np.zeros(n, np.float32)
        check(self.L.kmpm_surrogate_tuples(self.h, net, n, seed, X.ctypes.data_as(_FP),
                                           Y.ctypes.data_as(_FP)), "surrogate_tuples")
        return X, Y

    def tuples_from_context(self, ctx, net, stride=1):
        """Tuples at the states of a context's particles (slots 0, stride,
        ...), labelled by this surrogate's plain functions; NaN rows where
        a particle has no such state."""
        n = (ctx.np + stride - 1) // stride
        X = np.zeros((n, 2), np.float32)
        Y = np.zeros(n, np.float32)
        check(self.L.kmpm_surrogate_tuples_from_ctx(self.h, ctx.h, net, stride,
                                                    X.ctypes.data_as(_FP), Y.ctypes.data_as(_FP),
                                                    n), "surrogate_tuples_from_ctx")
        return X, Y

    def set_tuples(self, net, offset, X, Y):
        """Replace rows [offset, offset + len(Y)) of a network's training set."""
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        check(self.L.kmpm_surrogate_set_tuples(self.h, net, offset, X.ctypes.data_as(_FP),
                                               Y.ctypes.data_as(_FP), len(Y)),
              "surrogate_set_tuples")

    def eval(self, net, X):
        X = _f32(X)
        out = np.zeros(len(X), np.float32)
        check(self.L.kmpm_surrogate_eval(self.h, net, X.ctypes.data_as(_FP), len(X),
                                         out.ctypes.data_as(_FP)), "surrogate_eval")
        return out

    def gradient(self, net, X, Y):
        X, Y = _f32(X), _f32(Y)
        g = np.zeros(SURROGATE_NPARAM, np.float32)
        loss = C.c_double()
        check(self.L.kmpm_surrogate_gradient(self.h, net, X.ctypes.data_as(_FP),
                                             Y.ctypes.data_as(_FP), len(X),
                                             g.ctypes.data_as(_FP), C.byref(loss)),
              "surrogate_gradient")
        return g, loss.value

    def params(self, net):
        p = np.zeros(SURROGATE_NPARAM, np.float32)
        check(self.L.kmpm_surrogate_params(self.h, net, p.ctypes.data_as(_FP), len(p), 0),
              "surrogate_params")
        return p

    def set_params(self, net, p):
        p = _f32(p).copy()
        check(self.L.kmpm_surrogate_params(self.h, net, p.ctypes.data_as(_FP), len(p), 1),
              "surrogate_params")

    def normalization(self, net):
        m = (C.c_double * 12)()
        check(self.L.kmpm_surrogate_normalization(self.h, net, m), "surrogate_normalization")
        return {"in_scale": np.array(m[0:2]), "in_shift": np.array(m[2:4]),
                "out_scale": m[4], "out_shift": m[5], "base": np.array(m[6:12])}

    def envelope(self):
        lo, hi = (C.c_double * 4)(), (C.c_double * 4)()
        check(self.L.kmpm_surrogate_envelope(self.h, lo, hi), "surrogate_envelope")
        return np.array(lo[:]), np.array(hi[:])


class CutNet:
    """The cut network: one 18 -> 64 x5 -> 9 GELU MLP on the BF16 matrix
    cores in place of both constitutive calls (kmpm_cutnet)."""

    def __init__(self, mat, dt, cellsize, eps_p_max=0.5, rate_min=1e-3, rate_max=1e4,
                 damage_max=0.5, damage_init_max=0.9, pressure_max=3.0, increment_min=1e-6,
                 increment_max=1e-2, samples=1 << 22, iterations=3000, batch=16384,
                 judge_stride=64, learning_rate=2e-3, seed=20261203):
        self.L = lib()
        c = CutNetConfig()
        c.eps_p_max, c.rate_min, c.rate_max = eps_p_max, rate_min, rate_max
        c.damage_max, c.damage_init_max, c.pressure_max = damage_max, damage_init_max, pressure_max
        c.increment_min, c.increment_max = increment_min, increment_max
        c.dt, c.cellsize = dt, cellsize
        c.samples, c.iterations, c.batch = samples, iterations, batch
        c.judge_stride, c.learning_rate, c.seed = judge_stride, learning_rate, seed
        self.cfg = c
        self._mat = mat
        h = C.c_void_p()
        check(self.L.kmpm_cutnet_create(C.byref(mat), C.byref(c), C.byref(h)),
              "kmpm_cutnet_create")
        self.h = h

    def close(self):
        if self.h:
            self.L.kmpm_cutnet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train(self, iterations=None):
        """Returns (mean loss of the last iterations, seconds)."""
        it = self.cfg.iterations if iterations is None else iterations
        loss, secs = C.c_double(), C.c_double()
        check(self.L.kmpm_cutnet_train(self.h, it, C.byref(loss), C.byref(secs)), "cutnet_train")
        return loss.value, secs.value

    def tuples(self, n, seed):
        X = np.zeros((n, CUTNET_NI), np.float32)
        Y = np.zeros((n, CUTNET_NO), np.float32)
        check(self.L.kmpm_cutnet_tuples(self.h, n, seed, X.ctypes.data_as(_FP),
                                        Y.ctypes.data_as(_FP)), "cutnet_tuples")
        return X, Y

    def eval(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros((len(X), CUTNET_NO), np.float32)
        check(self.L.kmpm_cutnet_eval(self.h, X.ctypes.data_as(_FP), len(X),
                                      out.ctypes.data_as(_FP)), "cutnet_eval")
        return out

    def gradient(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        g = np.zeros(CUTNET_NPARAM, np.float32)
        loss = C.c_double()
        check(self.L.kmpm_cutnet_gradient(self.h, X.ctypes.data_as(_FP), Y.ctypes.data_as(_FP),
                                          len(X), g.ctypes.data_as(_FP), C.byref(loss)),
              "cutnet_gradient")
        return g, loss.value

    def params(self):
        p = np.zeros(CUTNET_NPARAM, np.float32)
        check(self.L.kmpm_cutnet_params(self.h, p.ctypes.data_as(_FP), len(p), 0), "cutnet_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32).copy()
        check(self.L.kmpm_cutnet_params(self.h, p.ctypes.data_as(_FP), len(p), 1), "cutnet_params")

    def maps(self):
        isc = np.zeros(CUTNET_NI, np.float32)
        ish = np.zeros(CUTNET_NI, np.float32)
        osc = np.zeros(CUTNET_NO, np.float32)
        sk = np.zeros((CUTNET_NO, CUTNET_NI + 1), np.float32)
        check(self.L.kmpm_cutnet_maps(self.h, isc.ctypes.data_as(_FP), ish.ctypes.data_as(_FP),
                                      osc.ctypes.data_as(_FP), sk.ctypes.data_as(_FP), 0),
              "cutnet_maps")
        return {"in_scale": isc, "in_shift": ish, "out_scale": osc, "skip": sk}

    def envelope(self):
        e = (C.c_double * 7)()
        check(self.L.kmpm_cutnet_envelope(self.h, e), "cutnet_envelope")
        return dict(zip(("eps_p", "rate", "damage", "damage_init", "pressure", "q_trial",
                         "increment"), e[:]))

#!/usr/bin/env python3
"""Throughput of the explicit MPM step on one MI355X.

Workload: 3D ULMPM tensile bar with Johnson-Cook strength, shock EOS and JC
damage (steel), cubic B-splines, USL, 2x2x2 particles per cell, cell 0.06 mm:
4.27M particles.  A step is one complete USL step over all particles
(sort, P2G, grid update with the end-velocity fixes, fused G2P + F + stress +
CFL, dt update), synthetic lattice inputs, state resident in HBM.

Paths (--paths, the last one is the headline `value`):
  plain-fp64  FP64 throughout, parity with the reference to 1e-10;
  mixed       FP32 state and transfers, FP64 position and deformation
              gradient, 64-bit fixed-point node sums (deterministic).
Each path runs the same steps from the same inputs; the headline's
displacement and stress are compared with the plain FP64 run's.

Prints ONE JSON line.  Progress goes to stderr.
"""
import argparse
import ctypes as C
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s: mixed+surrogate vs plain FP64 GPU at ≤1e-3 rel L2 vs FP64"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle of each timed phase, 3D, 8 particles per
# cell (node data divided by 8).
#
# plain FP64:
#   g2p  fused USL G2P/F/stress: reads x3 v3 F9 sigma6 eps_p epsdot_p
#        damage_init damage mass vol0 (27 doubles = 216 B), writes x v F sigma
#        eps_p epsdot_p damage_init damage (25 doubles = 200 B), reads v_update
#        and v of the grid (6 doubles per node -> 6 B)                 = 422 B
#   p2g  reads x3 v3 mass vol sigma6 (14 doubles = 112 B), writes node mass,
#        momentum and force (7 doubles per node -> 7 B)                 = 119 B
#   grid reads node mass, momentum, force (7 doubles), writes mass, v,
#        v_update (7 doubles): 14 doubles per node -> 14 B               = 14 B
#   sort reads the cell key and particle id, writes the permutation
#        (3 x 4 B)                                                      = 12 B
# mixed (x and F in FP64, the rest FP32, int64 fixed-point node sums):
#   g2p  reads x 24, F 72, v 12, sigma 24, 4 scalars 16, mass + vol0 8
#        (156 B), writes x v F sigma and the 4 scalars (148 B), reads
#        v_update and v (6 floats per node -> 3 B)                      = 307 B
#   p2g  reads x 24, v 12, mass 4, vol 4, sigma 24 (68 B), writes 7 int64
#        node sums (7 B)                                                = 75 B
#   grid reads 7 int64, writes 7 floats per node                        = 10.5 B
#   sort as plain                                                       = 12 B
ALG_BYTES_PER_PARTICLE = {
    "plain-fp64": {"g2p": 422, "p2g": 119, "grid": 14, "sort": 12},
    "mixed": {"g2p": 307, "p2g": 75, "grid": 10.5, "sort": 12},
}
KERNEL_NAMES = {"g2p": "k_g2p_usl", "p2g": "k_p2g", "grid": "k_grid_update", "sort": "sort"}
# Whole step, same counting (state once in, once out, grid once).
STEP_ALG_BYTES_PER_PARTICLE = {"plain-fp64": 442, "mixed": 324}
DTYPE = {"plain-fp64": "f64", "mixed": "f32+f64(x,F)+i64(grid sums)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def device_sync():
    hip = C.CDLL("libamdhip64.so")
    if hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def read_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (profiles/pmc_<kernel>_<workload>.txt), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_%s_%s.txt" % (kernel, workload))
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def run_path(prob, path, steps, warmup):
    """Times `steps` steps of one path after `warmup` steps; returns the
    timing record and the downloaded particle state."""
    from karamelo_amd import native as N
    from karamelo_amd import problems as P
    cfg = N.Config.from_buffer_copy(prob.cfg)
    cfg.precision = N.PRECISION_MIXED if path == "mixed" else N.PRECISION_FP64
    ctx = P.context_for(dataclasses.replace(prob, cfg=cfg))
    log("[%s] warm-up %d steps ..." % (path, warmup))
    ctx.step(warmup)
    st, tag = ctx.synchronize()
    if st != 0:
        raise RuntimeError("device error %d (particle %d) during warm-up" % (st, tag))
    device_sync()
    log("[%s] timed %d steps ..." % (path, steps))
    ctx.timing_enable(True)
    device_sync()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.synchronize()
    device_sync()
    t1 = time.perf_counter()
    phases = ctx.timing_read()
    ctx.timing_enable(False)
    st, tag = ctx.synchronize()
    if st != 0:
        raise RuntimeError("device error %d (particle %d) in the timed region" % (st, tag))
    state = ctx.download()
    ctx.close()
    secs = t1 - t0
    return {"secs": secs, "phases": phases}, state


def roofline_of(path, phases, npart, workload):
    dom = max(phases, key=lambda k: phases[k][0])
    ms_tot, launches = phases[dom]
    avg_s = ms_tot / max(launches, 1) * 1e-3
    alg = ALG_BYTES_PER_PARTICLE[path].get(dom, 0) * npart or None
    kname = KERNEL_NAMES.get(dom, dom)
    achieved = alg / avg_s / 1e9 if alg else None
    wl = workload if path == "plain-fp64" else workload + "_mixed"
    return {"bound": "hbm", "kernel": kname,
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": read_traffic(kname, wl),
            "alg_bytes_per_launch": alg,
            "avg_launch_ms": avg_s * 1e3}


def rel_l2(a, b):
    import numpy as np
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0))


def cpu_baseline(steps=2):
    """The oracle (a single-threaded FP64 port of the reference's loops,
    -O3 build) on a bounded sample of the same workload: the same bar
    cross-section, cell size, particles per cell, materials and fixes, but a
    2 mm long section (about 10% of the particles), one warm-up step then
    `steps` timed steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from karamelo_amd import problems as P
    prob = P.tensile_bar(length=0.002)
    o = O.oracle_for(prob, variant="liboracle_perf.so")
    o.step(1)
    t0 = time.perf_counter()
    st = o.step(steps)
    t1 = time.perf_counter()
    if st != 0:
        raise RuntimeError("oracle step failed: %d" % st)
    value = prob.np * steps / (t1 - t0)
    o.close()
    return {"value": value, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "oracle -O3, 1 thread: tensile bar section 2 mm long (%d particles, "
                      "h=0.06 mm, 2x2x2 ppc, JC+shock+damage, cubic, USL), 1 warm-up + %d "
                      "timed steps (%.1f s)" % (prob.np, steps, t1 - t0)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--h", type=float, default=6e-5, help="cell size (m)")
    ap.add_argument("--paths", default="plain-fp64,mixed",
                    help="comma-separated paths to time; the last one is the headline")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from karamelo_amd import problems as P

    paths = args.paths.split(",")
    log("building workload ...")
    prob = P.tensile_bar(h=args.h)
    npart = prob.np
    workload = "tensile_bar_jc_4.27M" if abs(args.h - 6e-5) < 1e-12 else "tensile_bar_jc_h%g" % args.h
    log("particles: %d" % npart)
    x0 = prob.particles["x"].copy()
    results, states = {}, {}
    for path in paths:
        results[path], states[path] = run_path(prob, path, args.steps, args.warmup)
    prob.particles = None

    head = paths[-1]
    r = results[head]
    value = npart * args.steps / r["secs"]
    ms_step = r["secs"] * 1e3 / args.steps
    per_path = {}
    for path in paths:
        rp = results[path]
        per_path[path] = {
            "value": npart * args.steps / rp["secs"],
            "ms_per_step": rp["secs"] * 1e3 / args.steps,
            "phases_ms_per_step": {k: v[0] / max(v[1], 1) for k, v in rp["phases"].items() if v[1]},
            "step_alg_gbs": STEP_ALG_BYTES_PER_PARTICLE[path] * npart / (rp["secs"] / args.steps) / 1e9,
            "roofline": roofline_of(path, rp["phases"], npart, workload),
        }
    accuracy = None
    if "plain-fp64" in states and head != "plain-fp64":
        ref, got = states["plain-fp64"], states[head]
        accuracy = {"vs": "plain FP64 GPU run, same inputs and steps",
                    "displacement_rel_l2": rel_l2(got["x"] - x0, ref["x"] - x0),
                    "stress_rel_l2": rel_l2(got["sigma"], ref["sigma"]),
                    "velocity_rel_l2": rel_l2(got["v"], ref["v"])}
    states = None

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        try:
            cpu = cpu_baseline(args.cpu_steps)
        except Exception as e:  # reported, not fatal
            log("cpu baseline failed: %r" % (e,))
    line = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": DTYPE[head],
        "data": "synthetic",
        "config": {"workload": workload, "particles": npart, "method": "ulmpm",
                   "scheme": "usl", "shape": "cubic-spline", "cellsize_m": args.h,
                   "materials": "steel JC strength + shock EOS + JC damage",
                   "path": head, "parallelism": "single-gpu"},
        "roofline": per_path[head]["roofline"],
        "speedup_vs_plain_fp64": (value / per_path["plain-fp64"]["value"]
                                  if "plain-fp64" in per_path else None),
        "accuracy_vs_fp64": accuracy,
        "paths": per_path,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

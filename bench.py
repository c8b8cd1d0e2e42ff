#!/usr/bin/env python3
"""Throughput of the explicit MPM step on one MI355X.

Workload: 3D ULMPM tensile bar with Johnson-Cook strength, shock EOS and JC
damage (steel), cubic B-splines, USL, 2x2x2 particles per cell, cell 0.06 mm:
4.27M particles.  A step is one complete USL step over all particles
(sort, P2G, grid update with the end-velocity fixes, fused G2P + F + stress +
CFL, dt update), synthetic lattice inputs, state resident in HBM.

Prints ONE JSON line.  Progress goes to stderr.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s: mixed+surrogate vs plain FP64 GPU at ≤1e-3 rel L2 vs FP64"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle of each timed phase (FP64, 3D, 8 particles
# per cell; node data divided by 8):
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
ALG_BYTES_PER_PARTICLE = {"g2p": 422, "p2g": 119, "grid": 14, "sort": 12}
KERNEL_NAMES = {"g2p": "k_g2p_usl", "p2g": "k_p2g", "grid": "k_grid_update", "sort": "sort"}
# Whole step, same counting (state once in, once out, grid once): 442 B.
STEP_ALG_BYTES_PER_PARTICLE = 442


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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from karamelo_amd import native as N
    from karamelo_amd import problems as P

    log("building workload ...")
    prob = P.tensile_bar(h=args.h)
    npart = prob.np
    workload = "tensile_bar_jc_4.27M" if abs(args.h - 6e-5) < 1e-12 else "tensile_bar_jc_h%g" % args.h
    log("particles: %d" % npart)
    ctx = P.context_for(prob)
    prob.particles = None  # host copy no longer needed
    log("warm-up %d steps ..." % args.warmup)
    ctx.step(args.warmup)
    st, tag = ctx.synchronize()
    if st != 0:
        raise RuntimeError("device error %d (particle %d) during warm-up" % (st, tag))
    device_sync()
    log("timed %d steps ..." % args.steps)
    ctx.timing_enable(True)
    device_sync()
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.synchronize()
    device_sync()
    t1 = time.perf_counter()
    phases = ctx.timing_read()
    ctx.timing_enable(False)
    st, tag = ctx.synchronize()
    if st != 0:
        raise RuntimeError("device error %d (particle %d) in the timed region" % (st, tag))
    secs = t1 - t0
    value = npart * args.steps / secs
    ms_step = secs * 1e3 / args.steps

    # dominant kernel: the phase with the largest total time
    dom = max(phases, key=lambda k: phases[k][0])
    ms_tot, launches = phases[dom]
    avg_s = ms_tot / max(launches, 1) * 1e-3
    alg = ALG_BYTES_PER_PARTICLE.get(dom, 0) * npart or None
    kname = KERNEL_NAMES.get(dom, dom)
    achieved = alg / avg_s / 1e9 if alg else None
    traffic = read_traffic(kname, workload)
    roofline = {"bound": "hbm", "kernel": kname,
                "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                "traffic": traffic,
                "alg_bytes_per_launch": alg,
                "avg_launch_ms": avg_s * 1e3}
    step_gbs = STEP_ALG_BYTES_PER_PARTICLE * npart / (secs / args.steps) / 1e9
    ctx.close()

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
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "particles": npart, "method": "ulmpm",
                   "scheme": "usl", "shape": "cubic-spline", "cellsize_m": args.h,
                   "materials": "steel JC strength + shock EOS + JC damage",
                   "path": "plain-fp64", "parallelism": "single-gpu"},
        "roofline": roofline,
        "step_alg_gbs": step_gbs,
        "phases_ms_per_step": {k: v[0] / max(v[1], 1) for k, v in phases.items() if v[1]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

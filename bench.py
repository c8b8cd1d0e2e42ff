#!/usr/bin/env python3
"""Throughput of the explicit MPM step on one MI355X.

Workload: 3D ULMPM tensile bar with Johnson-Cook strength, shock EOS and JC
damage (steel), cubic B-splines, USL, 2x2x2 particles per cell, cell 0.06 mm:
4.27M particles.  A step is one complete USL step over all particles
(sort, P2G, grid update with the end-velocity fixes, fused G2P + F + stress +
CFL, dt update), synthetic lattice inputs, state resident in HBM.  Every
path first runs --preload-steps untimed steps from rest (1000 steps, ~9 us of
pulling), which strain the whole bar past yield, then --warmup untimed steps:
the timed steps run the plastic branch of the constitutive update whatever
--warmup is, and the headline fails unless the networks evaluated about
every particle in every timed step.

Paths (--paths, the last one is the headline `value`):
  plain-fp64       FP64 throughout, parity with the reference to 1e-10;
  mixed            FP32 state and transfers, FP64 position and deformation
                   gradient, 64-bit fixed-point node sums (deterministic);
  mixed+cutnet     mixed, with both constitutive calls (JC strength and JC
                   damage) replaced by one 18 -> 64 x5 -> 9 GELU network on
                   the BF16 matrix cores (trained on the device before the
                   timed region on 50M tuples of the plain functions; the
                   plain functions judge a sample of its updates);
  mixed+surrogate  mixed, with the JC flow stress and failure strain from
                   the learned BF16 networks (trained on the device before
                   the timed region, on tuples of the plain functions).
Each path runs the same steps from the same inputs; the headline's
displacement and stress are compared with the plain FP64 run's.  The timed
steps are replays of a captured HIP graph of --graph-steps steps; the
per-phase launch times (phases, roofline) come from as many eager steps
again, with HIP events on the library's stream around each phase.

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
# cell (node data divided by 8): every state field a kernel must read or
# write once (the particle state is x, v, F, sigma, strain_el, eps_p,
# epsdot_p, damage_init, damage, mass, vol0, vol and the particle id and
# material index), node data once.  The permutation (the sorted order a
# kernel reads its particles in) and the re-reads of node blocks at tile
# borders are overhead, not counted.
#
# plain FP64 (34 doubles of state):
#   g2p  fused USL G2P + F + stress: reads x3 v3 F9 sigma6 strain_el6 eps_p
#        epsdot_p damage_init damage mass vol0 (33 doubles) and pid, mat
#        (272 B); writes the same plus vol and the next cell key (284 B);
#        reads v_update and v of the grid (6 doubles per node -> 6 B)  = 562 B
#   p2g  reads x3 v3 mass vol sigma6 (14 doubles = 112 B), writes node mass,
#        momentum and force (7 doubles per node -> 7 B)                 = 119 B
#   grid reads node mass, momentum, force (7 doubles), writes mass, v,
#        v_update (7 doubles): 14 doubles per node -> 14 B               = 14 B
#   sort reads the cell key and particle id, writes the permutation
#        (3 x 4 B)                                                      = 12 B
# mixed (x and F in FP64, the other 22 fields FP32, int64 fixed-point node
# sums):
#   g2p  reads x 24, F 72, v 12, sigma 24, strain_el 24, 4 scalars 16, mass
#        + vol0 8, pid + mat 8 (188 B); writes those, vol and the key
#        (196 B); reads v_update and v (6 floats per node -> 3 B)      = 387 B
#   p2g  reads x 24, v 12, mass 4, vol 4, sigma 24 (68 B), writes 7 int64
#        node sums (7 B)                                                = 75 B
#   grid reads 7 int64, writes 7 floats per node                        = 10.5 B
#   sort as plain                                                       = 12 B
# mixed + a learned constitutive update (the surrogate's networks or the
# cut network): the G2P kernel leaves the stress state to the constitutive
# kernel and stages its inputs (D, pressure, wave speed: 8 floats)
#   g2p  reads x 24, F 72, v 12, mass, vol0, damage, pid, mat 20 (128 B);
#        writes x, F, v, mass, vol0, vol, key, pid, mat (132 B) and the
#        staged 32 B; grid 3 B                                          = 295 B
#   learned  reads sigma, strain_el, eps_p, epsdot_p, damage_init, damage
#        (64 B) and the staged 32 B, writes the 64 B                   = 160 B
ALG_BYTES_PER_PARTICLE = {
    "plain-fp64": {"g2p": 562, "p2g": 119, "grid": 14, "sort": 12},
    "mixed": {"g2p": 387, "p2g": 75, "grid": 10.5, "sort": 12},
    "mixed+cutnet": {"g2p": 295, "p2g": 75, "grid": 10.5, "sort": 12, "learned": 160},
}
ALG_BYTES_PER_PARTICLE["mixed+surrogate"] = ALG_BYTES_PER_PARTICLE["mixed+cutnet"]
# Matrix-core work of one cut network evaluation, algorithmic
# (2 (18 x 64 + 4 x 64 x 64 + 64 x 9) flops), and per training sample
# (forward, the input-gradient products of layers 2..6, the weight
# gradients of all six layers).
CUTNET_FLOPS_PER_EVAL = 2 * (18 * 64 + 4 * 64 * 64 + 64 * 9)
CUTNET_FLOPS_PER_TRAIN_SAMPLE = CUTNET_FLOPS_PER_EVAL + 2 * (4 * 64 * 64 + 64 * 9) + \
    CUTNET_FLOPS_PER_EVAL
# Matrix-core work of one network evaluation: the 32x32 hidden layer as two
# 32x32x16 BF16 products per 32 samples = 2 * 32 * 32 flops per sample.
MFMA_FLOPS_PER_EVAL = 2 * 32 * 32
MFMA_BF16_PEAK_TFLOPS = 2500.0
KERNEL_NAMES = {"g2p": "k_g2p_usl", "p2g": "k_p2g", "grid": "k_grid_update", "sort": "sort",
                "learned": "k_stress_learned"}
# Whole step, same counting: the state read once and written once, node
# data once per pass (FP64: 272 + 284 B of state, 7 + 14 + 6 doubles of node
# traffic per node; mixed: 188 + 196 B, 7 + 7 int64 + 7 + 6 floats per node).
STEP_ALG_BYTES_PER_PARTICLE = {"plain-fp64": 583, "mixed": 405, "mixed+surrogate": 405,
                               "mixed+cutnet": 405}
DTYPE = {"plain-fp64": "f64", "mixed": "f32+f64(x,F)+i64(grid sums)",
         "mixed+surrogate": "f32+f64(x,F)+i64(grid sums)+bf16(mfma, f32 acc)",
         "mixed+cutnet": "f32+f64(x,F)+i64(grid sums)+bf16(mfma, f32 acc)"}


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


def train_surrogate(mat, iters=6000):
    """Trains the JC networks on the device; returns the surrogate and a
    record of the training (not part of the timed step)."""
    from karamelo_amd import native as N
    t0 = time.perf_counter()
    s = N.Surrogate(mat, samples=1 << 18, iterations=iters, batch=4096)
    loss = s.train()
    device_sync()
    t1 = time.perf_counter()
    # accuracy on fresh tuples of the plain functions
    err = {}
    import numpy as np
    for net, name in ((0, "flow_stress"), (1, "failure_strain")):
        X, Y = s.tuples(net, 20000, seed=99)
        out = s.eval(net, X)
        rel = np.abs(out - Y) / np.abs(Y) if net == 0 else np.abs(np.exp(out - Y) - 1)
        err[name] = {"max_rel": float(rel.max()), "rms_rel": float(np.sqrt(np.mean(rel ** 2)))}
    batch = 4096
    # training MFMA work: forward (1 product) + backward (2 products) per sample
    flops = 2 * iters * batch * 3 * MFMA_FLOPS_PER_EVAL
    rec = {"seconds": t1 - t0, "final_loss": list(loss), "validation": err,
           "iterations_per_net": iters, "batch": batch,
           "mfma_tflops": flops / (t1 - t0) / 1e12}
    return s, rec


def train_cutnet(mat, dt, h, iterations, samples):
    """Trains the cut network on the device; returns it and a record of the
    training (not part of the timed step)."""
    from karamelo_amd import native as N
    t0 = time.perf_counter()
    net = N.CutNet(mat, dt=dt, cellsize=h, samples=samples, iterations=iterations,
                   increment_min=1e-5, increment_max=0.05, batch=16384)
    device_sync()
    t1 = time.perf_counter()
    loss, secs = net.train()
    import numpy as np
    X, Y = net.tuples(20000, seed=99)
    out = net.eval(X)
    plastic = Y[:, 6] > 0
    s_err = float(np.linalg.norm(out[plastic, :6] - Y[plastic, :6]) /
                  np.linalg.norm(Y[plastic, :6]))
    rec = {"tuples": samples, "tuple_and_fit_seconds": t1 - t0, "iterations": iterations,
           "batch": 16384, "train_seconds": secs, "final_loss": loss,
           "validation": {"deviatoric_stress_rel_l2_plastic": s_err,
                          "branch_agreement": float(np.mean((out[:, 6] > 0) == plastic))},
           "mfma_tflops": iterations * 16384 * CUTNET_FLOPS_PER_TRAIN_SAMPLE / secs / 1e12}
    log("[cutnet] trained: %s" % rec)
    return net, rec


def run_path(prob, path, steps, warmup, surrogate=None, autotune_steps=5, graph_steps=10,
             preload=1000, cutnet=None):
    """Times `steps` steps of one path after `preload` + `warmup` steps;
    returns the timing record and the downloaded particle state."""
    from karamelo_amd import native as N
    from karamelo_amd import problems as P
    cfg = N.Config.from_buffer_copy(prob.cfg)
    cfg.precision = N.PRECISION_FP64 if path == "plain-fp64" else N.PRECISION_MIXED
    ctx = P.context_for(dataclasses.replace(prob, cfg=cfg))
    if path == "mixed+surrogate":
        ctx.set_surrogate(surrogate)
    if path == "mixed+cutnet":
        ctx.set_cutnet(cutnet)
    log("[%s] pre-load %d steps (pulling the bar past yield) ..." % (path, preload))
    ctx.step(preload)
    log("[%s] warm-up %d steps ..." % (path, warmup))
    ctx.step(warmup)
    st, tag = ctx.synchronize()
    if st != 0:
        raise RuntimeError("device error %d (particle %d) during warm-up" % (st, tag))
    # the autotuner times every tile variant from the warmed-up state, which
    # it restores, and keeps the fastest for the timed steps
    if autotune_steps > 0:
        variant, tune_ms = ctx.autotune(autotune_steps)
    else:
        variant, tune_ms = 0, [0.0] * 3
    log("[%s] autotune: tile %s, ms/step per variant %s" % (path, ctx.tile_shape(), tune_ms))
    # the steps are replays of a graph of `graph_steps` captured steps; the
    # first call captures it (and replays it once), outside the timed region
    ctx.step_graph(graph_steps, graph_steps)
    device_sync()
    ctx.cutnet_stats()             # resets the counters: the timed steps only
    log("[%s] timed %d steps (graph replays of %d) ..." % (path, steps, graph_steps))
    t0 = time.perf_counter()
    ctx.step_graph(steps, graph_steps)
    ctx.synchronize()
    device_sync()
    t1 = time.perf_counter()
    st, tag = ctx.synchronize()
    if st != 0:
        raise RuntimeError("device error %d (particle %d) in the timed region" % (st, tag))
    evals, fallbacks, judged, judged_err = ctx.cutnet_stats()
    # per-phase launch times: as many eager steps again, with HIP events
    # recorded on the library's stream around each phase
    log("[%s] %d eager steps with phase events ..." % (path, steps))
    ctx.timing_enable(True)
    ctx.step(steps)
    ctx.synchronize()
    phases = ctx.timing_read()
    ctx.timing_enable(False)
    st, tag = ctx.synchronize()
    if st != 0:
        raise RuntimeError("device error %d (particle %d) in the phase-timed steps" % (st, tag))
    state = ctx.download()
    tile = ctx.tile_shape()
    ctx.close()
    secs = t1 - t0
    return {"secs": secs, "phases": phases, "surrogate_evals": evals,
            "surrogate_fallbacks": fallbacks, "judged": judged, "judged_stress_rel_l2": judged_err,
            "autotune": {"tile": list(tile), "variant": variant,
                         "ms_per_step_per_variant": [round(float(v), 4) for v in tune_ms]}}, state


def run_batched(steps, warmup, graph_steps=10, n_inst=256):
    """The batched workload: n_inst independent 2D plane-strain Taylor
    impacts (particle counts log-uniform in 10k-200k, JC A, B, n, C within
    +-20% of copper or steel, v0 uniform in 50-300 m/s), mixed precision, on
    one grid, each with its own dt; the timed steps are replays of a
    captured graph of `graph_steps` steps."""
    from karamelo_amd import problems as P
    log("[batched] building %d instances ..." % n_inst)
    batch = P.taylor_plates(n_inst=n_inst, precision="mixed")
    npart = batch.problem.np
    ctx = P.context_for_batch(batch)
    batch.problem.particles = None
    log("[batched] %d particles; warm-up %d steps ..." % (npart, warmup))
    ctx.step(warmup)
    st, tag = ctx.synchronize()
    if st != 0:
        raise RuntimeError("batched: device error %d (particle %d) during warm-up" % (st, tag))
    variant, tune_ms = ctx.autotune(5)
    log("[batched] autotune: tile %s, ms/step per variant %s" % (ctx.tile_shape(), tune_ms))
    ctx.timing_enable(True)
    ctx.step(graph_steps)
    phases = ctx.timing_read()
    ctx.timing_enable(False)
    ctx.step_graph(graph_steps, graph_steps)       # capture and a first replay
    device_sync()
    ctx.synchronize()
    log("[batched] timed %d steps (graph replays of %d) ..." % (steps, graph_steps))
    t0 = time.perf_counter()
    ctx.step_graph(steps, graph_steps)
    ctx.synchronize()
    device_sync()
    t1 = time.perf_counter()
    st, tag = ctx.synchronize()
    if st != 0:
        raise RuntimeError("batched: device error %d (particle %d)" % (st, tag))
    dt, tm = ctx.instance_time()
    tile = ctx.tile_shape()
    streamed = stream_tuples(ctx)
    ctx.close()
    secs = t1 - t0
    return {"workload": "taylor_plates_2d_%dx" % n_inst, "instances": n_inst, "particles": npart,
            "particles_per_instance": {"min": min(q["np"] for q in batch.params),
                                       "max": max(q["np"] for q in batch.params)},
            "precision": "mixed", "steps": steps, "warmup": warmup,
            "graph_steps_per_launch": graph_steps,
            "value": npart * steps / secs, "unit": "particle-steps/s",
            "ms_per_step": secs * 1e3 / steps,
            "phases_ms_per_step": {k: v[0] / max(v[1], 1) for k, v in phases.items() if v[1]},
            "instance_dt_range_s": [float(dt.min()), float(dt.max())],
            "autotune": {"tile": list(tile), "variant": variant,
                         "ms_per_step_per_variant": [round(float(v), 4) for v in tune_ms]},
            "streamed_tuples": streamed}


def stream_tuples(ctx, stride=16, rows=1 << 15, iterations=6000):
    """The batched run's states as surrogate training data (outside the
    timed region): every `stride`-th particle's state gives a tuple of each
    network of a steel JC surrogate, labelled by the plain functions; the
    tuples inside the networks' envelope replace rows of its training set,
    and training continues on them."""
    import numpy as np
    from karamelo_amd import native as N
    from karamelo_amd import problems as P
    t0 = time.perf_counter()
    s = N.Surrogate(P.steel_jc(), samples=1 << 16, iterations=iterations, batch=4096)
    lo, hi = s.envelope()
    out = {"stride": stride, "rows": {}, "in_envelope": {}}
    for net in (0, 1):
        X, Y = s.tuples_from_context(ctx, net, stride)
        ok = np.isfinite(Y)
        inside = ok & np.all((X >= lo[2 * net:2 * net + 2]) & (X <= hi[2 * net:2 * net + 2]), axis=1)
        k = int(min(inside.sum(), rows))
        if k:
            s.set_tuples(net, 0, X[inside][:k], Y[inside][:k])
        out["rows"]["net%d" % net] = int(ok.sum())
        out["in_envelope"]["net%d" % net] = int(inside.sum())
    t1 = time.perf_counter()
    loss = s.train()
    t2 = time.perf_counter()
    s.close()
    out.update({"stream_seconds": t1 - t0, "retrain_iterations": iterations,
                "retrain_seconds": t2 - t1, "final_loss": [float(v) for v in loss]})
    log("[batched] streamed tuples: %s" % out)
    return out


def roofline_of(path, phases, npart, workload):
    dom = max(phases, key=lambda k: phases[k][0])
    ms_tot, launches = phases[dom]
    avg_s = ms_tot / max(launches, 1) * 1e-3
    alg = ALG_BYTES_PER_PARTICLE[path].get(dom, 0) * npart or None
    kname = KERNEL_NAMES.get(dom, dom)
    achieved = alg / avg_s / 1e9 if alg else None
    wl = workload + {"plain-fp64": "", "mixed": "_mixed", "mixed+surrogate": "_mixed_surrogate",
                     "mixed+cutnet": "_mixed_cutnet"}[path]
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


def host_cpu():
    """CPU model, physical cores of the host, and the CPUs this process may
    run on (the box's share)."""
    model, cores = None, set()
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count()
    return {"model": model, "physical_cores": len(cores) or None, "cpus_available": share}


def cpu_baseline(steps=20):
    """The oracle (an FP64 restatement of the reference's loops: per-step
    neighbour lists, node-major gather P2G) built -O3 with OpenMP over its
    particle and node loops, on every CPU this process may use, on a bounded
    sample of the same workload: the same bar cross-section, cell size,
    particles per cell, materials and fixes, but a 2 mm long section (about
    10% of the particles), one warm-up step then `steps` timed steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from karamelo_amd import problems as P
    cpu = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or cpu["cpus_available"]
    os.environ["OMP_NUM_THREADS"] = str(threads)
    prob = P.tensile_bar(length=0.002)
    o = O.oracle_for(prob, variant="liboracle_perf.so")
    try:
        threads = C.CDLL("libgomp.so.1").omp_get_max_threads()
    except OSError:
        pass
    o.step(1)
    t0 = time.perf_counter()
    st = o.step(steps)
    t1 = time.perf_counter()
    if st != 0:
        raise RuntimeError("oracle step failed: %d" % st)
    value = prob.np * steps / (t1 - t0)
    o.close()
    return {"value": value, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "cpu": cpu,
            "sample": "oracle -O3 + OpenMP, %d threads on %s (%s physical cores on the host, "
                      "%s CPUs available to this process): tensile bar section 2 mm long (%d "
                      "particles, h=0.06 mm, 2x2x2 ppc, JC+shock+damage, cubic, USL), 1 warm-up "
                      "+ %d timed steps (%.1f s)" % (threads, cpu["model"], cpu["physical_cores"],
                                                    cpu["cpus_available"], prob.np, steps,
                                                    t1 - t0)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10,
                    help="untimed steps after the pre-load")
    ap.add_argument("--preload-steps", type=int, default=1000,
                    help="untimed steps from rest before the warm-up; 1000 strain the whole "
                         "bar past yield, so the timed steps are plastic")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--h", type=float, default=6e-5, help="cell size (m)")
    ap.add_argument("--paths", default="plain-fp64,mixed,mixed+cutnet,mixed+surrogate",
                    help="comma-separated paths to time; the last one is the headline")
    ap.add_argument("--autotune-steps", type=int, default=5,
                    help="trial steps per tile variant (0: keep the default tiles)")
    ap.add_argument("--prestrain", type=float, default=0.0,
                    help="initial effective plastic strain of every particle (counter runs: "
                         "the surrogate is in use from the first step without a warm-up)")
    ap.add_argument("--train-iterations", type=int, default=6000)
    ap.add_argument("--cutnet-iterations", type=int, default=20000)
    ap.add_argument("--cutnet-tuples", type=int, default=50_000_000,
                    help="training tuples of the cut network (device-generated)")
    ap.add_argument("--graph-steps", type=int, default=10,
                    help="steps per captured graph in the timed region")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-instances workload")
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--batched-warmup", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the step has no multi-GPU path (a run "
              "would shard the tiles across ranks); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from karamelo_amd import problems as P

    paths = args.paths.split(",")
    log("building workload ...")
    prob = P.tensile_bar(h=args.h)
    if args.prestrain:
        prob.particles["eff_plastic_strain"][:] = args.prestrain
    npart = prob.np
    workload = "tensile_bar_jc_4.27M" if abs(args.h - 6e-5) < 1e-12 else "tensile_bar_jc_h%g" % args.h
    log("particles: %d" % npart)
    x0 = prob.particles["x"].copy()
    results, states = {}, {}
    surrogate, training = None, None
    if "mixed+surrogate" in paths:
        log("training the surrogate ...")
        surrogate, training = train_surrogate(prob.mats[0], args.train_iterations)
    cutnet, cut_training = None, None
    if "mixed+cutnet" in paths:
        m = prob.mats[0]
        # the CFL time step of the bar at rest (ulmpm.cpp:525-551, solid.cpp:1376-1385)
        dt = prob.cfg.dt_factor * args.h / float(np.sqrt((m.K + 4.0 / 3.0 * m.G) / m.rho0))
        log("training the cut network (dt %.3g s) ..." % dt)
        cutnet, cut_training = train_cutnet(m, dt, args.h, args.cutnet_iterations,
                                            args.cutnet_tuples)
    for path in paths:
        results[path], states[path] = run_path(prob, path, args.steps, args.warmup, surrogate,
                                               autotune_steps=args.autotune_steps,
                                               graph_steps=args.graph_steps,
                                               preload=args.preload_steps, cutnet=cutnet)
    if cutnet is not None:
        cutnet.close()
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
            "autotune": rp["autotune"],
        }
        if path == "mixed+cutnet":
            cut_ms, cut_n = rp["phases"]["learned"]
            ev = rp["surrogate_evals"]
            if ev < 0.5 * args.steps * npart:
                raise RuntimeError("mixed+cutnet: %d network evaluations in %d timed steps of "
                                   "%d particles" % (ev, args.steps, npart))
            mf = ev * CUTNET_FLOPS_PER_EVAL / (cut_ms * 1e-3) / 1e12
            per_path[path]["cutnet"] = {
                "evaluations": ev, "fallbacks": rp["surrogate_fallbacks"],
                "fallback_fraction": rp["surrogate_fallbacks"] / max(1, args.steps * npart),
                "judged_updates": rp["judged"],
                "judged_deviatoric_stress_rel_l2": rp["judged_stress_rel_l2"],
                "mfma_roofline": {
                    "bound": "mfma", "kernel": "k_stress_cutnet",
                    "flops_per_eval": CUTNET_FLOPS_PER_EVAL, "achieved": mf,
                    "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": mf / MFMA_BF16_PEAK_TFLOPS,
                    "avg_launch_ms": cut_ms / max(cut_n, 1),
                    "note": "algorithmic network flops over the cut kernel's launch time"},
                "training": cut_training}
        if path == "mixed+surrogate":
            # surrogate evaluations happen in the constitutive kernel after G2P
            g2p_ms, g2p_n = rp["phases"]["learned"]
            ev = rp["surrogate_evals"]
            # a surrogate number is only a surrogate number if the networks ran
            if ev < 0.5 * args.steps * npart:
                raise RuntimeError("mixed+surrogate: %d network evaluations in %d timed steps of "
                                   "%d particles; the timed steps are not plastic (raise "
                                   "--preload-steps)" % (ev, args.steps, npart))
            mf = ev * MFMA_FLOPS_PER_EVAL / (g2p_ms * 1e-3) / 1e12
            per_path[path]["surrogate"] = {
                "evaluations": ev, "fallbacks": rp["surrogate_fallbacks"],
                "fallback_fraction": rp["surrogate_fallbacks"] / max(1, args.steps * npart),
                "mfma_roofline": {
                    "bound": "mfma", "flops_per_eval": MFMA_FLOPS_PER_EVAL, "achieved": mf,
                    "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": mf / MFMA_BF16_PEAK_TFLOPS, "kernel": "k_stress_learned",
                    "avg_launch_ms": g2p_ms / max(g2p_n, 1),
                    "note": "network flops over the constitutive kernel's launch time"},
                "training": training}
    accuracy = None
    if "plain-fp64" in states and head != "plain-fp64":
        ref = states["plain-fp64"]
        accuracy = {"vs": "plain FP64 GPU run, same inputs and steps"}
        for path in paths:
            if path == "plain-fp64":
                continue
            got = states[path]
            accuracy[path] = {"displacement_rel_l2": rel_l2(got["x"] - x0, ref["x"] - x0),
                              "stress_rel_l2": rel_l2(got["sigma"], ref["sigma"]),
                              "velocity_rel_l2": rel_l2(got["v"], ref["v"])}
    states = None

    batched = None
    if not args.no_batched:
        batched = run_batched(args.batched_steps, args.batched_warmup)

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
        "batched": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

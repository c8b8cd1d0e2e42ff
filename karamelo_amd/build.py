"""Build libkmpm.so (gfx950 HIP kernels + C ABI) in-tree.

Usage: python -m karamelo_amd.build   (or karamelo_amd.build.build())
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libkmpm.so")
BUILD_DIR = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

# -ffp-contract=off keeps the shape-function and cell-index roundings equal to
# the reference expression order; hot loops use explicit fma().
# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]

HIP_SOURCES = ["kmpm_api.hip", "kmpm_kernels.hip", "kmpm_surrogate.hip", "kmpm_reforder.hip",
               "kmpm_cutnet.hip", "kmpm_io.hip", "kmpm_cpdi.hip"]
HOST_SOURCES = ["kmpm_host.cpp", "kmpm_expr.cpp"]
# kmpm_tiles.hip is compiled once per (dimension, shape function, tile
# variant); the variants build in parallel
TILE_SOURCE = "kmpm_tiles.hip"
TILE_VARIANTS = [(d, s, v) for d in (3, 2) for s in (4, 2) for v in range(3)]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build(verbose=False):
    """Compile every source and link libkmpm.so; always recompiles."""
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    os.makedirs(BUILD_DIR, exist_ok=True)
    jobs = []
    tile_jobs = []
    for s in HIP_SOURCES:
        obj = os.path.join(BUILD_DIR, s + ".o")
        jobs.append(([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", obj], obj))
    for d, s1, v in TILE_VARIANTS:
        obj = os.path.join(BUILD_DIR, "kmpm_tiles_%d%d%d.o" % (d, s1, v))
        defs = ["-DKMPM_TILE_DIM=%d" % d, "-DKMPM_TILE_S1=%d" % s1, "-DKMPM_TILE_V=%d" % v]
        tile_jobs.append(([HIPCC] + HIP_FLAGS + defs + ["-c", os.path.join(CSRC, TILE_SOURCE), "-o",
                                                        obj], obj))
    for s in HOST_SOURCES:
        obj = os.path.join(BUILD_DIR, s + ".o")
        jobs.append((["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, s), "-o", obj], obj))
    jobs = tile_jobs + jobs          # the longest compiles first
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(lambda j: _run(j[0]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [j[1] for j in jobs])
    os.replace(tmp, OUT)
    return OUT


def clean():
    shutil.rmtree(BUILD_DIR, ignore_errors=True)
    if os.path.exists(OUT):
        os.remove(OUT)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

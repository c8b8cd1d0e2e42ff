"""The C ABI consumed from C++: tests/abi/kmpm_abi_driver.cpp includes
include/kmpm.h, links libkmpm.so with g++ and calls the entry points as a
Method subclass in the reference would (set-up, the scheme loops' phase
order, the error path).  Host calls run here; the device part runs on the
GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "abi", "kmpm_abi_driver.cpp")
LIBDIR = os.path.join(ROOT, "karamelo_amd")


def _build(tmp_path):
    exe = str(tmp_path / "kmpm_abi_driver")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           SRC, "-o", exe, "-L", LIBDIR, "-lkmpm", "-Wl,-rpath," + LIBDIR]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    return exe


def _run(exe, mode, timeout):
    r = subprocess.run([exe, mode], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout)
    print(r.stdout, flush=True)
    sys.stdout.flush()
    assert r.returncode == 0, r.stdout
    return r.stdout


def test_header_compiles_as_c():
    """include/kmpm.h is plain C: a C translation unit includes it."""
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c", "-",
                        "-I", os.path.join(ROOT, "include")],
                       input='#include "kmpm.h"\nint main(void) { return KMPM_OK; }\n',
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


def test_cpp_driver_host_calls(tmp_path):
    out = _run(_build(tmp_path), "host", 120)
    assert "host checks: ok" in out


@pytest.mark.gpu
def test_cpp_driver_device_calls(tmp_path):
    out = _run(_build(tmp_path), "device", 240)
    assert "device checks: ok" in out

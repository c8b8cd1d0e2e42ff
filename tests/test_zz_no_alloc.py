"""kmpm_cutnet_gradient_noalloc, the cut network's gradient without any
allocation (block partial sums in a fixed device array), against
kmpm_cutnet_gradient on the same tuples: two calls in a row, the second
with fewer tuples (and fewer partial sums) than the first, both within
FP32 summation-order differences of the allocating entry point (the two
group the tuples into different numbers of block partial sums)."""
import ctypes as C

import numpy as np
import pytest

from karamelo_amd import native as N
from karamelo_amd import problems as P

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def _dev(h, arr):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), arr.nbytes) == 0
    assert h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0      # host to device
    return p


def test_gradient_noalloc_matches_gradient():
    L = N.lib()
    L.kmpm_cutnet_gradient_noalloc.restype = C.c_int
    L.kmpm_cutnet_gradient_noalloc.argtypes = [C.c_void_p] * 3 + [C.c_int64] + [C.c_void_p] * 2
    h = _hip()
    net = N.CutNet(P.steel_jc(), dt=9e-8, cellsize=6e-4, samples=1 << 16, iterations=0,
                   increment_min=1e-5, increment_max=0.2)
    net.train(20)                       # parameters away from the initial ones
    X, Y = net.tuples(40000, seed=21)
    dX, dY = _dev(h, X), _dev(h, Y)
    g = np.zeros(N.CUTNET_NPARAM, np.float32)
    loss = np.zeros(1, np.float32)
    dg, dl = _dev(h, g), _dev(h, loss)
    try:
        for n in (40000, 3000):         # the second call reads fewer partial sums
            print("noalloc gradient, n = %d" % n, flush=True)
            st = L.kmpm_cutnet_gradient_noalloc(net.h, dX, dY, n, dg, dl)
            assert st == 0, st
            assert h.hipMemcpy(g.ctypes.data, dg, g.nbytes, 2) == 0          # device to host
            assert h.hipMemcpy(loss.ctypes.data, dl, loss.nbytes, 2) == 0
            gr, lr = net.gradient(X[:n], Y[:n])
            assert abs(loss[0] - lr) <= 1e-5 * abs(lr), (loss[0], lr)
            e = np.linalg.norm(g - gr) / np.linalg.norm(gr)
            print("  rel L2 against kmpm_cutnet_gradient: %.2e" % e, flush=True)
            assert e <= 1e-5
        # sizes beyond the maximum and a non-positive count are rejected
        assert L.kmpm_cutnet_gradient_noalloc(net.h, dX, dY, 0, dg, dl) == N.ERR_CAPACITY
        assert L.kmpm_cutnet_gradient_noalloc(net.h, dX, dY, (1 << 26) + 1, dg, dl) == \
            N.ERR_CAPACITY
    finally:
        for p in (dX, dY, dg, dl):
            h.hipFree(p)
        net.close()

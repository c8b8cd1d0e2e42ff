"""Pins of the surrogate network's CPU oracle (oracle/mlp_oracle.py) before
it is used to check the device kernels:
  * its analytic gradient equals central finite differences of its own
    float64 loss (a dropped term, a wrong sign or a transposed W2 in the
    backward pass fails this);
  * known answers: with zero weights the output is b3, and the gradient of
    b3 is 2 mean(b3 - t);
  * BF16 rounding is round-to-nearest-even on the 8-bit mantissa.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import mlp_oracle as M  # noqa: E402

NORM = {"in_scale": np.array([0.8, 1.3]), "in_shift": np.array([-0.1, 0.2]),
        "out_scale": 0.7, "out_shift": 1.1, "base": np.array([0.3, -0.2, 0.1, 0.05, 0.02, -0.04])}


def _random_case(seed, n=40):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=0.4, size=M.NPARAM)
    F = rng.uniform(-1, 1, size=(n, 2))
    T = rng.normal(loc=1.0, size=n)
    return p, F, T


def test_gradient_matches_finite_differences():
    p, F, T = _random_case(1)
    loss, g = M.loss_and_grad(p, F, T, NORM, bf16_round=False)
    eps = 1e-6
    num = np.zeros_like(g)
    for k in range(M.NPARAM):
        e = np.zeros_like(p)
        e[k] = eps
        lp, _ = M.loss_and_grad(p + e, F, T, NORM, bf16_round=False)
        lm, _ = M.loss_and_grad(p - e, F, T, NORM, bf16_round=False)
        num[k] = (lp - lm) / (2 * eps)
    assert np.linalg.norm(g - num) <= 1e-7 * np.linalg.norm(num)
    # every parameter block is exercised
    u = M.unpack(num)
    for name in ("w1", "b1", "w2", "b2", "w3"):
        assert np.abs(u[name]).max() > 1e-6, name


def test_known_answers_zero_weights():
    p = np.zeros(M.NPARAM)
    p[-1] = 0.37                       # b3
    F = np.array([[0.1, 0.2], [0.3, -0.4], [0.0, 0.9]])
    T = np.array([1.0, 2.0, 0.5])
    z, _ = M.forward(p, F, NORM)
    assert np.allclose(z, 0.37)
    loss, g = M.loss_and_grad(p, F, T, NORM)
    t = (T - M.baseline(M.inputs(F, NORM), NORM) - NORM["out_shift"]) / NORM["out_scale"]
    assert np.isclose(g[-1], 2 * np.mean(0.37 - t))
    assert np.isclose(loss, np.mean((0.37 - t) ** 2))
    # h2 = 0 -> no gradient for w3; w3 = 0 -> no gradient below it
    assert np.all(g[:-1] == 0)


def test_baseline_is_the_quadratic():
    F = np.array([[0.5, -0.25], [0.0, 0.0], [1.0, 1.0]])
    X = M.inputs(F, NORM).astype(np.float64)
    c = NORM["base"]
    want = c[0] + c[1] * X[:, 0] + c[2] * X[:, 1] + c[3] * X[:, 0] ** 2 + \
        c[4] * X[:, 0] * X[:, 1] + c[5] * X[:, 1] ** 2
    assert np.allclose(M.baseline(M.inputs(F, NORM), NORM), want, rtol=1e-6, atol=1e-7)


def test_bf16_rounding():
    assert M.bf16(1.0) == 1.0
    assert M.bf16(1.0 + 2 ** -8) == 1.0                  # tie -> even
    assert M.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6    # tie -> even (up)
    assert M.bf16(1.0 + 2 ** -7) == 1.0 + 2 ** -7        # representable
    assert M.bf16(-3.0) == -3.0
    x = np.random.default_rng(0).normal(size=1000)
    assert np.all(np.abs(M.bf16(x) - x) <= np.abs(x) * 2 ** -8)   # unit roundoff, 8-bit significand

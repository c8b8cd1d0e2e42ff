"""Pins of the cut network's CPU restatement (oracle/cutnet_oracle.py), which
the device's BF16 matrix-core forward and backward are checked against:

- the backward pass against central differences of the loss (float64, no
  BF16 rounding), parameter by parameter in every layer, so a dropped term,
  a transposed operand or a wrong activation derivative fails;
- the sigmoid form of the activation against the textbook tanh form of
  GELU, and its derivative against differences;
- the split BF16 encoding keeps the inputs to 2^-16 relative;
- the decoder: the elastic branch returns the trial deviator bit for bit,
  the plastic branch the network's outputs, and the damage_init reset at
  triaxiality <= -3.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import cutnet_oracle as M  # noqa: E402


def _norm(rng):
    return {"in_scale": rng.uniform(0.5, 2.0, M.NI).astype(np.float32),
            "in_shift": rng.normal(scale=0.3, size=M.NI).astype(np.float32),
            "out_scale": rng.uniform(0.5, 2.0, M.NO).astype(np.float32),
            "skip": rng.normal(scale=0.2, size=(M.NO, M.NI + 1)).astype(np.float32)}


def test_gelu_is_the_tanh_form():
    z = np.linspace(-8, 8, 4001)
    ref = 0.5 * z * (1 + np.tanh(np.sqrt(2 / np.pi) * (z + 0.044715 * z ** 3)))
    assert np.max(np.abs(M.gelu(z) - ref)) <= 1e-13
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(M.gelu_grad(z) - fd)) <= 1e-8
    assert M.gelu(0.0) == 0.0 and abs(M.gelu(30.0) - 30.0) < 1e-12 and abs(M.gelu(-30.0)) < 1e-12


def test_split_encoding_keeps_16_bits():
    x = np.random.default_rng(1).normal(scale=3.0, size=10000).astype(np.float32).astype(np.float64)
    e = M.encode(x)
    assert np.max(np.abs(e - x) / np.abs(x)) <= 2.0 ** -15
    assert np.max(np.abs(M.bf16(x) - x) / np.abs(x)) > 2.0 ** -10   # one half alone does not


def test_backward_matches_central_differences():
    rng = np.random.default_rng(2)
    norm = _norm(rng)
    p = rng.normal(scale=0.2, size=M.NPARAM)
    F = rng.normal(size=(7, M.NI))
    Y = rng.normal(size=(7, M.NO))
    loss, g = M.loss_and_grad(p, F, Y, norm, bf16_round=False)
    q = M.unpack(np.arange(M.NPARAM))          # flat index of every parameter
    picks = []
    for name, _ in M.SHAPES:
        idx = np.ravel(q[name]).astype(int)
        picks += list(rng.choice(idx, size=min(6, len(idx)), replace=False))
    h = 1e-6
    for k in picks:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        lp, _ = M.loss_and_grad(pp, F, Y, norm, bf16_round=False)
        lm, _ = M.loss_and_grad(pm, F, Y, norm, bf16_round=False)
        fd = (lp - lm) / (2 * h)
        # central differences carry O(h^2) truncation: 1e-5 relative
        assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, g[k])


def test_predict_is_skip_plus_scaled_output():
    rng = np.random.default_rng(3)
    norm = _norm(rng)
    p = np.zeros(M.NPARAM)
    F = rng.normal(size=(5, M.NI))
    # zero weights: z6 = b6, so the output is skip(x) + out_scale b6
    q = M.unpack(p.copy())
    b6 = rng.normal(size=M.NO)
    q["b6"] = b6
    y = M.predict(M.pack(q), F, norm, bf16_round=True)
    X = M.inputs(F, norm)
    S = norm["skip"].astype(np.float64)
    want = X @ S[:, :M.NI].T + S[:, M.NI] + norm["out_scale"].astype(np.float64) * b6
    assert np.allclose(y, want, rtol=1e-12, atol=1e-12)


def test_decoder_branches():
    A, G = 792e6, 76.9e9
    trial = np.array([[1e8, -5e7, -5e7, 2e7, 0, 0], [3e8, -2e8, -1e8, 0, 1e7, 0],
                      [1e6, -1e6, 0, 0, 0, 0]])
    y = np.zeros((3, 9))
    y[0, 6] = -0.1                               # elastic: trial kept exactly
    y[1, :6] = [0.3, -0.2, -0.1, 0, 0.01, 0]
    y[1, 6] = 0.05                               # plastic: network outputs
    y[1, 7] = 0.02
    y[2, 6] = 0.01
    y[2, :6] = [1e-3, -1e-3, 0, 0, 0, 0]
    state = {"damage_init": np.array([0.2, 0.3, 0.4]), "damage": np.zeros(3),
             "pH": np.array([0.0, -1e8, 1e9])}   # third: triax <= -3, reset
    sd, dps, di, dm = M.decode(y, trial, A, G, state)
    assert np.array_equal(sd[0], trial[0]) and dps[0] == 0 and di[0] == 0.2
    assert np.allclose(sd[1], A * y[1, :6]) and np.isclose(dps[1], A * 0.05 / (3 * G))
    assert np.isclose(di[1], 0.3 + A * 0.02 / (3 * G))
    vm = 1.224744871 * np.sqrt(np.sum(sd[2, :3] ** 2) + 2 * np.sum(sd[2, 3:] ** 2))
    assert -1e9 / (vm + 1e6) <= -3 and di[2] == 0.0
    assert np.array_equal(dm, np.zeros(3))

"""CPU restatement of the surrogate network (forward and backward).

TEST INFRASTRUCTURE ONLY: the checker the device's BF16 matrix-core
training and inference kernels (karamelo_amd/csrc/kmpm_mlp.h,
kmpm_surrogate.hip) are compared against.  Nothing in the product imports it.

The network is 2 -> 32 -> 32 -> 1 with ISRU activations,
act(u) = u / sqrt(1 + u^2) (act' = (1 - act^2)^(3/2)), over a quadratic
baseline:
    x  = f * in_scale + in_shift                 (features -> inputs, FP32)
    h1 = act(W1 x + b1)
    h2 = act(W2 h1 + b2)
    z  = w3 . h2 + b3
    output = base(x) + out_shift + out_scale * z
    base(x) = c0 + x0 (c1 + c3 x0 + c4 x1) + x1 (c2 + c5 x1)   (FP32, this order)
loss L = mean((z - (t - base(x) - out_shift) / out_scale)^2).

With bf16=True the values the device feeds its matrix instructions are
rounded to BF16 (round to nearest even) exactly where the device rounds
them: W2 and h1 in the forward product, the layer-2 error g2 and h1 in the
weight-gradient product, W2 and g2 in the input-gradient product.  All other
arithmetic is float64 (the device accumulates in FP32).

The parameters use the flat order of the C ABI (include/kmpm.h,
KMPM_SURROGATE_NPARAM): w1[32][2], b1[32], w2[32][32] ([out][in]), b2[32],
w3[32], b3.
"""
import numpy as np

H, NI = 32, 2
NPARAM = H * NI + H + H * H + H + H + 1


def act(u):
    """The inverse square root unit u / sqrt(1 + u^2)."""
    return u / np.sqrt(1.0 + u * u)


def act_grad(h):
    """d act / du from the activation h: (1 - h^2)^(3/2)."""
    t = 1.0 - h * h
    return t * np.sqrt(t)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    assert p.shape == (NPARAM,)
    o = 0
    out = {}
    for name, shape in (("w1", (H, NI)), ("b1", (H,)), ("w2", (H, H)), ("b2", (H,)),
                        ("w3", (H,)), ("b3", ())):
        n = int(np.prod(shape)) if shape else 1
        out[name] = p[o:o + n].reshape(shape) if shape else p[o]
        o += n
    return out


def pack(g):
    return np.concatenate([np.ravel(g["w1"]), g["b1"], np.ravel(g["w2"]), g["b2"], g["w3"],
                           [g["b3"]]])


def bf16(x):
    """Round float values to the nearest BF16 (ties to even), as float64."""
    f = np.asarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16 << 16
    return (b.astype(np.uint32)).view(np.float32).astype(np.float64)


def _r(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


def inputs(F, norm):
    """Network inputs as the device forms them: one FP32 rounding of
    f * in_scale + in_shift."""
    F = np.asarray(F, np.float32).astype(np.float64)
    sc = np.asarray(norm["in_scale"], np.float32).astype(np.float64)
    sh = np.asarray(norm["in_shift"], np.float32).astype(np.float64)
    return (F * sc + sh).astype(np.float32)


def baseline(X32, norm):
    """Quadratic baseline in FP32 arithmetic, in the device's order."""
    c = np.asarray(norm["base"], np.float32)
    x0, x1 = X32[:, 0], X32[:, 1]
    return (c[0] + x0 * (c[1] + c[3] * x0 + c[4] * x1) + x1 * (c[2] + c[5] * x1)).astype(np.float64)


def forward(p, F, norm, bf16_round=True):
    """Raw outputs z for features F [n][2]; returns (z, cache)."""
    q = unpack(p)
    X = inputs(F, norm).astype(np.float64)
    h1 = act(X @ q["w1"].T + q["b1"])
    z2 = _r(h1, bf16_round) @ _r(q["w2"], bf16_round).T + q["b2"]
    h2 = act(z2)
    z = h2 @ q["w3"] + q["b3"]
    return z, (q, X, h1, h2)


def predict(p, F, norm, bf16_round=True):
    z, _ = forward(p, F, norm, bf16_round)
    return baseline(inputs(F, norm), norm) + norm["out_shift"] + z * norm["out_scale"]


def loss_and_grad(p, F, T, norm, bf16_round=True):
    """Mean squared error of the raw output against normalised targets and
    its gradient with respect to the flat parameters."""
    z, (q, X, h1, h2) = forward(p, F, norm, bf16_round)
    t = (np.asarray(T, np.float64) - baseline(inputs(F, norm), norm) - norm["out_shift"]) \
        / norm["out_scale"]
    n = len(z)
    r = z - t
    loss = float(np.mean(r * r))
    g = 2.0 * r / n                                   # dL/dz
    grad = {"b3": g.sum(), "w3": h2.T @ g}
    g2 = g[:, None] * q["w3"][None, :] * act_grad(h2)
    grad["b2"] = g2.sum(axis=0)
    grad["w2"] = _r(g2, bf16_round).T @ _r(h1, bf16_round)          # [out][in]
    g1 = (_r(g2, bf16_round) @ _r(q["w2"], bf16_round)) * act_grad(h1)
    grad["b1"] = g1.sum(axis=0)
    grad["w1"] = g1.T @ X
    return loss, pack(grad)

"""CPU oracle for the learned (k, p, q) proposer's network (test
infrastructure only: imported by tests and never by the library).

The reference has no learned component, so this restates the library's own
network definition: a 34 -> 128 -> 128 -> 3 MLP with exact (erf) GELU,
parameters packed W1[128][34], b1[128], W2[128][128], b2[128], W3[3][128],
b3[3], loss sum_s sum_o wo[o] (Y_so - T_so)^2 / B.

``emulate=True`` reproduces the GPU's rounding points (weights, inputs,
activations and back-propagated signals rounded to BF16 with
round-to-nearest-even through FP32; biases, pre-activations, GELU'
and all sums kept wide), so that the GPU and the oracle differ only by
summation order.  ``emulate=False`` is the plain FP64 function, pinned by
finite differences in the tests.
"""
import numpy as np
from scipy.special import erf

NIN, NH, NOUT = 34, 128, 3
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
         ("W3", (NOUT, NH)), ("b3", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])


def bf16(x):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    f = np.asarray(x, dtype=np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)
    return (u << np.uint64(16)).astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def init_params(seed):
    """He-style initialisation used by the proposer (weights N(0, 2/fan_in)
    scaled by 0.5 for the last layer, zero biases)."""
    r = np.random.default_rng(seed)
    d = {"W1": r.standard_normal((NH, NIN)) * np.sqrt(2.0 / NIN), "b1": np.zeros(NH),
         "W2": r.standard_normal((NH, NH)) * np.sqrt(2.0 / NH), "b2": np.zeros(NH),
         "W3": r.standard_normal((NOUT, NH)) * 0.5 * np.sqrt(1.0 / NH), "b3": np.zeros(NOUT)}
    return pack(d)


def forward(p, X, emulate=True):
    """Returns (Y, cache)."""
    q = unpack(p)
    rd = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    X = rd(np.asarray(X, dtype=np.float64))
    W1, W2, W3 = rd(q["W1"]), rd(q["W2"]), rd(q["W3"])
    Z1 = X @ W1.T + q["b1"]
    H1 = rd(gelu(Z1))
    Z2 = H1 @ W2.T + q["b2"]
    H2 = rd(gelu(Z2))
    Y = H2 @ W3.T + q["b3"]
    return Y, (X, W1, W2, W3, Z1, H1, Z2, H2)


def loss_grad(p, X, T, wo, emulate=True):
    """(loss, gradient as a packed vector)."""
    rd = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    Y, (X, W1, W2, W3, Z1, H1, Z2, H2) = forward(p, X, emulate)
    T = np.asarray(T, dtype=np.float64)
    wo = np.asarray(wo, dtype=np.float64)
    B = X.shape[0]
    E = Y - T
    loss = float(np.sum(wo * E * E) / B)
    dY = rd(2.0 * wo * E / B)
    g = {"W3": dY.T @ H2, "b3": dY.sum(axis=0)}
    dZ2 = rd((dY @ W3) * gelu_d(Z2))
    g["W2"] = dZ2.T @ H1
    g["b2"] = dZ2.sum(axis=0)
    dZ1 = rd((dZ2 @ W2) * gelu_d(Z1))
    g["W1"] = dZ1.T @ X
    g["b1"] = dZ1.sum(axis=0)
    return loss, pack(g)

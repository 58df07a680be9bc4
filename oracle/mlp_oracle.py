"""CPU restatement of the calibration surrogate (test infrastructure only).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

The 16-128-128-128-1 GELU perceptron in FP64.  With bf16=True the operands are
rounded to BF16 (round to nearest even) exactly where the GPU kernels round them
(weights, inputs, hidden activations, back-propagated deltas), so the GPU's FP32
accumulation is the only remaining difference; with bf16=False the network is
plain FP64, which finite differences can check.
"""
import math

import numpy as np

HID, NIN = 128, 16
SIZES = [("W1", (HID, NIN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, HID)),
         ("b3", (HID,)), ("w4", (HID,)), ("b4", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unflatten(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def flatten(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def bf16(a):
    """Round to BF16 (nearest, ties to even) via the FP32 bit pattern."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


_erf = np.vectorize(math.erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(p, x, bf=False):
    w = unflatten(p)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    xb = r(x)
    W1, W2, W3, w4 = r(w["W1"]), r(w["W2"]), r(w["W3"]), r(w["w4"])
    z1 = xb @ W1.T + w["b1"]
    h1 = r(gelu(z1))
    z2 = h1 @ W2.T + w["b2"]
    h2 = r(gelu(z2))
    z3 = h2 @ W3.T + w["b3"]
    h3 = r(gelu(z3))
    y = h3 @ w4 + w["b4"][0]
    return y, dict(xb=xb, W1=W1, W2=W2, W3=W3, w4=w4, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3)


def loss_and_grad(p, x, t, bf=False):
    """Mean squared error and its gradient (flat, parameter order)."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    y, c = forward(p, x, bf)
    n = len(t)
    loss = float(np.mean((y - t) ** 2))
    dy = 2.0 * (y - t) / n
    g = {"w4": dy @ c["h3"], "b4": np.array([dy.sum()])}
    dz3 = dy[:, None] * c["w4"][None, :] * gelu_grad(c["z3"])
    g["b3"] = dz3.sum(0)
    dz3b = r(dz3)
    g["W3"] = dz3b.T @ c["h2"]
    dz2 = (dz3b @ c["W3"]) * gelu_grad(c["z2"])
    g["b2"] = dz2.sum(0)
    dz2b = r(dz2)
    g["W2"] = dz2b.T @ c["h1"]
    dz1 = (dz2b @ c["W2"]) * gelu_grad(c["z1"])
    g["b1"] = dz1.sum(0)
    g["W1"] = r(dz1).T @ c["xb"]
    return loss, flatten(g)

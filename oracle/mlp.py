"""CPU oracle of the learned pose-increment prior (test infrastructure only: tests/, smoke() and bench.py's
cpu_baseline leg use it as the checker; the product runs the HIP kernels of lw_mlp.hip).

The network has no counterpart in the reference. What it must compute is fixed by its own specification: a
152 -> 256 -> 256 -> 256 -> 6 MLP with exact (erf) GELU and MSE loss, BF16 matrix operands with FP32
accumulation, trained by Adam. This restatement runs in float64 and rounds to BF16 (round to nearest even)
exactly where the device does: the input, each hidden activation, the weights used in products, d loss / d y
and each dZ. With bf16=False it is the plain float64 network; its backward pass is pinned by central finite
differences (tests/test_oracle_mlp.py), so a dropped term, a transposed operand or a wrong GELU derivative
fails there.

Parameter vector layout (float32, row-major [in][out]): W1 160x256, b1 256, W2 256x256, b2 256, W3 256x256,
b3 256, W4 256x16, b4 16; W1 rows 152..159 and W4 columns 6..15 are zero.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

IN, IN_PAD, HID, OUT, OUT_PAD, MOTION = 152, 160, 256, 6, 16, 24
SHAPES = [("W1", (IN_PAD, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, HID)),
          ("b3", (HID,)), ("W4", (HID, OUT_PAD)), ("b4", (OUT_PAD,))]
PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def init_params(rng, out_scale: float = 0.1) -> np.ndarray:
    """He-normal hidden weights, zero biases, a small last layer, padding zero."""
    p = np.zeros(PARAMS, np.float32)
    v = unpack(p)
    v["W1"][:IN] = rng.normal(0, math.sqrt(2.0 / IN), (IN, HID))
    v["W2"][:] = rng.normal(0, math.sqrt(2.0 / HID), (HID, HID))
    v["W3"][:] = rng.normal(0, math.sqrt(2.0 / HID), (HID, HID))
    v["W4"][:, :OUT] = rng.normal(0, out_scale * math.sqrt(1.0 / HID), (HID, OUT))
    return p


def bf16(x) -> np.ndarray:
    """float32 -> nearest BF16 (ties to even), returned as float64 values."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _r(x, on: bool):
    return bf16(x) if on else np.asarray(x, np.float64)


def forward(p: np.ndarray, x: np.ndarray, bf16_ops: bool = True):
    """y (n x 6) and the activations the backward pass needs."""
    v = unpack(np.asarray(p, np.float32 if bf16_ops else np.float64))
    n = len(x)
    X = np.zeros((n, IN_PAD))
    X[:, :IN] = _r(x, bf16_ops)
    h, Z, H = X, [], []
    for l in range(3):
        W = _r(v[f"W{l + 1}"], bf16_ops)
        z = h @ W + v[f"b{l + 1}"].astype(np.float64)
        Z.append(z)
        h = _r(gelu(z), bf16_ops)
        H.append(h)
    y = h @ _r(v["W4"], bf16_ops)[:, :OUT] + v["b4"][:OUT].astype(np.float64)
    return y, {"X": X, "Z": Z, "H": H}


def mse(y: np.ndarray, t: np.ndarray, bf16_ops: bool = True):
    """(loss, d loss / d y as the backward pass consumes it: n x 16, zero beyond column 6)."""
    n = len(y)
    d = np.asarray(y, np.float64) - np.asarray(t, np.float64)
    dy = np.zeros((n, OUT_PAD))
    dy[:, :OUT] = _r(2.0 * d / (6 * n), bf16_ops)
    return float(np.mean(d * d)), dy


def backward(p: np.ndarray, cache: dict, dy: np.ndarray, bf16_ops: bool = True) -> np.ndarray:
    """The gradient of the loss (through dy, n x 16) with respect to the parameter vector, float64."""
    v = unpack(np.asarray(p, np.float32 if bf16_ops else np.float64))
    g = np.zeros(PARAMS)
    gv = unpack(g)
    ins = [cache["X"]] + cache["H"][:2]
    dout = dy
    for l in range(3, -1, -1):
        h_in = cache["H"][2] if l == 3 else ins[l]
        gv[f"W{l + 1}"][:] = h_in.T @ dout
        gv[f"b{l + 1}"][:] = dout.sum(axis=0)
        if l == 0:
            break
        W = _r(v[f"W{l + 1}"], bf16_ops)
        dh = dout @ W.T
        dout = _r(dh * gelu_grad(cache["Z"][l - 1]), bf16_ops)
    return g


def adam(w, m, v, g, lr, b1, b2, eps, step):
    """One bias-corrected Adam step (float64); returns (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return w, m, v


def features(xyz: np.ndarray, r_max: float, e_max: float, motion) -> np.ndarray:
    """The 152 inputs of one frame: motion (24, as given), 64 range bins over [0, r_max) and 64 elevation bins
    over [-e_max, e_max) as count * 64 / m, end bins taking what falls outside; float32 as the device forms it."""
    p = np.asarray(xyz, np.float32).reshape(-1, 3)
    m = len(p)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    h = x * x
    h = h + y * y
    r = np.sqrt(h + z * z)
    e = np.arctan2(z, np.sqrt(h))
    rs = np.float32(64.0) / np.float32(r_max)
    es = np.float32(32.0) / np.float32(e_max)
    br = np.clip(np.floor(r * rs), 0, 63).astype(int)
    be = np.clip(np.floor((e + np.float32(e_max)) * es), 0, 63).astype(int)
    out = np.zeros(IN, np.float32)
    out[:MOTION] = np.asarray(motion, np.float32)
    scale = np.float32(64.0) / np.float32(m) if m else np.float32(0.0)
    out[MOTION:MOTION + 64] = np.bincount(br, minlength=64).astype(np.float32) * scale
    out[MOTION + 64:] = np.bincount(be, minlength=64).astype(np.float32) * scale
    return out

"""The MLP oracle pinned without a GPU: its float64 backward pass against central finite differences, BF16
rounding against hand-computed values, GELU's derivative, Adam's first step, and the feature histograms
against closed forms."""

import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def om():
    import mlp

    return mlp


def test_backward_matches_finite_differences(om):
    rng = np.random.default_rng(0)
    p = om.init_params(rng, out_scale=1.0).astype(np.float64)
    v = om.unpack(p)
    for name in ("b1", "b2", "b3"):
        v[name][:] = rng.normal(0, 0.3, v[name].shape)
    v["b4"][:om.OUT] = rng.normal(0, 0.3, om.OUT)
    x = rng.normal(0, 1, (5, om.IN))
    t = rng.normal(0, 1, (5, om.OUT))

    def loss(q):
        y, _ = om.forward(q, x, bf16_ops=False)
        return om.mse(y, t, bf16_ops=False)[0]

    y, cache = om.forward(p, x, bf16_ops=False)
    _, dy = om.mse(y, t, bf16_ops=False)
    g = om.backward(p, cache, dy, bf16_ops=False)
    picks, o = [], 0
    for name, shape in om.SHAPES:
        size = int(np.prod(shape))
        if name == "W1":
            picks += [o + int(i) * shape[1] + int(j) for i, j in zip(rng.integers(0, om.IN, 4), rng.integers(0, shape[1], 4))]
        elif name == "W4":
            picks += [o + int(i) * shape[1] + int(j) for i, j in zip(rng.integers(0, shape[0], 4), rng.integers(0, om.OUT, 4))]
        elif name == "b4":
            picks += [o + j for j in range(om.OUT)]
        else:
            picks += [o + int(k) for k in rng.integers(0, size, 4)]
        o += size
    h = 1e-6
    for k in picks:
        q = p.copy()
        q[k] += h
        lp = loss(q)
        q[k] -= 2 * h
        lm = loss(q)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-7, (k, fd, g[k])
    gv = om.unpack(g)
    assert np.all(gv["W1"][om.IN:] == 0) and np.all(gv["W4"][:, om.OUT:] == 0) and np.all(gv["b4"][om.OUT:] == 0)


def test_bf16_rounding(om):
    cases = [(1.0, 1.0), (1.0 + 2 ** -8, 1.0), (1.0 + 3 * 2 ** -8, 1.0 + 2 ** -6), (1.0 + 2 ** -8 + 2 ** -20, 1.0 + 2 ** -7),
             (-2.0 - 2 ** -7, -2.0), (0.0, 0.0)]
    for x, want in cases:
        assert float(om.bf16(np.array([x], np.float32))[0]) == want, x
    for x in np.random.default_rng(2).normal(0, 10, 200).astype(np.float32):
        got = float(om.bf16(np.array([x]))[0])
        ulp = 2.0 ** (math.floor(math.log2(abs(float(x)))) - 7)
        assert abs(got - float(x)) <= ulp / 2


def test_gelu_grad(om):
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (om.gelu(z + h) - om.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - om.gelu_grad(z))) < 1e-8
    assert om.gelu(np.array([0.0]))[0] == 0.0 and abs(om.gelu(np.array([10.0]))[0] - 10.0) < 1e-12


def test_adam_first_step_is_signed_lr(om):
    g = np.array([0.5, -2.0, 1e-3])
    w, m, v = om.adam(np.zeros(3), np.zeros(3), np.zeros(3), g, 0.01, 0.9, 0.999, 1e-12, 1)
    assert np.allclose(w, -0.01 * np.sign(g), rtol=1e-9)


def test_features_closed_form(om):
    rng = np.random.default_rng(1)
    n = 1000
    az = rng.uniform(0, 2 * math.pi, n)
    d = np.column_stack([np.cos(az), np.sin(az), np.zeros(n)])  # level rays: elevation 0
    xyz = (d * 5.1).astype(np.float32)                         # all at range 5.1 m
    f = om.features(xyz, 20.0, math.radians(20.0), np.arange(24))
    assert np.array_equal(f[:24], np.arange(24, dtype=np.float32))
    r, e = f[24:88], f[88:]
    assert r[16] == 64.0 and r.sum() == 64.0  # floor(5.1 * 64 / 20) = 16
    assert e[32] == 64.0 and e.sum() == 64.0  # floor((0 + e_max) * 32 / e_max) = 32
    far = np.array([[100.0, 0.0, 0.0], [0.0, 0.0, 50.0]], np.float32)  # beyond r_max, straight up: end bins
    g = om.features(far, 20.0, math.radians(20.0), np.zeros(24))
    assert g[24 + 63] == 64.0 and g[88 + 63] == 32.0 and g[88 + 32] == 32.0

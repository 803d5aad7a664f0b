"""The learned prior's kernels (lw_mlp_*, lw_frame_features) against the CPU oracle (oracle/mlp.py), which
rounds to BF16 where the device does: forward at ragged batch sizes, loss and every gradient block, one Adam
step, the frame histograms; and a training run that must learn a known mapping."""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def om():
    import mlp

    return mlp


def _params(om, seed):
    rng = np.random.default_rng(seed)
    p = om.init_params(rng, out_scale=1.0)
    v = om.unpack(p)
    for name in ("b1", "b2", "b3"):
        v[name][:] = rng.normal(0, 0.2, v[name].shape)
    v["b4"][:om.OUT] = rng.normal(0, 0.2, om.OUT)
    return p


def _blocks(om, g):
    return om.unpack(np.asarray(g, np.float64))


@pytest.mark.parametrize("n", [1, 7, 64, 300])
def test_forward_matches_oracle(lw, om, n, gpu_ok):
    p = _params(om, 1)
    net = lw.prior.MLP(batch_cap=512)
    net.set_params(p)
    x = np.random.default_rng(n).normal(0, 1, (n, om.IN)).astype(np.float32)
    print(f"lw_mlp_forward n={n}", flush=True)
    y = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    yo, _ = om.forward(p, x)
    err = np.abs(y - yo).max()
    assert err <= 2e-3 * np.abs(yo).max() + 1e-5, (err, np.abs(yo).max())


def test_strided_input(lw, om, gpu_ok):
    """x with a row stride larger than 152 (a column slice of a wider table)."""
    p = _params(om, 2)
    net = lw.prior.MLP(batch_cap=64)
    net.set_params(p)
    wide = np.random.default_rng(5).normal(0, 1, (33, 200)).astype(np.float32)
    xt = torch.from_numpy(wide).cuda()[:, 10:10 + om.IN]
    y = net.forward(xt).cpu().numpy()
    yo, _ = om.forward(p, wide[:, 10:10 + om.IN])
    assert np.abs(y - yo).max() <= 2e-3 * np.abs(yo).max() + 1e-5


@pytest.mark.parametrize("n", [37, 1000, 4096])
def test_loss_and_gradients_match_oracle(lw, om, n, gpu_ok):
    """Every gradient block; at 1000 and 4096 rows the weight gradients run split over K (the batch)."""
    p = _params(om, 3)
    rng = np.random.default_rng(4)
    x = rng.normal(0, 1, (n, om.IN)).astype(np.float32)
    t = rng.normal(0, 1, (n, om.OUT)).astype(np.float32)
    net = lw.prior.MLP(batch_cap=max(1024, n))
    net.set_params(p)
    print(f"lw_mlp_loss/backward n={n}", flush=True)
    y = net.forward(torch.from_numpy(x).cuda())
    loss = float(net.loss(y, torch.from_numpy(t).cuda()).item())
    net.backward(n)
    g = _blocks(om, net.grads())
    yo, cache = om.forward(p, x)
    lo, dy = om.mse(yo, t)
    go = _blocks(om, om.backward(p, cache, dy))
    assert abs(loss - lo) <= 1e-4 * lo
    for name, ref in go.items():
        got = g[name]
        rel = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
        assert rel < 1e-2, (name, rel)
        assert np.abs(got - ref).max() <= 0.05 * np.abs(ref).max() + 1e-9, name
    assert np.all(g["W1"][om.IN:] == 0) and np.all(g["W4"][:, om.OUT:] == 0) and np.all(g["b4"][om.OUT:] == 0)


def test_adam_step_matches_oracle(lw, om, gpu_ok):
    p = _params(om, 6)
    rng = np.random.default_rng(7)
    n = 128
    x = torch.from_numpy(rng.normal(0, 1, (n, om.IN)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.normal(0, 1, (n, om.OUT)).astype(np.float32)).cuda()
    net = lw.prior.MLP(batch_cap=n)
    net.set_params(p)
    w, m, v = p.astype(np.float64), np.zeros(len(p)), np.zeros(len(p))
    for step in (1, 2, 3):
        net.forward(x)
        net.loss(net._y[:n], t)
        net.backward(n)
        g = net.grads().astype(np.float64)
        net.adam(1e-3)
        w, m, v = om.adam(w, m, v, g, 1e-3, 0.9, 0.999, 1e-8, step)
        got = net.params()
        assert np.abs(got - w).max() <= 1e-6, step
    assert np.all(om.unpack(got)["W1"][om.IN:] == 0) and np.all(om.unpack(got)["W4"][:, om.OUT:] == 0)


def test_graph_training_step_equals_the_eager_steps(lw, om, gpu_ok):
    """lw_mlp_train_step replayed from its captured graph against the same step run launch by launch, four
    steps (Adam's bias correction changes every step), with other work and synchronisations between the
    replays: parameters and loss equal bit for bit after every step."""
    p = _params(om, 8)
    rng = np.random.default_rng(9)
    n = 2048
    x = torch.from_numpy(rng.normal(0, 1, (n, om.IN)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.normal(0, 1, (n, om.OUT)).astype(np.float32)).cuda()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    eager, graph = lw.prior.MLP(batch_cap=n, stream=s), lw.prior.MLP(batch_cap=n, stream=s)
    try:
        eager.set_params(p)
        graph.set_params(p)
        for step in range(4):
            le_t = eager.train_step(x, t, 1e-3, graph=False)
            lg_t = graph.train_step(x, t, 1e-3, graph=True)
            s.synchronize()  # the losses are written on the networks' stream
            le, lg = float(le_t.item()), float(lg_t.item())
            float((x * 2.0).sum().item())  # the program's own work between replays
            torch.cuda.synchronize()
            assert le == lg and np.array_equal(eager.params(), graph.params()), step
    finally:
        eager.close()
        graph.close()


def test_frame_features_match_oracle(lw, om, gpu_ok):
    ss = lw.synth.single_scan(seed=3, n=9000, map_points=20_000)
    eng = lw.Engine(reg_slots=1 << 16, frame_cap=1 << 14)
    m = eng.load_frame(ss.scan_xyz, ss.scan_attr, ss.scan_t)
    motion = np.linspace(-1, 1, 24).astype(np.float32)
    out = torch.zeros(om.IN, dtype=torch.float32, device="cuda")
    r_max, e_max = 20.0, math.radians(20.0)
    lw._abi.check(eng.lib.lw_frame_features(eng.ctx, m, r_max, e_max, lw._abi.ptr(motion, lw._abi._f32p),
                                            out.data_ptr()), "lw_frame_features")
    eng.sync()
    got = out.cpu().numpy()
    want = om.features(ss.scan_xyz, r_max, e_max, motion)
    assert np.array_equal(got[:24], want[:24])
    # counts may differ only where a point sits within a rounding of a bin edge (atan2 / sqrt last bits)
    scale = 64.0 / m
    for a, b in ((24, 88), (88, 152)):
        moved = np.abs(np.round((got[a:b] - want[a:b]) / scale)).sum()
        assert moved <= 4, moved
        assert abs(got[a:b].sum() - 64.0) < 1e-3


def test_training_learns_a_mapping(lw, om, gpu_ok):
    """A smooth target of the motion inputs: the loss must fall by 20x within 300 Adam steps."""
    rng = np.random.default_rng(8)
    n = 4096
    X = rng.normal(0, 1, (n, om.IN)).astype(np.float32)
    A = rng.normal(0, 0.5, (om.IN, om.OUT))
    Y = np.tanh(X @ A / math.sqrt(om.IN) * 3).astype(np.float32)
    Xd, Yd = torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda()
    params, hist, _ = lw.prior.fit(Xd, Yd, epochs=75, batch=1024, lr=2e-3)
    print("loss per epoch", [round(h, 5) for h in hist[::10]], flush=True)
    assert hist[-1] < hist[0] / 20

"""GPU frame cutting, deskew and thinning against the oracle: appends in input order (bit-exact, count past the
cap as the reference counts), deskew within float32 rounding of the oracle and of the reference's float64 host
restatement (prior_session.deskew), thinning winners and order bit-exact."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dt).cuda()


def test_frame_append_order_and_cap(lw, orc, gpu_ok):
    rng = np.random.default_rng(0)
    cap = 3000
    eng = lw.Engine(1 << 10, cap)
    o_xyz, o_attr, o_t, o_frame = (np.zeros((cap, 3), np.float32), np.zeros(cap, np.uint32),
                                   np.zeros(cap, np.float32), np.zeros(cap, np.int32))
    count = 0
    eng.frame_reset()
    for b in range(4):
        n = 1500
        xyz = rng.normal(size=(n, 3)).astype(np.float32)
        attr = rng.integers(0, 1 << 20, n).astype(np.uint32)
        t = np.sort(rng.uniform(0.05 * b, 0.05 * b + 0.12, n)).astype(np.float32)
        t[::17] = -0.01  # max(t, 0) -> frame 0
        k_open, k_min = 1, 0
        eng.frame_append(_dev(xyz, torch.float32), _dev(attr.view(np.int32), torch.int32), _dev(t, torch.float32),
                         10.0, k_min, k_open)
        count = orc.frame_append(xyz, attr, t, 10.0, k_min, k_open, cap, count, o_xyz, o_attr, o_t, o_frame)
    assert eng.frame_count() == count and count > cap
    torch.cuda.synchronize()
    assert np.array_equal(eng.f_raw.cpu().numpy(), o_xyz)
    assert np.array_equal(eng.f_attr.cpu().numpy().view(np.uint32), o_attr)
    assert np.array_equal(eng.f_t.cpu().numpy(), o_t)
    assert np.array_equal(eng.f_frame.cpu().numpy(), o_frame)


def test_deskew(lw, orc, gpu_ok):
    rng = np.random.default_rng(1)
    m = 20000
    raw = rng.uniform(-10, 10, (m, 3)).astype(np.float32)
    t = rng.uniform(0.3, 0.4, m).astype(np.float32)
    eng = lw.Engine(1 << 10, m)
    eng.load_frame(raw, np.zeros(m, np.uint32), t)
    for xi in (np.zeros(6), np.array([0.05, -0.02, 0.01, 0.02, -0.05, 0.09]), np.array([0.3, 0, 0, 0, 0, 0.6])):
        eng.deskew(m, 0.35, 10.0, xi)
        torch.cuda.synchronize()
        got = eng.f_xyz[:m].cpu().numpy()
        ref32 = orc.deskew(raw, t, 0.35, 10.0, xi)
        s = np.clip((t.astype(np.float64) - 0.35) * 10.0, -1, 1)
        ref64 = orc.deskew_fp64(raw, s, xi.astype(np.float32))
        if not xi.any():
            assert np.array_equal(got, raw)  # the identity is exact
        assert np.abs(got - ref32).max() < 1e-5
        assert np.abs(got - ref64).max() < 2e-5


@pytest.mark.parametrize("xi", [np.zeros(6), np.array([0.04, 0.0, 0.01, 0.0, 0.0, 0.08])])
def test_thinning_matches_serial(lw, orc, xi, gpu_ok):
    seq = lw.synth.Sequence(seed=23, seconds=1.0)
    xyz, attr, t = next(seq.frames(1))
    attr = attr.copy()
    attr[::40] |= np.uint32(1 << 8)
    m = len(t)
    eng = lw.Engine(1 << 12, 1 << 14)
    eng.load_frame(xyz, attr, t)
    eng.deskew(m, 0.05, 10.0, xi)
    torch.cuda.synchronize()
    desk = eng.f_xyz[:m].cpu().numpy()
    for voxel in (0.15, 0.2, 0.5):
        n = eng.thin(m, 2.0 / voxel, 0.3, 60.0, 0b1111)
        got = eng.thin_read(n)
        ref = orc.ds_insert(desk, attr, 2.0 / voxel, 0.3, 60.0, 0b1111, 1 << 14)
        assert n == len(ref) and np.array_equal(got, ref)
    # the table is left clean: the same call again gives the same result
    n2 = eng.thin(m, 2.0 / 0.2, 0.3, 60.0, 0b1111)
    assert np.array_equal(eng.thin_read(n2), orc.ds_insert(desk, attr, 10.0, 0.3, 60.0, 0b1111, 1 << 14))

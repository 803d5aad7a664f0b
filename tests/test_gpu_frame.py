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


@pytest.mark.parametrize("case", ["scan", "scattered", "large"])
def test_frame_chain_preparation(lw, orc, case, gpu_ok):
    """The deskew and thinning lw_odom_frame runs before its first registration: for a frame of up to 32768
    points one launch over many workgroups does both (claims in the global thinning table, two grid barriers;
    "scattered" holds more thinning voxels than a one-workgroup LDS table), for a larger frame ("large") the
    grid kernels. The deskewed frame is the grid deskew kernel's bit for bit, the thinned points are the
    oracle's serial thinning (winners and order) bit for bit, and the global thinning table is left clean."""
    import ctypes

    rng = np.random.default_rng({"scan": 1, "scattered": 2, "large": 3}[case])
    if case == "scan":
        seq = lw.synth.Sequence(seed=23, seconds=1.0)
        xyz, attr, t = next(seq.frames(1))
        t = (t - t.min()).astype(np.float32) * np.float32(0.9)
        cap = 1 << 15
    else:
        n = 20000 if case == "scattered" else 40000
        xyz = rng.uniform(-45.0, 45.0, size=(n, 3)).astype(np.float32)
        attr = (rng.integers(0, 4, n).astype(np.uint32) << np.uint32(16))
        t = np.sort(rng.uniform(0.0, 0.0999, n)).astype(np.float32)
        cap = 1 << 15 if case == "scattered" else 1 << 16
    attr = attr.copy()
    attr[::37] |= np.uint32(1 << 8)  # tagged points take no part in the thinning
    m = len(t)
    assert (m <= 32768) == (case != "large")
    eng = lw.Engine(1 << 18, cap)
    eng.insert(xyz, attr, np.eye(4), 0.2, max_range=60.0)  # something to register against
    eng.frame_reset()
    eng.frame_append(_dev(xyz, torch.float32), _dev(attr.view(np.int32), torch.int32), _dev(t, torch.float32),
                     10.0, -1, 0)
    xi0 = np.array([0.05, -0.02, 0.01, 0.01, -0.02, 0.06], np.float32)
    p = eng.params(reg_voxel=0.2, max_iter=3, min_corr=10)
    fi, fo = lw._abi.FrameIn(), lw._abi.FrameOut()
    np.ctypeslib.as_array(fi.T_pred)[:] = np.eye(4).ravel()
    np.ctypeslib.as_array(fi.T_prev_inv)[:] = np.eye(4).ravel()
    np.ctypeslib.as_array(fi.T_key_inv)[:] = np.eye(4).ravel()
    np.ctypeslib.as_array(fi.xi)[:] = xi0
    fi.kf_dist, fi.kf_angle, fi.redeskew_rot, fi.redeskew_trans = 1e9, 1e9, 1e9, 1e9
    fi.t_mid, fi.inv_dt = 0.05, 10.0
    fi.min_points, fi.has_prev, fi.has_key, fi.redeskew = 1, 0, 1, 0
    lw._abi.check(eng.lib.lw_odom_frame(eng.ctx, ctypes.byref(p), ctypes.byref(fi), ctypes.byref(fo)),
                  "lw_odom_frame")
    assert fo.m == m and fo.registrations == 1 and not fo.keyframe
    torch.cuda.synchronize()
    desk = eng.f_xyz[:m].cpu().numpy().copy()
    n_ds = int(fo.reg[0].n_ds)
    got = eng.thin_read(n_ds)
    ds_cap = 1 << (cap - 1).bit_length()
    ref = orc.ds_insert(desk, attr, 2.0 / 0.2, p.min_range, p.max_range, p.ret_mask, ds_cap)
    print(f"{case}: {m} points, {n_ds} thinned")
    assert n_ds == len(ref) and np.array_equal(got, ref)
    # the grid deskew from the same raw points, and the grid thinning on a clean table
    eng.deskew(-1, 0.05, 10.0, xi0)
    torch.cuda.synchronize()
    assert np.array_equal(eng.f_xyz[:m].cpu().numpy(), desk)
    n2 = eng.thin(m, 2.0 / 0.2, p.min_range, p.max_range, p.ret_mask)
    assert n2 == n_ds and np.array_equal(eng.thin_read(n2), ref)
    if case == "scattered":
        assert n_ds > 8192  # more thinning voxels than the workgroup's table
    eng.close()

"""LVXR recordings decoded on the device (lw_lvxr_decode through lw.lvxr.read_firings) against the oracle's
serial reader: every data type, the Mid-40 dual profile, GPS stamps, clock resets and jumps, malformed,
short, IMU and truncated datagrams, legacy LVXR1 files, and firings of different data types at the same
instants (ties ordered as the reference concatenates them). Times, cartesian points, reflectivity, tags and
validity match exactly; spherical points and directions within float32 sin/cos rounding. Then a recording of
a synthetic walk replays through the Odometry and tracks the walk."""

import numpy as np
import pytest
import torch

import lvxr_oracle as X
import lvxr_write as W

pytestmark = pytest.mark.gpu


def _files(tmp_path):
    rng = np.random.default_rng(9)
    out = {}
    for dt in (0, 1, 2, 3, 4, 5, 7, 8):
        out[f"type{dt}"] = (1, 0, W.stream(rng, dt, 12, 96 if dt < 4 else 48, 1_000_000, 10_000), {})
    out["dual"] = (1, 1, W.stream(rng, 0, 6, 100, 0, 10_000) + W.stream(rng, 1, 4, 100, 50 * 6 * 10_000, 10_000)
                   + [W.header(0, 0) + W.records(rng, 0, 98)], {})
    out["gps_fast"] = (3, 0, W.stream(rng, 2, 10, 96, 3_600_000_000_000 * 7 + 5_000, 4_167, ts_type=3), {})
    clock = (W.stream(rng, 2, 4, 96, 2_000_000_000, 10_000) + W.stream(rng, 2, 3, 96, 9_000_000_000, 10_000)
             + W.stream(rng, 2, 3, 96, 500_000, 10_000, jitter=3000))
    out["clock"] = (1, 0, clock, {})
    bad = (W.stream(rng, 4, 3, 48, 0, 10_000) + [W.header(4, 0) + W.records(rng, 4, 10)[:-5], bytes(12),
                                                  W.header(6, 0) + W.records(rng, 6, 1)]
           + W.stream(rng, 4, 3, 48, 1_440_000, 10_000))
    out["malformed"] = (1, 0, bad, {"truncate": 9})
    out["legacy"] = (1, None, W.stream(rng, 0, 5, 96, 0, 10_000), {})
    ties = W.stream(rng, 2, 4, 96, 0, 10_000) + W.stream(rng, 0, 4, 96, 0, 10_000) + \
        W.stream(rng, 7, 2, 30, 0, 10_000)
    out["ties"] = (1, 0, ties, {})
    paths = {}
    for name, (dev, prof, dg, kw) in out.items():
        p = tmp_path / f"{name}.lvxr"
        W.write(p, dev, prof, dg, **kw)
        paths[name] = p
    return paths


@pytest.mark.parametrize("name", ["type0", "type1", "type2", "type3", "type4", "type5", "type7", "type8", "dual",
                                  "gps_fast", "clock", "malformed", "legacy", "ties"])
def test_decode_matches_oracle(lw, gpu_ok, tmp_path, name):
    p = _files(tmp_path)[name]
    f = lw.lvxr.read_firings(p, "cuda:0")
    dev, prof, t, xyz, refl, tag, valid, d = X.read(p.read_bytes())
    assert (f.dev_type, f.profile) == (dev, prof)
    assert np.array_equal(f.t.cpu().numpy(), t)
    spherical = name in ("type1", "type3", "type5", "type8") or name == "dual"
    gx, gd = f.xyz.cpu().numpy(), f.direction.cpu().numpy()
    assert gx.shape == xyz.shape
    if spherical:
        assert np.abs(gx - xyz).max() <= 2e-6 * max(1.0, np.abs(xyz).max())
    else:
        assert np.array_equal(gx, xyz)
    assert np.array_equal(f.refl.cpu().numpy(), refl) and np.array_equal(f.tag.cpu().numpy(), tag)
    assert np.array_equal(f.valid.cpu().numpy(), valid)
    assert np.array_equal(np.isnan(gd), np.isnan(d))
    ok = ~np.isnan(d)
    assert np.abs(gd[ok] - d[ok]).max(initial=0.0) <= 2e-6
    for clean in (True, False):
        px, pa, pt, pf = (a.cpu().numpy() for a in f.points(clean=clean))
        ox, oa, ot, of = X.points(t, xyz, refl, tag, valid, clean=clean)
        if not spherical:  # spherical returns can sit on the min-range edge within rounding
            assert np.array_equal(pf, of) and np.array_equal(pa, oa) and np.array_equal(pt, ot)
            assert np.array_equal(px, ox)


def test_replay_through_odometry(lw, gpu_ok, tmp_path):
    """A synthetic walk written as extended-cartesian packets (millimetres, as a sensor sends them), read back
    on the device and pushed through the Odometry batch by batch: it tracks the walk."""
    import struct

    seq = lw.synth.Sequence(seed=21, seconds=4.0)
    batch = 2000
    dgrams = []
    for b in range(int(4.0 * seq.rate / batch)):
        xyz, attr, t = seq.batch(b * batch / seq.rate, batch, b)
        mm = np.round(xyz * 1000).astype(np.int32)
        for a in range(0, batch, 96):  # 96 records a packet, stamped with the first record's time
            sl = slice(a, min(a + 96, batch))
            body = b"".join(struct.pack("<iiiBB", *m.tolist(), int(r) & 255, (int(r) >> 8) & 255)
                            for m, r in zip(mm[sl], attr[sl]))
            dgrams.append(W.header(2, int(round(float(t[sl][0]) * 1e9))) + body)
    p = tmp_path / "walk.lvxr"
    W.write(p, 1, 0, dgrams)
    f = lw.lvxr.read_firings(p, "cuda:0")
    sx, sa, st64, _ = f.points(clean=True, min_range=0.3)
    st = st64.to(torch.float32)
    n = sx.shape[0]
    odo = lw.Odometry(lw.PoseTable(), "cuda:0", reg_slots=1 << 20, frame_cap=1 << 15)
    odo.reg_voxel, odo.max_iter = 0.2, 20
    odo.reset(seq.pose(0.05))
    t_host = st64.cpu().numpy()
    for a in range(0, n, batch):
        b = min(a + batch, n)
        odo.push(sx[a:b].contiguous(), sa[a:b].contiguous(), st[a:b].contiguous(), b - a, float(t_host[a]),
                 float(t_host[b - 1]), lambda *_: None)
    torch.cuda.synchronize()
    traj = odo.traj
    T0 = seq.pose(traj[0][1])
    err = np.array([np.linalg.norm((np.linalg.inv(traj[0][2]) @ T)[:3, 3] - (np.linalg.inv(T0) @ seq.pose(tm))[:3, 3])
                    for _, tm, T in traj])
    rmse = float(np.sqrt(np.mean(err ** 2)))
    print(f"{n} points replayed, {len(traj)} frames, rmse {rmse * 100:.2f} cm, final {err[-1] * 100:.2f} cm")
    assert len(traj) >= 30 and rmse < 0.1


def test_dual_known_answer(lw, gpu_ok, tmp_path):
    """check_mid40_dual.py:33-41's expectations on the device reader: 200 returns, return ids 0, 1, 0, 1, ...,
    both returns of a firing at one time, the second packet at 0.5 ms, and each return mask keeps 100."""
    p = tmp_path / "dual.lvxr"
    W.dual_known_answer(p)
    f = lw.lvxr.read_firings(p, "cuda:0")
    assert f.profile == 1
    xyz, attr, t, _ = (a.cpu().numpy() for a in f.points(clean=False))
    assert len(t) == 200
    ret = (attr >> 16) & 255
    assert np.array_equal(ret, np.tile([0, 1], 100))
    assert np.array_equal(t[::2], t[1::2]) and np.isclose(t[100], 0.0005, rtol=1e-9, atol=0)
    assert all(int(((ret == 0) & (mask & 1 > 0) | (ret == 1) & (mask & 2 > 0)).sum()) == want
               for mask, want in ((3, 200), (1, 100), (2, 100)))


def test_recordings_without_points_raise(lw, gpu_ok, tmp_path):
    """A recording of IMU packets only, or a header alone, holds no firings: read_firings raises."""
    rng = np.random.default_rng(3)
    p = tmp_path / "imu.lvxr"
    W.write(p, 1, 0, [W.header(6, 1000 * k) + W.records(rng, 6, 1) for k in range(5)])
    with pytest.raises(ValueError):
        lw.lvxr.read_firings(p, "cuda:0")
    q = tmp_path / "empty.lvxr"
    W.write(q, 1, 0, [])
    with pytest.raises(ValueError):
        lw.lvxr.read_firings(q, "cuda:0")
    with pytest.raises(ValueError):
        lw.lvxr._framing(b"LVXR9\n\x01\x00")  # not a recording

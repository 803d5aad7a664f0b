"""The drop-in Odometry on a seeded office walk, fed batch by batch like the reference's tests feed
Pipeline.stage -> Odometry.push (check_odometry.py:44-51), gated like check_odometry.py:92-97."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rot_angle(R):
    return np.degrees(np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1)))


def run_sequence(lw, seconds=8.0, seed=21, batch=2000, **attrs):
    seq = lw.synth.Sequence(seed=seed, seconds=seconds)
    pipe = lw.PoseTable()
    odo = lw.Odometry(pipe, "cuda:0", reg_slots=1 << 20, frame_cap=1 << 15)
    odo.reg_voxel = 0.2
    odo.max_iter = 20
    for k, v in attrs.items():
        setattr(odo, k, v)
    odo.reset(seq.pose(0.05))
    frames = []

    def on_frame(k, T, fx, fa, ft, ff, m):
        frames.append((k, m, dict(odo.stats)))

    n_batches = int(seconds * seq.rate / batch)
    for b in range(n_batches):
        xyz, attr, t = seq.batch(b * batch / seq.rate, batch, b)
        sx = torch.from_numpy(xyz).cuda()
        sa = torch.from_numpy(attr.view(np.int32)).cuda()
        st = torch.from_numpy(t).cuda()
        odo.push(sx, sa, st, batch, float(t[0]), float(t[-1]), on_frame)
    torch.cuda.synchronize()
    return seq, odo, frames


def test_walk_tracks_ground_truth(lw, gpu_ok):
    seq, odo, frames = run_sequence(lw)
    traj = odo.traj
    assert len(traj) >= 70
    T0 = seq.pose(traj[0][1])
    err_t, err_r, dist = [], [], 0.0
    for (k, tm, T), nxt in zip(traj, traj[1:] + [None]):
        gt = np.linalg.inv(T0) @ seq.pose(tm)
        est = np.linalg.inv(traj[0][2]) @ T
        err_t.append(np.linalg.norm(est[:3, 3] - gt[:3, 3]))
        err_r.append(rot_angle(est[:3, :3].T @ gt[:3, :3]))
        if nxt is not None:
            dist += np.linalg.norm(seq.pose(nxt[1])[:3, 3] - seq.pose(tm)[:3, 3])
    err_t, err_r = np.array(err_t), np.array(err_r)
    post = [s for _, _, s in frames[odo.warmup + 2:]]
    reg = np.mean([s["iters"] > 0 and s["corr"] >= odo.min_corr for s in post])
    rmse = float(np.sqrt(np.mean(err_t ** 2)))
    print(f"walk {dist:.1f} m: rmse {rmse * 100:.2f} cm, final {err_t[-1] * 100:.2f} cm, "
          f"rot rmse {np.sqrt(np.mean(err_r ** 2)):.3f} deg, registered {reg:.0%}, map {odo.stats['map_voxels']}")
    assert reg > 0.9 and dist > 3.0
    assert rmse < 0.15 and err_t[-1] < 0.02 * dist + 0.15 and np.sqrt(np.mean(err_r ** 2)) < 2.0


def test_pose_table_and_callbacks(lw, gpu_ok):
    seq, odo, frames = run_sequence(lw, seconds=2.0)
    ks = [k for k, _, _ in frames]
    assert ks == sorted(ks) and len(ks) >= 15
    for k, tm, T in odo.traj:
        P, owner, epoch = odo.pipe.pose(k)
        assert owner == k and np.allclose(P, T.astype(np.float32))
    assert all(m > 0 for _, m, _ in frames)


def test_frame_chain_matches_step_by_step(lw, gpu_ok):
    """The one-chain frame (lw_odom_frame: the re-deskew and keyframe decisions taken on the device) against
    the step-by-step path over a walk. Two runs build their maps with float atomics whose order varies, so
    the trajectories agree to within the run-to-run spread (a fraction of a centimetre), not bit for bit; the
    decisions (registrations, keyframes) agree frame for frame."""
    runs = {}
    for chain in (True, False):
        seq, odo, frames = run_sequence(lw, seconds=4.0, chain=chain)
        runs[chain] = (odo, frames)
    (oa, fa), (ob, fb) = runs[True], runs[False]
    assert len(oa.traj) == len(ob.traj) >= 35
    worst = max(float(np.abs(Ta[:3, 3] - Tb[:3, 3]).max()) for (_, _, Ta), (_, _, Tb) in zip(oa.traj, ob.traj))
    print(f"chain vs step-by-step: {len(oa.traj)} frames, worst position difference {worst * 1e3:.3f} mm, "
          f"keyframes {oa.stats['keyframes']}/{ob.stats['keyframes']}, registrations "
          f"{oa.totals['registrations']}/{ob.totals['registrations']}")
    assert worst < 0.01
    assert oa.stats["keyframes"] == ob.stats["keyframes"] and oa.totals["registrations"] == ob.totals["registrations"]
    for (ka, ma, sa), (kb, mb, sb) in zip(fa, fb):
        assert ka == kb and ma == mb and sa["skipped"] == sb["skipped"] and sa["keyframes"] == sb["keyframes"]


def test_frame_chain_decisions(lw, gpu_ok):
    """lw_odom_frame on one engine state against the same steps taken one by one with the host's decisions:
    the re-deskew fires, the second registration starts from the first's pose, the keyframe test fires, and
    the chain's pose, twist and decisions are the step-by-step ones."""
    import ctypes

    from importlib import import_module

    od = import_module("livox-warp_amd.odom")
    w = lw.synth.single_scan(seed=16)
    voxel = 0.25
    dev = torch.device("cuda", 0)
    sx = torch.from_numpy(np.asarray(w.scan_xyz, np.float32).reshape(-1, 3)).to(dev)
    sa = torch.from_numpy(np.asarray(w.scan_attr).astype(np.uint32).view(np.int32)).to(dev)
    st = torch.from_numpy(np.asarray(w.scan_t, np.float32)).to(dev)
    eng = lw.Engine(1 << 21, 1 << 15)
    eng.insert(w.map_xyz, None, np.eye(4), voxel, max_range=1e30)
    eng.frame_reset()
    eng.frame_append(sx, sa, st, 10.0, -1, 0)
    p = eng.params(reg_voxel=voxel, max_iter=20)
    xi0 = np.zeros(6, np.float32)
    T_prev = w.T_true.copy()
    T_prev[:3, 3] -= np.array([0.06, 0.0, 0.0])  # 6 cm behind: the re-deskew test fires
    T_key = np.eye(4)
    # step by step, on the host's decisions
    eng.deskew(-1, 0.05, 10.0, xi0)
    r1 = eng.register(-1, p, w.T_init)
    T1 = np.array(r1.T).reshape(4, 4)
    xi_new = od._body_twist(np.linalg.inv(T_prev) @ T1)
    assert r1.ok and (np.linalg.norm(xi_new[:3] - xi0[:3]) > od.REDESKEW_TRANS
                      or np.linalg.norm(xi_new[3:] - xi0[3:]) > od.REDESKEW_ROT)
    eng.deskew(-1, 0.05, 10.0, xi_new)
    r2 = eng.register(-1, p, T1)
    T2 = np.array(r2.T).reshape(4, 4)
    d = np.linalg.inv(T_key) @ T2
    key_host = bool(r2.ok) and (np.linalg.norm(d[:3, 3]) >= 0.05 or
                                np.degrees(np.linalg.norm(od._rotvec(d))) >= 1.0)
    # the chain from the same state (nothing was inserted above)
    v0 = eng.stats()[0]
    fi, fo = lw._abi.FrameIn(), lw._abi.FrameOut()
    np.ctypeslib.as_array(fi.T_pred)[:] = np.asarray(w.T_init, np.float64).ravel()
    np.ctypeslib.as_array(fi.T_prev_inv)[:] = np.linalg.inv(T_prev).ravel()
    np.ctypeslib.as_array(fi.T_key_inv)[:] = np.linalg.inv(T_key).ravel()
    np.ctypeslib.as_array(fi.xi)[:] = xi0
    fi.kf_dist, fi.kf_angle = 0.05, 1.0
    fi.redeskew_rot, fi.redeskew_trans = od.REDESKEW_ROT, od.REDESKEW_TRANS
    fi.t_mid, fi.inv_dt = 0.05, 10.0
    fi.min_points, fi.has_prev, fi.has_key, fi.redeskew = 100, 1, 1, 1
    lw._abi.check(eng.lib.lw_odom_frame(eng.ctx, ctypes.byref(p), ctypes.byref(fi), ctypes.byref(fo)),
                  "lw_odom_frame")
    assert fo.m == len(w.scan_t) and not fo.skipped and fo.registrations == 2
    assert np.array_equal(np.array(fo.xi, np.float32), xi_new)
    assert np.array_equal(np.array(fo.reg[0].T).reshape(4, 4), T1)
    assert np.array_equal(np.array(fo.T).reshape(4, 4), T2) and bool(fo.ok) == bool(r2.ok)
    assert fo.reg[1].iters == r2.iters and fo.reg[1].corr == r2.corr
    assert bool(fo.keyframe) == key_host and fo.keyframe
    assert fo.map_voxels == eng.stats()[0] and fo.map_voxels >= v0  # the counts after the chain's insert
    eng.close()

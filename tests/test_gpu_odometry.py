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

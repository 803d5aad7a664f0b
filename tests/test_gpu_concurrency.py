"""Two contexts at once. The persistent registration loop is a plain launch whose workgroups wait on each other,
which needs them resident together; its grid is at most half of what the device holds, so it becomes resident
beside any other work once that work's workgroups retire. Here it runs while batched-localisation launches fill
the device from another stream, and two odometry contexts step alternately on two streams. Every launch
finishes (no LW_ELOOP), the localisation results equal a launch alone bit for bit, and every registration of
each odometry equals the oracle's on the same inputs (its thinned points, start pose and the device's own map
at that point) at the parity bar. (Whole trajectories are not compared: two runs alone already differ by
millimetres after an ill-conditioned frame, from the map's float atomics.)"""

import numpy as np
import pytest
import seqcheck
import torch

pytestmark = pytest.mark.gpu

FRAMES = 60


@pytest.fixture(scope="module")
def walk(lw):
    seq = lw.synth.Sequence(seed=21, seconds=60.0)
    X, A, T = seq.frames_device(FRAMES, "cuda:0")
    torch.cuda.synchronize()
    return seq, X, A, T


def _odometry(lw, seq, stream):
    odo = lw.Odometry(lw.PoseTable(), "cuda:0", reg_slots=1 << 20, frame_cap=1 << 15, stream=stream)
    odo.reg_voxel, odo.max_iter = 0.2, 20
    odo.reset(seq.pose(0.05))
    return odo


def _push(odo, rec, X, A, T, k):
    per = X.shape[1]
    odo.push(X[k], A[k], T[k], per, float(T[k, 0]), float(T[k, -1]), rec.on_frame)


def _replayed(orc, odo, rec):
    """Every registration the odometry ran, replayed by the oracle: the worst (m, deg) and the count."""
    odo.synchronize()
    out, _ = seqcheck.replay(orc, rec.events, seqcheck.params(orc, odo), 1 << 20)
    errs = np.array([e for _, e in out.values()])
    return (float(errs[:, 0].max()), float(errs[:, 1].max())), len(out)


def test_odometry_beside_localisation_launches(lw, orc, walk, gpu_ok):
    seq, X, A, T = walk
    ss = lw.synth.single_scan(seed=5, n=6000, map_points=150_000)
    scans = [ss.scan_xyz[:3000], ss.scan_xyz[3000:6000]]
    rng = np.random.default_rng(3)
    hyps = np.zeros((2, 1024, 4, 4))
    for q in range(2):
        for h in range(1024):
            hyps[q, h] = np.eye(4)
            hyps[q, h, :3, :3] = ss.T_true[:3, :3]
            hyps[q, h, :3, 3] = ss.T_true[:3, 3] + rng.uniform(-0.3, 0.3, 3)
    loc = lw.loc.BatchLocalizer(ss.map_xyz, voxel=0.25, reg_slots=1 << 20, chunk=1 << 18)
    try:
        staged = loc._stage(scans, hyps)
        tile = (512, 8, 7)  # as many workgroups as the device holds
        alone = [x.clone() for x in loc.run(scans, hyps, tile=tile, staged=staged)]
        loc.stream.synchronize()
        odo = _odometry(lw, seq, torch.cuda.Stream())
        rec = seqcheck.Recorder(lw, odo)
        results = []
        try:
            for k in range(FRAMES):
                if k % 5 == 0:  # localisation launches queued on their own stream, running under the next frames
                    for _ in range(3):
                        results.append([x.clone() for x in loc.run(scans, hyps, tile=tile, staged=staged)])
                _push(odo, rec, X, A, T, k)
            worst, n = _replayed(orc, odo, rec)
            frames = len(odo.traj)
        finally:
            odo.close()
        loc.stream.synchronize()
    finally:
        loc.close()
    for r in results:
        for a, b in zip(r, alone):
            assert torch.equal(a, b)
    print(f"odometry beside {len(results)} localisation launches: {frames} frames, {n} registrations vs the oracle: "
          f"max {worst[0] * 1e3:.4f} mm / {worst[1]:.5f} deg", flush=True)
    assert frames == FRAMES - 1 and n >= FRAMES - 12 and worst[0] < 1e-3 and worst[1] < 0.01


def test_two_odometries_on_two_streams(lw, orc, walk, gpu_ok):
    seq, X, A, T = walk
    odos = [_odometry(lw, seq, torch.cuda.Stream()) for _ in range(2)]
    recs = [seqcheck.Recorder(lw, o) for o in odos]
    try:
        for k in range(FRAMES):  # each frame pushed to both, one after the other
            for o, r in zip(odos, recs):
                _push(o, r, X, A, T, k)
        res = [_replayed(orc, o, r) for o, r in zip(odos, recs)]
        frames = [len(o.traj) for o in odos]
    finally:
        for o in odos:
            o.close()
    worst = (max(r[0][0] for r in res), max(r[0][1] for r in res))
    print(f"two odometries at once: {[r[1] for r in res]} registrations vs the oracle: max {worst[0] * 1e3:.4f} mm / "
          f"{worst[1]:.5f} deg", flush=True)
    assert frames == [FRAMES - 1] * 2 and min(r[1] for r in res) >= FRAMES - 12
    assert worst[0] < 1e-3 and worst[1] < 0.01

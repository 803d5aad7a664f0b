"""The batched odometry (BatchOdometry, lw_batch_frame): S independent sequences stepped together, each frame's
whole chain -- deskew, thinning, the registrations, the re-deskew test, the keyframe decision, the insert into
that sequence's map -- one launch per step for the batch, a sequence's registration in one workgroup.

Checked as the single-stream path is (the sequence tests): every registration of two watched sequences equals
the oracle's on the same inputs (the sequence's thinned points, its start pose, its map as the device holds it),
the frames' thinning equals the oracle's on the device's deskewed points, the warm-up, keyframe and skip
decisions are the single-stream Odometry's on the same frames, and each sequence tracks its ground truth."""

import numpy as np
import pytest
import seqcheck
import torch

pytestmark = pytest.mark.gpu

S = 6
FRAMES = 80


@pytest.fixture(scope="module")
def run(lw, orc):
    seq = lw.synth.Sequence(seed=21, seconds=60.0)
    frames = [seq.frames_device(FRAMES, "cuda:0", seed=100 + s) for s in range(S)]  # same walk, own noise
    X = torch.stack([f[0] for f in frames]).transpose(0, 1).contiguous()  # (frame, sequence, point, 3)
    A = torch.stack([f[1] for f in frames]).transpose(0, 1).contiguous()
    Tt = torch.stack([f[2] for f in frames]).transpose(0, 1).contiguous()
    # sequence 2's frame 30 is cut to 150 points: a skipped frame inside a batch
    M = np.full((FRAMES, S), X.shape[2], np.int32)
    M[30, 2] = 150
    bo = lw.BatchOdometry(S, "cuda:0", reg_slots=1 << 18, frame_cap=1 << 15)
    bo.reg_voxel, bo.max_iter = 0.2, 20
    rec = seqcheck.BatchRecorder(bo, [0, 2])
    try:
        bo.reset(np.stack([seq.pose(0.05)] * S))
        outs = []
        for k in range(FRAMES):
            bo.step(X[k], A[k], Tt[k], k, M[k])
            rec.after_step(X[k], A[k], Tt[k], k)
            outs.append(bo.last.copy())
        P = seqcheck.params(orc, bo)
        traj = [list(t) for t in bo.traj]
        totals = dict(bo.totals)
    finally:
        bo.close()
    return seq, X, A, Tt, M, rec, outs, P, traj, totals


def test_batch_registrations_match_the_oracle(run, orc, gpu_ok):
    seq, X, A, Tt, M, rec, outs, P, traj, totals = run
    worst, n, cond_rel = (0.0, 0.0), 0, 0.0
    for s, ev in rec.events.items():
        out, _ = seqcheck.replay(orc, ev, P, 1 << 20)
        errs = np.array([e for _, e in out.values()])
        worst = (max(worst[0], errs[:, 0].max()), max(worst[1], errs[:, 1].max()))
        n += len(out)
        # the condition number the registration kernel computes from its last H (odom.py:797-800)
        cond_rel = max(cond_rel, max(abs(ev[i][5] - res.cond) / res.cond for i, (res, _) in out.items()))
    print(f"batch: {n} registrations of 2 watched sequences vs the oracle: max {worst[0] * 1e3:.4f} mm / "
          f"{worst[1]:.5f} deg, condition numbers within {cond_rel:.2e} relative; totals {totals}", flush=True)
    assert n >= 2 * (FRAMES - 12) and worst[0] < 1e-3 and worst[1] < 0.01 and cond_rel < 1e-3


def test_batch_thinning_and_deskew_match_the_oracle(run, orc, gpu_ok):
    seq, X, A, Tt, M, rec, outs, P, traj, totals = run
    for xyz, attr, ds in rec.thin_checks:
        assert np.array_equal(orc.ds_insert(xyz, attr, 2.0 / 0.2, 0.3, 60.0, 0b1111, 1 << 15), ds)
    worst = max(float(np.abs(orc.deskew(raw, t, t_mid, 10.0, xi) - xyz).max())
                for raw, t, t_mid, xi, xyz in rec.deskew_checks[::5])
    assert worst < 1e-5 and len(rec.thin_checks) >= 2 * (FRAMES - 12)


def test_batch_decisions_match_the_single_stream_odometry(run, lw, gpu_ok):
    """Sequence 2 (with its skipped frame) through the single-stream Odometry, frame by frame: the same warm-up
    inserts, keyframes, skips and registration counts."""
    seq, X, A, Tt, M, rec, outs, P, traj, totals = run
    odo = lw.Odometry(lw.PoseTable(), "cuda:0", reg_slots=1 << 18, frame_cap=1 << 15)
    odo.reg_voxel, odo.max_iter = 0.2, 20
    rows = []
    try:
        odo.reset(seq.pose(0.05))
        for k in range(FRAMES):
            m = int(M[k, 2])
            regs0, key0 = odo.totals["registrations"], odo.stats["keyframes"]
            odo.push(X[k, 2, :m].contiguous(), A[k, 2, :m].contiguous(), Tt[k, 2, :m].contiguous(), m,
                     float(Tt[k, 2, 0]), float(Tt[k, 2, m - 1]),
                     lambda *f: rows.append((f[0], f[-1], odo.stats["keyframes"], odo.stats["skipped"])))
    finally:
        odo.close()
    # the single-stream odometry finalises frame k when frame k + 1 arrives: rows[k] is frame k
    keys_b = np.cumsum([o["keyframe"][2] for o in outs])
    skip_b = np.cumsum([o["skipped"][2] for o in outs])
    for k, (kk, m, keys, skips) in enumerate(rows):
        assert kk == k and m == outs[k]["m"][2]
        assert keys == keys_b[k] and skips == skip_b[k], (k, keys, keys_b[k], skips, skip_b[k])
    assert skip_b[-1] == 1 and len(rows) == FRAMES - 1


def test_batch_sequences_track_their_ground_truth(run, gpu_ok):
    seq, X, A, Tt, M, rec, outs, P, traj, totals = run
    for s in range(S):
        et, er, dist = seqcheck.gt_errors(seq, traj[s])
        rmse = float(np.sqrt((et ** 2).mean()))
        print(f"sequence {s}: {dist:.1f} m, rmse {rmse * 100:.2f} cm, final {et[-1] * 100:.2f} cm", flush=True)
        assert dist > 3.0 and rmse < 0.15 and et[-1] < 0.02 * dist + 0.15
    assert totals["skipped"] == 1

"""Whole sequences through the drop-in Odometry on its default path (one device chain per frame), against the
oracle and the reference's ground-truth gates:

  * the plain odometry sequence (600 frames of the seeded office walk, seed 21): every registration the chain
    runs is registered again by the oracle on the same inputs -- the chain's thinned points (read back after
    the frame; the first registration's set is rebuilt by the same deskew and thinning), its start pose, and a
    map kept in lock step by inserting the device's own keyframe points at the device's poses -- at 1 mm /
    0.01 deg; every frame's thinning equals the oracle's on the device's deskewed points; the final map equals
    the oracle's voxel for voxel; and the trajectory passes the reference's walk gates (check_odometry.py:92-97)
    over their 20 s, and a wider final-error gate over the whole 50 m;
  * the reference's static and turn cases (check_odometry.py:89-92) in the office;
  * the lifecycle paths against OracleOdometry fed the same batches: a gap over max_gap (odom.py:605-610), a
    frame under min_points (odom.py:667-669), a jump over 1.5 m resetting the velocity (odom.py:684-694), and a
    64-slot map that overflows (odom.py:225-230, 698-702).
"""

import numpy as np
import pytest
import torch
from seqcheck import Recorder as _Recorder
from seqcheck import err as _err
from seqcheck import gt_errors as _gt_errors
from seqcheck import odometry as _odometry
from seqcheck import params as _params
from seqcheck import read_map as _read_map
from seqcheck import replay as _replay

pytestmark = pytest.mark.gpu

C2_FRAMES = 600


@pytest.fixture(scope="module")
def c2(lw, orc):
    seq = lw.synth.Sequence(seed=21, seconds=60.0)
    X, A, Tt = seq.frames_device(C2_FRAMES, "cuda:0")
    per = X.shape[1]
    odo = _odometry(lw)
    rec = _Recorder(lw, odo)
    odo.reset(seq.pose(0.05))
    for k in range(C2_FRAMES):
        odo.push(X[k], A[k], Tt[k], per, float(Tt[k, 0]), float(Tt[k, -1]), rec.on_frame)
    odo.synchronize()
    keys, rsum, router, rcnt = _read_map(odo)
    stats = dict(odo.stats)
    traj = list(odo.traj)
    P = _params(orc, odo)
    odo.close()
    out, om = _replay(orc, rec.events, P, 1 << 22)
    return seq, rec, out, om, (keys, rsum, router, rcnt), stats, traj


def test_c2_every_registration_matches_the_oracle(c2, gpu_ok):
    seq, rec, out, om, gmap, stats, traj = c2
    errs = np.array([e for _, e in out.values()])
    frames = {rec.events[i][4] for i in out}
    print(f"C2: {len(traj)} frames, {len(out)} registrations on {len(frames)} frames vs the oracle: max "
          f"{errs[:, 0].max() * 1e3:.4f} mm / {errs[:, 1].max():.5f} deg, rms {np.sqrt((errs[:, 0] ** 2).mean()) * 1e3:.4f}"
          f" mm", flush=True)
    assert len(traj) == C2_FRAMES - 1 and len(frames) >= C2_FRAMES - 40
    assert len(out) >= len(frames)
    assert errs[:, 0].max() < 1e-3 and errs[:, 1].max() < 0.01


def test_c2_thinning_and_deskew_match_the_oracle(c2, orc, gpu_ok):
    seq, rec, out, om, gmap, stats, traj = c2
    worst = 0.0
    for xyz, attr, ds in rec.thin_checks:
        ref = orc.ds_insert(xyz, attr, 2.0 / 0.2, 0.3, 60.0, 0b1111, 1 << 15)
        assert np.array_equal(ref, ds)
    for raw, t, t_mid, xi, xyz in rec.deskew_checks[::7]:
        ref = orc.deskew(raw, t, t_mid, 10.0, xi)
        worst = max(worst, float(np.abs(ref - xyz).max()))
    print(f"C2 thinning: {len(rec.thin_checks)} frames bit-exact; deskew worst {worst * 1e6:.3f} um", flush=True)
    assert worst < 1e-5


def test_c2_map_matches_the_oracle(c2, gpu_ok):
    seq, rec, out, om, gmap, stats, traj = c2
    keys, rsum, router, rcnt = gmap
    occ = keys != -1
    assert np.array_equal(occ, om.keys != -1)  # the same occupied slots in a table of the same size
    gk = {int(k): s for s, k in enumerate(keys) if k != -1}
    ok = {int(k): s for s, k in enumerate(om.keys) if k != -1}
    assert gk.keys() == ok.keys() and len(gk) == stats["map_voxels"] > 20_000
    gs = np.array([gk[k] for k in ok]), np.array([ok[k] for k in ok])
    assert np.array_equal(rcnt[gs[0]], om.rcnt[gs[1]])
    c = om.rcnt[gs[1]].astype(np.float64)[:, None]
    mean_g, mean_o = rsum[gs[0]] / c, om.rsum[gs[1]] / c
    assert np.abs(mean_g - mean_o).max() < 1e-5 * 0.2
    cov_g = router[gs[0]].reshape(-1, 9) / c
    cov_o = om.router[gs[1]].reshape(-1, 9) / c
    scale = np.abs(cov_o).max(axis=1, keepdims=True) + 1e-12
    assert (np.abs(cov_g - cov_o) / scale).max() < 1e-4
    print(f"C2 map: {len(gk)} voxels, counts exact, means within {np.abs(mean_g - mean_o).max() * 1e6:.3f} um",
          flush=True)


def test_c2_ground_truth_gates(c2, gpu_ok):
    """check_odometry.py:92-97's walk gates over the first 20 s (the length of the reference's walk case): rms
    under 15 cm, the final error within 2% of the distance + 15 cm, the rotation rms under 2 deg. Over the whole
    50 m the final error is reported and held to 3% of the distance + 15 cm: it varies from run to run with the
    float-atomic order of the map's moments (0.60-1.07 m measured over this sequence), which the reference's
    2% gate, written for 20 s, does not allow for."""
    seq, rec, out, om, gmap, stats, traj = c2
    et, er, dist = _gt_errors(seq, traj)
    et20, er20, dist20 = _gt_errors(seq, traj[:200])
    rmse, rot = float(np.sqrt((et ** 2).mean())), float(np.sqrt((er ** 2).mean()))
    rmse20, rot20 = float(np.sqrt((et20 ** 2).mean())), float(np.sqrt((er20 ** 2).mean()))
    print(f"C2 drift: {dist:.1f} m, rmse {rmse:.3f} m, final {et[-1]:.3f} m ({et[-1] / dist:.2%}), rot rmse "
          f"{rot:.3f} deg; first 20 s: {dist20:.1f} m, rmse {rmse20:.3f} m, final {et20[-1]:.3f} m, rot rmse "
          f"{rot20:.3f} deg", flush=True)
    assert dist20 > 10.0 and rmse20 < 0.15 and et20[-1] < 0.02 * dist20 + 0.15 and rot20 < 2.0
    assert dist > 40.0 and et[-1] < 0.03 * dist + 0.15 and rot < 2.0
    assert stats["map_overflow"] == 0 and stats["skipped"] == 0


@pytest.mark.parametrize("motion,seconds", [("static", 6.0), ("turn", 15.0)])
def test_reference_static_and_turn_cases(lw, motion, seconds, gpu_ok):
    """check_odometry.py:89-92: static rms < 2 cm and worst rotation < 0.2 deg; turn rms < 10 cm and rotation rms
    < 0.5 deg over a swing of more than 20 deg; registration after the warm-up on more than 90% of frames."""
    seq = lw.synth.Sequence(seed=21, seconds=seconds, motion=motion)
    n = int(round(seconds / 0.1))
    X, A, Tt = seq.frames_device(n, "cuda:0")
    odo = _odometry(lw, 1 << 20)
    stats = []
    odo.reset(seq.pose(0.05))
    try:
        for k in range(n):
            odo.push(X[k], A[k], Tt[k], X.shape[1], float(Tt[k, 0]), float(Tt[k, -1]),
                     lambda *_: stats.append(dict(odo.stats)))
        traj = list(odo.traj)
    finally:
        odo.close()
    et, er, _ = _gt_errors(seq, traj)
    span = max(_err(seq.pose(traj[0][1]), seq.pose(tm))[1] for _, tm, _ in traj)
    reg = np.mean([s["iters"] > 0 and s["corr"] >= 60 for s in stats[8 + 2:]])
    rmse = float(np.sqrt((et ** 2).mean()))
    print(f"{motion}: {len(traj)} frames, rmse {rmse * 100:.2f} cm, rot max {er.max():.3f} deg, rot rmse "
          f"{np.sqrt((er ** 2).mean()):.3f} deg, span {span:.1f} deg, registered {reg:.0%}", flush=True)
    assert reg > 0.9
    if motion == "static":
        assert rmse < 0.02 and er.max() < 0.2
    else:
        assert span > 20.0 and rmse < 0.10 and np.sqrt((er ** 2).mean()) < 0.5


# ---- lifecycle against OracleOdometry ---------------------------------------------------------------------


def _lifecycle_run(lw, orc, frames, reg_slots, events):
    """The same whole-frame batches into the device Odometry and OracleOdometry; `events` maps a batch index to
    an action taken on both before that batch is pushed: "drop" (the batch is never pushed: a gap), "sparse"
    (only 150 of its points), "jump" (the previous pose is moved 3 m sideways, as after a jump of that size)."""
    dev = _odometry(lw, reg_slots)
    rec = _Recorder(lw, dev)
    ora = orc.OracleOdometry(lw.PoseTable(), reg_slots=reg_slots, frame_cap=1 << 15)
    ora.reg_voxel, ora.max_iter = 0.2, 20
    rows = {"dev": [], "ora": []}
    ora_inserts = []
    try:
        for od, name in ((dev, "dev"), (ora, "ora")):
            od.reset(np.eye(4))
            for b, (x, a, t) in enumerate(frames):
                ev = events.get(b)
                if ev == "drop":
                    continue
                if ev == "sparse":
                    x, a, t = x[:150], a[:150], t[:150]
                if ev == "jump" and od.T_prev is not None:
                    od.T_prev = od.T_prev.copy()
                    od.T_prev[:3, 3] += [0.0, 3.0, 0.0]  # sideways: the walk itself cannot cancel it
                def cb(*f, od=od, name=name):
                    prev = rows[name][-1][2]["keyframes"] if rows[name] else 0
                    rows[name].append((f[0], f[-1], dict(od.stats), od.delta.copy(), od.map_frames))
                    if name == "dev":
                        rec.on_frame(*f)
                    elif od.stats["keyframes"] > prev:  # the oracle inserted this frame
                        m = f[-1]
                        ora_inserts.append((f[2][:m].copy(), f[3][:m].copy(), f[1].copy()))
                if name == "dev":
                    od.push(torch.from_numpy(x).cuda(), torch.from_numpy(a.view(np.int32)).cuda(),
                            torch.from_numpy(t).cuda(), len(t), float(t[0]), float(t[-1]), cb)
                else:
                    od.push(x, a, t, len(t), float(t[0]), float(t[-1]), cb)
        dev.synchronize()
        traj_d, traj_o = list(dev.traj), list(ora.traj)
        stats_d = dict(dev.stats)
        counts_d = dev.map_stats()
        keys_d = _read_map(dev)[0]
    finally:
        dev.close()
    ora.inserts = ora_inserts
    return traj_d, traj_o, rows, stats_d, counts_d, ora, rec, keys_d


def _frames(lw, seconds, seed=21):
    seq = lw.synth.Sequence(seed=seed, seconds=seconds + 1)
    return seq, list(seq.frames(int(round(seconds / 0.1))))


def test_lifecycle_gap_sparse_and_jump_match_the_oracle(lw, orc, gpu_ok):
    """The same batches, with a frame of 150 points (frame 20), a 3 m jump of the previous pose (seen by frame 29)
    and 25 missing frames (45-69), into both classes. The decisions are the oracle's frame for frame (frame ids,
    skips, gaps, keyframes, the velocity reset after the jump), every device registration equals the oracle's on
    the same inputs (the replay of the sequence test), and the two trajectories agree to 1 mm / 0.01 deg until the
    jump. After it they are compared only through the per-registration replay: the jumped frame's re-deskew uses a
    3 m twist and its registration runs to the iteration cap, so the rounding of two deskews (float32 trigonometry)
    moves thinned points across voxel faces and the runs part by millimetres, as two device runs would."""
    seq, frames = _frames(lw, 7.5)
    events = {20: "sparse", 30: "jump"}
    events.update({b: "drop" for b in range(45, 70)})  # 25 frames missing: more than max_gap (20)
    traj_d, traj_o, rows, stats_d, counts_d, ora, rec, _ = _lifecycle_run(lw, orc, frames, 1 << 20, events)
    assert [k for k, _, _ in traj_d] == [k for k, _, _ in traj_o]
    so = ora.stats
    assert stats_d["gaps"] == so["gaps"] == 1
    assert stats_d["skipped"] == so["skipped"] == 1
    # the jump is made before batch 30 is pushed, which finalizes frame 29: frames up to 28 never see it
    before = [i for i, (k, _, _) in enumerate(traj_d) if k < 29]
    for i in before:
        rd, ro = rows["dev"][i], rows["ora"][i]
        assert rd[0] == ro[0] and rd[1] == ro[1] and rd[4] == ro[4]
        assert rd[2]["skipped"] == ro[2]["skipped"] and rd[2]["keyframes"] == ro[2]["keyframes"]
        # the two maps are built apart, and the device's moments are float atomics whose order varies from run to
        # run (two device runs of this sequence end a frame with residuals 2e-4 apart): a step's size can fall on
        # either side of the convergence test (1e-4 rad and 1e-4 m), so the counts may part by one step, the
        # velocities by up to two such steps, and the next frame's deskew by as much, which can carry a point
        # across a thinning voxel's face; the replay below holds every device registration to the oracle's on
        # the device's own inputs
        assert abs(rd[2]["ds"] - ro[2]["ds"]) <= 2 and abs(rd[2]["iters"] - ro[2]["iters"]) <= 1
        assert np.abs(rd[3] - ro[3]).max() < 2e-4
    pairs = [_err(traj_d[i][2], traj_o[i][2]) for i in before]
    worst = (max(e[0] for e in pairs), max(e[1] for e in pairs))
    # frame 29's velocity update sees a 3 m move and resets the model, on both
    j = next(i for i, r in enumerate(rows["ora"]) if r[0] == 29)
    assert np.array_equal(rows["ora"][j][3], np.eye(4)) and np.array_equal(rows["dev"][j][3], np.eye(4))
    assert not np.array_equal(rows["ora"][j - 1][3], np.eye(4))
    # every registration the device ran, on its own inputs, through the oracle
    out, _ = _replay(orc, rec.events, _params(orc, rec.odo), 1 << 20)
    errs = np.array([e for _, e in out.values()])
    print(f"lifecycle: {len(traj_d)} frames, gaps {stats_d['gaps']}, skipped {stats_d['skipped']}, keyframes "
          f"{stats_d['keyframes']}/{so['keyframes']}; before the jump {worst[0] * 1e3:.4f} mm / {worst[1]:.5f} deg; "
          f"{len(out)} registrations vs the oracle: max {errs[:, 0].max() * 1e3:.4f} mm / {errs[:, 1].max():.5f} deg",
          flush=True)
    assert worst[0] < 1e-3 and worst[1] < 0.01
    assert len(out) >= 40 and errs[:, 0].max() < 1e-3 and errs[:, 1].max() < 0.01


def _overflow_of(orc, inserts, keyset, voxel=0.2, max_range=60.0, ret_mask=0b1111):
    """Points of the inserts that pass the insert's filters and whose voxel key is not in keyset: with every slot
    taken, exactly these find no slot within the probe window (odom.py:225-230)."""
    n = 0
    for xyz, attr, T in inserts:
        x = np.asarray(xyz, np.float32)
        r = np.sqrt((x[:, 0] * x[:, 0] + x[:, 1] * x[:, 1]) + x[:, 2] * x[:, 2])
        ret = (attr >> 16) & 255
        ok = (r <= np.float32(max_range)) & (((attr >> 8) & 255) == 0) & (ret < 31) & \
             (((np.int64(1) << np.minimum(ret, 30).astype(np.int64)) & ret_mask) != 0)
        keys = orc.voxel_key_np(orc.transform_points_f32(T, x[ok]), 1.0 / voxel)
        n += int((~np.isin(keys, np.fromiter(keyset, np.int64))).sum())
    return n


def test_lifecycle_full_map_overflows_like_the_oracle(lw, orc, gpu_ok):
    """A 64-slot map fills on the first insert. Which 64 keys win is the claim order (index order in the oracle,
    the device's race on the GPU), so each side is checked against its own key set: its overflow counter equals
    the filtered points of its inserts whose key it does not hold, and both hold 64 voxels."""
    seq, frames = _frames(lw, 1.2)
    traj_d, traj_o, rows, stats_d, counts_d, ora, rec, keys_d = _lifecycle_run(lw, orc, frames, 64, {})
    assert [k for k, _, _ in traj_d] == [k for k, _, _ in traj_o]
    assert counts_d[0] == int(ora.map.stats[0]) == 64 and (keys_d != -1).sum() == 64
    ins_d = [(x, a, T) for kind, x, a, T in (e for e in rec.events if e[0] == "insert")]
    assert len(ins_d) == stats_d["keyframes"] >= 8
    over_d = _overflow_of(orc, ins_d, {int(k) for k in keys_d if k != -1})
    assert stats_d["map_overflow"] == counts_d[1] == over_d > 0
    over_o = _overflow_of(orc, ora.inserts, {int(k) for k in ora.map.keys if k != -1})
    assert len(ora.inserts) == ora.stats["keyframes"] >= 8 and int(ora.map.stats[1]) == over_o > 0
    print(f"64-slot map: overflow device {counts_d[1]} / oracle {over_o}", flush=True)

"""Precision study: the office sequence translated by (1000, -750, 30) m, run with the three accumulation forms
(float64, compensated float32, naive float32; residuals float32 in all three) and compared frame by frame with
the all-float64 oracle (every device-side step in float64, the reference's numpy host step)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

OFFSET = (1000.0, -750.0, 30.0)
FRAMES = 60


def _pose_err(A, B):
    d = np.linalg.inv(A) @ B
    ang = np.degrees(np.arccos(np.clip((np.trace(d[:3, :3]) - 1) / 2, -1, 1)))
    return float(np.linalg.norm(A[:3, 3] - B[:3, 3])), float(ang)


@pytest.fixture(scope="module")
def translated(lw, orc):
    seq = lw.synth.Sequence(seed=21, seconds=FRAMES / 10 + 1, offset=OFFSET)
    frames = list(seq.frames(FRAMES))
    od = orc.OracleOdometry(lw.PoseTable(), reg_slots=1 << 20, frame_cap=1 << 15, fp64=True)
    od.reg_voxel = 0.2
    od.max_iter = 20
    od.reset(seq.pose(0.05))
    for x, a, t in frames:
        od.push(x, a, t, len(t), float(t[0]), float(t[-1]))
    return seq, frames, {k: T for k, _, T in od.traj}


@pytest.mark.parametrize("mode,bar_m,bar_deg", [("fp64", 0.01, 0.1), ("kahan32", 0.01, 0.1), ("fp32", 1.0, 10.0)])
def test_forms_against_fp64_oracle(lw, translated, mode, bar_m, bar_deg, gpu_ok):
    seq, frames, ref = translated
    odo = lw.Odometry(lw.PoseTable(), "cuda:0", reg_slots=1 << 20, frame_cap=1 << 15)
    odo.reg_voxel = 0.2
    odo.max_iter = 20
    odo.acc_mode = mode
    odo.reset(seq.pose(0.05))
    for x, a, t in frames:
        odo.push(torch.from_numpy(x).cuda(), torch.from_numpy(a.view(np.int32)).cuda(), torch.from_numpy(t).cuda(),
                 len(t), float(t[0]), float(t[-1]), lambda *_: None)
    errs = np.array([_pose_err(T, ref[k]) for k, _, T in odo.traj if k in ref])
    print(f"{mode}: vs float64 oracle, per frame max {errs[:, 0].max() * 1e3:.3f} mm / {errs[:, 1].max():.4f} deg, "
          f"rms {np.sqrt(np.mean(errs[:, 0] ** 2)) * 1e3:.3f} mm")
    assert len(errs) >= FRAMES - 2
    assert errs[:, 0].max() < bar_m and errs[:, 1].max() < bar_deg

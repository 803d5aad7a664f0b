"""The odometry worker (OdomWorker: the reference's slam_worker.py contract) on the device: batches submitted
from the caller's stream come back as frames with the same poses, point counts and deskewed points as the
Odometry driven directly; a reset drops what is queued for the old generation and restarts at its pose;
a full input queue and an exhausted output pool drop and count (own-or-drop), as check_slam_worker.py asks
of the reference."""

import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _batches(lw, seconds=3.0, seed=21, batch=2000):
    seq = lw.synth.Sequence(seed=seed, seconds=seconds)
    out = []
    for b in range(int(seconds * seq.rate / batch)):
        xyz, attr, t = seq.batch(b * batch / seq.rate, batch, b)
        out.append((xyz, attr.view(np.int32), t))
    return seq, out


def _configure(odo):
    odo.reg_voxel = 0.2
    odo.max_iter = 20


def _direct(lw, seq, batches):
    odo = lw.Odometry(lw.PoseTable(), "cuda:0", reg_slots=1 << 20, frame_cap=1 << 15)
    _configure(odo)
    odo.reset(seq.pose(0.05))
    frames = []

    def on_frame(k, T, fx, fa, ft, ff, m):
        frames.append((k, T.copy(), m, fx[:m].cpu().numpy().copy()))

    for xyz, attr, t in batches:
        odo.push(torch.from_numpy(xyz).cuda(), torch.from_numpy(attr).cuda(), torch.from_numpy(t).cuda(), len(t),
                 float(t[0]), float(t[-1]), on_frame)
    torch.cuda.synchronize()
    return frames


def _wait_idle(w, timeout=60.0):
    t0 = time.time()
    while not w.idle() and time.time() - t0 < timeout:
        time.sleep(0.01)
    assert w.idle(), w.status()


def test_worker_matches_direct(lw, gpu_ok):
    seq, batches = _batches(lw)
    ref = _direct(lw, seq, batches)
    w = lw.worker.OdomWorker(lw.PoseTable(), "cuda:0", frame_cap=1 << 15, reg_slots=1 << 20, out_slots=256)
    try:
        w.configure(reg_voxel=0.2, max_iter=20)
        w.reset(seq.pose(0.05))
        got = []
        for xyz, attr, t in batches:  # staged on the caller's stream, as the render thread does
            sx, sa, st = torch.from_numpy(xyz).cuda(), torch.from_numpy(attr).cuda(), torch.from_numpy(t).cuda()
            assert w.submit(sx, sa, st, len(t), float(t[0]), float(t[-1]))
            got += w.poll()
        _wait_idle(w)
        got += w.poll()
        torch.cuda.current_stream().synchronize()
        st = w.status()
        print(f"{len(got)} frames from the worker, {len(ref)} direct; {st}")
        assert st["drops"] == {"input_full": 0, "input_own": 0, "output_pool": 0, "stale": 0, "error": 0}
        assert len(got) == len(ref) >= 20
        worst = 0.0
        for r, (k, T, m, fx) in zip(got, ref):
            assert r.gen == 1 and r.k == k and r.m == m
            worst = max(worst, float(np.abs(r.T[:3, 3] - T[:3, 3]).max()))
            xyz = r.arrays()[0].cpu().numpy()
            assert xyz.shape == (m, 3) and np.abs(xyz - fx).max() < 5e-3
        print(f"worst position difference {worst * 1e3:.3f} mm")
        assert worst < 0.01  # two maps built with float atomics in different orders
        w.release_consumed()
    finally:
        w.close()


def test_worker_generations_and_drops(lw, gpu_ok):
    seq, batches = _batches(lw, seconds=2.0)
    w = lw.worker.OdomWorker(lw.PoseTable(), "cuda:0", frame_cap=1 << 15, reg_slots=1 << 20, max_batches=2,
                             out_slots=1)
    try:
        w.configure(reg_voxel=0.2, max_iter=20)
        w.reset(seq.pose(0.05))
        sent = 0
        for xyz, attr, t in batches:  # faster than the worker: the queue of 2 fills
            sx, sa, st = torch.from_numpy(xyz).cuda(), torch.from_numpy(attr).cuda(), torch.from_numpy(t).cuda()
            sent += bool(w.submit(sx, sa, st, len(t), float(t[0]), float(t[-1])))
        _wait_idle(w)
        st = w.status()
        print(st)
        assert st["drops"]["input_full"] > 0 and st["drops"]["input_full"] + sent == len(batches)
        assert st["processed"] == sent
        # one output slot per array and nothing released: the second frame finds the pool empty
        assert st["drops"]["output_pool"] > 0
        res = w.poll()
        assert res and all(r.gen == 1 for r in res)
        w.release_consumed()
        # a reset: results still queued for generation 1 are dropped, the next frames start from the new pose
        T0 = seq.pose(0.05).copy()
        T0[:3, 3] += np.array([10.0, 0.0, 0.0])
        w.reset(T0)
        for xyz, attr, t in batches[:6]:
            sx, sa, st_ = torch.from_numpy(xyz).cuda(), torch.from_numpy(attr).cuda(), torch.from_numpy(t).cuda()
            w.submit(sx, sa, st_, len(t), float(t[0]), float(t[-1]))
            _wait_idle(w)
        res = w.poll()
        assert res and all(r.gen == 2 for r in res)
        assert np.abs(res[0].T[:3, 3] - T0[:3, 3]).max() < 0.5  # the frame after the reset is posed near T0
        w.release_consumed()
        with pytest.raises(AttributeError):
            w.configure(no_such_tuning=1)
    finally:
        w.close()
    assert not w._thread.is_alive()

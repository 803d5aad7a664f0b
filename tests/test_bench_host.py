"""The benchmark's host side without a GPU: more than one GPU is refused with a message before anything
touches a device, and the synthetic LVXR recording the replay leg writes reads back through the oracle's
reader as the points it was written from (millimetre coordinates, reflectivity, tags, times on the packet
clock)."""

import os
import subprocess
import sys

import numpy as np

import lvxr_oracle as X

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_replay_recording_round_trip(tmp_path):
    sys.path.insert(0, ROOT)
    import bench

    rng = np.random.default_rng(5)
    n = 96 * 40 + 17  # a partial last packet is not written
    xyz = rng.uniform(-20, 20, (n, 3)).astype(np.float32)
    attr = (rng.integers(0, 256, n) | (rng.integers(0, 4, n) << 8)).astype(np.uint32)
    t = (np.arange(n) * 1e-5 + 3.0).astype(np.float32)
    p = tmp_path / "w.lvxr"
    m = bench._write_lvxr(p, xyz, attr, t)
    assert m == 96 * 40
    dev, prof, tt, x, refl, tag, valid, d = X.read(p.read_bytes())
    assert (dev, prof) == (1, 0) and len(tt) == m
    # times: the base is the first stamp; each packet's stamp is its first point's time, 10 us a firing after
    assert np.abs(tt - (t[:m].astype(np.float64) - float(t[0]))).max() < 1e-6
    assert np.abs(x[:, 0] - xyz[:m]).max() <= 5e-4 + 1e-6
    assert np.array_equal(refl[:, 0], attr[:m] & 255) and np.array_equal(tag[:, 0], (attr[:m] >> 8) & 255)

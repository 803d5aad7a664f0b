"""The pose sink Odometry writes to: frame id -> sensor-to-world pose and body twist.

Mirrors the pose part of livox_warp.gpu.Pipeline (python/livox_warp/gpu.py:1317-1344): set_pose and
set_pose_range stage host-side edits under a lock (the odometry may run on a worker thread), keyed by
frame id modulo POSE_SLOTS, with the owning (frame, epoch) recorded per slot.
"""

from __future__ import annotations

import threading

import numpy as np

POSE_SLOTS = 1 << 16


class PoseTable:
    def __init__(self):
        self._lock = threading.Lock()
        self.poses = np.tile(np.eye(4, dtype=np.float32), (POSE_SLOTS, 1, 1))
        self.twist = np.zeros((POSE_SLOTS, 6), np.float32)
        self.owner = np.full((POSE_SLOTS, 2), -1, np.int64)  # (frame, epoch)
        self.dirty: set[int] = set()

    def set_pose(self, frame: int, mat: np.ndarray, twist=None, epoch: int = 0):
        j = frame & (POSE_SLOTS - 1)
        m = np.asarray(mat, dtype=np.float32)
        tw = np.zeros(6, np.float32) if twist is None else np.asarray(twist, dtype=np.float32)
        with self._lock:
            self.poses[j] = m
            self.twist[j] = tw
            self.owner[j] = (frame, epoch)
            self.dirty.add(j)

    def set_pose_range(self, k0: int, k1: int, mat: np.ndarray, twist=None, epoch: int = 0):
        if k1 <= k0:
            return
        k0 = max(k0, k1 - POSE_SLOTS)
        ks = np.arange(k0, k1, dtype=np.int64)
        js = ks & (POSE_SLOTS - 1)
        m = np.asarray(mat, dtype=np.float32)
        tw = np.zeros(6, np.float32) if twist is None else np.asarray(twist, dtype=np.float32)
        with self._lock:
            self.poses[js] = m
            self.twist[js] = tw
            self.owner[js, 0] = ks
            self.owner[js, 1] = epoch
            self.dirty.update(js.tolist())

    def drain(self):
        """The slots edited since the last drain, with their poses, and the edit set cleared: the host half of
        Pipeline.flush_poses (gpu.py:1346-1365), for a device consumer to upload."""
        with self._lock:
            if not self.dirty:
                return np.zeros(0, np.int32), np.zeros((0, 4, 4), np.float32)
            idx = np.fromiter(self.dirty, dtype=np.int32, count=len(self.dirty))
            self.dirty.clear()
            return idx, self.poses[idx].copy()

    def pose(self, frame: int):
        """(4x4 float32 pose, frame id that wrote it, epoch) of a frame id's slot."""
        j = frame & (POSE_SLOTS - 1)
        with self._lock:
            return self.poses[j].copy(), int(self.owner[j, 0]), int(self.owner[j, 1])

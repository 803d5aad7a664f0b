"""Many independent odometry sequences on one GPU at once: the throughput form of the odometry.

A single sequence is a chain of dependent frames, each a few hundred microseconds of latency-bound work on a
device that could run hundreds of them side by side. BatchOdometry runs S sequences in lock step: one call to
step() takes one whole frame of every sequence through Odometry._finalize's work (python/livox_warp/odom.py:
642-727) -- deskew with the constant-velocity twist, thinning, registration, the re-deskew test and second
registration, the keyframe test, the insert into that sequence's own map -- as one launch per step for the
whole batch (lw_batch_frame: a sequence's registration is one workgroup running every Gauss-Newton pass),
with one upload and one readback. The host bookkeeping of odom.py:642-717 (warm-up, prediction, velocity
model, keyframes) runs over all sequences at once in numpy.

Frames come whole: frame k of every sequence is already cut from its stream (the reference's frame cutting,
k_frame_append, is streaming work that Odometry.push does per batch). The maps are independent: sequence s
owns context s (reg_slots slots), readable between steps with any per-context entry point.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _abi
from .odom import ACC_MODES, MAX_FRAME_JUMP, REDESKEW_ROT, REDESKEW_TRANS, WIDE_ITERS


def _torch():
    import torch

    return torch


def rotvec_batch(R: np.ndarray) -> np.ndarray:
    """Rotation vectors of rotations R (..., 3, 3) by the method of odom.py's host helper (scipy's
    Rotation.from_matrix / as_rotvec: the quaternion from the largest of R00, R11, R22 and the trace, w >= 0,
    then 2 atan2(|v|, w) with a series below 1e-3), over a batch."""
    R = np.asarray(R, np.float64)
    r00, r01, r02 = R[..., 0, 0], R[..., 0, 1], R[..., 0, 2]
    r10, r11, r12 = R[..., 1, 0], R[..., 1, 1], R[..., 1, 2]
    r20, r21, r22 = R[..., 2, 0], R[..., 2, 1], R[..., 2, 2]
    tr = r00 + r11 + r22
    top = np.maximum(np.maximum(r00, r11), np.maximum(r22, tr))
    c0, c1, c2 = r00 == top, (r11 == top) & (r00 != top), (r22 == top) & (r00 != top) & (r11 != top)
    c3 = ~(c0 | c1 | c2)
    x = np.select([c0, c1, c2, c3], [1.0 - tr + 2.0 * r00, r01 + r10, r02 + r20, r21 - r12])
    y = np.select([c0, c1, c2, c3], [r10 + r01, 1.0 - tr + 2.0 * r11, r12 + r21, r02 - r20])
    z = np.select([c0, c1, c2, c3], [r20 + r02, r21 + r12, 1.0 - tr + 2.0 * r22, r10 - r01])
    w = np.select([c0, c1, c2, c3], [r21 - r12, r02 - r20, r10 - r01, 1.0 + tr])
    n = np.sqrt(x * x + y * y + z * z + w * w)
    n = np.where(w < 0.0, -n, n)
    x, y, z, w = x / n, y / n, z / n, w / n
    ang = 2.0 * np.arctan2(np.sqrt(x * x + y * y + z * z), w)
    a2 = ang * ang
    small = ang <= 1e-3
    scale = np.where(small, 2.0 + a2 / 12.0 + 7.0 * a2 * a2 / 2880.0,
                     ang / np.sin(np.where(small, 1.0, ang) / 2.0))
    return np.stack([scale * x, scale * y, scale * z], axis=-1)


def rotmat_batch(v: np.ndarray) -> np.ndarray:
    """Rotations of rotation vectors v (..., 3) through the unit quaternion (scipy's from_rotvec / as_matrix
    method, as odom.py's host helper), over a batch."""
    v = np.asarray(v, np.float64)
    ang = np.sqrt((v * v).sum(-1))
    a2 = ang * ang
    small = ang <= 1e-3
    scale = np.where(small, 0.5 - a2 / 48.0 + a2 * a2 / 3840.0, np.sin(ang / 2.0) / np.where(small, 1.0, ang))
    x, y, z, w = scale * v[..., 0], scale * v[..., 1], scale * v[..., 2], np.cos(ang / 2.0)
    x2, y2, z2, w2 = x * x, y * y, z * z, w * w
    xy, zw, xz, yw, yz, xw = x * y, z * w, x * z, y * w, y * z, x * w
    return np.stack([np.stack([x2 - y2 - z2 + w2, 2.0 * (xy - zw), 2.0 * (xz + yw)], -1),
                     np.stack([2.0 * (xy + zw), -x2 + y2 - z2 + w2, 2.0 * (yz - xw)], -1),
                     np.stack([2.0 * (xz - yw), 2.0 * (yz + xw), -x2 - y2 + z2 + w2], -1)], -2)


class BatchOdometry:
    """S odometry sequences stepped together, frame by frame."""

    def __init__(self, n_seq: int, device=None, reg_slots: int = 1 << 18, frame_cap: int = 1 << 15, stream=None):
        torch = _torch()
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        assert reg_slots & (reg_slots - 1) == 0
        self.S = int(n_seq)
        self.reg_cap = reg_slots
        self.frame_cap = frame_cap
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        self._lib = _abi.lib()
        L = self._lib
        self._ctx = []
        self._bufs = []
        sp = ctypes.c_void_p(self.stream.cuda_stream)
        with torch.cuda.device(self.device):
            for _ in range(self.S):
                c = ctypes.c_void_p()
                _abi.check(L.lw_create(ctypes.byref(c), reg_slots, frame_cap, sp), "lw_create")
                self._ctx.append(c)
                b = (torch.zeros((frame_cap, 3), dtype=torch.float32, device=self.device),
                     torch.zeros((frame_cap, 3), dtype=torch.float32, device=self.device),
                     torch.zeros(frame_cap, dtype=torch.int32, device=self.device),
                     torch.zeros(frame_cap, dtype=torch.float32, device=self.device),
                     torch.zeros(frame_cap, dtype=torch.int32, device=self.device))
                _abi.check(L.lw_bind_frame(c, *(t.data_ptr() for t in b)), "lw_bind_frame")
                self._bufs.append(b)
            arr = (ctypes.c_void_p * self.S)(*[c.value for c in self._ctx])
            self._b = ctypes.c_void_p()
            _abi.check(L.lw_batch_create(ctypes.byref(self._b), arr, self.S, sp), "lw_batch_create")
        self._fin = (_abi.FrameIn * self.S)()
        self._fout = (_abi.FrameOut * self.S)()
        self._src = (_abi.BatchSrc * self.S)()
        self._vin = np.ctypeslib.as_array(self._fin)
        self._vout = np.ctypeslib.as_array(self._fout)
        self._vsrc = np.ctypeslib.as_array(self._src)
        self._params = _abi.RegParams()

        # tuning (odom.py:499-525), shared by every sequence
        self.frame_dt = 0.1
        self.reg_voxel = 0.15
        self.max_iter = 16
        self.warmup = 8
        self.min_points = 200
        self.min_corr = 60
        self.min_range = 0.3
        self.max_range = 60.0
        self.noise = 0.02
        self.ret_mask = 0b1111
        self.alpha = 0.05
        self.damping = 1.0e-3
        self.velocity_smooth = 0.85
        self.degen_thr = 3000.0
        self.prior_info = 2000.0
        self.kf_dist = 0.05
        self.kf_angle = 1.0
        self.deskew = True
        self.redeskew = True
        self.acc_mode = "fp64"
        self.wide_iters = WIDE_ITERS
        self.reset(np.broadcast_to(np.eye(4), (self.S, 4, 4)))

    # ---- lifecycle ---------------------------------------------------------------------------------------

    def close(self):
        """Wait for the batch's work and free it and its contexts (idempotent)."""
        b, self._b = getattr(self, "_b", None), None
        if b is not None and b.value:
            _abi.check(self._lib.lw_batch_destroy(b), "lw_batch_destroy")
        ctxs, self._ctx = getattr(self, "_ctx", []), []
        for c in ctxs:
            if c.value:
                _abi.check(self._lib.lw_destroy(c), "lw_destroy")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def context(self, s: int):
        """Sequence s's context (its map and buffers), for the per-context entry points."""
        return self._ctx[s]

    def reset(self, T0):
        """Forget every map and trajectory; sequence s's next frame is posed at T0[s] (odom.py:561-579)."""
        for c in self._ctx:
            _abi.check(self._lib.lw_map_reset(c), "lw_map_reset")
        S = self.S
        self.T = np.array(T0, dtype=np.float64).reshape(S, 4, 4).copy()
        self.T_prev = np.tile(np.eye(4), (S, 1, 1))
        self.has_prev = np.zeros(S, bool)
        self.delta = np.tile(np.eye(4), (S, 1, 1))
        self.xi = np.zeros((S, 6), np.float32)
        self.xi_s = np.zeros((S, 6))
        self.T_key = np.tile(np.eye(4), (S, 1, 1))
        self.has_key = np.zeros(S, bool)
        self.map_frames = np.zeros(S, np.int64)
        self.traj = [[] for _ in range(S)]
        self.totals = {"frames": 0, "registrations": 0, "gn_iters": 0, "keyframes": 0, "skipped": 0}
        self.last = None

    def reg_params(self) -> _abi.RegParams:
        p = self._params
        p.damping, p.degen_thr, p.prior_info = self.damping, self.degen_thr, self.prior_info
        p.reg_voxel, p.min_range, p.max_range = self.reg_voxel, self.min_range, self.max_range
        p.noise, p.alpha, p.ret_mask = self.noise, self.alpha, int(self.ret_mask)
        p.max_iter, p.min_corr, p.wide_iters = int(self.max_iter), int(self.min_corr), int(self.wide_iters)
        p.acc_mode = ACC_MODES[self.acc_mode]
        return p

    # ---- one frame of every sequence ---------------------------------------------------------------------

    def step(self, xyz, attr, t, k, m=None):
        """Frame k of every sequence: xyz (S, P, 3) float32, attr (S, P) int32 (the uint32 attribute bits), t
        (S, P) float32 -- device tensors, each sequence's frame in the first m[s] rows (default P). Returns the
        S poses (S, 4, 4)."""
        S = self.S
        P = xyz.shape[1]
        m = np.full(S, P, np.int32) if m is None else np.asarray(m, np.int32)
        for a in (xyz, attr, t):
            if not a.is_contiguous() or a.device != self.device or a.shape[0] != S:
                raise ValueError("step() takes contiguous (S, P, ...) device tensors")
        warm = self.map_frames < self.warmup
        t_mid = (k + 0.5) * self.frame_dt
        # the prediction and the deskew twist (odom.py:659-666): a warm-up frame keeps the pose, still
        T_pred = np.where(warm[:, None, None], self.T, self.T @ self.delta)
        xi0 = np.where((warm | (not self.deskew))[:, None], 0.0, self.xi).astype(np.float32)
        vin, src = self._vin, self._vsrc
        vin["T_pred"][:] = T_pred.reshape(S, 16)
        has_prev = self.has_prev & ~warm
        vin["has_prev"][:] = has_prev
        vin["T_prev_inv"][:] = np.linalg.inv(self.T_prev).reshape(S, 16)
        vin["has_key"][:] = self.has_key
        vin["T_key_inv"][:] = np.linalg.inv(self.T_key).reshape(S, 16)
        vin["xi"][:] = xi0
        vin["kf_dist"], vin["kf_angle"] = float(self.kf_dist), float(self.kf_angle)
        vin["redeskew_rot"], vin["redeskew_trans"] = REDESKEW_ROT, REDESKEW_TRANS
        vin["t_mid"], vin["inv_dt"] = float(t_mid), 1.0 / self.frame_dt
        vin["min_points"] = int(self.min_points)
        vin["redeskew"] = 1 if (self.deskew and self.redeskew) else 0
        base = np.arange(S, dtype=np.uint64)
        src["xyz"] = np.uint64(xyz.data_ptr()) + base * np.uint64(xyz.stride(0) * 4)
        src["attr"] = np.uint64(attr.data_ptr()) + base * np.uint64(attr.stride(0) * 4)
        src["t"] = np.uint64(t.data_ptr()) + base * np.uint64(t.stride(0) * 4)
        src["m"] = m
        src["warm"] = warm
        _abi.check(self._lib.lw_batch_frame(self._b, ctypes.byref(self.reg_params()), self._fin, self._src,
                                            self._fout), "lw_batch_frame")
        out = self._vout
        # a frame under min_points keeps the prediction (odom.py:667-669); the device leaves its pose unset
        T = np.where(out["skipped"].astype(bool)[:, None, None], T_pred, out["T"].reshape(S, 4, 4))
        key = out["keyframe"].astype(bool)
        regs = out["registrations"]
        self.totals["frames"] += S
        self.totals["registrations"] += int(regs.sum())
        self.totals["gn_iters"] += int(np.where(regs >= 1, out["reg"]["iters"][:, 0], 0).sum()
                                       + np.where(regs >= 2, out["reg"]["iters"][:, 1], 0).sum())
        self.totals["keyframes"] += int(key.sum())
        self.totals["skipped"] += int(out["skipped"].sum())
        # odom.py:648-657, 719-727: keyframes, the velocity model (not on warm-up frames), the pose
        self.map_frames += key
        self.T_key = np.where(key[:, None, None], T, self.T_key)
        self.has_key |= key
        upd = self.has_prev & ~warm
        if upd.any():
            d = np.linalg.inv(self.T_prev) @ T
            moved = np.linalg.norm(d[:, :3, 3], axis=1) < MAX_FRAME_JUMP
            xi_d = np.concatenate([d[:, :3, 3], rotvec_batch(d[:, :3, :3])], 1).astype(np.float32)
            a = self.velocity_smooth
            keep = upd & moved
            reset = upd & ~moved
            xi_s = a * self.xi_s + (1.0 - a) * xi_d.astype(np.float64) if a > 0.0 else self.xi_s
            delta = d.copy()
            if a > 0.0:
                delta = np.tile(np.eye(4), (S, 1, 1))
                delta[:, :3, :3] = rotmat_batch(xi_s[:, 3:])
                delta[:, :3, 3] = xi_s[:, :3]
            self.delta = np.where(keep[:, None, None], delta, np.where(reset[:, None, None], np.eye(4), self.delta))
            self.xi = np.where(keep[:, None], xi_d, np.where(reset[:, None], 0.0, self.xi)).astype(np.float32)
            self.xi_s = np.where(keep[:, None], xi_s, np.where(reset[:, None], 0.0, self.xi_s))
        self.T_prev = T
        self.has_prev[:] = True
        self.T = T
        for s in range(S):
            self.traj[s].append((k, t_mid, T[s]))
        self.last = out
        return T

    # ---- measurement -------------------------------------------------------------------------------------

    def prof_enable(self, on: bool):
        _abi.check(self._lib.lw_batch_prof_enable(self._b, 1 if on else 0), "lw_batch_prof_enable")

    def prof_reset(self):
        _abi.check(self._lib.lw_batch_prof_reset(self._b), "lw_batch_prof_reset")

    def prof_read(self) -> np.ndarray:
        out = np.zeros(8)
        _abi.check(self._lib.lw_batch_prof_read(self._b, _abi.ptr(out, _abi._f64p)), "lw_batch_prof_read")
        return out

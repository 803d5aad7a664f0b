"""LiDAR-only scan-to-map odometry on MI355X: the drop-in counterpart of livox_warp.odom.Odometry.

Same constructor, tuning attributes, methods and read-back state as the reference class
(python/livox_warp/odom.py:458-803); the host bookkeeping (frame cutting, warm-up, keyframes, velocity,
re-deskew) is the reference's, line for line in behaviour. Every device step goes through the C ABI of
liblwamd.so on the instance's stream:

    push            -> lw_frame_reset / lw_frame_append / lw_frame_count      (odom.py:596-638)
    _deskew         -> lw_deskew                                               (odom.py:707-710)
    _register       -> lw_register: thinning + the whole Gauss-Newton loop on the device, one readback
                       (odom.py:731-803: the reference reads the accumulator back every iteration)
    _insert         -> lw_map_insert                                           (odom.py:719-727)
    reset           -> lw_map_reset                                            (odom.py:561-579)

Device arrays are torch tensors (ROCm): the staged batch passed to push() and the frame buffers handed to
on_frame (f_xyz float32 (cap, 3), f_attr int32 holding the uint32 attribute bits, f_t float32, f_frame
int32), borrowed until the next frame as in the reference.
"""

from __future__ import annotations

import math
import time

import numpy as np
from . import _abi

ACC_SLOTS = 64
ACC_STRIDE = 64
ACC_H, ACC_B, ACC_COUNT, ACC_ERR, ACC_RMS, ACC_WEIGHT, ACC_RANGE, ACC_HG = 0, 21, 27, 28, 29, 30, 31, 32
REG_CAP = 4096
WIDE_ITERS = 3
KAPPA_START, KAPPA_DECAY, KAPPA_MIN = 6.0, 0.7, 2.0
STEP_TOL = 1e-4
TIGHT_FRACTION = 0.6
WEAK_COND = 1e-3
REDESKEW_ROT = 2e-3
REDESKEW_TRANS = 5e-3
MAX_FRAME_JUMP = 1.5

ACC_MODES = {"fp64": 0, "kahan32": 1, "fp32": 2}


def _rotvec(R) -> tuple[float, float, float]:
    """Rotation vector of a 3x3 rotation (any indexable of rows), by scipy's Rotation.from_matrix /
    as_rotvec method -- the unit quaternion from the largest of (R00, R11, R22, trace), then
    2 atan2(|v|, w) -- in plain float arithmetic: a few microseconds per frame instead of scipy's object
    overhead. Agrees with scipy to a few ulp."""
    r00, r01, r02 = float(R[0][0]), float(R[0][1]), float(R[0][2])
    r10, r11, r12 = float(R[1][0]), float(R[1][1]), float(R[1][2])
    r20, r21, r22 = float(R[2][0]), float(R[2][1]), float(R[2][2])
    tr = r00 + r11 + r22
    top = max(r00, r11, r22, tr)
    if r00 == top:
        x, y, z, w = 1.0 - tr + 2.0 * r00, r10 + r01, r20 + r02, r21 - r12
    elif r11 == top:
        x, y, z, w = r01 + r10, 1.0 - tr + 2.0 * r11, r21 + r12, r02 - r20
    elif r22 == top:
        x, y, z, w = r02 + r20, r12 + r21, 1.0 - tr + 2.0 * r22, r10 - r01
    else:
        x, y, z, w = r21 - r12, r02 - r20, r10 - r01, 1.0 + tr
    n = math.sqrt(x * x + y * y + z * z + w * w)
    if w < 0.0:
        n = -n
    x, y, z, w = x / n, y / n, z / n, w / n
    ang = 2.0 * math.atan2(math.sqrt(x * x + y * y + z * z), w)
    if ang <= 1e-3:
        a2 = ang * ang
        scale = 2.0 + a2 / 12.0 + 7.0 * a2 * a2 / 2880.0
    else:
        scale = ang / math.sin(ang / 2.0)
    return scale * x, scale * y, scale * z


def _rotmat(wx: float, wy: float, wz: float) -> np.ndarray:
    """3x3 rotation of a rotation vector, by scipy's Rotation.from_rotvec / as_matrix method (through the
    unit quaternion), in plain float arithmetic."""
    ang = math.sqrt(wx * wx + wy * wy + wz * wz)
    if ang <= 1e-3:
        a2 = ang * ang
        scale = 0.5 - a2 / 48.0 + a2 * a2 / 3840.0
    else:
        scale = math.sin(ang / 2.0) / ang
    x, y, z, w = scale * wx, scale * wy, scale * wz, math.cos(ang / 2.0)
    x2, y2, z2, w2 = x * x, y * y, z * z, w * w
    xy, zw, xz, yw, yz, xw = x * y, z * w, x * z, y * w, y * z, x * w
    return np.array([[x2 - y2 - z2 + w2, 2.0 * (xy - zw), 2.0 * (xz + yw)],
                     [2.0 * (xy + zw), -x2 + y2 - z2 + w2, 2.0 * (yz - xw)],
                     [2.0 * (xz - yw), 2.0 * (yz + xw), -x2 - y2 + z2 + w2]])


def _twist_between(T_from: np.ndarray, T_to: np.ndarray) -> np.ndarray:
    M = T_to @ np.linalg.inv(T_from)
    return np.array([M[0, 3], M[1, 3], M[2, 3], *_rotvec(M)])


def _apply_twist(T: np.ndarray, delta: np.ndarray) -> np.ndarray:
    Rz = _rotmat(float(delta[3]), float(delta[4]), float(delta[5]))
    out = T.copy()
    out[:3, :3] = Rz @ T[:3, :3]
    out[:3, 3] = Rz @ T[:3, 3] + delta[:3]
    return out


def _body_twist(d: np.ndarray) -> np.ndarray:
    return np.array([d[0, 3], d[1, 3], d[2, 3], *_rotvec(d)], dtype=np.float32)


def _twist_matrix(xi: np.ndarray) -> np.ndarray:
    out = np.eye(4)
    out[:3, :3] = _rotmat(float(xi[3]), float(xi[4]), float(xi[5]))
    out[:3, 3] = xi[:3]
    return out


def _torch():
    import torch

    return torch


class Odometry:
    """Frame cutting, the registration map and scan-to-map ICP on one stream (odom.py:458-803)."""

    def __init__(self, pipe, device=None, reg_slots: int = 1 << 21, frame_cap: int = 1 << 18, stream=None):
        torch = _torch()
        self.pipe = pipe
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError(f"Odometry runs on a GPU device, got {self.device}")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        assert reg_slots & (reg_slots - 1) == 0
        self.reg_cap = reg_slots
        self.frame_cap = frame_cap
        d = self.device
        self.f_raw = torch.zeros((frame_cap, 3), dtype=torch.float32, device=d)  # as measured
        self.f_xyz = torch.zeros((frame_cap, 3), dtype=torch.float32, device=d)  # deskewed to mid-frame
        self.f_attr = torch.zeros(frame_cap, dtype=torch.int32, device=d)
        self.f_t = torch.zeros(frame_cap, dtype=torch.float32, device=d)
        self.f_frame = torch.zeros(frame_cap, dtype=torch.int32, device=d)
        self._lib = _abi.lib()
        self._ctx = ctypes_ctx = _abi.ctypes.c_void_p()
        with torch.cuda.device(d):
            _abi.check(self._lib.lw_create(_abi.ctypes.byref(ctypes_ctx), reg_slots, frame_cap,
                                           _abi.ctypes.c_void_p(self.stream.cuda_stream)), "lw_create")
        _abi.check(self._lib.lw_bind_frame(self._ctx, self.f_raw.data_ptr(), self.f_xyz.data_ptr(),
                                           self.f_attr.data_ptr(), self.f_t.data_ptr(), self.f_frame.data_ptr()),
                   "lw_bind_frame")
        self._params = _abi.RegParams()
        self._result = _abi.RegResult()
        # one frame's whole device chain (lw_odom_frame); numpy views onto its inputs
        self._fin = _abi.FrameIn()
        self._fout = _abi.FrameOut()
        self._fin_T = np.ctypeslib.as_array(self._fin.T_pred)
        self._fin_prev = np.ctypeslib.as_array(self._fin.T_prev_inv)
        self._fin_key = np.ctypeslib.as_array(self._fin.T_key_inv)
        self._fin_xi = np.ctypeslib.as_array(self._fin.xi)
        self.chain = True  # False: the step-by-step path (also taken when the device has no cooperative launch)
        self._i32 = (_abi.ctypes.c_int32 * 2)()

        # tuning (odom.py:499-525)
        self.frame_dt = 0.1
        self.reg_voxel = 0.15
        self.max_iter = 16
        self.warmup = 8
        self.max_gap = 20
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
        self.epoch = 0
        self.acc_mode = "fp64"  # accumulation form: fp64 (reference), kahan32, fp32
        self.wide_iters = WIDE_ITERS
        # learned pose-increment prior (prior.LearnedPrior) in place of the constant-velocity guess; None: off
        self.prior = None
        self._map_counts = None  # (voxels, overflow) as last read back; None after a map change
        self._inv_memo = []  # (pose, its inverse), the latest first (_inv)
        # running totals beside the reference's per-frame stats: registrations, Gauss-Newton steps, prior use
        self.totals = {"registrations": 0, "gn_iters": 0, "prior_used": 0, "prior_fallbacks": 0}

        self.reset(np.eye(4, dtype=np.float64))

    def close(self):
        """Wait for the instance's queued work and free its context (map, frame work buffers, pinned readback
        memory). Idempotent; the instance is unusable afterwards. Call it before the interpreter exits rather than
        leaving the context to the garbage collector during shutdown."""
        prior = getattr(self, "prior", None)
        if prior is not None and hasattr(prior, "close"):
            prior.close()  # its network shares the instance's stream
        ctx = getattr(self, "_ctx", None)
        self._ctx = None
        if ctx is not None and ctx.value:
            _abi.check(self._lib.lw_destroy(ctx), "lw_destroy")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- stream helpers --------------------------------------------------------------------

    def synchronize(self):
        _abi.check(self._lib.lw_synchronize(self._ctx), "lw_synchronize")

    def _frame_count(self) -> int:
        _abi.check(self._lib.lw_frame_count(self._ctx, self._i32), "lw_frame_count")
        return int(self._i32[0])

    def _zero_count(self):
        _abi.check(self._lib.lw_frame_reset(self._ctx), "lw_frame_reset")

    # ---- state -----------------------------------------------------------------------------

    @staticmethod
    def empty_stats() -> dict:
        return {"frames": 0, "iters": 0, "corr": 0, "ds": 0, "rms": 0.0, "cond": 1.0, "ms": 0.0,
                "weak": False, "map_voxels": 0, "map_full": 0.0, "map_overflow": 0, "skipped": 0, "warm": True,
                "speed": 0.0, "turn": 0.0, "gaps": 0, "degen": 0, "keyframes": 0}

    def reset(self, T0: np.ndarray):
        """Forget the map and the trajectory; the next frame is posed at T0."""
        _abi.check(self._lib.lw_map_reset(self._ctx), "lw_map_reset")
        self._map_counts = None
        self.T = np.array(T0, dtype=np.float64)
        self.T_prev = None
        self.delta = np.eye(4)
        self.xi = np.zeros(6, np.float32)
        self.xi_s = np.zeros(6)
        self.k_open = None
        self.k_pred = -1
        self.map_frames = 0
        self.T_key = None
        self.traj = []
        self.stats = self.empty_stats()
        self._zero_count()

    def restart_clock(self):
        self.k_open = None
        self.k_pred = -1
        self.T_prev = None
        self.delta = np.eye(4)
        self.xi = np.zeros(6, np.float32)
        self.xi_s = np.zeros(6)
        self._zero_count()

    def _inv(self, T: np.ndarray) -> np.ndarray:
        """np.linalg.inv(T), remembered for the last two poses asked for: the previous frame's pose and the last
        keyframe's are inverted once each (and once in all when they are the same pose), not at every use."""
        for M, Mi in self._inv_memo:
            if M is T or np.array_equal(M, T):
                return Mi
        Mi = np.linalg.inv(T)
        self._inv_memo = [(T.copy(), Mi), self._inv_memo[0]] if self._inv_memo else [(T.copy(), Mi)]
        return Mi

    def predict(self) -> np.ndarray:
        return self.T @ self.delta

    # ---- per batch -------------------------------------------------------------------------

    @staticmethod
    def _dev_ptr(a, dtype, n: int, width: int = 1) -> int:
        torch = _torch()
        if not isinstance(a, torch.Tensor) or a.device.type != "cuda":
            raise TypeError("push() takes device tensors")
        if a.dtype not in dtype or not a.is_contiguous() or a.numel() < n * width:
            raise ValueError(f"bad staged array: {a.dtype}, {tuple(a.shape)}")
        return a.data_ptr()

    def push(self, sx, sa, st, n: int, t_min: float, t_max: float, on_frame):
        """Feed one staged batch (points, attributes, timestamps: device tensors). on_frame(k, T, f_xyz, f_attr,
        f_t, f_frame, m) runs per completed frame, with the frame's points deskewed to mid-frame."""
        if n <= 0:
            return
        torch = _torch()
        p_x = self._dev_ptr(sx, (torch.float32,), n, 3)
        p_a = self._dev_ptr(sa, (torch.int32, torch.uint32), n)
        p_t = self._dev_ptr(st, (torch.float32,), n)
        inv_dt = 1.0 / self.frame_dt
        k_lo = int(math.floor(max(t_min, 0.0) * inv_dt))
        k_hi = int(math.floor(max(t_max, 0.0) * inv_dt))
        k_prev = self.k_pred
        if self.k_open is not None and k_hi - self.k_open > self.max_gap:
            if self._frame_count() > 0:
                self._finalize(self.k_open, on_frame)
            self.restart_clock()
            self.stats["gaps"] += 1
        if self.k_open is None:
            self.k_open = k_lo if k_hi - k_lo <= self.max_gap else k_hi
            self._zero_count()
            self.pipe.set_pose(self.k_open, self.T, self.xi, self.epoch)
            lo = (k_prev + 1) if 0 <= k_prev < self.k_open else k_lo
            if lo < self.k_open:
                self.pipe.set_pose_range(lo, self.k_open, self.T, None, self.epoch)
            self.k_pred = self.k_open
        for k in range(max(self.k_pred + 1, self.k_open + 1), k_hi + 1):
            self.pipe.set_pose(k, self.predict(), self.xi, self.epoch)
            self.k_pred = k
        k_min = self.k_open - 1
        while True:
            _abi.check(self._lib.lw_frame_append(self._ctx, p_x, p_a, p_t, n, inv_dt, k_min, self.k_open),
                       "lw_frame_append")
            if k_hi <= self.k_open:
                break
            self._finalize(self.k_open, on_frame)
            self.k_open += 1
            k_min = self.k_open
            self._zero_count()
            self.pipe.set_pose(self.k_open, self.predict(), self.xi, self.epoch)
            self.k_pred = max(self.k_pred, self.k_open)

    # ---- per frame -------------------------------------------------------------------------

    def _finalize(self, k: int, on_frame):
        t0 = time.perf_counter()
        warm = self.map_frames < self.warmup
        t_mid = (k + 0.5) * self.frame_dt
        chained = False
        # warm-up frames, and a prior that cannot take the device chain (one that only records), go step by step
        stepwise = warm or (self.prior is not None and not hasattr(self.prior, "chain_inputs"))
        if not stepwise and self.chain:  # deskew, registrations, keyframe test and insert queued at once
            res = self._frame_chain(t_mid)
            if res is not None:
                T, xi, ok, m = res
                chained = True
        if not chained:
            if stepwise or self.prior is not None:
                m = min(self._frame_count(), self.frame_cap)
                T, xi, ok = self._estimate(m, t_mid, warm)
            else:  # the frame's point count comes back with the registration: no read of it before
                T, xi, ok, m = self._estimate_counted(t_mid)
        if not chained and ok and (warm or self._is_keyframe(T)):
            self._insert(m, T)
        if self.T_prev is not None and not warm:
            self._update_velocity(self._inv(self.T_prev) @ T)
            if self.prior is not None:
                self.prior.observe(self.T_prev, T)
        self.T_prev = self.T = T
        self.pipe.set_pose(k, T, xi, self.epoch)
        self.traj.append((k, t_mid, T.copy()))
        self._update_stats(t0)
        on_frame(k, T, self.f_xyz, self.f_attr, self.f_t, self.f_frame, m)

    def _estimate(self, m: int, t_mid: float, warm: bool):
        guess = None
        if self.prior is not None and not warm and m >= self.min_points:
            guess = self.prior.guess(self, m)  # (T, deskew twist) from the learned increment, or None
        if guess is not None:
            T, xi = guess[0], (guess[1] if self.deskew else np.zeros(6, np.float32))
        else:
            xi = self.xi if (self.deskew and not warm) else np.zeros(6, np.float32)
            T = self.T if warm else self.predict()
        if m > 0:
            self._deskew(m, t_mid, xi)
        if m < self.min_points:
            self.stats["skipped"] += 1
            return T, xi, False
        if warm:
            return T, xi, True
        T, ok = self._register(m, T)
        if guess is not None:
            self.totals["prior_used"] += 1
            if not self.prior.accept(float(self._result.rms0)):
                # the learned start left the first residual above threshold: plain ICP from the constant-
                # velocity guess decides the frame
                self.totals["prior_fallbacks"] += 1
                xi = self.xi if self.deskew else np.zeros(6, np.float32)
                self._deskew(m, t_mid, xi)
                T, ok = self._register(m, self.predict())
        if ok and self.deskew and self.redeskew and self.T_prev is not None:
            xi_new = _body_twist(self._inv(self.T_prev) @ T)
            if (np.linalg.norm(xi_new[3:] - xi[3:]) > REDESKEW_ROT
                    or np.linalg.norm(xi_new[:3] - xi[:3]) > REDESKEW_TRANS):
                xi = xi_new
                self._deskew(m, t_mid, xi)
                T, ok = self._register(m, T)
        return T, xi, ok

    def _frame_chain(self, t_mid: float):
        """_estimate and the keyframe insert of a frame after the warm-up as one device chain (lw_odom_frame, or
        lw_odom_frame_prior with the learned prior): the same decisions -- skip under min_points, the learned
        guess and its residual test, the re-deskew test and second registration, the keyframe test -- taken on
        the device between the steps. Returns (T, xi, ok, m), or None when the device has no persistent launch
        (the step-by-step path then runs this and every later frame)."""
        fi, fo = self._fin, self._fout
        xi = self.xi if self.deskew else np.zeros(6, np.float32)
        T_pred = self.predict()
        self._fin_T[:] = T_pred.ravel()
        fi.has_prev = 1 if self.T_prev is not None else 0
        if fi.has_prev:
            self._fin_prev[:] = self._inv(self.T_prev).ravel()
        fi.has_key = 1 if self.T_key is not None else 0
        if fi.has_key:
            self._fin_key[:] = self._inv(self.T_key).ravel()
        self._fin_xi[:] = xi
        fi.kf_dist, fi.kf_angle = float(self.kf_dist), float(self.kf_angle)
        fi.redeskew_rot, fi.redeskew_trans = REDESKEW_ROT, REDESKEW_TRANS
        fi.t_mid, fi.inv_dt = float(t_mid), 1.0 / self.frame_dt
        fi.min_points = int(self.min_points)
        fi.redeskew = 1 if (self.deskew and self.redeskew) else 0
        fp = self.prior.chain_inputs(self) if self.prior is not None else None
        if fp is None:
            rc = self._lib.lw_odom_frame(self._ctx, _abi.ctypes.byref(self.reg_params()), _abi.ctypes.byref(fi),
                                         _abi.ctypes.byref(fo))
        else:  # the learned guess, its residual test and the fallback, in the chain
            rc = self._lib.lw_odom_frame_prior(self._ctx, _abi.ctypes.byref(self.reg_params()),
                                               _abi.ctypes.byref(fi), _abi.ctypes.byref(fp), _abi.ctypes.byref(fo))
            if rc == _abi.LW_ERANGE:  # a frame beyond the one-launch preparation: this one goes step by step
                return None
        if rc == _abi.LW_ENOLOOP:
            self.chain = False
            return None
        _abi.check(rc, "lw_odom_frame")
        if fp is not None:
            self.prior.chain_result(self, fo)
        m = int(fo.m)
        if fo.skipped:
            self.stats["skipped"] += 1
            return T_pred, xi, False, m
        T_from = T_pred
        for k in range(int(fo.registrations)):  # the stats of each registration, in order
            T_from, _ = self._register_take(fo.reg[k], T_from)
        T = np.array(fo.T, dtype=np.float64).reshape(4, 4)
        if fo.keyframe:  # Odometry._insert's bookkeeping; the insert itself ran in the chain
            self.map_frames += 1
            self.T_key = T.copy()
            self.stats["keyframes"] += 1
        self._map_counts = (int(fo.map_voxels), int(fo.map_overflow))
        return T, np.array(fo.xi, dtype=np.float32), bool(fo.ok), m

    def _estimate_counted(self, t_mid: float):
        """_estimate after the warm-up without the learned prior, with the frame's point count left on the
        device: deskew and registration read it there, and it returns with the registration's result. Same
        decisions in the same order: when the count is below min_points the registration's result is set aside
        and the frame keeps the prediction, as if it had not run."""
        xi = self.xi if self.deskew else np.zeros(6, np.float32)
        T = self.predict()
        self._deskew(-1, t_mid, xi)
        r = self._register_run(-1, T)
        m = int(r.m)
        if m < self.min_points:
            self.stats["skipped"] += 1
            return T, xi, False, m
        T, ok = self._register_take(r, T)
        if ok and self.deskew and self.redeskew and self.T_prev is not None:
            xi_new = _body_twist(self._inv(self.T_prev) @ T)
            if (np.linalg.norm(xi_new[3:] - xi[3:]) > REDESKEW_ROT
                    or np.linalg.norm(xi_new[:3] - xi[:3]) > REDESKEW_TRANS):
                xi = xi_new
                self._deskew(m, t_mid, xi)
                T, ok = self._register(m, T)
        return T, xi, ok, m

    def _update_velocity(self, d: np.ndarray):
        if np.linalg.norm(d[:3, 3]) < MAX_FRAME_JUMP:
            self.delta = d
            self.xi = _body_twist(d)
            if self.velocity_smooth > 0.0:
                a = self.velocity_smooth
                self.xi_s = a * self.xi_s + (1.0 - a) * self.xi.astype(np.float64)
                self.delta = _twist_matrix(self.xi_s)
        else:
            self.delta = np.eye(4)
            self.xi = np.zeros(6, np.float32)
            self.xi_s = np.zeros(6)

    def map_stats(self) -> tuple[int, int]:
        _abi.check(self._lib.lw_map_stats(self._ctx, self._i32), "lw_map_stats")
        return int(self._i32[0]), int(self._i32[1])

    def _update_stats(self, t0: float):
        # the map's counts change only with an insert (or a reset): read them back only then
        if self._map_counts is None:
            self._map_counts = self.map_stats()
        nvox, overflow = self._map_counts
        self.stats["frames"] += 1
        self.stats["warm"] = self.map_frames < self.warmup
        self.stats["map_voxels"], self.stats["map_full"] = nvox, nvox / self.reg_cap
        self.stats["map_overflow"] = overflow
        self.stats["speed"] = float(np.linalg.norm(self.xi[:3]) / self.frame_dt)
        self.stats["turn"] = float(math.degrees(np.linalg.norm(self.xi[3:])) / self.frame_dt)
        self.stats["ms"] = (time.perf_counter() - t0) * 1e3

    def _deskew(self, m: int, t_mid: float, xi: np.ndarray):
        x = np.ascontiguousarray(xi, dtype=np.float32)
        _abi.check(self._lib.lw_deskew(self._ctx, m, float(t_mid), 1.0 / self.frame_dt, _abi.ptr(x, _abi._f32p)),
                   "lw_deskew")

    def _is_keyframe(self, T: np.ndarray) -> bool:
        if self.T_key is None:
            return True
        d = self._inv(self.T_key) @ T
        ang = math.degrees(np.linalg.norm(_rotvec(d)))
        return np.linalg.norm(d[:3, 3]) >= self.kf_dist or ang >= self.kf_angle

    def _insert(self, m: int, T: np.ndarray):
        T32 = np.ascontiguousarray(T, dtype=np.float32)
        _abi.check(self._lib.lw_map_insert(self._ctx, None, None, m, _abi.ptr(T32, _abi._f32p), 1.0 / self.reg_voxel,
                                           self.reg_voxel, self.max_range, int(self.ret_mask)), "lw_map_insert")
        self.map_frames += 1
        self._map_counts = None
        self.T_key = T.copy()
        self.stats["keyframes"] += 1

    # ---- registration ----------------------------------------------------------------------

    def reg_params(self) -> _abi.RegParams:
        p = self._params
        p.damping = self.damping
        p.degen_thr = self.degen_thr
        p.prior_info = self.prior_info
        p.reg_voxel = self.reg_voxel
        p.min_range = self.min_range
        p.max_range = self.max_range
        p.noise = self.noise
        p.alpha = self.alpha
        p.ret_mask = int(self.ret_mask)
        p.max_iter = int(self.max_iter)
        p.min_corr = int(self.min_corr)
        p.wide_iters = int(self.wide_iters)
        p.acc_mode = ACC_MODES[self.acc_mode]
        return p

    def _register(self, m: int, T_pred: np.ndarray):
        """Scan-to-map ICP of the deskewed frame, starting from T_pred. Returns (T, ok)."""
        return self._register_take(self._register_run(m, T_pred), T_pred)

    def _register_run(self, m: int, T_pred: np.ndarray):
        """lw_register on the open frame (m < 0: its count on the device); the result, nothing recorded."""
        Tp = np.ascontiguousarray(T_pred, dtype=np.float64)
        r = self._result
        _abi.check(self._lib.lw_register(self._ctx, m, _abi.ctypes.byref(self.reg_params()), _abi.ptr(Tp, _abi._f64p),
                                         _abi.ctypes.byref(r)), "lw_register")
        return r

    def _register_take(self, r, T_pred: np.ndarray):
        """Record a registration's result in the stats; returns (T, ok)."""
        self.stats["ds"] = int(r.n_ds)
        self.totals["registrations"] += 1
        if r.n_ds < self.min_corr:
            return T_pred, False
        self.totals["gn_iters"] += int(r.iters)
        T = np.array(r.T, dtype=np.float64).reshape(4, 4)
        self.stats.update(iters=int(r.iters), corr=int(r.corr), rms=float(r.rms), cond=float(r.cond),
                          weak=bool(r.weak), degen=int(r.degen))
        return T, bool(r.ok)

"""One context of liblwamd.so driven directly: map build, thinning, accumulation and registration on
device tensors or host arrays, without Odometry's frame bookkeeping (tests, benchmarks, batch drivers)."""

from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _abi
from .odom import ACC_MODES, TIGHT_FRACTION


def _torch():
    import torch

    return torch


class Engine:
    def __init__(self, reg_slots: int = 1 << 21, frame_cap: int = 1 << 18, device=None, stream=None):
        torch = _torch()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if self.device.type != "cuda":
            raise ValueError("Engine runs on a GPU device")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        self.cap = reg_slots
        self.frame_cap = frame_cap
        self.lib = _abi.lib()
        self.ctx = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _abi.check(self.lib.lw_create(ctypes.byref(self.ctx), reg_slots, frame_cap,
                                          ctypes.c_void_p(self.stream.cuda_stream)), "lw_create")
        d = self.device
        self.f_raw = torch.zeros((frame_cap, 3), dtype=torch.float32, device=d)
        self.f_xyz = torch.zeros((frame_cap, 3), dtype=torch.float32, device=d)
        self.f_attr = torch.zeros(frame_cap, dtype=torch.int32, device=d)
        self.f_t = torch.zeros(frame_cap, dtype=torch.float32, device=d)
        self.f_frame = torch.zeros(frame_cap, dtype=torch.int32, device=d)
        _abi.check(self.lib.lw_bind_frame(self.ctx, self.f_raw.data_ptr(), self.f_xyz.data_ptr(),
                                          self.f_attr.data_ptr(), self.f_t.data_ptr(), self.f_frame.data_ptr()),
                   "lw_bind_frame")
        self._i32 = (ctypes.c_int32 * 2)()

    def close(self):
        """Wait for the context's queued work and free it (idempotent)."""
        ctx = getattr(self, "ctx", None)
        self.ctx = None
        if ctx is not None and ctx.value:
            _abi.check(self.lib.lw_destroy(ctx), "lw_destroy")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- helpers ----------------------------------------------------------------------------
    def to_dev(self, a, dtype):
        torch = _torch()
        if isinstance(a, torch.Tensor):
            return a.to(self.device).contiguous()
        return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).to(self.device).contiguous()

    def sync(self):
        _abi.check(self.lib.lw_synchronize(self.ctx), "lw_synchronize")

    # ---- map --------------------------------------------------------------------------------
    def reset(self):
        _abi.check(self.lib.lw_map_reset(self.ctx), "lw_map_reset")

    def insert(self, xyz, attr, T, voxel: float, max_range: float = 60.0, ret_mask: int = 0b1111):
        """k_reg_insert of n points (host or device arrays) posed by T."""
        torch = _torch()
        x = self.to_dev(np.asarray(xyz, np.float32).reshape(-1, 3) if not isinstance(xyz, torch.Tensor) else xyz,
                        torch.float32)
        if attr is None:
            a = torch.zeros(x.shape[0], dtype=torch.int32, device=self.device)
        else:
            a = self.to_dev(np.asarray(attr).astype(np.uint32).view(np.int32) if not isinstance(attr, torch.Tensor)
                            else attr, torch.int32)
        T32 = np.ascontiguousarray(np.asarray(T, np.float64).reshape(4, 4), dtype=np.float32)
        _abi.check(self.lib.lw_map_insert(self.ctx, x.data_ptr(), a.data_ptr(), x.shape[0], _abi.ptr(T32, _abi._f32p),
                                          1.0 / voxel, voxel, max_range, int(ret_mask)), "lw_map_insert")
        self._keep = (x, a)  # alive until the stream has consumed them
        return self

    def stats(self) -> tuple[int, int]:
        _abi.check(self.lib.lw_map_stats(self.ctx, self._i32), "lw_map_stats")
        return int(self._i32[0]), int(self._i32[1])

    def read_map(self):
        """(keys i64[cap], rsum f32[cap,3], router f32[cap,3,3], rcnt i32[cap]) in the reference's layout."""
        keys = np.empty(self.cap, np.int64)
        rsum = np.empty((self.cap, 3), np.float32)
        router = np.empty((self.cap, 3, 3), np.float32)
        rcnt = np.empty(self.cap, np.int32)
        _abi.check(self.lib.lw_map_read(self.ctx, _abi.ptr(keys, _abi._i64p), _abi.ptr(rsum, _abi._f32p),
                                        _abi.ptr(router, _abi._f32p), _abi.ptr(rcnt, _abi._i32p)), "lw_map_read")
        return keys, rsum, router, rcnt

    def find(self, keys) -> np.ndarray:
        keys = np.ascontiguousarray(keys, np.int64)
        out = np.empty(len(keys), np.int32)
        _abi.check(self.lib.lw_map_find(self.ctx, _abi.ptr(keys, _abi._i64p), len(keys), _abi.ptr(out, _abi._i32p)),
                   "lw_map_find")
        return out

    # ---- frame ------------------------------------------------------------------------------
    def frame_reset(self):
        _abi.check(self.lib.lw_frame_reset(self.ctx), "lw_frame_reset")

    def frame_append(self, sx, sa, st, inv_dt: float, k_min: int, k_open: int):
        torch = _torch()
        _abi.check(self.lib.lw_frame_append(self.ctx, sx.data_ptr(), sa.data_ptr(), st.data_ptr(), st.shape[0], inv_dt,
                                            k_min, k_open), "lw_frame_append")

    def frame_count(self) -> int:
        _abi.check(self.lib.lw_frame_count(self.ctx, self._i32), "lw_frame_count")
        return int(self._i32[0])

    def load_frame(self, xyz, attr, t):
        """Put n points straight into the frame buffers (f_raw, f_attr, f_t)."""
        n = len(t)
        torch = _torch()
        self.f_raw[:n] = torch.from_numpy(np.asarray(xyz, np.float32).reshape(-1, 3)).to(self.device)
        self.f_attr[:n] = torch.from_numpy(np.asarray(attr).astype(np.uint32).view(np.int32)).to(self.device)
        self.f_t[:n] = torch.from_numpy(np.asarray(t, np.float32)).to(self.device)
        torch.cuda.synchronize(self.device)
        return n

    def deskew(self, m: int, t_mid: float, inv_dt: float, xi):
        x = np.ascontiguousarray(xi, dtype=np.float32)
        _abi.check(self.lib.lw_deskew(self.ctx, m, t_mid, inv_dt, _abi.ptr(x, _abi._f32p)), "lw_deskew")

    def thin(self, m: int, inv_v: float, min_range: float = 0.3, max_range: float = 60.0, ret_mask: int = 0b1111) -> int:
        _abi.check(self.lib.lw_thin(self.ctx, m, inv_v, min_range, max_range, int(ret_mask), self._i32), "lw_thin")
        return int(self._i32[0])

    def thin_read(self, n: int) -> np.ndarray:
        out = np.empty((n, 3), np.float32)
        _abi.check(self.lib.lw_thin_read(self.ctx, _abi.ptr(out, _abi._f32p), n), "lw_thin_read")
        return out

    def thin_load(self, xyz) -> int:
        x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        _abi.check(self.lib.lw_thin_load(self.ctx, _abi.ptr(x, _abi._f32p), len(x)), "lw_thin_load")
        return len(x)

    # ---- registration -----------------------------------------------------------------------
    def icp_accum(self, n: int, T, voxel: float, reach: int, max_d2: float, noise2: float, tight2: float,
                  alpha: float, kappa: float, acc_mode: str = "fp64", with_slots: bool = False):
        T32 = np.ascontiguousarray(np.asarray(T, np.float64).reshape(4, 4), dtype=np.float32)
        out = np.zeros(64, np.float64)
        slots = np.empty(n, np.int32) if with_slots else None
        _abi.check(self.lib.lw_icp_accum(self.ctx, n, _abi.ptr(T32, _abi._f32p), 1.0 / voxel, voxel, reach, max_d2,
                                         noise2, tight2, alpha, kappa, ACC_MODES[acc_mode], _abi.ptr(out, _abi._f64p),
                                         _abi.ptr(slots, _abi._i32p) if slots is not None else None), "lw_icp_accum")
        return (out, slots) if with_slots else out

    @staticmethod
    def params(reg_voxel=0.15, max_iter=16, min_corr=60, min_range=0.3, max_range=60.0, noise=0.02, ret_mask=0b1111,
               alpha=0.05, damping=1e-3, degen_thr=3000.0, prior_info=2000.0, wide_iters=3, acc_mode="fp64"):
        p = _abi.RegParams()
        p.damping, p.degen_thr, p.prior_info = damping, degen_thr, prior_info
        p.reg_voxel, p.min_range, p.max_range, p.noise, p.alpha = reg_voxel, min_range, max_range, noise, alpha
        p.ret_mask, p.max_iter, p.min_corr, p.wide_iters = int(ret_mask), int(max_iter), int(min_corr), int(wide_iters)
        p.acc_mode = ACC_MODES[acc_mode]
        return p

    def register_thinned(self, n: int, p: _abi.RegParams, T_pred) -> _abi.RegResult:
        Tp = np.ascontiguousarray(np.asarray(T_pred, np.float64).reshape(4, 4))
        r = _abi.RegResult()
        _abi.check(self.lib.lw_register_thinned(self.ctx, n, ctypes.byref(p), _abi.ptr(Tp, _abi._f64p),
                                                ctypes.byref(r)), "lw_register_thinned")
        return r

    def register(self, m: int, p: _abi.RegParams, T_pred) -> _abi.RegResult:
        Tp = np.ascontiguousarray(np.asarray(T_pred, np.float64).reshape(4, 4))
        r = _abi.RegResult()
        _abi.check(self.lib.lw_register(self.ctx, m, ctypes.byref(p), _abi.ptr(Tp, _abi._f64p), ctypes.byref(r)),
                   "lw_register")
        return r


def tight2_of(voxel: float) -> float:
    return float(np.float32((TIGHT_FRACTION * voxel / math.sqrt(12.0)) ** 2))

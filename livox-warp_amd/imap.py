"""The integration map: the consumer of the odometry's frames in the reference's Pipeline
(python/livox_warp/gpu.py:373-418 k_map_insert, 420-595 local_plane and k_carve, 1172-1219 the map's arrays,
1304-1312 clear_map, 1346-1372 flush_poses and rebase_clock, 1421-1452 map_insert and carve).

Voxel records keyed and probed like the registration map: hits, carving misses, float32 sums of x, y, z and
reflectivity over the first 4096 hits, and the first and last time a voxel was hit. The points come in through
their frames' poses, which the map uploads from a PoseTable (the odometry's pose sink) before each insert or
carve, as Pipeline.flush_poses does. Every step runs on the device through liblwamd.so; there is no other
implementation.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi

# one device record (lw_imap_read): the layout of include/livox_warp_amd.h
RECORD = np.dtype([("key", np.int64), ("cnt", np.int32), ("miss", np.int32), ("acc", np.float32, 4),
                   ("first", np.float32), ("last", np.float32), ("pad", np.int32, 6)])
assert RECORD.itemsize == 64


@dataclass
class View:
    """The fields of the reference's View (gpu.py:77-99) that a map insert reads (keep_raw, gpu.py:156-167)."""
    min_range: float = 0.3
    max_range: float = 60.0
    min_refl: int = 0
    noise_mask: int = 0b1111  # bit k keeps points whose spatial-noise confidence (tag bits 0-1) == k
    ret_mask: int = 0b1111    # bit k keeps return k

    def c(self) -> _abi.MapView:
        v = _abi.MapView()
        v.min_range, v.max_range = float(self.min_range), float(self.max_range)
        v.min_refl, v.noise_mask, v.ret_mask = int(self.min_refl), int(self.noise_mask), int(self.ret_mask)
        return v


def _torch():
    import torch

    return torch


class IntegrationMap:
    """Pipeline's map state and calls (map_insert, carve, clear_map, rebase_clock, map_stats) on one stream."""

    def __init__(self, poses, device=None, map_capacity: int = 1 << 23, stream=None):
        torch = _torch()
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError(f"IntegrationMap runs on a GPU device, got {self.device}")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        assert map_capacity & (map_capacity - 1) == 0
        self.map_cap = int(map_capacity)
        self.poses = poses
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        self._lib = _abi.lib()
        self._m = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _abi.check(self._lib.lw_imap_create(ctypes.byref(self._m), self.map_cap,
                                                ctypes.c_void_p(self.stream.cuda_stream)), "lw_imap_create")
        self._i32 = (ctypes.c_int32 * 2)()
        self.map_occupied = self.map_dropped = 0

    def close(self):
        """Wait for the map's queued work and free it (idempotent)."""
        m = getattr(self, "_m", None)
        self._m = None
        if m is not None and m.value:
            _abi.check(self._lib.lw_imap_destroy(m), "lw_imap_destroy")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _dev(a, dtypes, n: int, width: int = 1) -> int:
        torch = _torch()
        if not isinstance(a, torch.Tensor) or a.device.type != "cuda":
            raise TypeError("the map takes device tensors")
        if a.dtype not in dtypes or not a.is_contiguous() or a.numel() < n * width:
            raise ValueError(f"bad staged array: {a.dtype}, {tuple(a.shape)}")
        return a.data_ptr()

    def flush_poses(self):
        """Upload every pose edited in the PoseTable since the last flush (gpu.py:1346-1365)."""
        idx, mats = self.poses.drain()
        if len(idx):
            mats = np.ascontiguousarray(mats, np.float32)
            _abi.check(self._lib.lw_imap_set_poses(self._m, _abi.ptr(idx, _abi._i32p), _abi.ptr(mats, _abi._f32p),
                                                   len(idx)), "lw_imap_set_poses")

    def clear_map(self):
        _abi.check(self._lib.lw_imap_clear(self._m), "lw_imap_clear")
        self.map_occupied = self.map_dropped = 0

    def map_insert(self, sx, sa, st, sf, n: int, view: View, voxel: float):
        """Accumulate n staged points (xyz, attribute bits, time, frame id: device tensors) into the map through
        their frames' poses (gpu.py:1421-1434)."""
        if n <= 0:
            return
        torch = _torch()
        p = (self._dev(sx, (torch.float32,), n, 3), self._dev(sa, (torch.int32, torch.uint32), n),
             self._dev(st, (torch.float32,), n), self._dev(sf, (torch.int32,), n))
        self.flush_poses()
        v = view.c()
        _abi.check(self._lib.lw_imap_insert(self._m, *p, int(n), ctypes.byref(v), float(voxel)), "lw_imap_insert")

    def carve(self, sx, sa, sf, n: int, voxel: float, every: int = 2, max_range: float = 40.0,
              max_steps: int = 4096):
        """Free-space carving (gpu.py:1436-1452): misses for the map voxels each traced ray passes through."""
        if n <= 0:
            return
        torch = _torch()
        p = (self._dev(sx, (torch.float32,), n, 3), self._dev(sa, (torch.int32, torch.uint32), n),
             self._dev(sf, (torch.int32,), n))
        self.flush_poses()
        _abi.check(self._lib.lw_imap_carve(self._m, *p, int(n), float(voxel), max(1, int(every)), float(max_range),
                                           int(max_steps)), "lw_imap_carve")

    def rebase_clock(self, shift: float):
        """The live clock restarted: first and last times move onto the new clock, new = old - shift."""
        _abi.check(self._lib.lw_imap_rebase_clock(self._m, float(shift)), "lw_imap_rebase_clock")

    def map_stats(self) -> tuple[int, int]:
        """(voxels claimed, points that found no slot), as Pipeline keeps them in map_stats."""
        _abi.check(self._lib.lw_imap_stats(self._m, self._i32), "lw_imap_stats")
        self.map_occupied, self.map_dropped = int(self._i32[0]), int(self._i32[1])
        return self.map_occupied, self.map_dropped

    def records(self) -> np.ndarray:
        """Every slot's record (RECORD dtype), copied to the host."""
        out = np.empty(self.map_cap, RECORD)
        _abi.check(self._lib.lw_imap_read(self._m, out.ctypes.data_as(ctypes.c_void_p), out.nbytes), "lw_imap_read")
        return out

    def load(self, recs: np.ndarray):
        """Replace every slot's record (RECORD dtype, map_cap long)."""
        recs = np.ascontiguousarray(recs, RECORD)
        if len(recs) != self.map_cap:
            raise ValueError("one record per slot")
        _abi.check(self._lib.lw_imap_load(self._m, recs.ctypes.data_as(ctypes.c_void_p), recs.nbytes), "lw_imap_load")

"""Bulk voxelisation of a prior map on MI355X (lw_vox_*): the reference's prior_map.voxel_downsample
(python/livox_warp/prior_map.py:169-203) -- mean position, colour and intensity per voxel of a huge cloud,
inserted in chunks into one open-addressed table (128 probes, k_vox_insert) and compacted in slot order
(k_vox_flag, array_scan, k_vox_emit).

VoxTable is the device-level object (torch device tensors in and out); voxel_downsample keeps the
reference's signature and numpy results.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _abi


def _torch():
    import torch

    return torch


class VoxTable:
    """A table of 2^slots_log2 voxels on one device, its work queued on `stream` (default: the current
    stream)."""

    def __init__(self, slots_log2: int = 27, device=None, stream=None):
        torch = _torch()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        self.lib = _abi.lib()
        self.cap = 1 << slots_log2
        self.h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _abi.check(self.lib.lw_vox_create(ctypes.byref(self.h), slots_log2,
                                              ctypes.c_void_p(self.stream.cuda_stream)), "lw_vox_create")

    def close(self):
        h = getattr(self, "h", None)
        self.h = None
        if h is not None and h.value:
            _abi.check(self.lib.lw_vox_destroy(h), "lw_vox_destroy")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        _abi.check(self.lib.lw_vox_clear(self.h), "lw_vox_clear")

    def insert(self, xyz, voxel: float, intensity=None, rgb=None):
        """k_vox_insert over the rows of xyz ((n, 3) float32 device tensor); intensity (n,) and rgb (n, 3)
        float32 device tensors or None (zeros). The tensors must stay alive until the table's stream has run
        the insert."""
        torch = _torch()
        n = int(xyz.shape[0])
        for a, w in ((xyz, 3), (intensity, 1), (rgb, 3)):
            if a is not None and (a.dtype != torch.float32 or not a.is_contiguous() or a.device != self.device or
                                  a.shape[0] != n or (w > 1 and a.shape[1] != w)):
                raise ValueError("insert() takes contiguous float32 device tensors of n rows")
        _abi.check(self.lib.lw_vox_insert(self.h, xyz.data_ptr(), 0 if intensity is None else intensity.data_ptr(),
                                          0 if rgb is None else rgb.data_ptr(), n, float(voxel)), "lw_vox_insert")

    def stats(self):
        """(voxels claimed, points dropped), after the queued inserts."""
        out = (ctypes.c_int32 * 2)()
        _abi.check(self.lib.lw_vox_stats(self.h, out), "lw_vox_stats")
        return int(out[0]), int(out[1])

    def emit(self, voxel: float):
        """The claimed voxels in slot order as device tensors: mean xyz (V, 3), mean r, g, b, intensity (V, 4),
        count (V,)."""
        torch = _torch()
        v, _ = self.stats()
        xyz = torch.empty((max(v, 1), 3), dtype=torch.float32, device=self.device)
        col = torch.empty((max(v, 1), 4), dtype=torch.float32, device=self.device)
        cnt = torch.empty(max(v, 1), dtype=torch.int32, device=self.device)
        got = ctypes.c_int64()
        _abi.check(self.lib.lw_vox_emit(self.h, float(voxel), xyz.data_ptr(), col.data_ptr(), cnt.data_ptr(), v,
                                        ctypes.byref(got)), "lw_vox_emit")
        return xyz[:v], col[:v], cnt[:v]

    def read(self):
        """The table as numpy arrays (keys, offset sums, colour + intensity sums, counts), slot order."""
        torch = _torch()
        keys = torch.empty(self.cap, dtype=torch.int64, device=self.device)
        rsum = torch.empty((self.cap, 3), dtype=torch.float32, device=self.device)
        csum = torch.empty((self.cap, 4), dtype=torch.float32, device=self.device)
        cnt = torch.empty(self.cap, dtype=torch.int32, device=self.device)
        _abi.check(self.lib.lw_vox_read(self.h, keys.data_ptr(), rsum.data_ptr(), csum.data_ptr(), cnt.data_ptr()),
                   "lw_vox_read")
        self.stream.synchronize()
        return keys.cpu().numpy(), rsum.cpu().numpy(), csum.cpu().numpy(), cnt.cpu().numpy()


def voxel_downsample(xyz: np.ndarray, intensity: np.ndarray, rgb: np.ndarray, voxel: float, device=None,
                     chunk: int = 1 << 25, slots_log2: int = 27):
    """prior_map.voxel_downsample (prior_map.py:169-203): mean position / colour / intensity per voxel of a
    (huge) cloud, inserted in chunks. Returns (xyz, rgb, intensity, count) per voxel, in the table's slot
    order; RuntimeError if points were dropped (the table is full: raise slots_log2)."""
    torch = _torch()
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    stream = torch.cuda.Stream(device=dev)
    with VoxTable(slots_log2, dev, stream) as t, torch.cuda.stream(stream):
        n = len(xyz)
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            px = torch.from_numpy(np.ascontiguousarray(xyz[a:b], np.float32)).to(dev)
            pi = torch.from_numpy(np.ascontiguousarray(intensity[a:b], np.float32)).to(dev)
            pc = torch.from_numpy(np.ascontiguousarray(rgb[a:b], np.float32)).to(dev)
            t.insert(px, voxel, pi, pc)
            stream.synchronize()  # the chunk's tensors are freed at the next iteration
        _, dropped = t.stats()
        if dropped:
            raise RuntimeError(f"voxel table full: {dropped} points dropped; raise slots_log2")
        out_xyz, col, cnt = t.emit(voxel)
        stream.synchronize()
        col = col.cpu().numpy()
        return out_xyz.cpu().numpy(), col[:, :3], col[:, 3], cnt.cpu().numpy()

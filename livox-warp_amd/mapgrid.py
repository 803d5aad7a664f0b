"""The prior map's dense grid on the device (the reference's python/livox_warp/mapgrid.py): a 5 cm grid over
the map's extent holding, per cell, the map point nearest the cell centre (jump-flooded over a band of about
0.8 m) and that distance in cm as a byte, the distance field global localisation scores hypotheses with; and
the map points sorted by cell with their prefix array, so a nearest-point query scans 27 cells.

prepare() is the host work (numpy, as mapgrid.py:215-241: the sort order fixes the point indices the seeds
refer to); build() uploads and runs lw_grid_build; distances() is lw_grid_distances. Arrays are torch device
tensors; every launch names its stream.
"""

from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _abi

CELL = 0.05  # mapgrid.py:19
IDX_BITS = 23  # mapgrid.py:20
MAX_POINTS = 1 << IDX_BITS
JFA_STEPS = (8, 4, 2, 1, 1)  # mapgrid.py:25
PLANAR = 0.5  # mapgrid.py:26


def _torch():
    import torch

    return torch


class PriorGrid:
    """A prior map on the device: points sorted by cell, unit normals, planarity, and the grid (the reference's
    PriorGrid, mapgrid.py:207-300). Construction allocates; build() fills on a stream and waits for it; after
    that nothing writes the arrays, so any stream may read them."""

    @staticmethod
    def prepare(host: dict, cell: float = CELL, pad: float = 0.6) -> dict:
        """Host preparation (mapgrid.py:215-241): float32 arrays sorted by grid cell, unit normals, the grid
        geometry and cell prefix. A map without normals gets zero normals."""
        if host.get("_prepared"):
            return host
        xyz = np.asarray(host["xyz"], dtype=np.float32)
        n = len(xyz)
        if n >= MAX_POINTS:
            raise ValueError(f"prior map has {n:,} points; the grid seeds hold at most {MAX_POINTS - 1:,}")
        nrm = np.asarray(host["normal"], dtype=np.float32) if "normal" in host else np.zeros((n, 3), np.float32)
        planar = np.asarray(host["planarity"] if "planarity" in host else np.ones(n), dtype=np.float32)
        lo, hi = xyz.min(0), xyz.max(0)
        origin = (lo - pad).astype(np.float32)
        dims = np.ceil((hi + pad - origin) / cell).astype(np.int64)
        if int(np.prod(dims)) >= (1 << 31) - 1:
            raise ValueError(f"grid of {dims.tolist()} cells exceeds 2^31 cells")
        if dims[0] >= (1 << 24) or dims[1] * dims[2] >= (1 << 24) or dims[1] > 65535 or dims[2] > 65535:
            raise ValueError(f"grid of {dims.tolist()} cells: the kernels take at most 2^24 - 1 cells in x, "
                             f"2^24 - 1 rows (y times z) and 65535 in y and in z")
        ijk = np.floor((xyz - origin) / cell).astype(np.int64)
        key = (ijk[:, 2] * dims[1] + ijk[:, 1]) * dims[0] + ijk[:, 0]
        order = np.argsort(key, kind="stable")
        ncell = int(np.prod(dims))
        counts = np.bincount(key, minlength=ncell)
        start = np.zeros(ncell + 1, np.int32)
        np.cumsum(counts, out=start[1:])
        xyz = np.ascontiguousarray(xyz[order])
        nrm = nrm[order]
        nrm = np.ascontiguousarray(nrm / np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-6), dtype=np.float32)
        planar = np.ascontiguousarray(planar[order])
        return {"_prepared": True, "xyz": xyz, "normal": nrm, "planarity": planar, "lo": lo, "hi": hi,
                "origin": origin, "cell": float(cell), "dims": dims, "start": start}

    def __init__(self, host: dict, device=None):
        torch = _torch()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        host = self.prepare(host)
        self._host = host
        self.n = len(host["xyz"])
        self.lo, self.hi = host["lo"], host["hi"]
        self.dims = host["dims"]
        self.ncell = int(np.prod(self.dims))
        self.origin = np.asarray(host["origin"], np.float32)
        self.cell = float(host["cell"])
        # geometry as the kernels take it: origin, cell and 1 / cell in float32; nx, ny, nz
        self.geo = np.array([*self.origin, self.cell, 1.0 / self.cell], np.float32)
        self.dims32 = np.ascontiguousarray(self.dims, np.int32)
        d = self.device
        self.xyz = torch.empty((self.n, 3), dtype=torch.float32, device=d)
        self.nrm = torch.empty((self.n, 3), dtype=torch.float32, device=d)
        self.planar = torch.empty(self.n, dtype=torch.float32, device=d)
        self.start = torch.empty(self.ncell + 1, dtype=torch.int32, device=d)
        self.nn = torch.empty(self.ncell, dtype=torch.int32, device=d)
        self.df = torch.empty(self.ncell, dtype=torch.uint8, device=d)
        self.ready = False
        self.build_s = 0.0
        self.build_kernel_ms = 0.0

    def gpu_bytes(self) -> int:
        return self.n * (12 + 12 + 4) + self.ncell * (4 + 4 + 1)

    def _geo_args(self):
        return _abi.ptr(self.geo, _abi._f32p), _abi.ptr(self.dims32, _abi._i32p)

    def build(self, stream=None, steps=JFA_STEPS):
        """Upload the points and build the grid on `stream` (the current stream if None); returns when done. The
        jump flood's two arrays of seed records (32 bytes a cell) live only for the build."""
        torch = _torch()
        t0 = time.perf_counter()
        h = self._host
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        with torch.cuda.stream(s):
            for dst, src in ((self.xyz, h["xyz"]), (self.nrm, h["normal"]), (self.planar, h["planarity"]),
                             (self.start, h["start"])):
                dst.copy_(torch.from_numpy(np.ascontiguousarray(src)).reshape(dst.shape))
            L = _abi.lib()
            wb = int(L.lw_grid_work_bytes(_abi.ptr(self.dims32, _abi._i32p)))
            tmp = torch.empty(wb, dtype=torch.uint8, device=self.device)
            st = np.ascontiguousarray(steps, np.int32)
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record(s)
            _abi.check(L.lw_grid_build(self.xyz.data_ptr(), self.n, *self._geo_args(), _abi.ptr(st, _abi._i32p),
                                       len(st), self.nn.data_ptr(), tmp.data_ptr(), wb, self.df.data_ptr(),
                                       ctypes.c_void_p(s.cuda_stream)), "lw_grid_build")
            ev[1].record(s)
        s.synchronize()
        del tmp
        self.build_s = time.perf_counter() - t0
        self.build_kernel_ms = ev[0].elapsed_time(ev[1])  # the build's kernels alone (no upload)
        self.ready = True

    def distances(self, pts, n: int, T_map_from_pts: np.ndarray, out, stream=None):
        """Distance to the map of the first n points of `pts` (device (>= n, 3) float32, in a frame that T maps
        into the map) into `out` (device float32), on `stream` (mapgrid.py:292-300); 1e3 beyond the band."""
        if n <= 0:
            return
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        T = np.ascontiguousarray(np.asarray(T_map_from_pts, dtype=np.float32).reshape(16))
        _abi.check(_abi.lib().lw_grid_distances(pts.data_ptr(), int(n), _abi.ptr(T, _abi._f32p), *self._geo_args(),
                                                self.start.data_ptr(), self.nn.data_ptr(), self.xyz.data_ptr(),
                                                out.data_ptr(), ctypes.c_void_p(s.cuda_stream)), "lw_grid_distances")

"""Batched global localisation on MI355X: thousands of pose hypotheses x query scans registered against one
prior map in one persistent launch (lw_loc_batch), with the tile (threads per workgroup, persistent
workgroups per compute unit, neighbour keys a lane resolves together) picked by an autotuner that skips any
tile whose compiled kernel spills to scratch, and the launch optionally replayed from a captured hipGraph.

The reference's counterpart is GlobalLocalizer._icp (python/livox_warp/localize.py:517-553), which refines up
to 16 hypotheses with one host solve per hypothesis and iteration.
"""

from __future__ import annotations

import ctypes
import json
import math
import os
import time

import numpy as np

from . import _abi
from .engine import Engine


def _torch():
    import torch

    return torch


LocParams = _abi.LocParams

CACHE = os.path.join(os.path.expanduser("~"), ".cache", "livox-warp_amd", "loc_tiles.json")
# (threads per workgroup, workgroups per compute unit, neighbour keys per chunk)
TILE_CANDIDATES = [(b, g, c) for c in (14, 7) for b in (64, 128, 256, 512) for g in (1, 2, 4, 8)]
DEFAULT_TILE = (256, 2, 7)


def tile_info(lib, block: int, chunk: int) -> dict:
    """The compiled kernel of a tile: scratch bytes per lane, registers per lane, static LDS bytes."""
    info = (ctypes.c_int32 * 4)()
    _abi.check(lib.lw_loc_tile_info(block, chunk, info), "lw_loc_tile_info")
    return {"scratch": int(info[0]), "regs": int(info[1]), "lds": int(info[2]), "max_threads": int(info[3])}


class BatchLocalizer:
    """A prior map (built once, in index order, like prior_map.voxel_downsample's bulk insert) and batched
    hypothesis registration against it."""

    def __init__(self, map_xyz, voxel: float = 0.1, reg_slots: int = 1 << 24, chunk: int = 1 << 22, device=None,
                 stream=None, iters: int = 10, noise: float = 0.02, alpha: float = 0.05):
        torch = _torch()
        self.voxel = voxel
        self.stream = stream if stream is not None else torch.cuda.Stream(device=device)
        self.eng = Engine(reg_slots, chunk, device=device, stream=self.stream)
        with torch.cuda.stream(self.stream):
            x = map_xyz if isinstance(map_xyz, torch.Tensor) else torch.from_numpy(
                np.ascontiguousarray(map_xyz, np.float32))
            x = x.to(self.eng.device)
            self.eng.insert(x, None, np.eye(4), voxel, max_range=1e30)
        self.eng.sync()
        self.n_map = int(x.shape[0])
        self.params = LocParams(damping=1e-3, reg_voxel=voxel, noise=noise, alpha=alpha, iters=iters, min_count=30,
                                block=128, grid=0, use_graph=0)
        self.cus = torch.cuda.get_device_properties(self.eng.device).multi_processor_count
        self.tiles = {}
        self._bufs = None

    def close(self):
        """Free the map's context, with the captured graph it holds (idempotent)."""
        eng = getattr(self, "eng", None)
        self._bufs = None
        if eng is not None:
            eng.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def map_stats(self):
        return self.eng.stats()

    def _stage(self, scans, hyps):
        torch = _torch()
        d = self.eng.device
        sizes = np.array([len(s) for s in scans], np.int64)
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        order = np.argsort(-sizes, kind="stable").astype(np.int32)  # longest scans first
        with torch.cuda.stream(self.stream):
            S = torch.from_numpy(np.ascontiguousarray(np.concatenate(scans), np.float32)).to(d)
            O = torch.from_numpy(off).to(d)
            R = torch.from_numpy(order).to(d)
            Hy = hyps if isinstance(hyps, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(hyps, np.float64))
            Hy = Hy.to(d).contiguous()
        return S, O, R, Hy

    def run(self, scans, hyps, tile=None, use_graph: bool = False, staged=None):
        """Register hyps[q, h] (4x4 float64) against the map with scans[q]; returns (T, score, count, iters) as
        device tensors of shape (Q, H, 4, 4), (Q, H), (Q, H), (Q, H). Queued on the localiser's stream, and the
        caller's current stream is made to wait for it. The tensors are reused by the next call."""
        torch = _torch()
        Q = len(scans)
        H = int(hyps.shape[1])
        S, O, R, Hy = staged if staged is not None else self._stage(scans, hyps)
        d = self.eng.device
        if self._bufs is None or self._bufs[0].shape[0] != Q * H:
            self._bufs = (torch.empty((Q * H, 16), dtype=torch.float64, device=d),
                          torch.empty(Q * H, dtype=torch.float64, device=d),
                          torch.empty(Q * H, dtype=torch.int32, device=d),
                          torch.empty(Q * H, dtype=torch.int32, device=d))
        oT, oS, oC, oI = self._bufs
        block, per_cu, chunk = tile if tile is not None else self.tile_for(scans)
        p = LocParams.from_buffer_copy(self.params)
        p.block, p.grid, p.chunk, p.use_graph = block, per_cu * self.cus, chunk, 1 if use_graph else 0
        _abi.check(self.eng.lib.lw_loc_batch(self.eng.ctx, S.data_ptr(), O.data_ptr(), R.data_ptr(), Q, Hy.data_ptr(),
                                             H, ctypes.byref(p), oT.data_ptr(), oS.data_ptr(), oC.data_ptr(),
                                             oI.data_ptr()), "lw_loc_batch")
        self._keep = (S, O, R, Hy)
        # the outputs are written on the localiser's stream: order the caller's stream after it
        torch.cuda.current_stream(d).wait_stream(self.stream)
        return oT.view(Q, H, 4, 4), oS.view(Q, H), oC.view(Q, H), oI.view(Q, H)

    # ---- tile autotuner ------------------------------------------------------------------------------
    @staticmethod
    def _bucket(scans) -> str:
        n = int(np.mean([len(s) for s in scans]))
        return f"pts{1 << max(int(math.log2(max(n, 1))), 0)}"

    def tile_for(self, scans):
        key = self._bucket(scans)
        if key not in self.tiles:
            cached = self._load_cache().get(key)
            self.tiles[key] = tuple(cached) if cached and len(cached) == 3 else DEFAULT_TILE
        return self.tiles[key]

    def autotune(self, scans, hyps, repeats: int = 3):
        """Time every candidate tile whose kernel does not spill on (scans, hyps) and keep the fastest for this
        scan-size bucket. Returns (tile, {tile: best ms}, [tiles skipped for spilling])."""
        staged = self._stage(scans, hyps)
        times, spilled = {}, []
        for tile in TILE_CANDIDATES:
            if tile_info(self.eng.lib, tile[0], tile[2])["scratch"] > 0:
                spilled.append(tile)
                continue
            self.run(scans, hyps, tile=tile, staged=staged)
            self.stream.synchronize()
            best = float("inf")
            for _ in range(repeats):
                t0 = time.perf_counter()
                self.run(scans, hyps, tile=tile, staged=staged)
                self.stream.synchronize()
                best = min(best, time.perf_counter() - t0)
            times[tile] = best
        tile = min(times, key=times.get)
        key = self._bucket(scans)
        self.tiles[key] = tile
        cache = self._load_cache()
        cache[key] = list(tile)
        try:
            os.makedirs(os.path.dirname(CACHE), exist_ok=True)
            with open(CACHE, "w") as f:
                json.dump(cache, f)
        except OSError:
            pass
        return tile, {f"{b}x{g}x{c}": round(t * 1e3, 3) for (b, g, c), t in times.items()}, spilled

    @staticmethod
    def _load_cache() -> dict:
        try:
            with open(CACHE) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}


def hypotheses(rng, n: int, T_true: np.ndarray, lo, hi) -> np.ndarray:
    """n poses with x, y uniform in [lo, hi] and yaw uniform in [0, 2 pi); z, roll and pitch from T_true."""
    from scipy.spatial.transform import Rotation

    ypr = Rotation.from_matrix(T_true[:3, :3]).as_euler("ZYX")
    out = np.tile(np.eye(4), (n, 1, 1))
    yaw = rng.uniform(0.0, 2 * math.pi, n)
    R = Rotation.from_euler("ZYX", np.column_stack([yaw, np.full(n, ypr[1]), np.full(n, ypr[2])])).as_matrix()
    out[:, :3, :3] = R
    out[:, 0, 3] = rng.uniform(lo[0], hi[0], n)
    out[:, 1, 3] = rng.uniform(lo[1], hi[1], n)
    out[:, 2, 3] = T_true[2, 3]
    return out

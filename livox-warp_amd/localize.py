"""Coarse stages of global localisation on the device: the reference GlobalLocalizer's search before its
batched ICP (python/livox_warp/localize.py:426-505 and the scoring kernels at :55-123).

Every free-space position on an xy grid at the given sensor heights, at every heading, is scored by how close
a gravity-aligned scan lands to the map (a soft count within tau of the prior map's distance field,
lw_score_grid); the best local maxima in (x, y, heading) are re-scored on a finer local grid with a tighter
tolerance and more points (lw_score_list), and greedy non-maximum suppression keeps the candidates the ICP
refines (loc.BatchLocalizer, or the reference's own). The host parts (grid layout, local maxima, the pool of
best scores, suppression) follow the reference step for step, so the same scores pick the same candidates.

The reference's scan preparation (RANSAC gravity candidates, normals, the height search over the map's levels)
and its whole localize() / refine() flow are not part of this path: a caller gives the sensor height and the
gravity-aligned scan.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _abi
from .mapgrid import PriorGrid


def _torch():
    import torch

    return torch


EVAL_RADII = (0.03, 0.05, 0.10)  # a pose's fit: fractions of the scan within these distances (m) of the map
_TRIU = np.triu_indices(6)  # the order the accumulator stores the upper triangle of the 6x6 normal matrix in
ACC_H, ACC_B, ACC_COUNT = 0, 21, 27


def twist_apply(T: np.ndarray, delta: np.ndarray) -> np.ndarray:
    """A Gauss-Newton step (translation, rotation vector) left-multiplied in the map frame (localize.py:227-233)."""
    from scipy.spatial.transform import Rotation

    Rz = Rotation.from_rotvec(delta[3:]).as_matrix()
    out = T.copy()
    out[:3, :3] = Rz @ T[:3, :3]
    out[:3, 3] = Rz @ T[:3, 3] + delta[:3]
    return out


def rotz(a: float) -> np.ndarray:
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


class CoarseSearch:
    """The search stages of the reference's GlobalLocalizer over a built PriorGrid, with its tuning names."""

    PTS_CAP = 1 << 17  # scan points after 5 cm thinning (localize.py:343)
    SCORE_CAP = 1 << 23  # (position, heading) hypotheses per coarse launch
    HYP_CAP = 1 << 17  # listed hypotheses per launch
    ICP_CAP = 16  # poses refined or evaluated together

    def __init__(self, grid: PriorGrid, device=None):
        torch = _torch()
        self.grid = grid
        self.device = torch.device(device) if device is not None else grid.device
        d = self.device
        self.d_pts = torch.empty((self.PTS_CAP, 3), dtype=torch.float32, device=d)
        self.d_pos = torch.empty((self.SCORE_CAP // 64, 3), dtype=torch.float32, device=d)
        self.d_score = torch.empty(self.SCORE_CAP, dtype=torch.float32, device=d)
        self.d_hyp = torch.empty((self.HYP_CAP, 4), dtype=torch.float32, device=d)
        self.d_T = torch.empty((self.ICP_CAP, 16), dtype=torch.float32, device=d)
        self.d_acc = torch.empty((self.ICP_CAP, 64), dtype=torch.float64, device=d)
        self.d_counts = torch.empty((self.ICP_CAP, 3), dtype=torch.int32, device=d)
        wb = int(_abi.lib().lw_plane_work_bytes(self.ICP_CAP))
        self.d_work = torch.empty(wb, dtype=torch.uint8, device=d)
        # tuning (localize.py:375-388)
        self.xy_step = 0.25
        self.yaw_step_deg = 2.0
        self.tau_coarse = 25.0  # cm
        self.tau_mid = 12.0
        self.free_cm = 10.0  # the sensor sits at least this far from any mapped surface
        self.n_coarse_pts = 1024
        self.n_mid_pts = 2048
        self.keep_coarse = 128  # coarse local maxima re-scored on the finer grid
        self.n_icp = 16  # medium winners handed to the ICP
        self.last_kernel_ms = 0.0

    def _stream(self, stream):
        torch = _torch()
        return stream if stream is not None else torch.cuda.current_stream(self.device)

    def _upload_pts(self, pts: np.ndarray, s):
        torch = _torch()
        m = len(pts)
        if m > self.PTS_CAP:
            raise ValueError(f"{m} scan points; at most {self.PTS_CAP}")
        with torch.cuda.stream(s):
            self.d_pts[:m].copy_(torch.from_numpy(np.ascontiguousarray(pts, np.float32)))
        return m

    def coarse(self, pts_g: np.ndarray, heights, stream=None):
        """Every position on the xy grid at the given heights x every heading (localize.py:442-465): (pos
        float32 (P, 3) in the reference's order (height, x, y), shape (heights, nx, ny, headings), the heading
        step, scores float32 (P * headings,) position-major). Positions go to the device x-fastest, so a wave's
        reads share cache lines; the scores come back in the reference's order."""
        torch = _torch()
        g = self.grid
        s = self._stream(stream)
        gx = np.arange(g.lo[0], g.hi[0] + 1e-6, self.xy_step)
        gy = np.arange(g.lo[1], g.hi[1] + 1e-6, self.xy_step)
        X, Y = np.meshgrid(gx, gy, indexing="ij")
        pos = np.concatenate([np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], 1) for z in heights])
        pos = pos.astype(np.float32)
        nz, nx, ny = len(heights), len(gx), len(gy)
        n_yaw = int(round(360.0 / self.yaw_step_deg))
        dyaw = math.radians(self.yaw_step_deg)
        m = self._upload_pts(pts_g, s)
        pos_x = np.ascontiguousarray(pos.reshape(nz, nx, ny, 3).transpose(0, 2, 1, 3).reshape(-1, 3))
        # each x-fastest position's row in the reference's (height, x, y) order, where the kernel writes it
        rows = np.ascontiguousarray(np.arange(len(pos), dtype=np.int32).reshape(nz, nx, ny).transpose(0, 2, 1)).reshape(-1)
        per = max(1, min(self.d_pos.shape[0], self.SCORE_CAP // n_yaw))
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ms = 0.0
        with torch.cuda.stream(s):
            out = torch.empty(len(pos) * n_yaw, dtype=torch.float32, device=self.device)
            d_rows = torch.from_numpy(rows).to(self.device)
        for a in range(0, len(pos_x), per):
            b = min(len(pos_x), a + per)
            with torch.cuda.stream(s):
                self.d_pos[: b - a].copy_(torch.from_numpy(pos_x[a:b]))
                ev[0].record(s)
                _abi.check(_abi.lib().lw_score_grid(self.d_pts.data_ptr(), m, self.d_pos.data_ptr(), b - a, n_yaw, 0.0,
                                                    dyaw, *g._geo_args(), g.df.data_ptr(), float(self.tau_coarse),
                                                    float(self.free_cm), d_rows[a:b].data_ptr(), out.data_ptr(),
                                                    ctypes.c_void_p(s.cuda_stream)), "lw_score_grid")
                ev[1].record(s)
            ev[1].synchronize()  # the next chunk's positions overwrite this one's
            ms += ev[0].elapsed_time(ev[1])
        self.last_kernel_ms = ms
        with torch.cuda.stream(s):
            sc = out.cpu().numpy()
        return pos, (nz, nx, ny, n_yaw), dyaw, sc

    @staticmethod
    def top(scores: np.ndarray, pool: int = 20000) -> np.ndarray:
        """Indices of the `pool` best scores, best first (localize.py:500-505)."""
        pool = min(pool, len(scores))
        idx = np.argpartition(-scores, pool - 1)[:pool] if pool < len(scores) else np.arange(len(scores))
        return idx[np.argsort(-scores[idx])]

    @classmethod
    def local_maxima(cls, sc: np.ndarray, shape, keep: int) -> np.ndarray:
        """Flat indices of the best `keep` coarse hypotheses that no grid neighbour in (x, y, heading) at their
        height beats (localize.py:467-497): from the pool of best positive scores, each compared with its 26
        neighbours (headings wrap, x and y stop at the grid's edge)."""
        nz, nx, ny, nyaw = shape
        pool = cls.top(sc)
        pool = pool[sc[pool] > 0]
        iyaw = pool % nyaw
        rest = pool // nyaw
        iy = rest % ny
        ix = (rest // ny) % nx
        iz = rest // (ny * nx)
        mine = sc[pool]
        keep_mask = np.ones(len(pool), bool)
        for ox in (-1, 0, 1):
            for oy in (-1, 0, 1):
                for oa in (-1, 0, 1):
                    if ox == 0 and oy == 0 and oa == 0:
                        continue
                    jx, jy = ix + ox, iy + oy
                    inside = (jx >= 0) & (jx < nx) & (jy >= 0) & (jy < ny)
                    j = ((iz * nx + np.clip(jx, 0, nx - 1)) * ny + np.clip(jy, 0, ny - 1)) * nyaw + (iyaw + oa) % nyaw
                    keep_mask &= ~inside | (sc[j] <= mine)
        return pool[keep_mask][:keep]

    @classmethod
    def nms(cls, hyps: np.ndarray, scores: np.ndarray, keep: int, dist: float, ang_deg: float) -> np.ndarray:
        """Greedy non-maximum suppression over (x, y, z, yaw) hypotheses (localize.py:507-528): indices of the
        survivors, best first; a hypothesis falls when a survivor is within `dist` and `ang_deg` of it."""
        chosen = []
        ang = math.radians(ang_deg)
        for i in cls.top(scores):
            if scores[i] <= 0:
                break
            h = hyps[i]
            clash = False
            for j in chosen:
                o = hyps[j]
                dyaw = abs((h[3] - o[3] + math.pi) % (2 * math.pi) - math.pi)
                if np.linalg.norm(h[:3] - o[:3]) < dist and dyaw < ang:
                    clash = True
                    break
            if not clash:
                chosen.append(i)
                if len(chosen) >= keep:
                    break
        return np.array(chosen, dtype=np.int64)

    def score_list(self, pts: np.ndarray, hyps: np.ndarray, tau: float, stream=None) -> np.ndarray:
        """The soft count of each (x, y, z, yaw) hypothesis for the scan points `pts` (localize.py:530-538)."""
        torch = _torch()
        g = self.grid
        s = self._stream(stream)
        m = self._upload_pts(pts, s)
        hyps = np.ascontiguousarray(hyps, np.float32)
        out = np.empty(len(hyps), np.float32)
        for a in range(0, len(hyps), self.HYP_CAP):
            b = min(len(hyps), a + self.HYP_CAP)
            with torch.cuda.stream(s):
                self.d_hyp[: b - a].copy_(torch.from_numpy(hyps[a:b]))
                _abi.check(_abi.lib().lw_score_list(self.d_pts.data_ptr(), m, self.d_hyp.data_ptr(), b - a,
                                                    *g._geo_args(), g.df.data_ptr(), float(tau),
                                                    self.d_score.data_ptr(), ctypes.c_void_p(s.cuda_stream)),
                           "lw_score_list")
                out[a:b] = self.d_score[: b - a].cpu().numpy()
        return out

    def medium(self, coarse_hyps: np.ndarray, pts_m: np.ndarray, stream=None):
        """The finer local grid around each coarse winner (localize.py:650-666): +-1 coarse step in x and y at a
        quarter step, +-4 cm in height, +-1 heading step at 0.5 deg, scored within tau_mid, then suppression
        within 0.1 m and 1.5 deg. Returns (all local hypotheses, their scores, the survivors' indices)."""
        off = np.arange(-self.xy_step, self.xy_step + 1e-6, self.xy_step / 4)
        offz = np.array([-0.04, 0.0, 0.04])
        offy = np.radians(np.arange(-self.yaw_step_deg, self.yaw_step_deg + 1e-6, 0.5))
        A, B, Cz, D = np.meshgrid(off, off, offz, offy, indexing="ij")
        local = np.stack([A.ravel(), B.ravel(), Cz.ravel(), D.ravel()], 1)
        hyps = (coarse_hyps[:, None, :] + local[None, :, :]).reshape(-1, 4)
        ms = self.score_list(pts_m, hyps, self.tau_mid, stream)
        return hyps, ms, self.nms(hyps, ms, self.n_icp, 0.1, 1.5)

    def coarse_candidate(self, pts_c: np.ndarray, heights, Rg: np.ndarray, stream=None) -> dict:
        """The coarse stage for one gravity candidate (localize.py:633-647): its kept local maxima as (x, y, z,
        yaw) hypotheses, their scores, and the best score per point."""
        pc = np.asarray(pts_c, float) @ Rg.T
        pos, shape, dyaw, sc = self.coarse(pc, heights, stream)
        n_yaw = shape[3]
        best = self.local_maxima(sc, shape, self.keep_coarse)
        hyps = np.concatenate([pos[best // n_yaw], ((best % n_yaw) * dyaw)[:, None]], 1) if len(best) else \
            np.zeros((0, 4))
        return {"Rg": Rg, "heights": heights, "coarse": hyps, "coarse_score": sc[best],
                "best_coarse": float(sc[best[0]]) / max(len(pc), 1) if len(best) else 0.0, "hypotheses": len(sc)}

    def search(self, pts_c: np.ndarray, pts_m: np.ndarray, heights, Rg: np.ndarray | None = None, stream=None):
        """Coarse, local maxima and medium stages for one gravity candidate: pts_c / pts_m are the coarse and
        medium subsets of the scan in the sensor frame, Rg rotates the sensor frame to the gravity-aligned one
        (identity for a level sensor). Returns (poses (k, 4, 4) sensor-to-map, medium scores per point, the
        coarse scores of the kept maxima per point), best first."""
        Rg = np.eye(3) if Rg is None else np.asarray(Rg, float)
        pc = np.asarray(pts_c, float) @ Rg.T
        pm = np.asarray(pts_m, float) @ Rg.T
        pos, shape, dyaw, sc = self.coarse(pc, heights, stream)
        n_yaw = shape[3]
        best = self.local_maxima(sc, shape, self.keep_coarse)
        if not len(best):
            return np.zeros((0, 4, 4)), np.zeros(0), np.zeros(0)
        coarse_hyps = np.concatenate([pos[best // n_yaw], ((best % n_yaw) * dyaw)[:, None]], 1)
        hyps, ms, win = self.medium(coarse_hyps, pm, stream)
        Ts = []
        for i in win:
            x, y, z, yaw = hyps[i]
            T = np.eye(4)
            T[:3, :3] = rotz(yaw) @ Rg
            T[:3, 3] = (x, y, z)
            Ts.append(T)
        return np.array(Ts).reshape(-1, 4, 4), ms[win] / max(len(pm), 1), sc[best] / max(len(pc), 1)

    # ---- point-to-plane ICP against the grid (localize.py:517-600) ----------------------------------------

    def _poses_up(self, Ts, s):
        torch = _torch()
        k = len(Ts)
        if k > self.ICP_CAP:
            raise ValueError(f"{k} poses; at most {self.ICP_CAP} together")
        with torch.cuda.stream(s):
            self.d_T[:k].copy_(torch.from_numpy(np.ascontiguousarray(np.array(Ts, np.float32).reshape(k, 16))))
        return k

    def icp(self, Ts, n_pts: int, schedule=((0.40, 8), (0.20, 8), (0.10, 8), (0.05, 10)), stream=None):
        """Point-to-plane ICP of every pose in Ts at once against the map, the scan already uploaded (its first
        n_pts points, see upload_scan): per stage (correspondence distance m, iterations) one accumulation on the
        device per iteration, then the damped 6x6 solve per pose on the host, until no pose moves more than
        2e-4 (rotation weighted x5) (localize.py:517-553)."""
        Ts = [np.array(T, dtype=np.float64) for T in Ts]
        for dmax, iters in schedule:
            for _ in range(iters):
                v = self.accumulate(Ts, n_pts, dmax, stream)
                k = len(Ts)
                moved = 0.0
                for h in range(k):
                    if v[h, ACC_COUNT] < 30:
                        continue
                    H = np.zeros((6, 6))
                    H[_TRIU] = v[h, ACC_H:ACC_H + 21]
                    H += np.triu(H, 1).T
                    b = v[h, ACC_B:ACC_B + 6]
                    A = H + np.diag(1e-3 * np.diag(H) + 1e-9)
                    try:
                        delta = np.linalg.solve(A, -b)
                    except np.linalg.LinAlgError:
                        continue
                    Ts[h] = twist_apply(Ts[h], delta)
                    moved = max(moved, float(np.linalg.norm(delta[:3])), float(np.linalg.norm(delta[3:])) * 5.0)
                if moved < 2e-4:
                    break
        return Ts

    def accumulate(self, Ts, n_pts: int, dmax: float, stream=None, alpha: float = 0.05) -> np.ndarray:
        """One point-to-plane accumulation for every pose in Ts against the map (correspondences within dmax,
        Geman-McClure scale dmax / 2): (k, 64) float64 rows in the accumulator layout, read back."""
        torch = _torch()
        g = self.grid
        s = self._stream(stream)
        k = self._poses_up(Ts, s)
        c2 = (dmax * 0.5) ** 2
        with torch.cuda.stream(s):
            _abi.check(_abi.lib().lw_plane_icp(self.d_pts.data_ptr(), n_pts, self.d_T.data_ptr(), k, *g._geo_args(),
                                               g.start.data_ptr(), g.nn.data_ptr(), g.xyz.data_ptr(), g.nrm.data_ptr(),
                                               g.planar.data_ptr(), float(dmax * dmax), float(c2), float(alpha),
                                               self.d_work.data_ptr(), self.d_work.numel(), self.d_acc.data_ptr(),
                                               ctypes.c_void_p(s.cuda_stream)), "lw_plane_icp")
            return self.d_acc[:k].cpu().numpy()

    def evaluate(self, Ts, n_pts: int, stream=None) -> np.ndarray:
        """Fraction of the uploaded scan within each of EVAL_RADII of the map, per pose (localize.py:555-565)."""
        torch = _torch()
        g = self.grid
        s = self._stream(stream)
        k = self._poses_up(Ts, s)
        radii = np.array(EVAL_RADII, np.float32)
        with torch.cuda.stream(s):
            _abi.check(_abi.lib().lw_plane_eval(self.d_pts.data_ptr(), n_pts, self.d_T.data_ptr(), k, *g._geo_args(),
                                                g.start.data_ptr(), g.nn.data_ptr(), g.xyz.data_ptr(),
                                                _abi.ptr(radii, _abi._f32p), self.d_counts.data_ptr(),
                                                ctypes.c_void_p(s.cuda_stream)), "lw_plane_eval")
            c = self.d_counts[:k].cpu().numpy()
        return c / max(n_pts, 1)

    def upload_scan(self, pts: np.ndarray, stream=None) -> int:
        """The scan (sensor frame) that icp() and evaluate() read: returns its point count."""
        return self._upload_pts(np.asarray(pts, np.float32), self._stream(stream))

GlobalLocalizer = CoarseSearch  # the reference's name for the search (localize.py:339)

"""Coarse stages of global localisation on the device: the reference GlobalLocalizer's search before its
batched ICP (python/livox_warp/localize.py:426-505 and the scoring kernels at :55-123).

Every free-space position on an xy grid at the given sensor heights, at every heading, is scored by how close
a gravity-aligned scan lands to the map (a soft count within tau of the prior map's distance field,
lw_score_grid); the best local maxima in (x, y, heading) are re-scored on a finer local grid with a tighter
tolerance and more points (lw_score_list), and greedy non-maximum suppression keeps the candidates the ICP
refines (loc.BatchLocalizer, or the reference's own). The host parts (grid layout, local maxima, the pool of
best scores, suppression) follow the reference step for step, so the same scores pick the same candidates.

ScanPrep here is the part of the reference's (localize.py:243-337) that the heights need: the 5 cm voxel
means, per-point normals from the 30 nearest neighbours within 0.2 m, and the height profile of the
horizontal points. The reference also finds "up" from RANSAC planes (open3d); here the caller gives it.
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


def rot_to(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """The rotation taking unit vector a onto unit vector b (localize.py:211-219)."""
    a = np.asarray(a, float) / np.linalg.norm(a)
    b = np.asarray(b, float) / np.linalg.norm(b)
    v = np.cross(a, b)
    c = float(a @ b)
    if np.linalg.norm(v) < 1e-9:
        return np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])  # a half turn about x
    K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + K + K @ K / (1.0 + c)


def pose_difference(A: np.ndarray, B: np.ndarray):
    """(translation m, rotation deg) between two 4x4 poses (localize.py:236-240)."""
    d = np.linalg.inv(A) @ B
    ang = math.degrees(math.acos(np.clip((np.trace(d[:3, :3]) - 1) / 2, -1, 1)))
    return float(np.linalg.norm(A[:3, 3] - B[:3, 3])), ang


class ScanPrep:
    """A scan in the sensor frame prepared for the height search (localize.py:243-337, less the RANSAC planes):
    one point per 5 cm voxel (the mean of its points, voxels counted from half a voxel below the scan's
    minimum as open3d's voxel_down_sample does) and each point's normal, the least-spread direction of its 30
    nearest neighbours within 0.2 m (sign arbitrary)."""

    def __init__(self, xyz: np.ndarray, voxel: float = 0.05, radius: float = 0.2, max_nn: int = 30,
                 max_tilt_deg: float = 45.0, seed: int = 0):
        from scipy.spatial import cKDTree

        p = np.asarray(xyz, np.float64)
        key = np.floor((p - (p.min(axis=0) - 0.5 * voxel)) / voxel).astype(np.int64)
        _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
        inv = inv.reshape(-1)
        sums = np.zeros((len(cnt), 3))
        np.add.at(sums, inv, p)
        self.pts = sums / cnt[:, None]
        d, nb = cKDTree(self.pts).query(self.pts, k=min(max_nn, len(self.pts)), distance_upper_bound=radius)
        d, nb = np.atleast_2d(d), np.atleast_2d(nb)
        ok = np.isfinite(d)
        q = self.pts[np.where(ok, nb, 0)] * ok[..., None]
        cnt = np.maximum(ok.sum(axis=1), 1)[:, None]
        mean = q.sum(axis=1) / cnt
        c = (q - mean[:, None, :]) * ok[..., None]
        cov = np.einsum("nki,nkj->nij", c, c)
        _, V = np.linalg.eigh(cov)
        self.normals = V[:, :, 0] * (ok.sum(axis=1) >= 3)[:, None]  # fewer than 3 neighbours: no normal
        # up to 8 planes, each the best of 1000 three-point samples (inliers within 3 cm) of the points no
        # earlier plane took, while one holds at least 150 points (localize.py:259-273)
        rng = np.random.default_rng(seed)
        self.planes = []  # (inliers, unit normal, offset, inlier indices)
        idx = np.arange(len(self.pts))
        for _ in range(8):
            if len(idx) < 300:
                break
            P = self.pts[idx]
            tri = P[rng.integers(0, len(P), (1000, 3))]
            nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
            ln = np.linalg.norm(nrm, axis=1)
            good = ln > 1e-12
            nrm = nrm[good] / ln[good, None]
            off = -np.einsum("ij,ij->i", nrm, tri[good, 0])
            cnt = np.zeros(len(nrm), np.int64)
            for a in range(0, len(nrm), 64):
                cnt[a:a + 64] = (np.abs(P @ nrm[a:a + 64].T + off[a:a + 64]) < 0.03).sum(axis=0)
            if not len(cnt):
                break
            b = int(np.argmax(cnt))
            inl = np.flatnonzero(np.abs(P @ nrm[b] + off[b]) < 0.03)
            if len(inl) < 150:
                break
            self.planes.append((len(inl), nrm[b], float(off[b]), idx[inl]))
            idx = np.delete(idx, inl)
        self.max_tilt = max_tilt_deg

    def up_candidates(self):
        """[(up unit vector, evidence)] ranked by plane support, both signs of each horizontal family: one normal
        fitted to all planes of a family together, the least-spread direction of the sum of their scatter
        matrices (localize.py:276-310); a level sensor either way up when no plane is horizontal enough."""
        cos_tilt = math.cos(math.radians(self.max_tilt))
        fams = []  # [support, [plane indices], seed normal]
        for i, (cnt, n, _, _) in enumerate(self.planes):
            if abs(n[2]) < cos_tilt:
                continue  # a wall, for a sensor that is roughly level (or upside down)
            for f in fams:
                if abs(float(n @ f[2])) > math.cos(math.radians(10.0)):
                    f[0] += cnt
                    f[1].append(i)
                    break
            else:
                fams.append([cnt, [i], n])
        out = []
        for support, members, n0 in sorted(fams, key=lambda f: -f[0])[:2]:
            S = np.zeros((3, 3))
            for i in members:
                P = self.pts[self.planes[i][3]]
                P = P - P.mean(0)
                S += P.T @ P
            _, V = np.linalg.eigh(S)
            n = V[:, 0]
            if n @ n0 < 0:
                n = -n
            out += [(n, support), (-n, support)]
        if not out:
            out = [(np.array([0.0, 0.0, 1.0]), 0), (np.array([0.0, 0.0, -1.0]), 0)]
        return out

    def height_profile(self, up: np.ndarray, lo: float, n_bins: int, bin_m: float):
        """Histogram of heights (along up) of the scan's horizontal points, relative to the sensor
        (localize.py:331-336)."""
        horiz = np.abs(self.normals @ up) > 0.97
        z = self.pts[horiz] @ up
        h, _ = np.histogram(z, bins=n_bins, range=(lo, lo + n_bins * bin_m))
        return h.astype(np.float32)


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
        self.seed = 0  # refine() draws its subsets from a fresh generator with this seed
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
        self.n_heights = 8  # sensor heights tried
        # unique: inliers (3 cm) at least accept_inliers, and the best answer elsewhere at least accept_gap lower
        # and below accept_ratio of it (localize.py:390-394)
        self.accept_inliers = 0.45
        self.accept_gap = 0.15
        self.accept_ratio = 0.75
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

    def refine(self, xyz: np.ndarray, T0: np.ndarray, stream=None, schedule=((0.30, 6), (0.12, 8), (0.05, 8)),
               starts=(), switch_margin: float = 0.05):
        """Point-to-plane ICP of a point set against the map from T0 (the pose-tracking step, localize.py:
        567-612): one point per 5 cm voxel, T0 and the offset starts (x, y, z m, optional heading deg about the
        vertical through the points' centre) refined together; an offset start wins only if its 3 cm fit beats
        the plain result by switch_margin. Returns (T, fit, info)."""
        rng = np.random.default_rng(self.seed)
        pts = np.asarray(xyz, dtype=np.float32)
        key = np.floor(pts / 0.05).astype(np.int64)
        _, first = np.unique(key, axis=0, return_index=True)
        pts = pts[np.sort(first)]
        if len(pts) > self.PTS_CAP:
            pts = pts[rng.choice(len(pts), self.PTS_CAP, replace=False)]
        n = len(pts)
        if n < 200:
            raise RuntimeError(f"too few points to refine ({n} after 5 cm thinning)")
        self.upload_scan(pts, stream)
        T0 = np.asarray(T0, dtype=np.float64)
        inits = [T0]
        c_map = T0[:3, :3] @ pts.mean(axis=0).astype(np.float64) + T0[:3, 3]
        for off in list(starts)[: self.ICP_CAP - 2]:
            Ti = T0.copy()
            off = np.asarray(off, dtype=np.float64)
            if len(off) > 3 and off[3] != 0.0:
                Rz = rotz(math.radians(off[3]))
                Ti[:3, :3] = Rz @ Ti[:3, :3]
                Ti[:3, 3] = Rz @ (Ti[:3, 3] - c_map) + c_map
            Ti[:3, 3] += off[:3]
            inits.append(Ti)
        Ts = self.icp(inits, n, schedule, stream)
        fits = self.evaluate(Ts + [T0], n, stream)
        fit_start = fits[-1]
        fits = fits[:-1]
        best = 0
        for i in range(1, len(Ts)):
            if fits[i, 0] > fits[best, 0] + (switch_margin if best == 0 else 0.0):
                best = i
        return Ts[best], fits[best], {"switched": best != 0, "plain_fit": float(fits[0, 0]), "n": n,
                                      "fit_start": float(fit_start[0])}

    # ---- the heights and the whole search for one gravity candidate (localize.py:426-446, 603-720) ---------

    def heights(self, prep: ScanPrep, up: np.ndarray):
        """Sensor heights where the scan's horizontal surfaces line up with the map's, best first
        (localize.py:426-446): the square roots of both height histograms, smoothed over +-2 bins, correlated;
        the best n_heights shifts at least 8 cm apart that put the sensor within the map's height +-0.5 m."""
        h_map, lo, bin_m = self.grid.levels
        span = 6.0
        nb = int(round(2 * span / bin_m))
        h_scan = prep.height_profile(up, -span, nb, bin_m)
        if h_scan.sum() < 50:
            return None
        k = np.array([0.25, 0.5, 1.0, 0.5, 0.25], np.float32)
        hm = np.convolve(np.sqrt(h_map), k, mode="same")
        hs = np.convolve(np.sqrt(h_scan), k, mode="same")
        corr = np.correlate(hm, hs, mode="full")
        shifts = np.arange(len(corr)) - (nb - 1)
        z_sensor = lo + shifts * bin_m + span
        ok = (z_sensor >= self.grid.lo[2] - 0.5) & (z_sensor <= self.grid.hi[2] + 0.5)
        corr = np.where(ok, corr, -1.0)
        out = []
        for i in np.argsort(-corr):
            if corr[i] <= 0 or len(out) >= self.n_heights:
                break
            z = float(z_sensor[i])
            if all(abs(z - o) > 0.08 for o in out):
                out.append(z)
        return out

    def _unique(self, best: float, alt: float) -> bool:
        return best >= self.accept_inliers and best - alt >= self.accept_gap and alt <= self.accept_ratio * best

    @staticmethod
    def _runner_up(results) -> float:
        best = results[0][3]
        for r in results[1:]:
            dt, ang = pose_difference(best, r[3])
            if dt > 1.0 or ang > 15.0:
                return r[0]
        return 0.0

    def localize(self, xyz: np.ndarray, up=None, up_prior=None, stream=None):
        """Sensor-to-map pose of a sensor-frame scan, and diagnostics: the reference's localize()
        (localize.py:614-720). 5 cm points (at most PTS_CAP), the coarse subset (1024, one per 15 cm voxel first)
        and the medium subset (2048); gravity candidates from the scan's planes (or `up`, sensor frame, alone),
        ordered by up_prior if given; for each, the heights (or every 0.25 m of the map's height when the scan
        shows no horizontal surface) and the coarse stage; then, best coarse score first, the medium stage, the
        grid ICP of its candidates and the 3 cm fit, until one candidate's best answer is unique."""
        rng = np.random.default_rng(self.seed)
        prep = ScanPrep(xyz, seed=self.seed)
        pts = prep.pts
        if len(pts) > self.PTS_CAP:
            pts = pts[rng.choice(len(pts), self.PTS_CAP, replace=False)]
        n_pts = len(pts)
        if n_pts < 200:
            raise RuntimeError(f"scan too sparse to localise ({n_pts} points after 5 cm thinning)")
        key = np.floor(pts / 0.15).astype(np.int64)
        _, first = np.unique(key, axis=0, return_index=True)
        spread = pts[first]
        sub_c = spread[rng.choice(len(spread), min(self.n_coarse_pts, len(spread)), replace=False)]
        sub_m = pts[rng.choice(n_pts, min(self.n_mid_pts, n_pts), replace=False)]
        if up is not None:
            ups = [(np.asarray(up, float) / np.linalg.norm(up), 0)]
        else:
            ups = prep.up_candidates()
            if up_prior is not None:
                upp = np.asarray(up_prior, float) / np.linalg.norm(up_prior)
                ups.sort(key=lambda u: -float(u[0] @ upp))
        cands = []
        for u, support in ups:
            Rg = rot_to(u, np.array([0.0, 0.0, 1.0]))  # sensor -> gravity-aligned sensor frame
            heights = self.heights(prep, u)
            if not heights:
                heights = list(np.arange(self.grid.lo[2], self.grid.hi[2] + 1e-6, self.xy_step))
            c = self.coarse_candidate(sub_c, heights, Rg, stream)
            c.update(up=u, support=support)
            cands.append(c)
        cands.sort(key=lambda c: -c["best_coarse"])
        results, refined = [], 0
        for ci, c in enumerate(cands):
            if not len(c["coarse"]):
                continue
            Rg = c["Rg"]
            hyps, ms, win = self.medium(c["coarse"], np.asarray(sub_m, float) @ Rg.T, stream)
            Ts = []
            for i in win:
                x, y, z, yaw = hyps[i]
                T = np.eye(4)
                T[:3, :3] = rotz(yaw) @ Rg
                T[:3, 3] = (x, y, z)
                Ts.append(T)
            if not Ts:
                continue
            self.upload_scan(pts, stream)
            Tr = self.icp(Ts, n_pts, stream=stream)
            fr = self.evaluate(Tr, n_pts, stream)
            for j, i in enumerate(win):
                results.append((float(fr[j, 0]), float(fr[j, 1]), float(fr[j, 2]), Tr[j], ci,
                                float(ms[i]) / max(len(sub_m), 1)))
            refined += 1
            results.sort(key=lambda r: -r[0])
            if self._unique(results[0][0], self._runner_up(results)):
                break  # this candidate's best answer is good and unique
        if not results:
            raise RuntimeError("no hypothesis survived (is the scan inside the map?)")
        results.sort(key=lambda r: -r[0])
        top = results[0]
        alt = self._runner_up(results)
        c = cands[top[4]]
        return top[3], {"inliers": top[0], "inliers5": top[1], "inliers10": top[2], "runner_up": alt,
                        "unique": self._unique(top[0], alt), "up": c["up"], "heights": c["heights"],
                        "gravity_candidates": len(cands), "refined_candidates": refined,
                        "hypotheses": sum(c["hypotheses"] for c in cands), "scan_points": n_pts}


GlobalLocalizer = CoarseSearch  # the reference's name for the search (localize.py:339)


class Localizer:
    """Single-stream convenience wrapper (localize.py:765-790): builds the grid from a map's points and normals
    and searches or refines on one stream."""

    def __init__(self, map_xyz: np.ndarray, map_normal: np.ndarray, map_planarity: np.ndarray | None = None,
                 device=None, stream=None):
        torch = _torch()
        host = {"xyz": map_xyz, "normal": map_normal}
        if map_planarity is not None:
            host["planarity"] = map_planarity
        self.grid = PriorGrid(host, device)
        self.loc = GlobalLocalizer(self.grid, self.grid.device)
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.grid.device)
        self.grid.build(self.stream)

    def search(self, xyz: np.ndarray, up_prior=None):
        """GlobalLocalizer.localize on this stream: the pose of a sensor-frame scan, and diagnostics."""
        return self.loc.localize(xyz, up_prior=up_prior, stream=self.stream)

    def refine(self, xyz: np.ndarray, T0: np.ndarray, **kw):
        """GlobalLocalizer.refine on this stream: ICP of a point set from the known pose T0."""
        return self.loc.refine(xyz, T0, self.stream, **kw)

"""Seeded synthetic inputs: analytic rooms, a rose-curve rosette scanner, smooth handheld walks.

Everything is numpy on the host and reproducible from a seed. Scenes are a room (the inside of an
axis-aligned box) plus solid axis-aligned boxes (furniture, pillars, interior walls). A scan fires
one ray per point along a rose-curve rosette (38.4 deg circular field of view, off-axis angle
19.2 deg * |sin(k theta)|) at 100 kHz, from the sensor pose at each point's own timestamp, and
returns the first hit with Gaussian range noise. Points are in the sensor frame (x forward), with
the attr word of the reference's sources (reflectivity bits 0-7, noise tag 8-15, return index 16-23,
sources.py:4-6).

The reference's own simulator (sources.py:297-304, 406-470) uses a two-prism rosette; this module
follows the workload descriptions instead (rose curve, seeded rooms and walks).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

HALF_FOV = math.radians(19.2)
RATE = 100_000.0  # points per second


@dataclass
class Scene:
    room_lo: np.ndarray
    room_hi: np.ndarray
    box_lo: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    box_hi: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    refl: np.ndarray = field(default_factory=lambda: np.zeros(0))

    def add_box(self, lo, hi, refl=40.0):
        self.box_lo = np.vstack([self.box_lo, np.asarray(lo, float)[None]])
        self.box_hi = np.vstack([self.box_hi, np.asarray(hi, float)[None]])
        self.refl = np.append(self.refl, refl)

    def translated(self, off) -> "Scene":
        off = np.asarray(off, float)
        return Scene(self.room_lo + off, self.room_hi + off, self.box_lo + off, self.box_hi + off, self.refl.copy())


def box_room(seed: int = 1) -> Scene:
    """An 8 x 6 x 3 m room with 6 axis-aligned boxes (edges 0.5-1.5 m) standing on the floor."""
    rng = np.random.default_rng(seed)
    S = Scene(np.array([0.0, 0.0, 0.0]), np.array([8.0, 6.0, 3.0]))
    placed = []
    while len(placed) < 6:
        e = rng.uniform(0.5, 1.5, 3)
        c = np.array([rng.uniform(0.8, 7.2), rng.uniform(0.8, 5.2)])
        lo = np.array([c[0] - e[0] / 2, c[1] - e[1] / 2, 0.0])
        hi = np.array([c[0] + e[0] / 2, c[1] + e[1] / 2, e[2]])
        if lo[0] < 0.1 or lo[1] < 0.1 or hi[0] > 7.9 or hi[1] > 5.9:
            continue
        if np.hypot(c[0] - 2.5, c[1] - 3.0) < 1.3:  # keep the sensor's spot free
            continue
        if any(not (hi[0] < q[0] - 0.2 or lo[0] > q[3] + 0.2 or hi[1] < q[1] - 0.2 or lo[1] > q[4] + 0.2)
               for q in placed):
            continue
        placed.append(np.concatenate([lo, hi]))
        S.add_box(lo, hi, refl=float(rng.uniform(20, 80)))
    return S


def office(seed: int = 2, path=None) -> Scene:
    """A 20 x 15 x 3 m office: random desks (0.75 m tall), partition walls and pillars, kept 1 m off `path`."""
    rng = np.random.default_rng(seed)
    S = Scene(np.array([0.0, 0.0, 0.0]), np.array([20.0, 15.0, 3.0]))

    def clear(lo, hi, margin=1.0):
        if path is None:
            return True
        dx = np.maximum(np.maximum(lo[0] - path[:, 0], path[:, 0] - hi[0]), 0.0)
        dy = np.maximum(np.maximum(lo[1] - path[:, 1], path[:, 1] - hi[1]), 0.0)
        return bool(np.min(np.hypot(dx, dy)) > margin)

    # pillars on a grid
    for x in (5.0, 10.0, 15.0):
        for y in (3.5, 11.5):
            lo, hi = (x - 0.25, y - 0.25, 0.0), (x + 0.25, y + 0.25, 3.0)
            if clear(lo, hi, 0.6):
                S.add_box(lo, hi, 45.0)
    # partition walls along the long sides
    for x0 in (2.0, 8.0, 14.0):
        lo, hi = (x0, 13.2, 0.0), (x0 + 3.0, 13.3, 2.2)
        if clear(lo, hi, 0.6):
            S.add_box(lo, hi, 35.0)
        lo, hi = (x0 + 1.0, 1.7, 0.0), (x0 + 4.0, 1.8, 1.6)
        if clear(lo, hi, 0.6):
            S.add_box(lo, hi, 35.0)
    # desks
    n, tries = 0, 0
    while n < 28 and tries < 4000:
        tries += 1
        w, d = rng.uniform(1.2, 1.8), rng.uniform(0.6, 0.9)
        if rng.random() < 0.5:
            w, d = d, w
        cx, cy = rng.uniform(0.6, 19.4), rng.uniform(0.6, 14.4)
        lo = np.array([cx - w / 2, cy - d / 2, 0.0])
        hi = np.array([cx + w / 2, cy + d / 2, 0.75])
        if lo[0] < 0.05 or lo[1] < 0.05 or hi[0] > 19.95 or hi[1] > 14.95 or not clear(lo, hi):
            continue
        ok = all(hi[0] < bl[0] - 0.3 or lo[0] > bh[0] + 0.3 or hi[1] < bl[1] - 0.3 or lo[1] > bh[1] + 0.3
                 for bl, bh in zip(S.box_lo, S.box_hi))
        if not ok:
            continue
        S.add_box(lo, hi, float(rng.uniform(30, 70)))
        n += 1
    return S


def multi_room_floor(seed: int = 3) -> Scene:
    """A 60 x 40 x 4 m floor split into rooms by interior walls with door gaps, plus furniture boxes."""
    rng = np.random.default_rng(seed)
    S = Scene(np.array([0.0, 0.0, 0.0]), np.array([60.0, 40.0, 4.0]))
    t = 0.15
    for x in (12.0, 24.0, 36.0, 48.0):  # walls of constant x, two door gaps each
        gaps = sorted(rng.uniform(2, 38, 2))
        y0 = 0.0
        for g in gaps:
            if g - 0.5 > y0:
                S.add_box((x, y0, 0.0), (x + t, g - 0.5, 4.0), 35.0)
            y0 = g + 0.5
        S.add_box((x, y0, 0.0), (x + t, 40.0, 4.0), 35.0)
    for y in (13.0, 26.0):  # walls of constant y, a door in every room
        x0 = 0.0
        for xs in (12.0, 24.0, 36.0, 48.0, 60.0):
            g = rng.uniform(x0 + 1.5, xs - 1.5)
            S.add_box((x0, y, 0.0), (g - 0.5, y + t, 4.0), 35.0)
            S.add_box((g + 0.5, y, 0.0), (xs, y + t, 4.0), 35.0)
            x0 = xs + t
    for _ in range(90):  # furniture
        e = rng.uniform(0.4, 2.0, 3) * np.array([1, 1, 0.6])
        c = rng.uniform([1, 1], [59, 39])
        S.add_box((c[0] - e[0] / 2, c[1] - e[1] / 2, 0.0), (c[0] + e[0] / 2, c[1] + e[1] / 2, e[2]),
                  float(rng.uniform(20, 80)))
    return S


# ---- ray casting ------------------------------------------------------------------------------


def cast(o: np.ndarray, d: np.ndarray, S: Scene, chunk: int = 8192):
    """First hit distance and reflectivity along rays o + t d (o inside the room)."""
    n = len(d)
    t_out = np.empty(n)
    r_out = np.empty(n)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        oo, dd = o[a:b], d[a:b]
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / np.where(np.abs(dd) < 1e-12, 1e-12, dd)
            ta = (S.room_lo[None] - oo) * inv
            tb = (S.room_hi[None] - oo) * inv
        t_room = np.max(np.stack([ta, tb]), axis=0).min(axis=1)
        wall = np.argmin(np.max(np.stack([ta, tb]), axis=0), axis=1)
        best = t_room
        refl = 20.0 + 8.0 * wall
        if len(S.box_lo):
            t1 = (S.box_lo[None] - oo[:, None]) * inv[:, None]
            t2 = (S.box_hi[None] - oo[:, None]) * inv[:, None]
            tmin = np.minimum(t1, t2).max(axis=2)
            tmax = np.maximum(t1, t2).min(axis=2)
            hit = (tmax >= np.maximum(tmin, 0.0)) & (tmin > 0.0)
            tb_ = np.where(hit, tmin, np.inf)
            j = np.argmin(tb_, axis=1)
            tj = tb_[np.arange(len(j)), j]
            closer = tj < best
            best = np.where(closer, tj, best)
            refl = np.where(closer, S.refl[j], refl)
        t_out[a:b] = best
        r_out[a:b] = refl
    return t_out, r_out


def cast_torch(o, d, S: Scene, chunk: int = 1 << 18):
    """cast() on torch float64 tensors (any device): the same slab tests, for bulk synthetic data."""
    import torch

    dev = o.device
    lo_r = torch.as_tensor(S.room_lo, dtype=torch.float64, device=dev)
    hi_r = torch.as_tensor(S.room_hi, dtype=torch.float64, device=dev)
    blo = torch.as_tensor(S.box_lo, dtype=torch.float64, device=dev)
    bhi = torch.as_tensor(S.box_hi, dtype=torch.float64, device=dev)
    brefl = torch.as_tensor(S.refl, dtype=torch.float64, device=dev)
    t_out = torch.empty(len(d), dtype=torch.float64, device=dev)
    r_out = torch.empty(len(d), dtype=torch.float64, device=dev)
    for a in range(0, len(d), chunk):
        oo, dd = o[a:a + chunk], d[a:a + chunk]
        inv = 1.0 / torch.where(dd.abs() < 1e-12, torch.full_like(dd, 1e-12), dd)
        ta, tb = (lo_r - oo) * inv, (hi_r - oo) * inv
        far = torch.maximum(ta, tb)
        best, wall = far.min(dim=1)
        refl = 20.0 + 8.0 * wall.to(torch.float64)
        if len(S.box_lo):
            t1 = (blo[None] - oo[:, None]) * inv[:, None]
            t2 = (bhi[None] - oo[:, None]) * inv[:, None]
            tmin = torch.minimum(t1, t2).amax(dim=2)
            tmax = torch.maximum(t1, t2).amin(dim=2)
            hit = (tmax >= tmin.clamp_min(0.0)) & (tmin > 0.0)
            tb_ = torch.where(hit, tmin, torch.full_like(tmin, float("inf")))
            tj, j = tb_.min(dim=1)
            closer = tj < best
            best = torch.where(closer, tj, best)
            refl = torch.where(closer, brefl[j], refl)
        t_out[a:a + chunk] = best
        r_out[a:a + chunk] = refl
    return t_out, r_out


# ---- poses ------------------------------------------------------------------------------------


def rot_zyx(yaw, pitch, roll) -> np.ndarray:
    """R = Rz(yaw) Ry(pitch) Rx(roll), vectorised over leading dims."""
    yaw, pitch, roll = np.broadcast_arrays(np.asarray(yaw, float), np.asarray(pitch, float), np.asarray(roll, float))
    cz, sz, cy, sy, cx, sx = np.cos(yaw), np.sin(yaw), np.cos(pitch), np.sin(pitch), np.cos(roll), np.sin(roll)
    R = np.empty(yaw.shape + (3, 3))
    R[..., 0, 0] = cz * cy
    R[..., 0, 1] = cz * sy * sx - sz * cx
    R[..., 0, 2] = cz * sy * cx + sz * sx
    R[..., 1, 0] = sz * cy
    R[..., 1, 1] = sz * sy * sx + cz * cx
    R[..., 1, 2] = sz * sy * cx - cz * sx
    R[..., 2, 0] = -sy
    R[..., 2, 1] = cy * sx
    R[..., 2, 2] = cy * cx
    return R


def pose_matrix(p, yaw, pitch, roll) -> np.ndarray:
    T = np.eye(4)
    T[:3, :3] = rot_zyx(yaw, pitch, roll)
    T[:3, 3] = p
    return T


class Walk:
    """A seeded smooth handheld walk around an elliptical loop: speed 0.5-1.2 m/s, yaw following the heading
    plus a look-around swing (|yaw rate| <= 60 deg/s), and smooth Gaussian jitter in every axis."""

    def __init__(self, seed: int = 4, centre=(10.0, 7.5), radii=(6.0, 4.2), height: float = 1.4, hold: float = 1.5):
        rng = np.random.default_rng(seed)
        self.hold = hold
        self.c = np.asarray(centre, float)
        self.ab = np.asarray(radii, float)
        self.h = height
        self.phase0 = rng.uniform(0, 2 * math.pi)
        self.v_f = rng.uniform(0.05, 0.12, 2)
        self.v_ph = rng.uniform(0, 2 * math.pi, 2)
        self.look_a = math.radians(rng.uniform(8, 15))
        self.look_f = rng.uniform(0.1, 0.2)
        # jitter: a few random sinusoids per axis with Gaussian amplitudes (x y z yaw pitch roll)
        self.j_amp = rng.normal(0, 1, (6, 4)) * np.array([0.006, 0.006, 0.008, 0.004, 0.006, 0.006])[:, None]
        self.j_f = rng.uniform(0.3, 2.5, (6, 4))
        self.j_ph = rng.uniform(0, 2 * math.pi, (6, 4))
        # arc length by integrating the speed on a fine grid
        self._tg = np.linspace(0.0, 400.0, 400_001)
        v = self.speed(self._tg)
        self._s = np.concatenate([[0.0], np.cumsum(0.5 * (v[1:] + v[:-1]) * np.diff(self._tg))])
        # the ellipse's arc length as a function of its parameter
        u = np.linspace(0, 2 * math.pi, 20001)
        du = np.hypot(self.ab[0] * np.sin(u), self.ab[1] * np.cos(u))
        self._u = u
        self._su = np.concatenate([[0.0], np.cumsum(0.5 * (du[1:] + du[:-1]) * np.diff(u))])
        self.perimeter = self._su[-1]

    def ramp(self, t):
        """0 while the sensor holds still (the odometry's warm-up frames assume it does, as the reference's
        simulator holds for 1.5 s, sources.py HOLD), then a smooth 1 s ramp up to 1."""
        u = np.clip((np.asarray(t, float) - self.hold) / 1.0, 0.0, 1.0)
        return u * u * (3.0 - 2.0 * u)

    def speed(self, t):
        t = np.asarray(t, float)
        s = 0.85 + 0.22 * np.sin(2 * math.pi * self.v_f[0] * t + self.v_ph[0]) \
            + 0.12 * np.sin(2 * math.pi * self.v_f[1] * t + self.v_ph[1])
        return np.clip(s, 0.5, 1.2) * self.ramp(t)

    def _param(self, t):
        s = np.interp(t, self._tg, self._s)
        lap, rem = np.divmod(s, self.perimeter)
        return np.interp(rem, self._su, self._u) + 2 * math.pi * lap + self.phase0

    def _jit(self, t, k):
        g = self.ramp(t)
        t = np.asarray(t, float)[..., None]
        return g * np.sum(self.j_amp[k] * np.sin(2 * math.pi * self.j_f[k] * t + self.j_ph[k]), axis=-1)

    def state(self, t):
        """(position (..., 3), yaw, pitch, roll) at times t."""
        t = np.asarray(t, float)
        u = self._param(t)
        x = self.c[0] + self.ab[0] * np.cos(u)
        y = self.c[1] + self.ab[1] * np.sin(u)
        heading = np.arctan2(self.ab[1] * np.cos(u), -self.ab[0] * np.sin(u))
        yaw = heading + self.ramp(t) * self.look_a * np.sin(2 * math.pi * self.look_f * t) + self._jit(t, 3)
        p = np.stack([x + self._jit(t, 0), y + self._jit(t, 1), self.h + self._jit(t, 2)], axis=-1)
        return p, yaw, self._jit(t, 4), self._jit(t, 5)

    def pose(self, t) -> np.ndarray:
        p, yaw, pitch, roll = self.state(t)
        T = np.zeros(np.shape(t) + (4, 4))
        T[..., :3, :3] = rot_zyx(yaw, pitch, roll)
        T[..., :3, 3] = p
        T[..., 3, 3] = 1.0
        return T

    def path(self, seconds: float = 60.0, step: float = 0.05) -> np.ndarray:
        return self.state(np.arange(0.0, seconds, step))[0]


class InPlace:
    """The reference simulator's other two motions about one spot of a walk (sources.py:197-200, 264-285):
    "static" holds the walk's start pose; "turn" holds it for 1.5 s, then turns in place with yaw
    0.7 sin(0.5 u) and pitch 0.03 sin(1.1 u) added to the start attitude, u = (t - 1.5) 2 pi / 40 (one slow
    swing of about 35 degrees over 15 s). Same interface as Walk."""

    HOLD = 1.5
    OMEGA = 2.0 * math.pi / 40.0

    def __init__(self, walk: Walk, motion: str):
        if motion not in ("static", "turn"):
            raise ValueError(f"unknown in-place motion {motion!r}")
        self.walk = walk
        self.turn = motion == "turn"
        p, yaw, pitch, roll = walk.state(0.0)
        self.p0, self.yaw0, self.pitch0, self.roll0 = np.asarray(p, float), float(yaw), float(pitch), float(roll)

    def state(self, t):
        t = np.asarray(t, float)
        u = np.maximum(t - self.HOLD, 0.0) * self.OMEGA
        dyaw = 0.7 * np.sin(0.5 * u) if self.turn else np.zeros_like(t)
        dpitch = 0.03 * np.sin(1.1 * u) if self.turn else np.zeros_like(t)
        p = np.broadcast_to(self.p0, np.shape(t) + (3,)).copy()
        return p, self.yaw0 + dyaw, self.pitch0 + dpitch, self.roll0 + np.zeros_like(t)

    pose = Walk.pose

    def path(self, seconds: float = 60.0, step: float = 0.05) -> np.ndarray:
        return self.walk.path(seconds, step)  # the scene is furnished around the walk, as for the walk itself


class Rosette:
    """Rose-curve scan directions: off-axis angle HALF_FOV * |sin(k theta)|, theta advancing with time, the whole
    pattern precessing slowly so successive frames fill the circle (k rational, from the seed)."""

    def __init__(self, seed: int = 5):
        rng = np.random.default_rng(seed)
        p = int(rng.integers(5, 12))
        q = int(rng.choice([2, 3, 4]))
        self.k = p / q
        self.omega = 2 * math.pi * float(rng.uniform(60.0, 90.0))
        self.prec = 2 * math.pi * float(rng.uniform(0.5, 1.5))

    def dirs(self, t: np.ndarray) -> np.ndarray:
        th = self.omega * t
        phi = HALF_FOV * np.abs(np.sin(self.k * th))
        az = th + self.prec * t
        return np.stack([np.cos(phi), np.sin(phi) * np.cos(az), np.sin(phi) * np.sin(az)], axis=-1)


def scan(S: Scene, rosette: Rosette, pose_fn, t0: float, n: int, seed: int, noise: float = 0.02,
         rate: float = RATE, tag_fraction: float = 0.0):
    """n points fired at t0 + i / rate, each from pose_fn(t_i) (4x4 per time), into scene S.

    Returns (xyz float32 (n, 3) in the sensor frame, attr uint32 (n,), t float32 (n,))."""
    rng = np.random.default_rng(seed)
    t = t0 + np.arange(n) / rate
    ds = rosette.dirs(t)
    T = pose_fn(t)
    if T.ndim == 2:
        T = np.broadcast_to(T, (n, 4, 4))
    dw = np.einsum("nij,nj->ni", T[:, :3, :3], ds)
    o = T[:, :3, 3]
    rng_true, refl = cast(o, dw, S)
    r = rng_true + rng.normal(0.0, noise, n)
    xyz = (ds * r[:, None]).astype(np.float32)
    attr = np.clip(refl + rng.normal(0, 3, n), 0, 255).astype(np.uint32)
    if tag_fraction > 0:
        tagged = rng.random(n) < tag_fraction
        attr = np.where(tagged, attr | np.uint32(1 << 8), attr).astype(np.uint32)
    return xyz, attr, t.astype(np.float32)


def sample_surfaces(S: Scene, spacing: float, seed: int = 7, prune: bool = True, normals: bool = False):
    """Points on every visible surface of the scene (room inside faces, box faces except the bottom) on a
    uniform grid of the given spacing with a seeded sub-spacing offset; world float32 (n, 3). prune=False keeps
    the few points hidden inside other boxes (faster for very large scenes). normals=True also returns each
    point's face normal (float32 (n, 3), unit, sign arbitrary): (points, normals)."""
    rng = np.random.default_rng(seed)
    out, nrm = [], []

    def face(origin, u, v, lu, lv):
        nu, nv = max(int(lu / spacing), 1), max(int(lv / spacing), 1)
        a = (np.arange(nu) + rng.uniform(0, 1)) * (lu / nu)
        b = (np.arange(nv) + rng.uniform(0, 1)) * (lv / nv)
        A, B = np.meshgrid(a, b, indexing="ij")
        out.append(origin[None] + A.reshape(-1, 1) * u[None] + B.reshape(-1, 1) * v[None])
        nrm.append(np.broadcast_to(np.cross(u, v), (A.size, 3)))

    ex, ey, ez = np.eye(3)

    def boxfaces(lo, hi, bottom: bool):
        L = hi - lo
        face(lo, ex, ey, L[0], L[1]) if bottom else None
        face(lo + ez * L[2], ex, ey, L[0], L[1])
        face(lo, ex, ez, L[0], L[2])
        face(lo + ey * L[1], ex, ez, L[0], L[2])
        face(lo, ey, ez, L[1], L[2])
        face(lo + ex * L[0], ey, ez, L[1], L[2])

    boxfaces(S.room_lo, S.room_hi, True)
    for lo, hi in zip(S.box_lo, S.box_hi):
        boxfaces(lo, hi, lo[2] > S.room_lo[2] + 1e-6)
    pts = np.concatenate(out)
    keep = np.ones(len(pts), bool)
    if prune:  # drop points hidden inside a box (faces that touch another box or the floor)
        for lo, hi in zip(S.box_lo, S.box_hi):
            keep &= ~np.all((pts > lo + 1e-6) & (pts < hi - 1e-6), axis=1)
    pts = (pts[keep] if prune else pts).astype(np.float32)
    if not normals:
        return pts
    n = np.concatenate(nrm)
    return pts, (n[keep] if prune else n).astype(np.float32)


@dataclass
class LocWorkload:
    """Batched global localisation: a multi-room floor prior map and query scans from seeded true poses."""

    map_xyz: np.ndarray
    scans: list
    T_true: np.ndarray  # (Q, 4, 4)
    scene: Scene


def loc_workload(seed: int = 31, n_queries: int = 4, spacing: float = 0.02, min_pts: int = 5000,
                 max_pts: int = 30000) -> LocWorkload:
    """The 60 x 40 x 4 m multi-room floor sampled at `spacing` (about 20M points at 20 mm), and n_queries
    rosette scans of log-uniform 5k-30k points from seeded poses: sensor at 1.4 m, away from walls, random yaw,
    level."""
    rng = np.random.default_rng(seed)
    S = multi_room_floor(seed)
    mp = sample_surfaces(S, spacing, seed + 1, prune=False)
    walls_x = (12.0, 24.0, 36.0, 48.0)
    walls_y = (13.0, 26.0)
    scans, poses = [], []
    ros = Rosette(seed + 2)
    while len(scans) < n_queries:
        x, y = rng.uniform(1.0, 59.0), rng.uniform(1.0, 39.0)
        if min(abs(x - w) for w in walls_x) < 0.6 or min(abs(y - w) for w in walls_y) < 0.6:
            continue
        T = pose_matrix(np.array([x, y, 1.4]), rng.uniform(0, 2 * math.pi), 0.0, 0.0)
        n = int(round(math.exp(rng.uniform(math.log(min_pts), math.log(max_pts)))))
        xyz, _, _ = scan(S, ros, lambda tt, T=T: T, 0.0, n, seed * 1000 + len(scans))
        scans.append(xyz)
        poses.append(T)
    return LocWorkload(mp, scans, np.stack(poses), S)


# ---- the workloads ------------------------------------------------------------------------------


@dataclass
class SingleScan:
    """One scan-to-map registration: a dense surface map and one 10k-point scan from a perturbed pose."""

    map_xyz: np.ndarray
    scan_xyz: np.ndarray
    scan_attr: np.ndarray
    scan_t: np.ndarray
    T_true: np.ndarray
    T_init: np.ndarray
    scene: Scene


def single_scan(seed: int = 11, n: int = 10_000, map_points: int = 600_000, spacing: float | None = None,
                offset=(0.0, 0.0, 0.0)) -> SingleScan:
    """The 8x6x3 m box-room registration: map sampled so it holds about `map_points` points (or at `spacing`),
    a rosette scan of n points from the true pose = origin + U(+-0.2 m) + U(+-5 deg) per Euler angle, range noise
    2 cm, timestamps over 0.1 s (the sensor holds still), registration starting at the unperturbed origin."""
    rng = np.random.default_rng(seed)
    S = box_room(seed)
    if spacing is None:
        area = sum(np.prod(np.sort(h - l)[1:]) for l, h in [(S.room_lo, S.room_hi)])  # rough, refined below
        probe = sample_surfaces(S, 0.05, seed)
        area = len(probe) * 0.05 * 0.05
        spacing = math.sqrt(area / map_points)
        del probe
    S = S.translated(offset)
    mp = sample_surfaces(S, spacing, seed + 1)
    origin = np.array([2.5, 3.0, 1.4]) + np.asarray(offset, float)
    yaw0 = rng.uniform(-0.5, 0.5)
    T_init = pose_matrix(origin, yaw0, 0.0, 0.0)
    dp = rng.uniform(-0.2, 0.2, 3)
    da = np.radians(rng.uniform(-5.0, 5.0, 3))
    T_true = pose_matrix(origin + dp, yaw0 + da[0], da[1], da[2])
    ros = Rosette(seed + 2)
    xyz, attr, t = scan(S, ros, lambda tt: T_true, 0.0, n, seed + 3)
    return SingleScan(mp, xyz, attr, t, T_true, T_init, S)


class Sequence:
    """A seeded odometry sequence: a walk through the office, scanned at 100 kHz, cut into batches on demand."""

    def __init__(self, seed: int = 21, seconds: float = 60.0, offset=(0.0, 0.0, 0.0), points_per_frame: int = 10_000,
                 frame_dt: float = 0.1, motion: str = "walk"):
        self.walk = Walk(seed) if motion == "walk" else InPlace(Walk(seed), motion)
        self.offset = np.asarray(offset, float)
        self.scene = office(seed + 1, path=self.walk.path(seconds + 1.0)).translated(self.offset)
        self.rosette = Rosette(seed + 2)
        self.seed = seed
        self.seconds = seconds
        self.rate = points_per_frame / frame_dt
        self.frame_dt = frame_dt

    def pose(self, t) -> np.ndarray:
        T = self.walk.pose(t)
        T[..., :3, 3] += self.offset
        return T

    def batch(self, t0: float, n: int, index: int):
        return scan(self.scene, self.rosette, self.pose, t0, n, self.seed * 100_003 + index, rate=self.rate)

    def frames_device(self, n_frames: int, device, seed: int | None = None):
        """All n_frames at once, generated on a torch device: (xyz (F, P, 3) float32, attr (F, P) int32, t (F, P)
        float32). Same scene, walk and rosette as frames(); per-point poses interpolated from a 1 kHz sampling of
        the walk, ray casting on the device, noise from a seeded torch generator (so the points differ from
        frames()'s, not their distribution)."""
        import torch

        per = int(round(self.rate * self.frame_dt))
        n = n_frames * per
        dev = torch.device(device)
        step = 1e-3
        tg = np.arange(0.0, n / self.rate + 3 * step, step)
        p, yaw, pitch, roll = self.walk.state(tg)
        g = np.column_stack([p + self.offset, np.unwrap(yaw), pitch, roll])
        G = torch.from_numpy(g).to(dev)
        t = torch.arange(n, dtype=torch.float64, device=dev) / self.rate
        u = t / step
        i0 = u.floor().long()
        f = (u - i0.to(torch.float64))[:, None]
        st = G[i0] * (1.0 - f) + G[i0 + 1] * f
        cz, sz = torch.cos(st[:, 3]), torch.sin(st[:, 3])
        cy, sy = torch.cos(st[:, 4]), torch.sin(st[:, 4])
        cx, sx = torch.cos(st[:, 5]), torch.sin(st[:, 5])
        R = torch.stack([cz * cy, cz * sy * sx - sz * cx, cz * sy * cx + sz * sx,
                         sz * cy, sz * sy * sx + cz * cx, sz * sy * cx - cz * sx,
                         -sy, cy * sx, cy * cx], dim=1).view(n, 3, 3)
        th = self.rosette.omega * t
        phi = HALF_FOV * torch.sin(self.rosette.k * th).abs()
        az = th + self.rosette.prec * t
        ds = torch.stack([torch.cos(phi), torch.sin(phi) * torch.cos(az), torch.sin(phi) * torch.sin(az)], dim=1)
        dw = torch.bmm(R, ds[:, :, None])[:, :, 0]
        rng_true, refl = cast_torch(st[:, :3].contiguous(), dw, self.scene)
        gen = torch.Generator(device=dev)
        gen.manual_seed(self.seed * 1_000_003 + 17 if seed is None else seed)
        r = rng_true + 0.02 * torch.randn(n, dtype=torch.float64, device=dev, generator=gen)
        xyz = (ds * r[:, None]).to(torch.float32).view(n_frames, per, 3)
        attr = (refl + 3.0 * torch.randn(n, dtype=torch.float64, device=dev, generator=gen)).clamp(0, 255)
        attr = attr.to(torch.int32).view(n_frames, per)
        return xyz.contiguous(), attr.contiguous(), t.to(torch.float32).view(n_frames, per).contiguous()

    def frames(self, n_frames: int | None = None):
        """Yield (xyz, attr, t) per frame_dt batch, n_frames of them (default: the whole sequence)."""
        n_frames = int(round(self.seconds / self.frame_dt)) if n_frames is None else n_frames
        per = int(round(self.rate * self.frame_dt))
        for k in range(n_frames):
            yield self.batch(k * self.frame_dt, per, k)

"""Oracle for the scan-to-map odometry path -- TEST INFRASTRUCTURE ONLY.

Only tests, the smoke check and the CPU-baseline leg of the benchmark may import this module, and
only as the checker: the product never routes through it.

Two layers:
  * `liblw_oracle.so` (lw_oracle.c): a serial C restatement of the reference's Warp kernels with
    Warp-CPU semantics (index order, float32 device math, float64 accumulators);
  * this file: ctypes bindings for it, a second, independent numpy int64 restatement of the key
    and hash functions, and the reference's host-side Gauss-Newton step restated with the same
    numpy/scipy routines the reference calls (odom.py:387-455 and 752-803).

Reference files (relative to the reference repository root): python/livox_warp/odom.py,
python/livox_warp/gpu.py, python/livox_warp/prior_session.py, python/livox_warp/localize.py.
"""

from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np
from scipy.spatial.transform import Rotation

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblw_oracle.so")

# odom.py:72-94
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
# gpu.py:59-64
VOXEL_KEY_BITS = 21
VOXEL_KEY_BIAS = 1 << (VOXEL_KEY_BITS - 1)
VOXEL_KEY_MASK = (1 << VOXEL_KEY_BITS) - 1
HASH_MULTIPLIER = -7046029254386353131  # 0x9E3779B97F4A7C15 as int64

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_libs = {}
LIB64_PATH = os.path.join(HERE, "liblw_oracle64.so")

# signatures; "R" is the device-side real (float32, or float64 in the all-float64 build), "RP" a pointer to it
_I, _L = ctypes.c_int, ctypes.c_int64
_SIGS = {
    "lwo_voxel_key_ijk": (ctypes.c_int64, [ctypes.c_int32] * 3),
    "lwo_slot_hash": (ctypes.c_int32, [ctypes.c_int64]),
    "lwo_find_slot": (ctypes.c_int32, [ctypes.c_int64, _i64p, ctypes.c_int32]),
    "lwo_frame_append": (None, ["RP", _u32p, "RP", _I, "R", _I, _I, _I, _i32p, "RP", _u32p, "RP", _i32p]),
    "lwo_deskew": (None, ["RP", "RP", _I, "R", "R", "RP", "RP", "RP"]),
    "lwo_ds_insert": (None, ["RP", _u32p, _I, "R", "R", "R", _I, _I, _i64p, _i32p, "RP"]),
    "lwo_reg_insert": (None, ["RP", _u32p, _I, "RP", "R", "R", "R", _I, _I, _i64p, "RP", "RP", _i32p, _i32p]),
    "lwo_eig3": (None, [_f64p, _f64p, _f64p]),
    "lwo_voxel_precision": (ctypes.c_int, ["RP", "RP", ctypes.c_int32, "R", "R", "R", "R", "RP", "RP", "RP"]),
    "lwo_icp": (None, ["RP", _I, "RP", "R", "R", _I, _i64p, "RP", "RP", _i32p, _I, "R", "R", "R", "R", "R", _f64p,
                       _i32p]),
    "lwo_reduce": (None, [_f64p, _f64p]),
    "lwo_map_insert": (None, [_f32p, _u32p, _f32p, _i32p, _I, _f32p, ctypes.c_float, ctypes.c_float, _I, _I, _I,
                              ctypes.c_float, ctypes.c_int32, _i64p, _f32p, _i32p, _f32p, _f32p, _i32p]),
    "lwo_sym_eig3": (None, [_f64p, _f64p, _f64p]),
    "lwo_local_plane": (None, [ctypes.c_int32] * 3 + [_f32p, _i64p, _f32p, _i32p, ctypes.c_int32, _f32p]),
    "lwo_carve": (None, [_f32p, _u32p, _i32p, _f32p, _I, _I, ctypes.c_float, ctypes.c_int32, ctypes.c_float, _I,
                         _i64p, _f32p, _i32p, _i32p]),
    "lwo_grid_build": (None, [_f32p, ctypes.c_int32, _f32p, _i32p, _i32p, ctypes.c_int32, _i32p, _i32p, _u8p]),
    "lwo_grid_nearest": (ctypes.c_int32, [_f32p, _f32p, _i32p, _i32p, _i32p, _f32p]),
    "lwo_grid_distances": (None, [_f32p, ctypes.c_int32, _f32p, _f32p, _i32p, _i32p, _i32p, _f32p, _f32p]),
    "lwo_score_grid": (None, [_f32p, ctypes.c_int32, _f32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                              ctypes.c_float, _f32p, _i32p, _u8p, ctypes.c_float, ctypes.c_float, _f32p]),
    "lwo_score_list": (None, [_f32p, ctypes.c_int32, _f32p, ctypes.c_int32, _f32p, _i32p, _u8p, ctypes.c_float,
                              _f32p]),
    "lwo_plane_icp": (None, [_f32p, ctypes.c_int32, _f32p, ctypes.c_int32, _f32p, _i32p, _i32p, _i32p, _f32p, _f32p,
                             _f32p, ctypes.c_float, ctypes.c_float, ctypes.c_float, _f64p]),
    "lwo_plane_eval": (None, [_f32p, ctypes.c_int32, _f32p, ctypes.c_int32, _f32p, _i32p, _i32p, _i32p, _f32p, _f32p,
                              _i32p]),
    "lwo_vox_insert": (None, [_f32p, _f32p, _f32p, _L, ctypes.c_float, ctypes.c_float, _L, _i64p, _f32p, _f32p,
                              _i32p, _i32p]),
    "lwo_vox_emit": (_L, [_i64p, _f32p, _f32p, _i32p, _L, ctypes.c_float, _f32p, _f32p, _i32p]),
}


def lib(fp64: bool = False):
    """The compiled restatement (built by `make -C oracle`): the float32 one (the reference's arithmetic), or
    with fp64=True the same code with every device-side step in float64."""
    if fp64 not in _libs:
        path = LIB64_PATH if fp64 else LIB_PATH
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        R = ctypes.c_double if fp64 else ctypes.c_float
        RP = _f64p if fp64 else _f32p
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = [R if a == "R" else RP if a == "RP" else a for a in args]
        _libs[fp64] = L
    return _libs[fp64]


def _p(a: np.ndarray, t):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(t)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _real(a, fp64: bool) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64 if fp64 else np.float32)


def _rp(fp64: bool):
    return _f64p if fp64 else _f32p


# ---- keys and hash: an independent numpy int64 restatement (gpu.py:231-252) -----------------------


def voxel_key_ijk_np(ix, iy, iz) -> np.ndarray:
    """gpu.py:231-240 in numpy int64 (wrapping two's-complement arithmetic)."""
    b, m = np.int64(VOXEL_KEY_BIAS), np.int64(VOXEL_KEY_MASK)
    x = (np.asarray(ix, np.int64) + b) & m
    y = (np.asarray(iy, np.int64) + b) & m
    z = (np.asarray(iz, np.int64) + b) & m
    return x | (y << np.int64(VOXEL_KEY_BITS)) | (z << np.int64(2 * VOXEL_KEY_BITS))


def slot_hash_np(k) -> np.ndarray:
    """gpu.py:247-252 in numpy int64: h = k * MUL (wrapping); h ^= h >> 29 (arithmetic); h & 0x7fffffff."""
    with np.errstate(over="ignore"):
        h = np.asarray(k, np.int64) * np.int64(HASH_MULTIPLIER)
    h = h ^ (h >> np.int64(29))
    return (h & np.int64(0x7FFFFFFF)).astype(np.int32)


def voxel_key_np(w: np.ndarray, inv: float) -> np.ndarray:
    """gpu.py:242-245: float32 multiply, floor, int32, then the key."""
    w = _f32(w)
    ijk = np.floor(w * np.float32(inv)).astype(np.int32)
    return voxel_key_ijk_np(ijk[..., 0], ijk[..., 1], ijk[..., 2])


def transform_points_f32(T, p: np.ndarray) -> np.ndarray:
    """wp.transform_point in float32 with the row order ((m0 p0 + m1 p1) + m2 p2) + m3 (no fused multiply-add)."""
    T = _f32(np.asarray(T).reshape(4, 4))
    p = _f32(p)
    out = np.empty_like(p)
    for r in range(3):
        a = T[r, 0] * p[:, 0]
        a = a + T[r, 1] * p[:, 1]
        a = a + T[r, 2] * p[:, 2]
        out[:, r] = a + T[r, 3]
    return out


# ---- thin wrappers over the C restatement ---------------------------------------------------------


def voxel_key_ijk(ix: int, iy: int, iz: int) -> int:
    return int(lib().lwo_voxel_key_ijk(ix, iy, iz))


def slot_hash(k: int) -> int:
    return int(lib().lwo_slot_hash(k))


class OracleMap:
    """The registration map of odom.py:468-474 (SoA, per slot) built by the serial k_reg_insert. fp64=True keeps
    the moments in float64 (the all-float64 oracle)."""

    def __init__(self, slots: int, fp64: bool = False):
        assert slots & (slots - 1) == 0
        self.cap = slots
        self.fp64 = fp64
        rt = np.float64 if fp64 else np.float32
        self.keys = np.full(slots, -1, np.int64)
        self.rsum = np.zeros((slots, 3), rt)
        self.router = np.zeros((slots, 3, 3), rt)
        self.rcnt = np.zeros(slots, np.int32)
        self.stats = np.zeros(2, np.int32)

    def insert(self, xyz, attr, T, voxel: float, max_range: float = 60.0, ret_mask: int = 0b1111):
        """odom.py:186-230, driven as Odometry._insert drives it (odom.py:719-727)."""
        fp = self.fp64
        xyz = _real(xyz, fp).reshape(-1, 3)
        attr = np.ascontiguousarray(attr, np.uint32)
        Tr = _real(np.asarray(T, np.float64).reshape(4, 4), fp)
        rp = _rp(fp)
        lib(fp).lwo_reg_insert(_p(xyz, rp), _p(attr, _u32p), len(xyz), _p(Tr, rp), 1.0 / voxel, voxel, max_range,
                               int(ret_mask), self.cap - 1, _p(self.keys, _i64p), _p(self.rsum, rp),
                               _p(self.router, rp), _p(self.rcnt, _i32p), _p(self.stats, _i32p))

    def find_slot(self, key: int) -> int:
        return int(lib().lwo_find_slot(key, _p(self.keys, _i64p), self.cap - 1))

    def occupied(self) -> np.ndarray:
        return self.keys != -1


def frame_append(sx, sa, st, inv_dt: float, k_min: int, k_open: int, cap: int, count: int,
                 f_xyz, f_attr, f_t, f_frame, fp64: bool = False) -> int:
    """odom.py:97-123; returns the new count (which, as in the reference, may exceed cap)."""
    sx, st = _real(sx, fp64).reshape(-1, 3), _real(st, fp64)
    sa = np.ascontiguousarray(sa, np.uint32)
    c = np.array([count], np.int32)
    rp = _rp(fp64)
    lib(fp64).lwo_frame_append(_p(sx, rp), _p(sa, _u32p), _p(st, rp), len(st), inv_dt, k_min, k_open, cap,
                               _p(c, _i32p), _p(f_xyz, rp), _p(f_attr, _u32p), _p(f_t, rp), _p(f_frame, _i32p))
    return int(c[0])


def deskew(raw, t, t_mid: float, inv_dt: float, xi, fp64: bool = False) -> np.ndarray:
    """odom.py:163-178 in float32 (or float64)."""
    raw, t = _real(raw, fp64).reshape(-1, 3), _real(t, fp64)
    xi = _real(xi, fp64)
    rho, theta = _real(xi[:3], fp64), _real(xi[3:], fp64)
    out = np.empty_like(raw)
    rp = _rp(fp64)
    lib(fp64).lwo_deskew(_p(raw, rp), _p(t, rp), len(t), t_mid, inv_dt, _p(rho, rp), _p(theta, rp), _p(out, rp))
    return out


def deskew_fp64(p: np.ndarray, s: np.ndarray, xi) -> np.ndarray:
    """prior_session.py:51-57, the reference's own host restatement of gpu.deskew, in float64."""
    p = np.asarray(p, np.float64)
    if xi is None or not np.any(xi):
        return p
    rv = s[:, None] * np.asarray(xi[3:], dtype=np.float64)[None, :]
    return Rotation.from_rotvec(rv).apply(p) + s[:, None] * np.asarray(xi[:3], dtype=np.float64)[None, :]


def ds_insert(xyz, attr, inv_v: float, min_range: float, max_range: float, ret_mask: int, ds_cap: int,
              fp64: bool = False):
    """odom.py:126-160 with a freshly cleared table of ds_cap slots (odom.py:731-737): the thinned points in
    index order (the serial semantics)."""
    xyz = _real(xyz, fp64).reshape(-1, 3)
    attr = np.ascontiguousarray(attr, np.uint32)
    keys = np.full(ds_cap, -1, np.int64)
    cnt = np.zeros(1, np.int32)
    out = np.zeros_like(xyz)
    rp = _rp(fp64)
    lib(fp64).lwo_ds_insert(_p(xyz, rp), _p(attr, _u32p), len(xyz), inv_v, min_range, max_range, int(ret_mask),
                            ds_cap - 1, _p(keys, _i64p), _p(cnt, _i32p), _p(out, rp))
    return out[: int(cnt[0])].copy()


def voxel_precision(rsum, router, c: int, voxel: float, noise2: float, tight2: float, alpha: float):
    """(shape flag, M6, Mg6, q) of one voxel: the eigen-weighting of odom.py:336-368."""
    rs, ro = _f32(rsum), _f32(router)
    M6, Mg6, q = np.zeros(6, np.float32), np.zeros(6, np.float32), np.zeros(3, np.float32)
    shape = lib().lwo_voxel_precision(_p(rs, _f32p), _p(ro, _f32p), int(c), voxel, noise2, tight2, alpha,
                                      _p(M6, _f32p), _p(Mg6, _f32p), _p(q, _f32p))
    return int(shape), M6, Mg6, q


def eig3(A) -> tuple[np.ndarray, np.ndarray]:
    A = np.ascontiguousarray(A, np.float64).reshape(3, 3)
    ev, Q = np.zeros(3), np.zeros((3, 3))
    lib().lwo_eig3(_p(A, _f64p), _p(ev, _f64p), _p(Q, _f64p))
    return ev, Q


def icp_accumulate(omap: OracleMap, ds, T, voxel: float, reach: int, max_d2: float, noise2: float, tight2: float,
                   alpha: float, kappa: float, with_slots: bool = False):
    """One k_icp pass + k_reduce (odom.py:275-384): the 64 reduced float64 entries of the ACC layout (device math
    in the map's arithmetic: float32, or float64 for an fp64 map)."""
    fp = omap.fp64
    rp = _rp(fp)
    ds = _real(ds, fp).reshape(-1, 3)
    Tr = _real(np.asarray(T, np.float64).reshape(4, 4), fp)
    acc = np.zeros(ACC_SLOTS * ACC_STRIDE, np.float64)
    best = np.zeros(len(ds), np.int32) if with_slots else None
    lib(fp).lwo_icp(_p(ds, rp), len(ds), _p(Tr, rp), 1.0 / voxel, voxel, omap.cap - 1, _p(omap.keys, _i64p),
                    _p(omap.rsum, rp), _p(omap.router, rp), _p(omap.rcnt, _i32p), reach, max_d2, noise2, tight2,
                    alpha, kappa, _p(acc, _f64p), _p(best, _i32p) if best is not None else None)
    out = np.zeros(ACC_STRIDE, np.float64)
    lib(fp).lwo_reduce(_p(acc, _f64p), _p(out, _f64p))
    return (out, best) if with_slots else out


# ---- host step (odom.py:387-455), restated with the routines the reference calls --------------------

_TRIU6 = np.triu_indices(6)


def sym6(upper: np.ndarray) -> np.ndarray:
    H = np.zeros((6, 6))
    H[_TRIU6] = upper
    return H + np.triu(H, 1).T


def twist_between(T_from: np.ndarray, T_to: np.ndarray) -> np.ndarray:
    M = T_to @ np.linalg.inv(T_from)
    theta = Rotation.from_matrix(M[:3, :3]).as_rotvec()
    return np.concatenate([M[:3, 3], theta])


def apply_twist(T: np.ndarray, delta: np.ndarray) -> np.ndarray:
    Rz = Rotation.from_rotvec(delta[3:]).as_matrix()
    out = T.copy()
    out[:3, :3] = Rz @ T[:3, :3]
    out[:3, 3] = Rz @ T[:3, 3] + delta[:3]
    return out


def body_twist(d: np.ndarray) -> np.ndarray:
    return np.concatenate([d[:3, 3], Rotation.from_matrix(d[:3, :3]).as_rotvec()]).astype(np.float32)


def twist_matrix(xi: np.ndarray) -> np.ndarray:
    out = np.eye(4)
    out[:3, :3] = Rotation.from_rotvec(np.asarray(xi[3:], dtype=np.float64)).as_matrix()
    out[:3, 3] = xi[:3]
    return out


def sensor_basis(T: np.ndarray, L: float) -> np.ndarray:
    c = T[:3, 3]
    J = np.eye(6)
    J[:3, 3:] = np.array([[0.0, -c[2], c[1]], [c[2], 0.0, -c[0]], [-c[1], c[0], 0.0]])
    B = np.zeros((6, 6))
    B[:3, :3] = B[3:, 3:] = T[:3, :3]
    return J @ B @ np.diag([1.0, 1.0, 1.0, 1.0 / L, 1.0 / L, 1.0 / L])


def constrained_step(A, rhs, Hg, T, L, thr, to_pred):
    J = sensor_basis(T, L)
    lam, V = np.linalg.eigh(J.T @ Hg @ J)
    good = lam >= thr
    if good.all():
        return np.linalg.solve(A, rhs), (lam, V, good)
    Vd = V[:, ~good]
    d0 = Vd @ (Vd.T @ np.linalg.solve(J, to_pred))
    if good.any():
        Vg = V[:, good]
        As = J.T @ A @ J
        d0 = d0 + Vg @ np.linalg.solve(Vg.T @ As @ Vg, Vg.T @ (J.T @ rhs - As @ d0))
    return J @ d0, (lam, V, good)


@dataclass
class Params:
    """The Odometry tuning attributes of odom.py:499-525 that registration reads."""

    reg_voxel: float = 0.15
    max_iter: int = 16
    min_corr: int = 60
    min_range: float = 0.3
    max_range: float = 60.0
    noise: float = 0.02
    ret_mask: int = 0b1111
    alpha: float = 0.05
    damping: float = 1.0e-3
    degen_thr: float = 3000.0
    prior_info: float = 2000.0
    wide_iters: int = WIDE_ITERS


@dataclass
class RegResult:
    T: np.ndarray
    ok: bool
    iters: int = 0
    corr: int = 0
    rms: float = 0.0
    cond: float = 1.0
    weak: bool = False
    degen: int = 0
    n_ds: int = 0
    accs: list = field(default_factory=list)


def gn_step(v: np.ndarray, T: np.ndarray, T_pred: np.ndarray, P: Params):
    """The body of one iteration of odom.py:768-793 after the accumulation, given the reduced accumulator v.
    Returns (delta or None on LinAlgError, n_degen, rms, H)."""
    corr = int(v[ACC_COUNT])
    H, b, Hg = sym6(v[ACC_H:ACC_H + 21]), v[ACC_B:ACC_B + 6], sym6(v[ACC_HG:ACC_HG + 21])
    L = max(v[ACC_RANGE] / corr, 1.0)
    rms = math.sqrt(max(v[ACC_RMS], 0.0) / max(v[ACC_WEIGHT], 1e-9))
    dg = np.diag(H)
    A = H + np.diag(1e-9 * dg.mean() + P.damping * dg)
    rhs = -b
    xi = twist_between(T, T_pred)
    if P.prior_info > 0.0:
        Ji = np.linalg.inv(sensor_basis(T, L))
        Pm = P.prior_info * (Ji.T @ Ji)
        A = A + Pm
        rhs = rhs + Pm @ xi
    n_degen = 0
    try:
        if P.degen_thr > 0.0:
            delta, (_, _, kept) = constrained_step(A, rhs, Hg, T, L, P.degen_thr, xi)
            n_degen = int((~kept).sum())
        else:
            delta = np.linalg.solve(A, rhs)
    except np.linalg.LinAlgError:
        return None, n_degen, rms, H
    return delta, n_degen, rms, H


def register(omap: OracleMap, ds: np.ndarray, T_pred: np.ndarray, P: Params, keep_accs: bool = False) -> RegResult:
    """odom.py:752-803 from the thinned points onward (the thinning itself is ds_insert)."""
    n_ds = len(ds)
    if n_ds < P.min_corr:
        return RegResult(T=T_pred.copy(), ok=False, n_ds=n_ds)
    T = T_pred.copy()
    voxel = P.reg_voxel
    rt = np.float64 if omap.fp64 else np.float32  # the kernel arguments' type (odom.py:744-746)
    noise2 = rt(P.noise * P.noise)
    tight2 = rt((TIGHT_FRACTION * voxel / math.sqrt(12.0)) ** 2)
    iters, corr, rms, cond, n_degen = 0, 0, 0.0, 1.0, 0
    H = np.eye(6)
    accs = []
    for it in range(P.max_iter):
        reach = 2 if it < P.wide_iters else 1
        max_d2 = ((reach + 0.5) * voxel) ** 2
        kappa = max(KAPPA_MIN, KAPPA_START * (KAPPA_DECAY ** it))
        v = icp_accumulate(omap, ds, T, voxel, reach, max_d2, float(noise2), float(tight2), P.alpha, kappa)
        if keep_accs:
            accs.append((T.copy(), reach, max_d2, kappa, v.copy()))
        corr = int(v[ACC_COUNT])
        if corr < P.min_corr:
            break
        delta, n_degen, rms, H = gn_step(v, T, T_pred, P)
        if delta is None:
            break
        T = apply_twist(T, delta)
        iters = it + 1
        if np.linalg.norm(delta[:3]) < STEP_TOL and np.linalg.norm(delta[3:]) < STEP_TOL:
            break
    if corr >= P.min_corr:
        dg = np.sqrt(np.maximum(np.diag(H), 1e-12))
        ev = np.linalg.eigvalsh(H / np.outer(dg, dg))
        cond = float(max(ev[0], 0.0) / max(ev[-1], 1e-12))
    return RegResult(T=T, ok=corr >= P.min_corr, iters=iters, corr=corr, rms=rms, cond=cond,
                     weak=cond < WEAK_COND or n_degen > 0, degen=n_degen, n_ds=n_ds, accs=accs)


def pose_error(A: np.ndarray, B: np.ndarray) -> tuple[float, float]:
    """(translation m, rotation deg) between two 4x4 poses."""
    d = np.linalg.inv(A) @ B
    ang = math.degrees(math.acos(np.clip((np.trace(d[:3, :3]) - 1) / 2, -1, 1)))
    return float(np.linalg.norm(A[:3, 3] - B[:3, 3])), ang


class OracleOdometry:
    """odom.py:458-803 over host arrays: frame cutting, warm-up, deskew, thinning, registration with the numpy
    host step every iteration, keyframe inserts, constant-velocity prediction. Poses go to `pipe` like the
    reference's (set_pose / set_pose_range)."""

    def __init__(self, pipe, reg_slots: int = 1 << 21, frame_cap: int = 1 << 18, fp64: bool = False):
        self.pipe = pipe
        self.fp64 = fp64
        rt = np.float64 if fp64 else np.float32
        self.reg_cap = reg_slots
        self.frame_cap = frame_cap
        self.ds_cap = 1 << (frame_cap - 1).bit_length()
        self.f_raw = np.zeros((frame_cap, 3), rt)
        self.f_xyz = np.zeros((frame_cap, 3), rt)
        self.f_attr = np.zeros(frame_cap, np.uint32)
        self.f_t = np.zeros(frame_cap, rt)
        self.f_frame = np.zeros(frame_cap, np.int32)
        self.f_count = 0
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
        self.reset(np.eye(4))

    def reset(self, T0):
        self.map = OracleMap(self.reg_cap, fp64=self.fp64)
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
        self.stats = {"frames": 0, "iters": 0, "corr": 0, "ds": 0, "rms": 0.0, "cond": 1.0, "weak": False,
                      "degen": 0, "skipped": 0, "gaps": 0, "keyframes": 0, "map_voxels": 0}
        self.f_count = 0

    def restart_clock(self):
        self.k_open = None
        self.k_pred = -1
        self.T_prev = None
        self.delta = np.eye(4)
        self.xi = np.zeros(6, np.float32)
        self.xi_s = np.zeros(6)
        self.f_count = 0

    def predict(self):
        return self.T @ self.delta

    def params(self) -> Params:
        return Params(reg_voxel=self.reg_voxel, max_iter=self.max_iter, min_corr=self.min_corr,
                      min_range=self.min_range, max_range=self.max_range, noise=self.noise, ret_mask=self.ret_mask,
                      alpha=self.alpha, damping=self.damping, degen_thr=self.degen_thr, prior_info=self.prior_info)

    def push(self, sx, sa, st, n, t_min, t_max, on_frame=None):
        if n <= 0:
            return
        inv_dt = 1.0 / self.frame_dt
        k_lo = int(math.floor(max(t_min, 0.0) * inv_dt))
        k_hi = int(math.floor(max(t_max, 0.0) * inv_dt))
        k_prev = self.k_pred
        if self.k_open is not None and k_hi - self.k_open > self.max_gap:
            if self.f_count > 0:
                self._finalize(self.k_open, on_frame)
            self.restart_clock()
            self.stats["gaps"] += 1
        if self.k_open is None:
            self.k_open = k_lo if k_hi - k_lo <= self.max_gap else k_hi
            self.f_count = 0
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
            self.f_count = frame_append(sx, sa, st, inv_dt, k_min, self.k_open, self.frame_cap, self.f_count,
                                        self.f_raw, self.f_attr, self.f_t, self.f_frame, fp64=self.fp64)
            if k_hi <= self.k_open:
                break
            self._finalize(self.k_open, on_frame)
            self.k_open += 1
            k_min = self.k_open
            self.f_count = 0
            self.pipe.set_pose(self.k_open, self.predict(), self.xi, self.epoch)
            self.k_pred = max(self.k_pred, self.k_open)

    def _finalize(self, k, on_frame):
        m = min(self.f_count, self.frame_cap)
        warm = self.map_frames < self.warmup
        t_mid = (k + 0.5) * self.frame_dt
        T, xi, ok = self._estimate(m, t_mid, warm)
        if ok and (warm or self._is_keyframe(T)):
            self._insert(m, T)
        if self.T_prev is not None and not warm:
            self._update_velocity(np.linalg.inv(self.T_prev) @ T)
        self.T_prev = self.T = T
        self.pipe.set_pose(k, T, xi, self.epoch)
        self.traj.append((k, t_mid, T.copy()))
        self.stats["frames"] += 1
        self.stats["map_voxels"] = int(self.map.stats[0])
        if on_frame is not None:
            on_frame(k, T, self.f_xyz, self.f_attr, self.f_t, self.f_frame, m)

    def _estimate(self, m, t_mid, warm):
        xi = self.xi if (self.deskew and not warm) else np.zeros(6, np.float32)
        if m > 0:
            self.f_xyz[:m] = deskew(self.f_raw[:m], self.f_t[:m], t_mid, 1.0 / self.frame_dt, xi, self.fp64)
        T = self.T if warm else self.predict()
        if m < self.min_points:
            self.stats["skipped"] += 1
            return T, xi, False
        if warm:
            return T, xi, True
        T, ok = self._register(m, T)
        if ok and self.deskew and self.redeskew and self.T_prev is not None:
            xi_new = body_twist(np.linalg.inv(self.T_prev) @ T)
            if (np.linalg.norm(xi_new[3:] - xi[3:]) > REDESKEW_ROT
                    or np.linalg.norm(xi_new[:3] - xi[:3]) > REDESKEW_TRANS):
                xi = xi_new
                self.f_xyz[:m] = deskew(self.f_raw[:m], self.f_t[:m], t_mid, 1.0 / self.frame_dt, xi, self.fp64)
                T, ok = self._register(m, T)
        return T, xi, ok

    def _update_velocity(self, d):
        if np.linalg.norm(d[:3, 3]) < MAX_FRAME_JUMP:
            self.delta = d
            self.xi = body_twist(d)
            if self.velocity_smooth > 0.0:
                a = self.velocity_smooth
                self.xi_s = a * self.xi_s + (1.0 - a) * self.xi.astype(np.float64)
                self.delta = twist_matrix(self.xi_s)
        else:
            self.delta = np.eye(4)
            self.xi = np.zeros(6, np.float32)
            self.xi_s = np.zeros(6)

    def _is_keyframe(self, T):
        if self.T_key is None:
            return True
        d = np.linalg.inv(self.T_key) @ T
        ang = math.degrees(np.linalg.norm(Rotation.from_matrix(d[:3, :3]).as_rotvec()))
        return np.linalg.norm(d[:3, 3]) >= self.kf_dist or ang >= self.kf_angle

    def _insert(self, m, T):
        self.map.insert(self.f_xyz[:m], self.f_attr[:m], T, self.reg_voxel, self.max_range, self.ret_mask)
        self.map_frames += 1
        self.T_key = T.copy()
        self.stats["keyframes"] += 1

    def _register(self, m, T_pred):
        ds = ds_insert(self.f_xyz[:m], self.f_attr[:m], 2.0 / self.reg_voxel, self.min_range, self.max_range,
                       self.ret_mask, self.ds_cap, self.fp64)
        self.stats["ds"] = len(ds)
        r = register(self.map, ds, T_pred, self.params())
        if len(ds) < self.min_corr:
            return T_pred, False
        self.stats.update(iters=r.iters, corr=r.corr, rms=r.rms, cond=r.cond, weak=r.weak, degen=r.degen)
        return r.T, r.ok


def loc_register(omap: OracleMap, scan, T0, voxel: float, iters: int = 10, min_count: int = 30,
                 damping: float = 1e-3, noise: float = 0.02, alpha: float = 0.05):
    """One hypothesis of the batched localisation: odom.py's accumulation at reach 1 (odom.py:275-384) with
    localize.py's step (localize.py:531-549: A = H + diag(damping diag H + 1e-9), solve, left twist; skipped
    below min_count matches; stop once the move max(|d_rho|, 5 |d_theta|) < 2e-4, here per hypothesis).
    Returns (T, score = robust weight sum of the last pass, matched count of the last pass, steps taken)."""
    T = np.array(T0, dtype=np.float64)
    noise2 = float(np.float32(noise * noise))
    tight2 = float(np.float32((TIGHT_FRACTION * voxel / math.sqrt(12.0)) ** 2))
    max_d2 = float(np.float32((1.5 * voxel) ** 2))
    score, count, steps = 0.0, 0, 0
    for it in range(iters):
        kappa = float(np.float32(max(KAPPA_MIN, KAPPA_START * (KAPPA_DECAY ** min(it, 15)))))
        v = icp_accumulate(omap, scan, T, voxel, 1, max_d2, noise2, tight2, alpha, kappa)
        score, count = float(v[ACC_WEIGHT]), int(v[ACC_COUNT])
        if v[ACC_COUNT] < min_count:
            continue
        H = sym6(v[ACC_H:ACC_H + 21])
        A = H + np.diag(damping * np.diag(H) + 1e-9)
        try:
            delta = np.linalg.solve(A, -v[ACC_B:ACC_B + 6])
        except np.linalg.LinAlgError:
            continue
        T = apply_twist(T, delta)
        steps = it + 1
        if max(float(np.linalg.norm(delta[:3])), float(np.linalg.norm(delta[3:])) * 5.0) < 2e-4:
            break
    return T, score, count, steps


# ---- the integration map (gpu.py:373-595) ------------------------------------------------------------------

MAP_COUNT_CAP = 4096  # gpu.py:55
POSE_SLOTS = 1 << 16  # gpu.py:56


@dataclass
class IntegrationMap:
    """The reference's integration-map arrays (gpu.py:1213-1219) for `cap` slots, cleared as clear_map
    (gpu.py:1304-1312) leaves them."""
    cap: int
    keys: np.ndarray = None
    acc: np.ndarray = None
    cnt: np.ndarray = None
    last: np.ndarray = None
    first: np.ndarray = None
    miss: np.ndarray = None
    stats: np.ndarray = None

    def __post_init__(self):
        self.keys = np.full(self.cap, -1, np.int64)
        self.acc = np.zeros((self.cap, 4), np.float32)
        self.cnt = np.zeros(self.cap, np.int32)
        self.last = np.zeros(self.cap, np.float32)
        self.first = np.full(self.cap, 1.0e30, np.float32)
        self.miss = np.zeros(self.cap, np.int32)
        self.stats = np.zeros(2, np.int32)


def map_insert(m: IntegrationMap, xyz, attr, t, frame, poses, voxel: float, min_range=0.0, max_range=1e30,
               min_refl=0, noise_mask=0b1111, ret_mask=0b1111):
    """k_map_insert (gpu.py:373-413) as Warp's CPU backend runs it: points in index order. poses: (65536, 4, 4)
    float32 by frame id & 65535."""
    xyz, t = _f32(xyz).reshape(-1, 3), _f32(t)
    attr = np.ascontiguousarray(attr, np.uint32)
    frame = np.ascontiguousarray(frame, np.int32)
    P = _f32(poses).reshape(POSE_SLOTS, 16)
    inv = np.float32(1.0) / np.float32(max(voxel, 1e-4))
    lib().lwo_map_insert(_p(xyz, _f32p), _p(attr, _u32p), _p(t, _f32p), _p(frame, _i32p), len(t), _p(P, _f32p),
                         min_range, max_range, int(min_refl), int(noise_mask), int(ret_mask), float(inv), m.cap - 1,
                         _p(m.keys, _i64p), _p(m.acc, _f32p), _p(m.cnt, _i32p), _p(m.last, _f32p),
                         _p(m.first, _f32p), _p(m.stats, _i32p))


def local_plane(m: IntegrationMap, ix: int, iy: int, iz: int, origin) -> np.ndarray:
    out = np.zeros(4, np.float32)
    o = _f32(origin)
    lib().lwo_local_plane(ix, iy, iz, _p(o, _f32p), _p(m.keys, _i64p), _p(m.acc, _f32p), _p(m.cnt, _i32p),
                          m.cap - 1, _p(out, _f32p))
    return out


def carve(m: IntegrationMap, xyz, attr, frame, poses, voxel: float, every: int = 2, max_range: float = 40.0,
          max_steps: int = 4096):
    """k_carve (gpu.py:470-595) in index order: misses added to m.miss."""
    xyz = _f32(xyz).reshape(-1, 3)
    attr = np.ascontiguousarray(attr, np.uint32)
    frame = np.ascontiguousarray(frame, np.int32)
    P = _f32(poses).reshape(POSE_SLOTS, 16)
    inv = np.float32(1.0) / np.float32(max(voxel, 1e-4))
    lib().lwo_carve(_p(xyz, _f32p), _p(attr, _u32p), _p(frame, _i32p), _p(P, _f32p), len(attr), max(1, int(every)),
                    float(inv), m.cap - 1, float(max_range), int(max_steps), _p(m.keys, _i64p), _p(m.acc, _f32p),
                    _p(m.cnt, _i32p), _p(m.miss, _i32p))


def identity_poses() -> np.ndarray:
    return np.tile(np.eye(4, dtype=np.float32), (POSE_SLOTS, 1, 1))


# ---- the prior-map grid and its coarse scoring (mapgrid.py, localize.py:55-123) ----------------------

GRID_JFA_STEPS = (8, 4, 2, 1, 1)  # mapgrid.py:25


def _geo(origin, cell: float):
    """The grid geometry as the restatement takes it: origin x, y, z, cell and 1 / cell in float32."""
    return np.array([*np.asarray(origin, np.float32), cell, 1.0 / cell], np.float32)


def grid_build(xyz, origin, cell: float, dims, steps=GRID_JFA_STEPS):
    """mapgrid.py:275-290 on the host: (nn int32 per cell, df uint8 per cell) of points already sorted by cell."""
    xyz = _f32(xyz).reshape(-1, 3)
    dims = np.ascontiguousarray(dims, np.int32)
    ncell = int(np.prod(dims.astype(np.int64)))
    nn, tmp, df = np.empty(ncell, np.int32), np.empty(ncell, np.int32), np.empty(ncell, np.uint8)
    st = np.ascontiguousarray(steps, np.int32)
    lib().lwo_grid_build(_p(xyz, _f32p), len(xyz), _p(_geo(origin, cell), _f32p), _p(dims, _i32p), _p(st, _i32p),
                         len(st), _p(nn, _i32p), _p(tmp, _i32p), _p(df, _u8p))
    return nn, df


def grid_distances(pts, T, origin, cell: float, dims, start, nn, xyz):
    """mapgrid.py:170-193 (k_changes_T): each point's distance to the nearest map point, 1e3 beyond the band."""
    pts = _f32(pts).reshape(-1, 3)
    out = np.empty(len(pts), np.float32)
    lib().lwo_grid_distances(_p(pts, _f32p), len(pts), _p(_f32(T).reshape(16), _f32p),
                             _p(_geo(origin, cell), _f32p), _p(np.ascontiguousarray(dims, np.int32), _i32p),
                             _p(np.ascontiguousarray(start, np.int32), _i32p),
                             _p(np.ascontiguousarray(nn, np.int32), _i32p), _p(_f32(xyz).reshape(-1, 3), _f32p),
                             _p(out, _f32p))
    return out


def grid_nearest(w, origin, cell: float, dims, start, nn, xyz) -> int:
    """mapgrid.py:65-100 nearest_point for one query."""
    return int(lib().lwo_grid_nearest(_p(_f32(w).reshape(3), _f32p), _p(_geo(origin, cell), _f32p),
                                      _p(np.ascontiguousarray(dims, np.int32), _i32p),
                                      _p(np.ascontiguousarray(start, np.int32), _i32p),
                                      _p(np.ascontiguousarray(nn, np.int32), _i32p),
                                      _p(_f32(xyz).reshape(-1, 3), _f32p)))


def score_grid(pts, pos, n_yaw: int, yaw0: float, dyaw: float, origin, cell: float, dims, df, tau_cm: float,
               free_cm: float):
    """localize.py:55-92 (k_score_grid): (len(pos) * n_yaw,) soft counts, position-major."""
    pts, pos = _f32(pts).reshape(-1, 3), _f32(pos).reshape(-1, 3)
    out = np.empty(len(pos) * n_yaw, np.float32)
    lib().lwo_score_grid(_p(pts, _f32p), len(pts), _p(pos, _f32p), len(pos), n_yaw, float(yaw0), float(dyaw),
                         _p(_geo(origin, cell), _f32p), _p(np.ascontiguousarray(dims, np.int32), _i32p),
                         _p(np.ascontiguousarray(df, np.uint8), _u8p), float(tau_cm), float(free_cm),
                         _p(out, _f32p))
    return out


def score_list(pts, hyp, origin, cell: float, dims, df, tau_cm: float):
    """localize.py:95-123 (k_score_list): the soft count of each (x, y, z, yaw) hypothesis."""
    pts, hyp = _f32(pts).reshape(-1, 3), _f32(hyp).reshape(-1, 4)
    out = np.empty(len(hyp), np.float32)
    lib().lwo_score_list(_p(pts, _f32p), len(pts), _p(hyp, _f32p), len(hyp), _p(_geo(origin, cell), _f32p),
                         _p(np.ascontiguousarray(dims, np.int32), _i32p), _p(np.ascontiguousarray(df, np.uint8), _u8p),
                         float(tau_cm), _p(out, _f32p))
    return out


EVAL_RADII = (0.03, 0.05, 0.10)  # localize.py:46


def plane_icp(pts, Ts, grid: dict, nn, max_d2: float, c2: float, alpha: float = 0.05):
    """localize.py:126-180 (k_icp_batch, k_reduce_batch): (k, 64) accumulator rows for the poses Ts (k, 4, 4).
    grid: a prepared prior map (origin, cell, dims, start, xyz, normal, planarity)."""
    pts = _f32(pts).reshape(-1, 3)
    T = _f32(Ts).reshape(-1, 16)
    out = np.zeros((len(T), 64))
    lib().lwo_plane_icp(_p(pts, _f32p), len(pts), _p(T, _f32p), len(T), _p(_geo(grid["origin"], grid["cell"]), _f32p),
                        _p(np.ascontiguousarray(grid["dims"], np.int32), _i32p),
                        _p(np.ascontiguousarray(grid["start"], np.int32), _i32p),
                        _p(np.ascontiguousarray(nn, np.int32), _i32p), _p(_f32(grid["xyz"]), _f32p),
                        _p(_f32(grid["normal"]), _f32p), _p(_f32(grid["planarity"]), _f32p), float(max_d2), float(c2),
                        float(alpha), _p(out, _f64p))
    return out


def plane_eval(pts, Ts, grid: dict, nn, radii=EVAL_RADII):
    """localize.py:183-206 (k_eval_batch): (k, 3) counts of points within each radius of the map."""
    pts = _f32(pts).reshape(-1, 3)
    T = _f32(Ts).reshape(-1, 16)
    counts = np.zeros((len(T), 3), np.int32)
    lib().lwo_plane_eval(_p(pts, _f32p), len(pts), _p(T, _f32p), len(T), _p(_geo(grid["origin"], grid["cell"]), _f32p),
                         _p(np.ascontiguousarray(grid["dims"], np.int32), _i32p),
                         _p(np.ascontiguousarray(grid["start"], np.int32), _i32p),
                         _p(np.ascontiguousarray(nn, np.int32), _i32p), _p(_f32(grid["xyz"]), _f32p),
                         _p(np.array(radii, np.float32), _f32p), _p(counts, _i32p))
    return counts


# ---- bulk voxelisation of a prior map (prior_map.py:55-123, 169-203) ----------------------------------------
VOX_PROBES = 128  # prior_map.py:36


class VoxTable:
    """k_vox_insert's table (prior_map.py:55-91) filled in index order by the C restatement: keys (-1 empty),
    offset sums from the voxel centres, colour + intensity sums, counts, stats [voxels, dropped points]."""

    def __init__(self, slots: int):
        assert slots & (slots - 1) == 0
        self.cap = slots
        self.keys = np.full(slots, -1, np.int64)
        self.rsum = np.zeros((slots, 3), np.float32)
        self.csum = np.zeros((slots, 4), np.float32)
        self.cnt = np.zeros(slots, np.int32)
        self.stats = np.zeros(2, np.int32)

    def insert(self, xyz, voxel: float, intensity=None, rgb=None):
        xyz = _f32(xyz).reshape(-1, 3)
        it = None if intensity is None else _f32(intensity)
        c = None if rgb is None else _f32(rgb).reshape(-1, 3)
        inv = np.float32(1.0 / voxel)  # the launch argument 1.0 / voxel rounded to float32 (prior_map.py:187)
        lib().lwo_vox_insert(_p(xyz, _f32p), None if it is None else _p(it, _f32p), None if c is None else _p(c, _f32p),
                             len(xyz), float(inv), float(voxel), self.cap - 1, _p(self.keys, _i64p),
                             _p(self.rsum, _f32p), _p(self.csum, _f32p), _p(self.cnt, _i32p), _p(self.stats, _i32p))

    def emit(self, voxel: float):
        """k_vox_flag / array_scan / k_vox_emit (prior_map.py:192-201): (xyz, colour + intensity, count) of the
        claimed slots in slot order."""
        v = int(self.stats[0])
        xyz, col, cnt = np.empty((v, 3), np.float32), np.empty((v, 4), np.float32), np.empty(v, np.int32)
        got = lib().lwo_vox_emit(_p(self.keys, _i64p), _p(self.rsum, _f32p), _p(self.csum, _f32p), _p(self.cnt, _i32p),
                                 self.cap, float(voxel), _p(xyz, _f32p), _p(col, _f32p), _p(cnt, _i32p))
        assert got == v
        return xyz, col, cnt

"""The oracle's accumulation and registration against known answers: a point on a planar voxel's mean plane has
no normal residual; a sparse voxel takes the weak isotropic branch; the symmetric 3x3 eigen-decomposition matches
numpy; a translated box room registers back to the offset; the identity stays the identity; a plane-only scene
holds three directions at the prediction."""

import math

import numpy as np


def _plane_map(orc, voxel=0.25, n=4000, seed=0):
    rng = np.random.default_rng(seed)
    xy = rng.uniform(0.01, 0.24, (n, 2))
    pts = np.column_stack([xy, np.full(n, 0.1)]).astype(np.float32)
    m = orc.OracleMap(1 << 10)
    m.insert(pts, np.zeros(n, np.uint32), np.eye(4), voxel, max_range=1e30)
    return m, pts


def test_eig3_matches_numpy(orc):
    rng = np.random.default_rng(7)
    for _ in range(300):
        B = rng.normal(size=(3, 3)) * 10 ** rng.uniform(-4, 1)
        A = B @ B.T
        ev, Q = orc.eig3(A)
        ref = np.linalg.eigvalsh(A)
        assert np.allclose(np.sort(ev), ref, rtol=1e-10, atol=1e-14 * np.abs(ref).max())
        assert np.allclose(A @ Q, Q * ev, atol=1e-10 * np.abs(ref).max())


def test_plane_point_has_no_normal_residual(orc):
    m, pts = _plane_map(orc)
    voxel = 0.25
    s = m.find_slot(orc.voxel_key_ijk(0, 0, 0))
    mu = 0.125 + m.rsum[s] / m.rcnt[s]
    shape, M6, Mg6, q = orc.voxel_precision(m.rsum[s], m.router[s], m.rcnt[s], voxel, np.float32(4e-4),
                                            np.float32((0.6 * voxel / math.sqrt(12)) ** 2), 0.05)
    assert shape == 1 and abs(abs(q[2]) - 1) < 1e-6  # the tight direction is the plane normal
    # Mg is the normal direction only: (1/noise^2) n n^T
    assert abs(Mg6[5] - 1.0 / (4e-4 + max(np.var(pts[:, 2]), 0))) < 1e-2 * Mg6[5]
    # a query on the mean plane: residual along the normal is zero, so b has no z component from it
    qpt = np.array([[mu[0] + 0.03, mu[1] - 0.02, mu[2]]], np.float32)
    v = orc.icp_accumulate(m, qpt, np.eye(4), voxel, 1, (1.5 * voxel) ** 2, float(np.float32(4e-4)),
                           float(np.float32((0.6 * voxel / math.sqrt(12)) ** 2)), 0.05, 6.0)
    assert v[orc.ACC_COUNT] == 1
    assert abs(v[orc.ACC_RMS]) < 1e-12  # rn = r . n = 0
    # float32 rounding of the covariance tilts the fitted normal by ~1e-8 rad: a residual of that order only
    assert abs(v[orc.ACC_B + 2]) < 1e-4 * abs(v[orc.ACC_B + 0])


def test_sparse_voxel_is_weak_isotropic(orc):
    pts = np.array([[0.1, 0.1, 0.1], [0.12, 0.1, 0.1], [0.1, 0.13, 0.1]], np.float32)
    m = orc.OracleMap(64)
    m.insert(pts, np.zeros(3, np.uint32), np.eye(4), 0.25, max_range=1e30)
    s = m.find_slot(orc.voxel_key_ijk(0, 0, 0))
    shape, M6, Mg6, q = orc.voxel_precision(m.rsum[s], m.router[s], m.rcnt[s], 0.25, np.float32(4e-4),
                                            np.float32(0.0019), 0.05)
    wa = np.float32(0.05) / (np.float32(0.25) * np.float32(0.25) * np.float32(0.25) + np.float32(4e-4))
    assert shape == 0 and np.allclose(M6, [wa, 0, 0, wa, 0, wa]) and not Mg6.any()


def test_translated_room_registers_back(orc, lw):
    w = lw.synth.single_scan(seed=15, map_points=300_000)
    voxel = 0.25
    m = orc.OracleMap(1 << 20)
    m.insert(w.map_xyz, np.zeros(len(w.map_xyz), np.uint32), np.eye(4), voxel, max_range=1e30)
    ds = orc.ds_insert(w.scan_xyz, w.scan_attr, 2.0 / voxel, 0.3, 60.0, 15, 1 << 15)
    P = orc.Params(reg_voxel=voxel, max_iter=20, prior_info=0.0, degen_thr=0.0)
    # start from the truth shifted by a pure translation: registration removes the offset
    T0 = w.T_true.copy()
    T0[:3, 3] += [0.08, -0.06, 0.04]
    r = orc.register(m, ds, T0, P)
    dt, da = orc.pose_error(r.T, w.T_true)
    assert r.ok and dt < 0.02 and da < 0.3, (dt, da)
    # the truth is (nearly) a fixed point
    r2 = orc.register(m, ds, w.T_true, P)
    dt2, da2 = orc.pose_error(r2.T, w.T_true)
    assert dt2 < 0.02 and da2 < 0.3


def test_plane_only_scene_holds_three_directions(orc):
    # a floor seen from 1.5 m: sliding along it (x, y) and yawing about its normal are unconstrained
    rng = np.random.default_rng(9)
    g = rng.uniform(-6, 6, (200000, 2))
    floor = np.column_stack([g, np.zeros(len(g))]).astype(np.float32)
    voxel = 0.25
    m = orc.OracleMap(1 << 18)
    m.insert(floor, np.zeros(len(floor), np.uint32), np.eye(4), voxel, max_range=1e30)
    T = np.eye(4)
    T[:3, 3] = [0.0, 0.0, 1.5]
    T[:3, :3] = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], float)  # looking straight down (sensor x = -z)
    d = rng.normal(size=(4000, 3)) * [0, 0.3, 0.3] + [1, 0, 0]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    scan = (d * (1.5 / d[:, 0:1])).astype(np.float32)  # hits z = 0
    ds = orc.ds_insert(scan, np.zeros(len(scan), np.uint32), 2.0 / voxel, 0.3, 60.0, 15, 1 << 13)
    P = orc.Params(reg_voxel=voxel, max_iter=5)
    r = orc.register(m, ds, T, P)
    assert r.degen == 3, r.degen

"""The C-ABI library without a GPU: it loads and exports every function include/livox_warp_amd.h declares, and
the host builds of the device's float64 step and per-voxel precision agree with the reference's numpy/scipy
routines (restated in the oracle) on accumulators from real registrations, including held directions."""

import ctypes
import math
import os
import re

import numpy as np

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "livox_warp_amd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|double)\s+\*?(lw_\w+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol(lw):
    names = declared()
    assert len(names) >= 25
    L = lw._abi.lib()
    for n in names:
        assert hasattr(L, n), n
        assert n in lw._abi.SIGNATURES, n
    assert set(lw._abi.SIGNATURES) == set(names)
    assert L.lw_version() >= 100


def _host_step(lw, v, T, Tp, P):
    A = lw._abi
    L = A.lib()
    delta = np.zeros(6)
    rms, nd = ctypes.c_double(), ctypes.c_int32()
    rc = L.lw_host_gn_step(A.ptr(np.ascontiguousarray(v), A._f64p), A.ptr(np.ascontiguousarray(T), A._f64p),
                           A.ptr(np.ascontiguousarray(Tp), A._f64p), P.damping, P.prior_info, P.degen_thr,
                           P.min_corr, A.ptr(delta, A._f64p), ctypes.byref(rms), ctypes.byref(nd))
    return rc, delta, rms.value, nd.value


def test_gn_step_matches_numpy(lw, orc):
    w = lw.synth.single_scan(seed=14, map_points=300_000)
    m = orc.OracleMap(1 << 20)
    m.insert(w.map_xyz, np.zeros(len(w.map_xyz), np.uint32), np.eye(4), 0.25, max_range=1e30)
    ds = orc.ds_insert(w.scan_xyz, w.scan_attr, 8.0, 0.3, 60.0, 15, 1 << 15)
    checked = 0
    for prior, degen in ((2000.0, 3000.0), (0.0, 0.0), (2000.0, 0.0)):
        P = orc.Params(reg_voxel=0.25, max_iter=12, prior_info=prior, degen_thr=degen)
        r = orc.register(m, ds, w.T_init, P, keep_accs=True)
        for T, _, _, _, v in r.accs:
            d_ref, nd_ref, rms_ref, _ = orc.gn_step(v, T, w.T_init, P)
            rc, d, rms, nd = _host_step(lw, v, T, w.T_init, P)
            assert rc == 0 and nd == nd_ref
            assert np.abs(d - d_ref).max() <= 1e-9 * np.abs(d_ref).max() + 1e-15
            assert abs(rms - rms_ref) <= 1e-12 * max(rms_ref, 1.0)
            checked += 1
    assert checked >= 10


def test_gn_step_holds_degenerate_directions(lw, orc):
    # a floor seen from above: three directions held at the prediction, the rest solved
    rng = np.random.default_rng(9)
    g = rng.uniform(-6, 6, (150000, 2))
    floor = np.column_stack([g, np.zeros(len(g))]).astype(np.float32)
    m = orc.OracleMap(1 << 18)
    m.insert(floor, np.zeros(len(floor), np.uint32), np.eye(4), 0.25, max_range=1e30)
    T = np.eye(4)
    T[:3, 3] = [0.0, 0.0, 1.5]
    T[:3, :3] = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], float)
    d = rng.normal(size=(3000, 3)) * [0, 0.3, 0.3] + [1, 0, 0]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    scan = (d * (1.5 / d[:, 0:1])).astype(np.float32)
    ds = orc.ds_insert(scan, np.zeros(len(scan), np.uint32), 8.0, 0.3, 60.0, 15, 1 << 13)
    Tp = T.copy()
    Tp[:3, 3] += [0.02, -0.01, 0.03]  # the prediction the held directions return to
    P = orc.Params(reg_voxel=0.25, max_iter=4)
    r = orc.register(m, ds, Tp, P, keep_accs=True)
    assert r.degen == 3
    for Ti, _, _, _, v in r.accs:
        d_ref, nd_ref, _, _ = orc.gn_step(v, Ti, Tp, P)
        rc, dd, _, nd = _host_step(lw, v, Ti, Tp, P)
        assert rc == 0 and nd == nd_ref == 3
        assert np.abs(dd - d_ref).max() <= 1e-8 * np.abs(d_ref).max() + 1e-14


def test_few_correspondences_stop(lw, orc):
    v = np.zeros(64)
    v[orc.ACC_COUNT] = 10
    rc, _, _, _ = _host_step(lw, v, np.eye(4), np.eye(4), orc.Params())
    assert rc == 1


def test_voxel_precision_matches_oracle(lw, orc):
    A = lw._abi
    L = A.lib()
    w = lw.synth.single_scan(seed=16, map_points=200_000)
    voxel = 0.25
    m = orc.OracleMap(1 << 19)
    m.insert(w.map_xyz, np.zeros(len(w.map_xyz), np.uint32), np.eye(4), voxel, max_range=1e30)
    noise2 = np.float32(0.02 * 0.02)
    tight2 = np.float32((0.6 * voxel / math.sqrt(12.0)) ** 2)
    shapes = 0
    for s in np.nonzero(m.keys != -1)[0]:
        c = int(m.rcnt[s])
        sh, M6, G6, q = orc.voxel_precision(m.rsum[s], m.router[s], c, voxel, noise2, tight2, 0.05)
        r6 = np.ascontiguousarray(m.router[s][[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]], np.float32)
        Mb, Gb, qb = np.zeros(6, np.float32), np.zeros(6, np.float32), np.zeros(3, np.float32)
        shb = L.lw_host_voxel_precision(A.ptr(np.ascontiguousarray(m.rsum[s]), A._f32p), A.ptr(r6, A._f32p), c,
                                        voxel, noise2, tight2, 0.05, A.ptr(Mb, A._f32p), A.ptr(Gb, A._f32p),
                                        A.ptr(qb, A._f32p))
        assert sh == shb
        assert np.abs(M6 - Mb).max() <= 1e-5 * np.abs(M6).max()
        assert np.abs(G6 - Gb).max() <= 1e-5 * max(np.abs(G6).max(), 1e-30)
        if sh:
            shapes += 1
            assert abs(abs(float(np.dot(q, qb))) - 1.0) < 1e-5
    assert shapes > 100


def test_twist_maps_match_scipy(lw, orc):
    from scipy.spatial.transform import Rotation

    A = lw._abi
    L = A.lib()
    rng = np.random.default_rng(0)
    for _ in range(300):
        T = np.eye(4)
        T[:3, :3] = Rotation.from_rotvec(rng.normal(size=3)).as_matrix()
        T[:3, 3] = rng.normal(size=3) * 5
        d = np.concatenate([rng.normal(size=3) * 0.1, rng.normal(size=3) * 10 ** rng.uniform(-9, 0.4)])
        T2 = T.copy()
        L.lw_host_apply_twist(A.ptr(T2, A._f64p), A.ptr(d, A._f64p))
        assert np.abs(T2 - orc.apply_twist(T, d)).max() < 1e-12
        xi = np.zeros(6)
        L.lw_host_twist_between(A.ptr(T, A._f64p), A.ptr(T2, A._f64p), A.ptr(xi, A._f64p))
        assert np.abs(xi - orc.twist_between(T, T2)).max() < 1e-9


def test_cond_matches_numpy(lw):
    A = lw._abi
    L = A.lib()
    rng = np.random.default_rng(1)
    for _ in range(100):
        B = rng.normal(size=(6, 6)) * 10 ** rng.uniform(-2, 3, 6)
        H = np.ascontiguousarray(B @ B.T)
        dg = np.sqrt(np.maximum(np.diag(H), 1e-12))
        ev = np.linalg.eigvalsh(H / np.outer(dg, dg))
        ref = max(ev[0], 0.0) / max(ev[-1], 1e-12)
        assert abs(L.lw_host_cond(A.ptr(H, A._f64p)) - ref) <= 1e-9 * max(ref, 1e-6)


def test_grid_and_replay_entry_points_reject_bad_arguments(lw):
    """The prior-grid, scoring and LVXR entry points refuse malformed calls before touching the device."""
    L, A = lw._abi.lib(), lw._abi
    E = A.LW_EINVAL
    geo = np.array([0, 0, 0, 0.05, 20.0], np.float32)
    dims = np.array([10, 10, 10], np.int32)
    g, d = A.ptr(geo, A._f32p), A.ptr(dims, A._i32p)
    steps = np.array([8, 4, 2, 1, 1], np.int32)
    st = A.ptr(steps, A._i32p)
    fake = 4096  # never dereferenced: each call below fails its checks first
    wb = L.lw_grid_work_bytes(d)
    assert wb == 2 * 16 * 1000
    assert L.lw_grid_work_bytes(A.ptr(np.array([0, 5, 5], np.int32), A._i32p)) == -1
    assert L.lw_grid_build(fake, 1 << 23, g, d, st, 5, fake, fake, wb, fake, None) == E  # too many points
    assert L.lw_grid_build(fake, 10, g, d, st, 5, fake, fake, wb - 1, fake, None) == E  # work too small
    assert L.lw_grid_build(fake, 10, g, d, st, 5, fake, fake + 8, wb, fake, None) == E  # work misaligned
    assert L.lw_grid_build(fake, 10, g, d, A.ptr(np.array([8, 0], np.int32), A._i32p), 2, fake, fake, wb, fake,
                           None) == E  # a zero step
    for bad in ([0, 10, 10], [1 << 24, 1, 1], [10, 70000, 1], [46341, 46341, 2]):
        assert L.lw_grid_build(fake, 10, g, A.ptr(np.array(bad, np.int32), A._i32p), st, 5, fake, fake, 1 << 40,
                               fake, None) == E, bad
    assert L.lw_grid_build(fake, 10, A.ptr(np.array([0, 0, 0, 0.0, 0.0], np.float32), A._f32p), d, st, 5, fake,
                           fake, wb, fake, None) == E  # zero cell
    assert L.lw_score_grid(fake, 10, fake, 5, 0, 0.0, 0.1, g, d, fake, 25.0, 10.0, None, fake, None) == E  # no headings
    assert L.lw_score_grid(fake, 10, fake, 5, 180, 0.0, 0.1, g, d, fake, 0.0, 10.0, None, fake, None) == E  # tau 0
    assert L.lw_score_grid(fake, 10, fake, 0, 180, 0.0, 0.1, g, d, fake, 25.0, 10.0, None, fake, None) == 0  # nothing
    assert L.lw_score_list(fake, 10, fake + 4, 5, g, d, fake, 12.0, fake, None) == E  # hypotheses misaligned
    assert L.lw_score_list(fake, 10, fake, 0, g, d, fake, 12.0, fake, None) == 0
    assert L.lw_grid_distances(fake, 0, A.ptr(np.eye(4, dtype=np.float32), A._f32p), g, d, fake, fake, fake, fake,
                               None) == 0
    args = [fake] * 6
    assert L.lw_lvxr_decode(*args, 3, 0, 10_000, 4, *[fake] * 6, None) == E  # four returns
    assert L.lw_lvxr_decode(*args, 3, 0, 0, 1, *[fake] * 6, None) == E  # no interval
    assert L.lw_lvxr_decode(*args, 0, 0, 10_000, 1, *[fake] * 6, None) == 0  # no packets

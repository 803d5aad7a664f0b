"""GPU registration (the whole Gauss-Newton loop on the device) against the oracle's loop with the reference's
numpy host step: the single-scan box-room registration agrees within 1 mm / 0.01 deg, with every schedule
option; lw_register (thinning from the frame buffers) equals the thinned path."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,prior,degen", [(14, 0.0, 0.0), (15, 2000.0, 3000.0), (17, 0.0, 3000.0)])
def test_single_scan_registration(lw, orc, seed, prior, degen, gpu_ok):
    w = lw.synth.single_scan(seed=seed)
    voxel = 0.25
    eng = lw.Engine(1 << 21, 1 << 18)
    eng.insert(w.map_xyz, None, np.eye(4), voxel, max_range=1e30)
    om = orc.OracleMap(1 << 21)
    om.insert(w.map_xyz, np.zeros(len(w.map_xyz), np.uint32), np.eye(4), voxel, max_range=1e30)
    ds = orc.ds_insert(w.scan_xyz, w.scan_attr, 2.0 / voxel, 0.3, 60.0, 15, 1 << 15)
    eng.thin_load(ds)
    P = orc.Params(reg_voxel=voxel, max_iter=20, prior_info=prior, degen_thr=degen)
    ro = orc.register(om, ds, w.T_init, P)
    p = eng.params(reg_voxel=voxel, max_iter=20, prior_info=prior, degen_thr=degen)
    rg = eng.register_thinned(len(ds), p, w.T_init)
    Tg = np.array(rg.T).reshape(4, 4)
    dt, da = orc.pose_error(Tg, ro.T)
    print(f"seed {seed}: gpu vs oracle {dt * 1e3:.4f} mm {da:.5f} deg; iters {rg.iters}/{ro.iters}, "
          f"corr {rg.corr}/{ro.corr}; vs truth {orc.pose_error(Tg, w.T_true)}")
    assert dt < 1e-3 and da < 0.01
    assert rg.ok == ro.ok and rg.corr == ro.corr and rg.degen == ro.degen
    assert abs(rg.iters - ro.iters) <= 1
    assert abs(rg.rms - ro.rms) <= 1e-6 + 1e-4 * ro.rms


def test_persistent_loop_matches_per_pass(lw, orc, gpu_ok):
    w = lw.synth.single_scan(seed=18)
    voxel = 0.25
    eng = lw.Engine(1 << 21, 1 << 18)
    eng.insert(w.map_xyz, None, np.eye(4), voxel, max_range=1e30)
    ds = orc.ds_insert(w.scan_xyz, w.scan_attr, 2.0 / voxel, 0.3, 60.0, 15, 1 << 15)
    eng.thin_load(ds)
    out = {}
    for on in (1, 2, 0):  # persistent (plain launch), persistent (cooperative launch), one launch per pass
        lw._abi.check(eng.lib.lw_set_persistent(eng.ctx, on), "lw_set_persistent")
        for mode in ("fp64", "kahan32"):
            p = eng.params(reg_voxel=voxel, max_iter=20, acc_mode=mode)
            out[(on, mode)] = eng.register_thinned(len(ds), p, w.T_init)
    for mode in ("fp64", "kahan32"):
        for on in (1, 2):
            a, b = out[(on, mode)], out[(0, mode)]
            assert a.iters == b.iters and a.corr == b.corr and a.degen == b.degen and a.ok == b.ok
            assert np.abs(np.array(a.T) - np.array(b.T)).max() < 1e-9


def test_register_from_frame_buffers(lw, orc, gpu_ok):
    w = lw.synth.single_scan(seed=16)
    voxel = 0.25
    eng = lw.Engine(1 << 21, 1 << 18)
    eng.insert(w.map_xyz, None, np.eye(4), voxel, max_range=1e30)
    m = eng.load_frame(w.scan_xyz, w.scan_attr, w.scan_t)
    eng.deskew(m, 0.05, 10.0, np.zeros(6))
    p = eng.params(reg_voxel=voxel, max_iter=20, prior_info=0.0, degen_thr=0.0)
    r1 = eng.register(m, p, w.T_init)
    ds = orc.ds_insert(w.scan_xyz, w.scan_attr, 2.0 / voxel, 0.3, 60.0, 15, 1 << 15)
    assert r1.n_ds == len(ds)
    eng.thin_load(ds)
    r2 = eng.register_thinned(len(ds), p, w.T_init)
    assert np.array_equal(np.array(r1.T), np.array(r2.T))


def test_device_frame_count(lw, gpu_ok):
    """m < 0 (the open frame's count left on the device) deskews and registers exactly as the count read back
    and passed in, the registration returns the count, and an insert of the frame with n < 0 adds the same
    voxels and counts as with the count passed in."""
    import torch

    w = lw.synth.single_scan(seed=16)
    voxel = 0.25
    dev = torch.device("cuda", 0)
    sx = torch.from_numpy(np.asarray(w.scan_xyz, np.float32).reshape(-1, 3)).to(dev)
    sa = torch.from_numpy(np.asarray(w.scan_attr).astype(np.uint32).view(np.int32)).to(dev)
    st = torch.from_numpy(np.asarray(w.scan_t, np.float32)).to(dev)
    xi = np.array([0.02, -0.01, 0.0, 0.001, 0.0, 0.004], np.float32)

    def open_frame(eng):
        eng.frame_reset()
        eng.frame_append(sx, sa, st, 10.0, -1, 0)  # every point in frame 0 ([0, 0.1) s)
        m = eng.frame_count()
        assert m == len(w.scan_t)
        return m

    # one map (its float moments depend on the atomics' order), both counts
    eng = lw.Engine(1 << 21, 1 << 15)
    eng.insert(w.map_xyz, None, np.eye(4), voxel, max_range=1e30)
    m = open_frame(eng)
    res = {}
    for mm in (m, -1):
        eng.deskew(mm, 0.05, 10.0, xi)
        r = eng.register(mm, eng.params(reg_voxel=voxel, max_iter=20), w.T_init)
        assert r.m == m
        res[mm] = (np.array(r.T).reshape(4, 4), r.iters, r.corr, r.n_ds, r.ok)
    assert np.array_equal(res[m][0], res[-1][0])
    assert res[m][1:] == res[-1][1:]
    T32 = np.ascontiguousarray(res[m][0], dtype=np.float32)
    eng.close()

    maps = {}
    for mm in (m, -1):
        eng = lw.Engine(1 << 21, 1 << 15)
        open_frame(eng)
        eng.deskew(mm, 0.05, 10.0, xi)
        lw._abi.check(eng.lib.lw_map_insert(eng.ctx, None, None, mm, lw._abi.ptr(T32, lw._abi._f32p), 1.0 / voxel,
                                            voxel, 60.0, 15), "lw_map_insert")
        maps[mm] = (eng.stats(),) + eng.read_map()
        eng.close()
    a, b = maps[m], maps[-1]
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[4], b[4])  # keys and counts
    np.testing.assert_allclose(a[2], b[2], rtol=1e-5, atol=1e-6)  # float atomics: order-dependent rounding


def _step_check(lw, v, T, Tp, P):
    A = lw._abi
    out = np.zeros(48)
    A.check(A.lib().lw_gn_step_check(A.ptr(np.ascontiguousarray(v, np.float64), A._f64p),
                                     A.ptr(np.ascontiguousarray(T, np.float64), A._f64p),
                                     A.ptr(np.ascontiguousarray(Tp, np.float64), A._f64p), P.damping, P.prior_info,
                                     P.degen_thr, P.min_corr, A.ptr(out, A._f64p)), "lw_gn_step_check")
    return out


def test_wave_step_equals_scalar_step(lw, orc, gpu_ok):
    """The wave form of the float64 step (the persistent registration's common case: the set-up and the
    factorisations an entry a lane, the substitutions a row a lane) and its pose update give the one-lane scalar
    form's step, rms and new pose bit for bit, on the accumulators of real registrations with the prior and the degeneracy test on and
    off; a floor seen from above, with held directions, is left to the scalar form."""
    w = lw.synth.single_scan(seed=14, map_points=300_000)
    m = orc.OracleMap(1 << 20)
    m.insert(w.map_xyz, np.zeros(len(w.map_xyz), np.uint32), np.eye(4), 0.25, max_range=1e30)
    ds = orc.ds_insert(w.scan_xyz, w.scan_attr, 8.0, 0.3, 60.0, 15, 1 << 15)
    handled = 0
    for prior, degen in ((2000.0, 3000.0), (0.0, 0.0), (2000.0, 0.0), (0.0, 3000.0)):
        P = orc.Params(reg_voxel=0.25, max_iter=12, prior_info=prior, degen_thr=degen)
        r = orc.register(m, ds, w.T_init, P, keep_accs=True)
        for T, _, _, _, v in r.accs:
            out = _step_check(lw, v, T, w.T_init, P)
            assert out[40] == 0 and out[32] in (0, 1)
            if out[32] == 0:  # a held direction: the scalar form's case
                continue
            assert np.array_equal(out[34:40], out[42:48]) and out[33] == out[41]
            assert np.array_equal(out[0:16], out[16:32])
            handled += 1
    assert handled >= 20
    # held directions: not the wave form's case
    rng = np.random.default_rng(9)
    g = rng.uniform(-6, 6, (150000, 2))
    floor = np.column_stack([g, np.zeros(len(g))]).astype(np.float32)
    mf = orc.OracleMap(1 << 18)
    mf.insert(floor, np.zeros(len(floor), np.uint32), np.eye(4), 0.25, max_range=1e30)
    T = np.eye(4)
    T[:3, 3] = [0.0, 0.0, 1.5]
    T[:3, :3] = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], float)
    d = rng.normal(size=(3000, 3)) * [0, 0.3, 0.3] + [1, 0, 0]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    scan = (d * (1.5 / d[:, 0:1])).astype(np.float32)
    dsf = orc.ds_insert(scan, np.zeros(len(scan), np.uint32), 8.0, 0.3, 60.0, 15, 1 << 13)
    P = orc.Params(reg_voxel=0.25, max_iter=4)
    r = orc.register(mf, dsf, T, P, keep_accs=True)
    for Ti, _, _, _, v in r.accs:
        assert _step_check(lw, v, Ti, T, P)[32] == 0

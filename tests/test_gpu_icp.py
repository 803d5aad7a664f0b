"""GPU accumulation (lw_icp_accum) against the oracle's k_icp + k_reduce on the same map moments (read back from
the device): the chosen voxel of every point bit-exact, the 64 reduced float64 entries within 1e-9 relative;
the compensated and naive float32 forms against the float64 one."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(lw, orc, seed=14):
    w = lw.synth.single_scan(seed=seed)
    voxel = 0.25
    eng = lw.Engine(1 << 21, 1 << 18)
    eng.insert(w.map_xyz, None, np.eye(4), voxel, max_range=1e30)
    om = orc.OracleMap(1 << 21)
    om.keys, om.rsum, om.router, om.rcnt = eng.read_map()
    ds = orc.ds_insert(w.scan_xyz, w.scan_attr, 2.0 / voxel, 0.3, 60.0, 15, 1 << 15)
    eng.thin_load(ds)
    return w, eng, om, ds, voxel


def test_accumulator_matches_oracle(lw, orc, gpu_ok):
    w, eng, om, ds, voxel = _setup(lw, orc)
    noise2 = float(np.float32(0.02 * 0.02))
    tight2 = float(np.float32((0.6 * voxel / math.sqrt(12.0)) ** 2))
    rng = np.random.default_rng(0)
    for j, (reach, kappa) in enumerate(((2, 6.0), (1, 2.94), (1, 2.0))):
        T = w.T_init.copy()
        T[:3, 3] += rng.normal(0, 0.03, 3)
        max_d2 = float(np.float32(((reach + 0.5) * voxel) ** 2))
        g, gs = eng.icp_accum(len(ds), T, voxel, reach, max_d2, noise2, tight2, 0.05, kappa, with_slots=True)
        o, os_ = orc.icp_accumulate(om, ds, T, voxel, reach, max_d2, noise2, tight2, 0.05, kappa, with_slots=True)
        assert np.array_equal(gs, os_), "chosen voxels differ"
        assert g[orc.ACC_COUNT] == o[orc.ACC_COUNT] > 100
        scale = np.abs(o[:53]).max()
        assert np.abs(g[:53] - o[:53]).max() <= 1e-9 * scale
        for k in (orc.ACC_COUNT, orc.ACC_WEIGHT, orc.ACC_RANGE, orc.ACC_RMS, orc.ACC_ERR):
            assert abs(g[k] - o[k]) <= 1e-9 * abs(o[k]) + 1e-12


def test_float32_forms(lw, orc, gpu_ok):
    w, eng, om, ds, voxel = _setup(lw, orc, seed=15)
    noise2 = float(np.float32(0.02 * 0.02))
    tight2 = float(np.float32((0.6 * voxel / math.sqrt(12.0)) ** 2))
    args = (len(ds), w.T_init, voxel, 1, float(np.float32((1.5 * voxel) ** 2)), noise2, tight2, 0.05, 3.0)
    ref = eng.icp_accum(*args, acc_mode="fp64")
    kah = eng.icp_accum(*args, acc_mode="kahan32")
    f32 = eng.icp_accum(*args, acc_mode="fp32")
    scale = np.abs(ref[:53]).max()
    ek = np.abs(kah[:53] - ref[:53]).max() / scale
    ef = np.abs(f32[:53] - ref[:53]).max() / scale
    assert ek < 1e-6 and ef < 1e-4, (ek, ef)
    assert ek <= ef

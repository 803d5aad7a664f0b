"""GPU map build (lw_map_insert) against the oracle's serial k_reg_insert: keys and occupied slots bit-exact,
per-voxel counts exact, moments within 1e-5 (of the voxel edge for the mean, of |cov| for the covariance)."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _compare_maps(orc, eng, om, voxel):
    keys, rsum, router, rcnt = eng.read_map()
    assert np.array_equal(np.sort(keys[keys != -1]), np.sort(om.keys[om.keys != -1]))
    assert np.array_equal(keys != -1, om.keys != -1), "occupied slots differ from the serial insert"
    st = eng.stats()
    assert st == (int(om.stats[0]), int(om.stats[1]))
    occ = np.nonzero(keys != -1)[0]
    # the same key may sit in a different slot of the same probe run; compare per key
    o_slot = {int(k): s for s, k in zip(np.nonzero(om.keys != -1)[0], om.keys[om.keys != -1])}
    worst_mu, worst_cov = 0.0, 0.0
    for s in occ:
        so = o_slot[int(keys[s])]
        assert rcnt[s] == om.rcnt[so]
        c = float(rcnt[s])
        mu_g, mu_o = rsum[s] / c, om.rsum[so] / c
        worst_mu = max(worst_mu, float(np.abs(mu_g - mu_o).max()) / voxel)
        cg = router[s] / c - np.outer(mu_g, mu_g)
        co = om.router[so] / c - np.outer(mu_o, mu_o)
        worst_cov = max(worst_cov, float(np.abs(cg - co).max() / max(np.abs(co).max(), 1e-12)))
    assert worst_mu <= 1e-5 and worst_cov <= 1e-4, (worst_mu, worst_cov)
    return keys


def test_single_scan_map(lw, orc, gpu_ok):
    w = lw.synth.single_scan(seed=14)
    voxel = 0.25
    eng = lw.Engine(1 << 21, 1 << 18)
    eng.insert(w.map_xyz, None, np.eye(4), voxel, max_range=1e30)
    om = orc.OracleMap(1 << 21)
    om.insert(w.map_xyz, np.zeros(len(w.map_xyz), np.uint32), np.eye(4), voxel, max_range=1e30)
    keys = _compare_maps(orc, eng, om, voxel)
    occ_keys = keys[keys != -1]
    slots = eng.find(occ_keys)
    assert np.array_equal(keys[slots], occ_keys)
    absent = orc.voxel_key_ijk_np(np.arange(1000) + 10_000, np.zeros(1000, int), np.zeros(1000, int))
    assert (eng.find(absent) == -1).all()


def test_posed_filtered_inserts(lw, orc, gpu_ok):
    seq = lw.synth.Sequence(seed=21, seconds=1.0)
    eng = lw.Engine(1 << 18, 1 << 16)
    om = orc.OracleMap(1 << 18)
    for k, (xyz, attr, t) in enumerate(seq.frames(4)):
        attr = attr.copy()
        attr[::50] |= np.uint32(1 << 8)  # noise-tagged
        attr[7::97] |= np.uint32(5 << 16)  # a return index outside the mask
        xyz[3::101] *= 100.0  # beyond max_range
        T = seq.pose(k * 0.1 + 0.05)
        eng.insert(xyz, attr, T, 0.2, max_range=60.0, ret_mask=0b1111)
        om.insert(xyz, attr, T, 0.2, max_range=60.0, ret_mask=0b1111)
    _compare_maps(orc, eng, om, 0.2)


def test_cap_keeps_first_points(lw, orc, gpu_ok):
    rng = np.random.default_rng(5)
    dense = rng.uniform(0.01, 0.24, (9000, 3)).astype(np.float32)  # one voxel at 0.25 m
    other = rng.uniform(-2, 2, (3000, 3)).astype(np.float32)
    pts = np.concatenate([other[:1000], dense[:5000], other[1000:], dense[5000:]])
    eng = lw.Engine(1 << 14, 1 << 14)
    om = orc.OracleMap(1 << 14)
    # two inserts: the second finds the voxel partly filled and takes only its quota, in index order
    for part in (pts[:4000], pts[4000:]):
        eng.insert(part, None, np.eye(4), 0.25, max_range=1e30)
        om.insert(part, np.zeros(len(part), np.uint32), np.eye(4), 0.25, max_range=1e30)
    keys = _compare_maps(orc, eng, om, 0.25)
    s = int(eng.find([orc.voxel_key_ijk(0, 0, 0)])[0])
    assert eng.read_map()[3][s] == 4096


def test_full_table_counts_overflow(lw, orc, gpu_ok):
    pts = np.stack([np.arange(80) * 1.0 + 0.5, np.zeros(80), np.zeros(80)], 1).astype(np.float32)
    eng = lw.Engine(64, 256)
    eng.insert(pts, None, np.eye(4), 1.0, max_range=1e30)
    assert eng.stats() == (64, 16)


def test_reset_empties_the_map(lw, orc, gpu_ok):
    eng = lw.Engine(1 << 12, 1 << 12)
    pts = np.random.default_rng(0).uniform(-1, 1, (2000, 3)).astype(np.float32)
    eng.insert(pts, None, np.eye(4), 0.3, max_range=1e30)
    assert eng.stats()[0] > 0
    eng.reset()
    assert eng.stats() == (0, 0)
    keys = eng.read_map()[0]
    assert (keys == -1).all()
    eng.insert(pts, None, np.eye(4), 0.3, max_range=1e30)
    om = orc.OracleMap(1 << 12)
    om.insert(pts, np.zeros(len(pts), np.uint32), np.eye(4), 0.3, max_range=1e30)
    _compare_maps(orc, eng, om, 0.3)


def test_one_chunk_larger_than_the_launch_grid_cap(lw, orc, gpu_ok):
    """17.3M points in one chunk (frame_cap 2^25): more than 65535 workgroups of 256 threads would cover, so the
    insert kernels' grid-stride loops must reach the tail. Counts, occupancy and drops equal the oracle's."""
    rng = np.random.default_rng(21)
    n = 65535 * 256 + 500_000
    xyz = rng.uniform(-40.0, 40.0, (n, 3)).astype(np.float32)
    with lw.Engine(1 << 22, 1 << 25) as eng:
        eng.insert(xyz, None, np.eye(4), 1.0, max_range=1e30)  # 512k voxels of about 34 points
        stats = eng.stats()
        keys, _, _, rcnt = eng.read_map()
    om = orc.OracleMap(1 << 22)
    om.insert(xyz, np.zeros(n, np.uint32), np.eye(4), 1.0, max_range=1e30)
    assert stats == (int(om.stats[0]), int(om.stats[1])) and stats[1] == 0
    assert np.array_equal(keys != -1, om.keys != -1)
    assert int(rcnt.sum()) == int(om.rcnt.sum()) == n  # every point counted, the tail included

"""The oracle's serial map build (k_reg_insert, odom.py:186-230) against invariants and a float64 group-by:
no duplicate keys, every key found by the probe, claimed count = distinct keys, occupied-slot set independent
of insertion order, the 4096-point cap keeping exactly the first points in index order."""

import numpy as np


def _keys_of(orc, pts, voxel):
    return orc.voxel_key_np(pts, 1.0 / voxel)


def test_insert_invariants(orc):
    rng = np.random.default_rng(2)
    pts = rng.uniform(-3, 3, (20000, 3)).astype(np.float32)
    voxel = 0.25
    m = orc.OracleMap(1 << 16)
    m.insert(pts, np.zeros(len(pts), np.uint32), np.eye(4), voxel, max_range=1e30)
    keys = m.keys[m.keys != -1]
    assert len(np.unique(keys)) == len(keys)
    distinct = np.unique(_keys_of(orc, pts, voxel))
    assert m.stats[0] == len(distinct) and m.stats[1] == 0
    assert set(keys.tolist()) == set(distinct.tolist())
    for k in distinct[::13]:
        s = m.find_slot(int(k))
        assert s >= 0 and m.keys[s] == k
    assert m.rcnt.sum() == len(pts)


def test_occupied_set_order_invariant(orc):
    rng = np.random.default_rng(3)
    pts = rng.normal(0, 1.5, (15000, 3)).astype(np.float32)
    a = orc.OracleMap(1 << 15)  # ~0.3 load: no key overflows the probe window
    b = orc.OracleMap(1 << 15)
    a.insert(pts, np.zeros(len(pts), np.uint32), np.eye(4), 0.2, max_range=1e30)
    perm = rng.permutation(len(pts))
    b.insert(pts[perm], np.zeros(len(pts), np.uint32), np.eye(4), 0.2, max_range=1e30)
    assert a.stats[1] == 0 and b.stats[1] == 0
    assert np.array_equal(a.keys != -1, b.keys != -1)  # linear probing without deletion
    assert np.array_equal(np.sort(a.keys), np.sort(b.keys))


def test_moments_match_float64_groupby(orc):
    rng = np.random.default_rng(4)
    pts = (rng.uniform(-2, 2, (30000, 3)) + np.array([0.05, 0.1, 0.0])).astype(np.float32)
    voxel = 0.4
    m = orc.OracleMap(1 << 15)
    m.insert(pts, np.zeros(len(pts), np.uint32), np.eye(4), voxel, max_range=1e30)
    ijk = np.floor(pts * np.float32(1.0 / voxel)).astype(np.int64)
    keys = orc.voxel_key_ijk_np(ijk[:, 0], ijk[:, 1], ijk[:, 2])
    r = pts.astype(np.float64) - (ijk + 0.5) * np.float64(np.float32(voxel))
    for k in np.unique(keys)[:200]:
        sel = keys == k
        s = m.find_slot(int(k))
        assert m.rcnt[s] == sel.sum()
        mean64 = r[sel].mean(0)
        mean32 = m.rsum[s] / m.rcnt[s]
        assert np.abs(mean32 - mean64).max() <= 1e-5 * voxel
        cov64 = np.cov(r[sel].T, bias=True)
        c32 = m.router[s] / m.rcnt[s] - np.outer(mean32, mean32)
        assert np.abs(c32 - cov64).max() <= 1e-5 * np.abs(cov64).max() + 1e-9


def test_reg_cap_first_points(orc):
    # 5000 points in one voxel: exactly the first 4096 (index order) contribute
    rng = np.random.default_rng(5)
    pts = (rng.uniform(0.01, 0.24, (5000, 3))).astype(np.float32)
    m = orc.OracleMap(1 << 10)
    m.insert(pts, np.zeros(len(pts), np.uint32), np.eye(4), 0.25, max_range=1e30)
    s = m.find_slot(orc.voxel_key_ijk(0, 0, 0))
    assert m.rcnt[s] == 4096
    r = pts[:4096].astype(np.float64) - 0.125
    assert np.allclose(m.rsum[s], r.sum(0), rtol=1e-4, atol=1e-3)


def test_filters_and_overflow(orc):
    pts = np.array([[1, 0, 0], [100, 0, 0], [2, 0, 0], [3, 0, 0]], np.float32)
    attr = np.array([0, 0, 1 << 8, 5 << 16], np.uint32)  # ok, too far, noise-tagged, return 5 not in mask
    m = orc.OracleMap(1 << 8)
    m.insert(pts, attr, np.eye(4), 0.5, max_range=60.0, ret_mask=0b1111)
    assert m.stats[0] == 1 and m.rcnt.sum() == 1
    # a table of 64 slots fills: the 65th distinct voxel finds no slot within the probe window
    pts = np.stack([np.arange(80) * 1.0 + 0.5, np.zeros(80), np.zeros(80)], 1).astype(np.float32)
    m = orc.OracleMap(64)
    m.insert(pts, np.zeros(80, np.uint32), np.eye(4), 1.0, max_range=1e30)
    assert m.stats[0] == 64 and m.stats[1] == 16

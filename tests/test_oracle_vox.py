"""The oracle's bulk voxelisation (prior_map.py:55-123 k_vox_insert / k_vox_flag / k_vox_emit, restated in
oracle/lw_vox_oracle.c) pinned by what the mathematics fixes: per-voxel counts and means against a numpy
group-by of the same points, the linear-probe invariant of the table (every key within 128 probes of its
home, no empty slot between home and key), the drop rule of a full table in index order, and a voxel whose
mean is its centre exactly. CPU only."""

import numpy as np

import oracle as O


def _ijk(xyz, voxel):
    inv = np.float32(1.0) / np.float32(voxel)
    return np.floor(xyz.astype(np.float32) * inv).astype(np.int64)


def test_means_and_counts_match_a_group_by():
    rng = np.random.default_rng(3)
    n = 60_000
    xyz = rng.uniform(-3.0, 3.0, (n, 3)).astype(np.float32)
    xyz[: n // 3] = np.round(xyz[: n // 3] * 5) / 5 + 0.013  # shared voxels
    inten = rng.uniform(0, 255, n).astype(np.float32)
    rgb = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    voxel = 0.1
    t = O.VoxTable(1 << 18)
    t.insert(xyz, voxel, inten, rgb)
    out_xyz, out_col, out_cnt = t.emit(voxel)
    occ = np.nonzero(t.keys != -1)[0]  # slot order = emit order
    ijk = _ijk(xyz, voxel)
    keys = O.voxel_key_ijk_np(ijk[:, 0], ijk[:, 1], ijk[:, 2])
    uk, inv, cnt = np.unique(keys, return_inverse=True, return_counts=True)
    assert t.stats.tolist() == [len(uk), 0] and len(occ) == len(uk)
    pos = np.searchsorted(uk, t.keys[occ])
    assert np.array_equal(uk[pos], t.keys[occ]) and np.array_equal(cnt[pos], out_cnt)
    msum = np.zeros((len(uk), 3))
    np.add.at(msum, inv, xyz.astype(np.float64))
    csum = np.zeros((len(uk), 4))
    np.add.at(csum, inv, np.c_[rgb, inten].astype(np.float64))
    assert np.abs(out_xyz - msum[pos] / cnt[pos, None]).max() < 2e-6 * 3.0
    assert np.abs(out_col - csum[pos] / cnt[pos, None]).max() < 1e-4


def test_table_keeps_the_linear_probe_invariant():
    rng = np.random.default_rng(4)
    xyz = rng.uniform(-20, 20, (40_000, 3)).astype(np.float32)
    t = O.VoxTable(1 << 16)  # about half full: long probe chains
    t.insert(xyz, 0.25)
    mask = t.cap - 1
    occ = np.nonzero(t.keys != -1)[0]
    assert t.stats[1] == 0 and len(occ) == t.stats[0] > 20_000
    home = O.slot_hash_np(t.keys[occ]) & mask
    dist = (occ - home) & mask
    assert dist.max() < O.VOX_PROBES and dist.max() > 4
    full = t.keys != -1
    for s, d in zip(occ[::97], dist[::97]):
        assert full[(s - np.arange(d + 1)) & mask].all()


def test_a_full_table_drops_later_points_in_index_order():
    voxel = 1.0
    centres = np.array([[i + 0.5, 0.5, 0.5] for i in range(12)], np.float32)
    xyz = np.repeat(centres, 3, axis=0)  # 12 voxels, 3 points each, in order
    t = O.VoxTable(8)
    t.insert(xyz, voxel)
    assert t.stats.tolist() == [8, 12]  # the first 8 voxels claim every slot; the last 4 voxels' 12 points drop
    kept = set(int(k) for k in t.keys)
    first8 = O.voxel_key_ijk_np(np.arange(8), np.zeros(8, np.int64), np.zeros(8, np.int64))
    assert kept == set(int(k) for k in first8)
    out_xyz, _, out_cnt = t.emit(voxel)
    assert np.all(out_cnt == 3) and np.allclose(np.sort(out_xyz[:, 0]), centres[:8, 0])


def test_a_symmetric_voxel_has_its_centre_as_mean():
    voxel = 0.5
    c = np.array([-1.25, 2.75, 0.25], np.float32)  # a voxel centre with negative coordinates
    d = np.array([[0.1, 0.0, 0.0], [-0.1, 0.0, 0.0], [0.0, 0.2, -0.05], [0.0, -0.2, 0.05]], np.float32)
    t = O.VoxTable(64)
    t.insert(c + d, voxel, np.full(4, 10.0, np.float32), np.full((4, 3), 0.5, np.float32))
    xyz, col, cnt = t.emit(voxel)
    assert cnt.tolist() == [4] and np.array_equal(xyz[0], c) and np.array_equal(col[0], [0.5, 0.5, 0.5, 10.0])

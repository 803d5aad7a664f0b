"""Bulk voxelisation of a prior map (lw_vox_*, prior_map.VoxTable / voxel_downsample) against the oracle's
serial k_vox_insert / k_vox_emit (prior_map.py:55-123, 169-203):

  * random points with shared voxels, colour and intensity: the same occupied slots (linear probing fills the
    same set in any arrival order), the same keys, every count exact, the sums within float32 summation
    order; the emitted means, colours and counts equal the oracle's voxel by voxel, in slot order;
  * scan-ordered input (runs of one voxel longer than a wave): the runs combined per wave keep the counts;
  * a full table: a voxel keeps all of its points or drops all of them, claims + drops account for every
    point;
  * at full size: the 60 x 40 m floor's 20.5M map points at 5 cm into 2^25 slots, against the oracle's serial
    build of the same table.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _compare(lw, orc, xyz, voxel, slots_log2, inten=None, rgb=None, chunk=None):
    t = lw.VoxTable(slots_log2)
    try:
        n = len(xyz)
        chunk = chunk or n
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            t.insert(_dev(xyz[a:b]), voxel, None if inten is None else _dev(inten[a:b]),
                     None if rgb is None else _dev(rgb[a:b]))
        stats = t.stats()
        keys, rsum, csum, cnt = t.read()
        exyz, ecol, ecnt = (x.cpu().numpy() for x in t.emit(voxel))
    finally:
        t.close()
    om = orc.VoxTable(1 << slots_log2)
    om.insert(xyz, voxel, inten, rgb)
    assert stats == tuple(int(x) for x in om.stats)
    occ = keys != -1
    assert np.array_equal(occ, om.keys != -1)
    g, o = np.argsort(keys[occ]), np.argsort(om.keys[occ])
    sg, so = np.nonzero(occ)[0][g], np.nonzero(occ)[0][o]
    assert np.array_equal(keys[sg], om.keys[so]) and np.array_equal(cnt[sg], om.cnt[so])
    c = om.cnt[so].astype(np.float64)[:, None]
    r_err = float(np.abs(rsum[sg] / c - om.rsum[so] / c).max())
    c_err = float(np.abs(csum[sg] / c - om.csum[so] / c).max()) if inten is not None else 0.0
    # the emit: the device's own slot order, compared by key with the oracle's
    oxyz, ocol, ocnt = om.emit(voxel)
    okeys = om.keys[om.keys != -1]
    pos = np.argsort(okeys)[np.searchsorted(np.sort(okeys), keys[occ])]
    assert np.array_equal(ecnt, ocnt[pos])
    e_err = float(np.abs(exyz - oxyz[pos]).max())
    assert np.abs(ecol - ocol[pos]).max() <= 1e-5 * max(1.0, float(np.abs(ocol).max()))
    return stats, r_err, c_err, e_err


def test_random_points_match_the_oracle(lw, orc, gpu_ok):
    rng = np.random.default_rng(3)
    n = 200_000
    xyz = rng.uniform(-8.0, 8.0, (n, 3)).astype(np.float32)
    xyz[: n // 2] = (np.round(xyz[: n // 2] * 4) / 4 + 0.013).astype(np.float32)  # shared voxels
    inten = rng.uniform(0, 255, n).astype(np.float32)
    rgb = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    stats, r_err, c_err, e_err = _compare(lw, orc, xyz, 0.1, 19, inten, rgb, chunk=70_001)
    print(f"vox: {n} points -> {stats[0]} voxels; offsets within {r_err:.2e} m, colour {c_err:.2e}, means "
          f"{e_err:.2e} m", flush=True)
    assert stats[1] == 0 and r_err < 1e-6 and c_err < 1e-3 and e_err < 1e-6


def test_scan_ordered_runs_keep_the_counts(lw, orc, gpu_ok):
    rng = np.random.default_rng(5)
    runs = rng.integers(1, 300, 2000)
    centres = rng.uniform(-30, 30, (len(runs), 3)).astype(np.float32)
    xyz = np.concatenate([np.repeat(c[None], k, 0) + rng.uniform(-0.01, 0.01, (k, 3)) for c, k in zip(centres, runs)])
    stats, r_err, _, e_err = _compare(lw, orc, xyz.astype(np.float32), 0.05, 16)
    assert stats[1] == 0 and r_err < 1e-6 and e_err < 1e-6


def test_a_full_table_drops_whole_voxels(lw, orc, gpu_ok):
    rng = np.random.default_rng(6)
    per = rng.integers(1, 9, 3000)
    centres = (rng.permutation(200_000)[:3000, None] * np.array([[0.1, 0.0, 0.0]]) + 0.05).astype(np.float32)
    xyz = np.repeat(centres, per, 0)
    key_of = np.repeat(np.arange(3000), per)
    rng.shuffle(order := np.arange(len(xyz)))
    xyz, key_of = xyz[order], key_of[order]
    with lw.VoxTable(10) as t:
        t.insert(_dev(xyz), 0.1)
        voxels, dropped = t.stats()
        exyz, _, ecnt = (x.cpu().numpy() for x in t.emit(0.1))
    assert voxels == 1024 and int(ecnt.sum()) + dropped == len(xyz)
    ix = np.floor(exyz[:, 0] / 0.1).astype(np.int64)
    want = dict(zip(np.floor(centres[:, 0] / 0.1).astype(np.int64), per))
    assert all(want[i] == c for i, c in zip(ix, ecnt))  # every claimed voxel kept all of its points


def test_floor_map_at_full_size(lw, orc, gpu_ok):
    W = lw.synth.loc_workload(seed=31, n_queries=1)
    xyz = W.map_xyz
    stats, r_err, _, e_err = _compare(lw, orc, xyz, 0.05, 25)
    print(f"vox full size: {len(xyz)} points -> {stats[0]} voxels of 5 cm in 2^25 slots; offsets within "
          f"{r_err:.2e} m, means {e_err:.2e} m", flush=True)
    assert stats[1] == 0 and stats[0] > 1_000_000 and r_err < 1e-6 and e_err < 2e-6


def test_voxel_downsample_matches_the_table(lw, orc, gpu_ok):
    rng = np.random.default_rng(8)
    xyz = rng.uniform(-2, 2, (50_000, 3)).astype(np.float32)
    inten = rng.uniform(0, 100, 50_000).astype(np.float32)
    rgb = rng.uniform(0, 1, (50_000, 3)).astype(np.float32)
    mx, mc, mi, cnt = lw.voxel_downsample(xyz, inten, rgb, 0.2, chunk=20_000, slots_log2=16)
    om = orc.VoxTable(1 << 16)
    om.insert(xyz, 0.2, inten, rgb)
    ox, oc, on = om.emit(0.2)
    a, b = np.lexsort(np.floor(mx / 0.2).T), np.lexsort(np.floor(ox / 0.2).T)  # by voxel index
    assert np.array_equal(cnt[a], on[b]) and np.abs(mx[a] - ox[b]).max() < 1e-6
    assert np.abs(mc[a] - oc[b, :3]).max() < 1e-5 and np.abs(mi[a] - oc[b, 3]).max() < 1e-3
    with pytest.raises(RuntimeError):
        lw.voxel_downsample(xyz, inten, rgb, 0.01, slots_log2=10)

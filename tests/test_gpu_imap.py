"""The integration map on the device (lw_imap_*: k_map_insert, k_carve and the pose table of gpu.py) against
the oracle: inserts give the same voxels, hit counts, first and last times and drop counts (the float32 sums
within rounding: atomics add in another order), carving on one map gives the oracle's misses voxel for
voxel, and on a scene whose object moves away the voxels it left are carved while the floor, seen at grazing
incidence, and the far wall are kept (the behaviour check_perception.py:84-115 asks of the reference)."""

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

V = 0.1


def _dev(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dt).cuda()


def _scene(rng, frames, ghost_until, n=6000):
    """A sensor at the origin (identity pose for every frame) looking down +x at a wall at x = 6, a floor at
    z = -1 and, in frames < ghost_until, a 0.4 m box centred at (3, 0, 0)."""
    xyz, attr, t, fid = [], [], [], []
    for k in range(frames):
        d = np.stack([np.ones(n), rng.uniform(-0.35, 0.35, n), rng.uniform(-0.45, 0.25, n)], 1)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        s = np.full(n, 6.0 / d[:, 0])  # the wall
        down = d[:, 2] < 0
        s[down] = np.minimum(s[down], -1.0 / d[down, 2])  # the floor
        if k < ghost_until:
            lo, hi = np.array([2.8, -0.2, -0.2]), np.array([3.2, 0.2, 0.2])
            with np.errstate(divide="ignore", invalid="ignore"):
                t0, t1 = (lo / d), (hi / d)
            tn, tf = np.minimum(t0, t1).max(1), np.maximum(t0, t1).min(1)
            hit = (tn <= tf) & (tn > 0)
            s[hit] = np.minimum(s[hit], tn[hit])
        p = d * s[:, None]
        xyz.append(p.astype(np.float32))
        attr.append(np.full(n, 40, np.uint32))
        t.append(np.full(n, 0.1 * k, np.float32) + rng.uniform(0, 0.1, n).astype(np.float32))
        fid.append(np.full(n, k, np.int32))
    return np.concatenate(xyz), np.concatenate(attr), np.concatenate(t), np.concatenate(fid)


def _records_to_oracle(recs, cap):
    m = O.IntegrationMap(cap)
    m.keys[:] = recs["key"]
    m.acc[:] = recs["acc"]
    m.cnt[:] = recs["cnt"]
    m.first[:] = recs["first"]
    m.last[:] = recs["last"]
    m.miss[:] = recs["miss"]
    return m


def test_insert_matches_oracle(lw, orc, gpu_ok):
    rng = np.random.default_rng(7)
    n = 40000
    xyz = rng.uniform(-6, 6, (n, 3)).astype(np.float32)
    xyz[: n // 2] = np.round(xyz[: n // 2] * 4) / 4  # shared voxels
    attr = (rng.integers(0, 256, n) | (rng.integers(0, 4, n) << 8) | (rng.integers(0, 4, n) << 16)).astype(np.uint32)
    t = rng.uniform(0, 20, n).astype(np.float32)
    fid = (rng.integers(0, 8, n) + 65530).astype(np.int32)  # the ids wrap onto pose slots 65530.. and 0..1
    pt = lw.PoseTable()
    P = O.identity_poses()
    for k in range(65530, 65538):
        a = rng.uniform(-np.pi, np.pi)
        T = np.eye(4, dtype=np.float32)
        T[:3, :3] = [[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]]
        T[:3, 3] = rng.uniform(-3, 3, 3)
        pt.set_pose(k, T)
        P[k & 0xFFFF] = T
    view = lw.imap.View(min_range=0.5, max_range=9.0, min_refl=5, noise_mask=0b1011, ret_mask=0b0111)
    cap = 1 << 16
    mp = lw.IntegrationMap(pt, "cuda:0", map_capacity=cap)
    mp.map_insert(_dev(xyz, torch.float32), _dev(attr.view(np.int32), torch.int32), _dev(t, torch.float32),
                  _dev(fid, torch.int32), n, view, 0.2)
    occ, dropped = mp.map_stats()
    recs = mp.records()
    ref = O.IntegrationMap(cap)
    O.map_insert(ref, xyz, attr, t, fid, P, 0.2, min_range=0.5, max_range=9.0, min_refl=5, noise_mask=0b1011,
                 ret_mask=0b0111)
    print(f"{n} points -> {occ} voxels, {dropped} dropped")
    assert (occ, dropped) == tuple(ref.stats) and occ > 5000
    g = {int(k): i for i, k in enumerate(recs["key"]) if k >= 0}
    r = {int(k): i for i, k in enumerate(ref.keys) if k >= 0}
    assert set(g) == set(r)
    gi = np.array([g[k] for k in r]), np.array(list(r.values()))
    assert np.array_equal(recs["cnt"][gi[0]], ref.cnt[gi[1]])
    assert np.array_equal(recs["first"][gi[0]], ref.first[gi[1]])
    assert np.array_equal(recs["last"][gi[0]], ref.last[gi[1]])
    assert np.all(recs["miss"] == 0)
    # float32 sums in another order: within a few ulps of the sum's magnitude per added term
    tol = 1e-6 * np.abs(ref.acc[gi[1]]).max(1, keepdims=True) * ref.cnt[gi[1], None] + 1e-6
    assert np.all(np.abs(recs["acc"][gi[0]] - ref.acc[gi[1]]) <= tol)
    # the clock moves by `shift` for occupied slots only
    mp.rebase_clock(2.5)
    rb = mp.records()
    o = rb["key"] >= 0
    assert np.array_equal(rb["first"][o], recs["first"][o] - np.float32(2.5))
    assert np.array_equal(rb["last"][~o], recs["last"][~o])
    mp.clear_map()
    assert mp.map_stats() == (0, 0) and np.all(mp.records()["key"] == -1)


def test_insert_runs_and_the_count_cap(lw, orc, gpu_ok):
    """Scan-ordered points: runs of one voxel longer than a wave, runs of every length 1..70 and a voxel hit
    9,000 times, the count cap (4096) crossed inside a run. Counts, times and occupancy equal the oracle's; the
    capped voxel's sums hold exactly 4096 of its (equal) points."""
    rng = np.random.default_rng(11)
    parts, lens = [], list(range(1, 71)) * 3
    rng.shuffle(lens)
    for k, ln in enumerate(lens):
        c = np.array([0.5 + 0.2 * (k % 30), 0.5 + 0.2 * (k // 30), 1.1], np.float32)
        parts.append(np.repeat(c[None], ln, 0) + rng.uniform(-0.05, 0.05, (ln, 3)).astype(np.float32))
    hot = np.repeat(np.array([[-2.05, 1.05, 0.55]], np.float32), 9000, 0)
    xyz = np.concatenate(parts[:100] + [hot] + parts[100:]).astype(np.float32)
    n = len(xyz)
    attr = np.full(n, 40, np.uint32)
    t = np.linspace(0.0, 1.0, n).astype(np.float32)
    fid = np.zeros(n, np.int32)
    pt = lw.PoseTable()
    pt.set_pose(0, np.eye(4, dtype=np.float32))
    view = lw.imap.View(min_range=0.0, max_range=100.0, min_refl=0, noise_mask=0b1111, ret_mask=0b1111)
    cap = 1 << 16
    mp = lw.IntegrationMap(pt, "cuda:0", map_capacity=cap)
    mp.map_insert(_dev(xyz, torch.float32), _dev(attr.view(np.int32), torch.int32), _dev(t, torch.float32),
                  _dev(fid, torch.int32), n, view, 0.2)
    recs = mp.records()
    ref = O.IntegrationMap(cap)
    O.map_insert(ref, xyz, attr, t, fid, O.identity_poses(), 0.2, min_range=0.0, max_range=100.0, min_refl=0,
                 noise_mask=0b1111, ret_mask=0b1111)
    assert mp.map_stats() == tuple(ref.stats)
    g = {int(k): i for i, k in enumerate(recs["key"]) if k >= 0}
    r = {int(k): i for i, k in enumerate(ref.keys) if k >= 0}
    assert set(g) == set(r)
    gi = np.array([g[k] for k in r]), np.array(list(r.values()))
    assert np.array_equal(recs["cnt"][gi[0]], ref.cnt[gi[1]]) and ref.cnt.max() == 9000
    assert np.array_equal(recs["first"][gi[0]], ref.first[gi[1]])
    assert np.array_equal(recs["last"][gi[0]], ref.last[gi[1]])
    capped = ref.cnt[gi[1]] > 4096
    tol = 1e-6 * np.abs(ref.acc[gi[1]]).max(1, keepdims=True) * np.minimum(ref.cnt[gi[1], None], 4096) + 1e-6
    assert np.all(np.abs(recs["acc"][gi[0]] - ref.acc[gi[1]]) <= tol)
    assert capped.sum() == 1
    hot_acc = recs["acc"][gi[0][capped]][0]
    assert np.allclose(hot_acc[:3], 4096 * hot[0], rtol=2e-5) and hot_acc[3] == 4096 * 40


@pytest.mark.parametrize("every", [1, 2])
def test_carve_matches_oracle(lw, orc, every, gpu_ok):
    rng = np.random.default_rng(11)
    xyz, attr, t, fid = _scene(rng, 12, ghost_until=4, n=3000)
    n = len(t)
    attr[::53] |= np.uint32(1 << 8)  # tagged returns: inserted (noise mask allows), never traced
    pt = lw.PoseTable()
    for k in range(12):
        pt.set_pose(k, np.eye(4))
    cap = 1 << 18
    mp = lw.IntegrationMap(pt, "cuda:0", map_capacity=cap)
    sx, sa, st, sf = (_dev(xyz, torch.float32), _dev(attr.view(np.int32), torch.int32), _dev(t, torch.float32),
                      _dev(fid, torch.int32))
    mp.map_insert(sx, sa, st, sf, n, lw.imap.View(min_range=0.1, max_range=40.0), V)
    ref = _records_to_oracle(mp.records(), cap)
    mp.carve(sx, sa, sf, n, V, every=every)
    got = mp.records()
    O.carve(ref, xyz, attr, fid, O.identity_poses(), V, every=every)
    print(f"every {every}: {int((got['key'] >= 0).sum())} voxels, misses {int(got['miss'].sum())}")
    assert got["miss"].sum() > 0
    assert np.array_equal(got["miss"], ref.miss)


def test_ghost_carved_floor_kept(lw, gpu_ok):
    rng = np.random.default_rng(13)
    xyz, attr, t, fid = _scene(rng, 40, ghost_until=5)
    n = len(t)
    pt = lw.PoseTable()
    for k in range(40):
        pt.set_pose(k, np.eye(4))
    mp = lw.IntegrationMap(pt, "cuda:0", map_capacity=1 << 18)
    per = n // 40
    for k in range(40):  # batch by batch, as the pipeline feeds it
        sl = slice(k * per, (k + 1) * per)
        sx, sa, st, sf = (_dev(xyz[sl], torch.float32), _dev(attr[sl].view(np.int32), torch.int32),
                          _dev(t[sl], torch.float32), _dev(fid[sl], torch.int32))
        mp.map_insert(sx, sa, st, sf, per, lw.imap.View(min_range=0.1, max_range=40.0), V)
        mp.carve(sx, sa, sf, per, V, every=1)
    r = mp.records()
    occ = r["key"] >= 0
    cen = r["acc"][occ, :3] / np.minimum(r["cnt"][occ], 4096)[:, None]
    cnt, miss = r["cnt"][occ], r["miss"][occ]
    ghost = np.linalg.norm(cen - np.array([2.8, 0.0, 0.0]), axis=1) < 0.25  # the box's front face
    floor = (cen[:, 2] < -0.95) & (cen[:, 0] > 2.0) & (cen[:, 0] < 5.5)
    wall = cen[:, 0] > 5.95
    carved = miss > 2 * cnt  # carve_ratio 2 (check_perception.py:36)
    print(f"ghost {ghost.sum()} voxels, carved {carved[ghost].mean():.2f}; floor {floor.sum()}, carved "
          f"{carved[floor].mean():.4f}; wall {wall.sum()}, carved {carved[wall].mean():.4f}")
    assert ghost.sum() > 10 and carved[ghost].mean() > 0.5
    assert floor.sum() > 300 and carved[floor].mean() < 0.01
    assert wall.sum() > 300 and carved[wall].mean() < 0.01

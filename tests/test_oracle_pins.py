"""Pins for the oracle functions that the key, map and known-answer tests leave open, each checked against a
restatement written independently of the C code (numpy or plain Python), chosen so that a plausible slip in the
oracle -- a dropped filter, a wrong window edge, a transposed Jacobian, a sign -- fails here:

  * frame_append (odom.py:97-123): numpy over the whole batch -- the frame of floor(max(t, 0) * inv_dt) in float32,
    the window k_min <= k <= k_open, input order, the count running past the cap;
  * ds_insert (odom.py:126-160): numpy np.unique first occurrence of the half-voxel keys after the range, tag and
    return filters, in index order (the drop-free case), and a plain-Python linear probe on a tiny table (the
    case where keys find no slot);
  * k_icp's per-point entries (odom.py:275-375): float64 numpy w J^T M J, w J^T M r, w J^T Mg J, the error, weight,
    rms and range terms, with M and Mg from numpy.linalg.eigh of the voxel covariance and the closest mean chosen
    by a brute-force scan of the neighbour voxels through a Python dict of the table;
  * loc_register (localize.py:517-553 with odom.py's accumulation): convergence back to the true pose from
    offset starts, and no step at all from a start that sees none of the map.
"""

import math

import numpy as np
import pytest

# ---- frame_append --------------------------------------------------------------------------------------


def frame_append_np(sx, sa, st, inv_dt, k_min, k_open, cap, count):
    t = np.maximum(np.asarray(st, np.float32), np.float32(0.0))
    k = np.floor(t * np.float32(inv_dt)).astype(np.int64)
    sel = np.nonzero((k >= k_min) & (k <= k_open))[0]
    room = max(cap - count, 0)
    take = sel[:room]
    return count + len(sel), sx[take], sa[take], np.asarray(st, np.float32)[take], count


def _run_append(orc, batches, inv_dt, cap):
    f_xyz = np.zeros((cap, 3), np.float32)
    f_attr = np.zeros(cap, np.uint32)
    f_t = np.zeros(cap, np.float32)
    f_frame = np.full(cap, -7, np.int32)
    count = 0
    ref_x, ref_a, ref_t, ref_f = [], [], [], []
    ref_count = 0
    for sx, sa, st, k_min, k_open in batches:
        count = orc.frame_append(sx, sa, st, inv_dt, k_min, k_open, cap, count, f_xyz, f_attr, f_t, f_frame)
        ref_count, x, a, t, _ = frame_append_np(sx, sa, st, inv_dt, k_min, k_open, cap, ref_count)
        ref_x.append(x), ref_a.append(a), ref_t.append(t), ref_f.append(np.full(len(t), k_open, np.int32))
    n = min(ref_count, cap)
    assert count == ref_count
    assert np.array_equal(f_xyz[:n], np.concatenate(ref_x)[:n])
    assert np.array_equal(f_attr[:n], np.concatenate(ref_a)[:n])
    assert np.array_equal(f_t[:n], np.concatenate(ref_t)[:n])
    assert np.array_equal(f_frame[:n], np.concatenate(ref_f)[:n])
    assert np.all(f_frame[n:] == -7)  # nothing written past the cap
    return count


def test_frame_append_window_and_order(orc):
    rng = np.random.default_rng(3)
    inv_dt = 10.0
    batches = []
    for b in range(6):
        n = 3000
        st = rng.uniform(b * 0.05 - 0.03, b * 0.05 + 0.12, n).astype(np.float32)
        st[:20] = -rng.uniform(0, 1, 20)  # negative times clamp to frame 0
        st[20:40] = (np.arange(20) // 4) * np.float32(0.1)  # exactly on frame boundaries
        st[40:60] = np.nextafter(st[20:40], np.float32(-1))  # just below them
        rng.shuffle(st)
        sx = rng.normal(0, 5, (n, 3)).astype(np.float32)
        sa = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
        k_open = int(b * 0.05 * inv_dt)
        batches.append((sx, sa, st, k_open - 1 if b % 2 else k_open, k_open))
    count = _run_append(orc, batches, inv_dt, cap=1 << 15)
    assert 0 < count < 6 * 3000


def test_frame_append_counts_past_the_cap(orc):
    rng = np.random.default_rng(4)
    n = 2500
    st = rng.uniform(0.0, 0.0999, n).astype(np.float32)
    sx = rng.normal(0, 5, (n, 3)).astype(np.float32)
    sa = np.arange(n, dtype=np.uint32)
    count = _run_append(orc, [(sx, sa, st, -1, 0), (sx, sa, st, -1, 0)], 10.0, cap=3000)
    assert count == 2 * n  # the count keeps running; only the first 3000 are stored


def test_frame_append_plants_a_mistake(orc):
    """The numpy restatement is sharp enough to see a window that includes k_open + 1."""
    rng = np.random.default_rng(5)
    st = rng.uniform(0.0, 0.3, 1000).astype(np.float32)
    sx = rng.normal(0, 5, (1000, 3)).astype(np.float32)
    sa = np.zeros(1000, np.uint32)
    good = frame_append_np(sx, sa, st, 10.0, 0, 1, 4096, 0)[0]
    wrong = frame_append_np(sx, sa, st, 10.0, 0, 2, 4096, 0)[0]
    assert good != wrong


# ---- ds_insert -------------------------------------------------------------------------------------------


def _filters(orc, xyz, attr, min_range, max_range, ret_mask):
    x = xyz.astype(np.float32)
    r = np.sqrt((x[:, 0] * x[:, 0] + x[:, 1] * x[:, 1]) + x[:, 2] * x[:, 2])
    tag = (attr >> 8) & 255
    ret = (attr >> 16) & 255
    ret_ok = (ret < 31) & (((np.int64(1) << np.minimum(ret, 30).astype(np.int64)) & ret_mask) != 0)
    return (r >= np.float32(min_range)) & (r <= np.float32(max_range)) & (tag == 0) & ret_ok


def ds_insert_np(orc, xyz, attr, inv_v, min_range, max_range, ret_mask):
    keep = np.nonzero(_filters(orc, xyz, attr, min_range, max_range, ret_mask))[0]
    keys = orc.voxel_key_np(xyz[keep], inv_v)
    _, first = np.unique(keys, return_index=True)
    return xyz[keep[np.sort(first)]]


def ds_insert_probe_py(orc, xyz, attr, inv_v, min_range, max_range, ret_mask, slots):
    """The reference's claim loop, serially, in plain Python ints: 64 probes from the key's home slot."""
    keep = _filters(orc, xyz, attr, min_range, max_range, ret_mask)
    keys = orc.voxel_key_np(xyz, inv_v)
    homes = orc.slot_hash_np(keys)
    table = [None] * slots
    out = []
    for i in range(len(xyz)):
        if not keep[i]:
            continue
        key, h = int(keys[i]), int(homes[i]) & (slots - 1)
        for probe in range(64):
            s = (h + probe) & (slots - 1)
            if table[s] is None:
                table[s] = key
                out.append(xyz[i])
                break
            if table[s] == key:
                break
    return np.array(out, np.float32).reshape(-1, 3)


def _thin_cloud(rng, n, inv_v):
    xyz = rng.normal(0, 6, (n, 3)).astype(np.float32)
    # points exactly on voxel faces, and one float32 step either side of them
    face = (rng.integers(-40, 40, (n // 8, 3)) / np.float32(inv_v)).astype(np.float32)
    xyz[: n // 8] = face
    xyz[n // 8: n // 4] = np.nextafter(face, np.float32(100.0))
    xyz[n // 4: 3 * n // 8] = np.nextafter(face, np.float32(-100.0))
    xyz[-50:] = xyz[:50]  # repeats of earlier points: never kept twice
    attr = np.zeros(n, np.uint32)
    attr[rng.random(n) < 0.05] |= np.uint32(1 << 8)  # noise-tagged
    attr[rng.random(n) < 0.05] |= np.uint32(2 << 16)  # third return
    return xyz, attr


@pytest.mark.parametrize("ret_mask", [0b1111, 0b0011])
def test_ds_insert_first_occurrence(orc, ret_mask):
    rng = np.random.default_rng(8)
    inv_v = 2.0 / 0.2
    xyz, attr = _thin_cloud(rng, 12_000, inv_v)
    got = orc.ds_insert(xyz, attr, inv_v, 0.3, 20.0, ret_mask, 1 << 15)
    ref = ds_insert_np(orc, xyz, attr, inv_v, 0.3, 20.0, ret_mask)
    assert got.shape == ref.shape and np.array_equal(got, ref)
    assert 1000 < len(ref) < 12_000


def test_ds_insert_full_table_drops_like_a_serial_probe(orc):
    rng = np.random.default_rng(9)
    inv_v = 2.0 / 0.2
    xyz, attr = _thin_cloud(rng, 3000, inv_v)
    slots = 256  # far fewer slots than voxels: keys past the 64-probe window are dropped
    got = orc.ds_insert(xyz, attr, inv_v, 0.3, 20.0, 0b1111, slots)
    ref = ds_insert_probe_py(orc, xyz, attr, inv_v, 0.3, 20.0, 0b1111, slots)
    assert len(ref) <= slots and np.array_equal(got, ref)
    # with room to spare, the serial probe and the first-occurrence restatement agree
    big = ds_insert_probe_py(orc, xyz, attr, inv_v, 0.3, 20.0, 0b1111, 1 << 14)
    assert np.array_equal(big, ds_insert_np(orc, xyz, attr, inv_v, 0.3, 20.0, 0b1111))


# ---- k_icp per point --------------------------------------------------------------------------------------

_TRIU = np.triu_indices(6)


def icp_point_np(orc, om, slot_of, p, T, voxel, reach, max_d2, noise2, tight2, alpha, kappa):
    """One point of k_icp in float64 numpy (float32 where the kernel's float32 fixes a value that decides
    something: the world point, the voxel index, the means and the distances they are chosen by)."""
    f32 = np.float32
    w = orc.transform_points_f32(T, p[None, :])[0]
    inv_v = f32(1.0 / voxel)
    b = np.floor(w * inv_v).astype(np.int64)
    best, best_d2, best_mu = None, f32(max_d2), None
    for dx in range(-reach, reach + 1):
        for dy in range(-reach, reach + 1):
            for dz in range(-reach, reach + 1):
                j = b + np.array([dx, dy, dz])
                s = slot_of.get(int(orc.voxel_key_ijk_np(j[0], j[1], j[2])))
                if s is None or om.rcnt[s] <= 0:
                    continue
                c = f32(om.rcnt[s])
                mu = ((j.astype(f32) + f32(0.5)) * f32(voxel)) + om.rsum[s] / c
                dd = (w - mu).astype(f32)
                d2 = (dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]
                if d2 < best_d2:
                    best, best_d2, best_mu = s, d2, mu
    if best is None:
        return None, None
    c = int(om.rcnt[best])
    m = (om.rsum[best] / f32(c)).astype(np.float64)
    cov = om.router[best].astype(np.float64) / c - np.outer(m, m)
    lam, Q = np.linalg.eigh(cov)
    tight = lam < tight2
    r = (w - best_mu).astype(np.float64)
    if c >= 5 and tight.sum() <= 1:
        wa = np.where(tight, 1.0, alpha) / (np.maximum(lam, 0.0) + noise2)
        M = (Q * wa) @ Q.T
        Mg = (Q * np.where(tight, wa, 0.0)) @ Q.T
        rn = float(r @ Q[:, int(np.argmin(lam))])
    else:
        M = np.eye(3) * alpha / (0.25 * voxel * voxel + noise2)
        Mg = np.zeros((3, 3))
        rn = float(np.linalg.norm(r))
    wv = w.astype(np.float64)
    J = np.zeros((3, 6))
    J[:, :3] = np.eye(3)
    J[:, 3:] = -np.array([[0, -wv[2], wv[1]], [wv[2], 0, -wv[0]], [-wv[1], wv[0], 0]])  # -[w]x
    d2 = float(r @ M @ r)
    g = (1.0 / (1.0 + d2 / (kappa * kappa))) ** 2
    out = np.zeros(64)
    out[0:21] = (g * J.T @ M @ J)[_TRIU]
    out[21:27] = g * J.T @ M @ r
    out[27] = 1.0
    out[28] = g * d2
    out[29] = g * rn * rn
    out[30] = g
    out[31] = float(np.linalg.norm(p.astype(np.float64)))
    out[32:53] = (g * J.T @ Mg @ J)[_TRIU]
    return best, out


@pytest.fixture(scope="module")
def room(lw, orc):
    w = lw.synth.single_scan(seed=14, map_points=150_000, n=4000)
    voxel = 0.25
    om = orc.OracleMap(1 << 18)
    om.insert(w.map_xyz, np.zeros(len(w.map_xyz), np.uint32), np.eye(4), voxel, max_range=1e30)
    ds = orc.ds_insert(w.scan_xyz, w.scan_attr, 2.0 / voxel, 0.3, 60.0, 15, 1 << 14)
    slot_of = {int(k): s for s, k in enumerate(om.keys) if k != -1}
    return w, om, ds, slot_of, voxel


@pytest.mark.parametrize("reach,kappa", [(1, 2.0), (2, 6.0)])
def test_icp_entries_per_point(orc, room, reach, kappa):
    w, om, ds, slot_of, voxel = room
    noise2 = float(np.float32(0.02 * 0.02))
    tight2 = float(np.float32((0.6 * voxel / math.sqrt(12.0)) ** 2))
    max_d2 = ((reach + 0.5) * voxel) ** 2
    T = w.T_true.copy()
    T[:3, 3] += [0.04, -0.03, 0.02]  # near the truth, as in the last iterations
    rng = np.random.default_rng(2)
    idx = rng.choice(len(ds), 150, replace=False)
    branches, matched, worst = set(), 0, 0.0
    for i in idx:
        p = ds[i]
        v, slot = orc.icp_accumulate(om, p[None, :], T, voxel, reach, max_d2, noise2, tight2, 0.05, kappa,
                                     with_slots=True)
        s_ref, ref = icp_point_np(orc, om, slot_of, p, T, voxel, reach, max_d2, noise2, tight2, 0.05, kappa)
        assert (slot[0] if slot[0] >= 0 else None) == s_ref
        if s_ref is None:
            assert not v.any()
            continue
        matched += 1
        branches.add(bool(ref[53 - 21:53].any()))  # Hg is zero on the isotropic branch
        for lo, hi in ((0, 21), (21, 27), (32, 53)):  # H, b, Hg: relative to the block's scale
            scale = max(np.abs(ref[lo:hi]).max(), 1e-30)
            err = np.abs(v[lo:hi] - ref[lo:hi]).max() / scale
            worst = max(worst, err)
            assert err < 1e-5, (i, lo, err)
        for k in (27, 28, 29, 30, 31):
            assert abs(v[k] - ref[k]) <= 1e-5 * max(abs(ref[k]), 1e-12), (i, k, v[k], ref[k])
    assert matched > 100 and branches == {True, False}
    print(f"reach {reach}: {matched} points, worst block error {worst:.2e}")


def test_icp_entries_catch_a_transposed_jacobian(orc, room):
    """A Jacobian with [w]x in place of -[w]x (a sign slip) moves the rotation entries well past the bar."""
    w, om, ds, slot_of, voxel = room
    p = ds[0]
    noise2 = float(np.float32(4e-4))
    tight2 = float(np.float32((0.6 * voxel / math.sqrt(12.0)) ** 2))
    v = orc.icp_accumulate(om, p[None, :], w.T_true, voxel, 1, (1.5 * voxel) ** 2, noise2, tight2, 0.05, 2.0)
    assert v[27] == 1.0
    # the b entries' rotation part flips sign under the slip: far beyond 1e-5 of the block
    s, ref = icp_point_np(orc, om, slot_of, p, w.T_true, voxel, 1, (1.5 * voxel) ** 2, noise2, tight2, 0.05, 2.0)
    slipped = ref[21:27].copy()
    slipped[3:] = -slipped[3:]
    assert np.abs(v[21:27] - slipped).max() > 1e-3 * np.abs(ref[21:27]).max()


# ---- loc_register -------------------------------------------------------------------------------------------


@pytest.fixture(scope="module")
def loc_room(lw, orc):
    w = lw.synth.single_scan(seed=12, map_points=300_000, n=6000)
    voxel = 0.1
    om = orc.OracleMap(1 << 20)
    om.insert(w.map_xyz, np.zeros(len(w.map_xyz), np.uint32), np.eye(4), voxel, max_range=1e30)
    return w, om, voxel


def _perturb(T, dp, deg):
    from scipy.spatial.transform import Rotation

    out = T.copy()
    R = Rotation.from_euler("z", deg, degrees=True).as_matrix()
    out[:3, :3] = R @ T[:3, :3]
    out[:3, 3] = T[:3, 3] + dp
    return out


STARTS = ([((s * 0.12, 0.0, 0.0), 0.0) for s in (1, -1)] + [((0.0, s * 0.12, 0.0), 0.0) for s in (1, -1)]
          + [((sx * 0.06, sy * 0.06, 0.0), d) for sx in (1, -1) for sy in (1, -1) for d in (1.0, -1.0)])


@pytest.mark.parametrize("dp,deg", STARTS)
def test_loc_register_converges_back(orc, loc_room, dp, deg):
    """Starts inside the reach-1 basin (a 0.1 m voxel's 27 neighbours see 0.15 m: 12 cm of translation, or 6 cm
    and 1 degree) return to the true pose within the scan's noise floor."""
    w, om, voxel = loc_room
    scan = w.scan_xyz
    T0 = _perturb(w.T_true, np.array(dp), deg)
    T, score, count, steps = orc.loc_register(om, scan, T0, voxel, iters=30)
    dt, da = orc.pose_error(w.T_true, T)
    assert dt < 0.01 and da < 0.1, (dp, deg, dt, da)
    assert steps > 0 and count > 0.5 * len(scan) and score > 0
    # the last step moved less than the stop threshold: a fixed point, not an iteration cap
    T2, _, _, steps2 = orc.loc_register(om, scan, T, voxel, iters=3)
    assert orc.pose_error(T, T2)[0] < 1e-3


def test_loc_register_no_overlap_takes_no_step(orc, loc_room):
    w, om, voxel = loc_room
    T0 = w.T_true.copy()
    T0[:3, 3] += [500.0, 500.0, 0.0]  # nowhere near the map
    T, score, count, steps = orc.loc_register(om, w.scan_xyz, T0, voxel, iters=10)
    assert steps == 0 and count == 0 and score == 0.0 and np.array_equal(T, T0)

"""The oracle's integration map (gpu.py:373-595: k_map_insert, local_plane, k_carve) pinned without a GPU:
the insert against an independent pure-Python restatement (dict of voxels, numpy float32 arithmetic, the
probe over a small table with collisions and drops), the 3x3 eigen solver against numpy.linalg.eigh, the
local plane on planes, lines and sparse sets with known normals, and carving on rays with closed-form
outcomes (a voxel on the ray is pierced, one beside it is not, the margin before the return is spared, the
every-th sampling scales the misses)."""

import numpy as np
import pytest

import oracle as O


def _key(ix, iy, iz):
    b, m = 1 << 20, (1 << 21) - 1
    return ((ix + b) & m) | (((iy + b) & m) << 21) | (((iz + b) & m) << 42)


def _hash(k):
    h = (k * 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
    h ^= h >> 29
    return h & 0x7FFFFFFF


def _insert_py(cap, xyz, attr, t, frame, poses, voxel, view):
    """k_map_insert in index order, restated with a Python dict for the table's contents."""
    f32 = np.float32
    keys = [-1] * cap
    rec = {}
    stats = [0, 0]
    inv = f32(1.0) / f32(voxel)
    for i in range(len(t)):
        p = xyz[i].astype(f32)
        a = int(attr[i])
        r = np.sqrt(f32(p[0] * p[0] + p[1] * p[1]) + f32(p[2] * p[2]), dtype=f32)
        if r < f32(view["min_range"]) or r > f32(view["max_range"]):
            continue
        if (a & 255) < view["min_refl"] or not (view["noise_mask"] >> (((a >> 8) & 255) & 3)) & 1:
            continue
        if not (view["ret_mask"] >> (((a >> 16) & 255) & 3)) & 1:
            continue
        T = poses[int(frame[i]) & 0xFFFF].astype(f32)
        w = [f32(f32(f32(T[r_, 0] * p[0]) + f32(T[r_, 1] * p[1])) + f32(T[r_, 2] * p[2])) + T[r_, 3] for r_ in range(3)]
        w = [f32(v) for v in w]
        key = _key(*[int(np.floor(f32(v * inv))) for v in w])
        h = _hash(key) & (cap - 1)
        for probe in range(64):
            s = (h + probe) & (cap - 1)
            if keys[s] == -1:
                keys[s] = key
                stats[0] += 1
            if keys[s] == key:
                e = rec.setdefault(key, {"cnt": 0, "acc": [f32(0)] * 4, "first": f32(1e30), "last": f32(0)})
                if e["cnt"] < O.MAP_COUNT_CAP:
                    e["acc"] = [f32(e["acc"][k] + v) for k, v in enumerate(w + [f32(a & 255)])]
                e["cnt"] += 1
                e["last"] = max(e["last"], f32(t[i]))
                e["first"] = min(e["first"], f32(t[i]))
                break
        else:
            stats[1] += 1
    return rec, stats


def _poses(rng, frames):
    P = O.identity_poses()
    for k in frames:
        a = rng.uniform(-np.pi, np.pi)
        P[k, :3, :3] = [[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]]
        P[k, :3, 3] = rng.uniform(-5, 5, 3)
    return P


@pytest.mark.parametrize("cap", [64, 1 << 12])
def test_map_insert_matches_python(orc, cap):
    rng = np.random.default_rng(cap)
    n = 3000
    xyz = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    xyz[: n // 2] = np.round(xyz[: n // 2], 1)  # many points on shared voxels
    attr = (rng.integers(0, 256, n) | (rng.integers(0, 4, n) << 8) | (rng.integers(0, 4, n) << 16)).astype(np.uint32)
    t = rng.uniform(0, 10, n).astype(np.float32)
    frame = rng.integers(0, 5, n).astype(np.int32) + 70000  # ids past 65535 wrap onto the pose slots
    poses = _poses(rng, [k & 0xFFFF for k in range(70000, 70005)])
    view = dict(min_range=0.5, max_range=4.5, min_refl=10, noise_mask=0b0111, ret_mask=0b1011)
    m = O.IntegrationMap(cap)
    O.map_insert(m, xyz, attr, t, frame, poses, 0.25, **view)
    ref, stats = _insert_py(cap, xyz, attr, t, frame, poses, 0.25, view)
    assert list(m.stats) == stats
    if cap == 64:
        assert stats[1] > 0  # the small table fills and drops points
    occ = np.flatnonzero(m.keys >= 0)
    assert sorted(int(k) for k in m.keys[occ]) == sorted(ref)
    for s in occ:
        e = ref[int(m.keys[s])]
        assert m.cnt[s] == e["cnt"] and m.first[s] == e["first"] and m.last[s] == e["last"]
        assert np.array_equal(m.acc[s], np.array(e["acc"], np.float32))


def test_count_cap_stops_sums_not_counts(orc):
    m = O.IntegrationMap(1 << 10)
    n = O.MAP_COUNT_CAP + 100
    xyz = np.tile(np.array([[1.05, 2.05, 0.05]], np.float32), (n, 1))
    attr = np.full(n, 7, np.uint32)
    t = np.arange(n, dtype=np.float32)
    O.map_insert(m, xyz, attr, t, np.zeros(n, np.int32), O.identity_poses(), 0.1)
    s = int(np.flatnonzero(m.keys >= 0)[0])
    assert m.cnt[s] == n and m.first[s] == 0.0 and m.last[s] == n - 1
    assert m.acc[s, 3] == 7.0 * O.MAP_COUNT_CAP  # exactly the first 4096 reflectivities


def test_sym_eig3_matches_eigh(orc):
    rng = np.random.default_rng(3)
    for _ in range(50):
        B = rng.normal(size=(3, 3))
        A = B @ B.T
        ev, V = np.zeros(3), np.zeros(9)
        orc.lib().lwo_sym_eig3(orc._p(A.ravel().copy(), orc._f64p), orc._p(ev, orc._f64p), orc._p(V, orc._f64p))
        V = V.reshape(3, 3)
        assert np.allclose(np.sort(ev), np.linalg.eigvalsh(A), rtol=1e-12, atol=1e-12)
        assert np.allclose(A @ V, V * ev, atol=1e-10)  # column i is eigenvector i


def _fill(m, centroids, voxel=0.1):
    """Records whose centroids are the given points (one hit each)."""
    pts = np.asarray(centroids, np.float32)
    O.map_insert(m, pts, np.zeros(len(pts), np.uint32), np.zeros(len(pts), np.float32), np.zeros(len(pts), np.int32),
                 O.identity_poses(), voxel)


def test_local_plane_normals(orc):
    rng = np.random.default_rng(5)
    v = 0.1
    # a tilted plane through voxel (10, 10, 10): centroids on it, one per voxel of the 3x3x3 block
    n = np.array([0.2, -0.3, 0.93])
    n /= np.linalg.norm(n)
    pts = []
    for gx in range(-1, 2):
        for gy in range(-1, 2):
            x, y = (10.5 + gx) * v + rng.uniform(-0.02, 0.02), (10.5 + gy) * v + rng.uniform(-0.02, 0.02)
            z = 1.05 - (n[0] * (x - 1.05) + n[1] * (y - 1.05)) / n[2]
            pts.append([x, y, z])
    m = O.IntegrationMap(1 << 10)
    _fill(m, pts, v)
    out = O.local_plane(m, 10, 10, 10, [1.05, 1.05, 1.05])
    assert out[3] == 2.0 and abs(abs(float(np.dot(out[:3], n))) - 1.0) < 1e-5
    # a 3D cross (the six face neighbours): occupied but no surface, every spread alike
    m2 = O.IntegrationMap(1 << 10)
    cross = [[1.05 + dx * v, 1.05 + dy * v, 1.05 + dz * v] for dx, dy, dz in
             ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1))]
    _fill(m2, cross, v)
    out2 = O.local_plane(m2, 10, 10, 10, [1.05, 1.05, 1.05])
    assert out2[3] == 1.0
    # fewer than 4 occupied voxels
    m3 = O.IntegrationMap(1 << 10)
    _fill(m3, [[1.05, 1.05, 1.05], [1.15, 1.05, 1.05]], v)
    assert np.array_equal(O.local_plane(m3, 10, 10, 10, [1.05, 1.05, 1.05]), np.array([0, 0, 1, 0], np.float32))


def _ray_case(voxel_centroids, end, every=1, n_rays=1):
    """Rays from the origin (identity pose) to `end`; the map holds one hit at each given centroid."""
    m = O.IntegrationMap(1 << 12)
    _fill(m, voxel_centroids, 0.1)
    xyz = np.tile(np.asarray(end, np.float32), (n_rays, 1))
    O.carve(m, xyz, np.zeros(n_rays, np.uint32), np.zeros(n_rays, np.int32), O.identity_poses(), 0.1, every=every)
    out = {}
    for c in voxel_centroids:
        ijk = np.floor(np.asarray(c, np.float32) / np.float32(0.1)).astype(int)
        out[tuple(ijk)] = int(m.miss[np.flatnonzero(m.keys == np.int64(_key(*ijk)))[0]])
    return out


def test_carve_closed_form(orc):
    # a lone voxel on the ray (centroid on it): pierced once; one 0.07 m beside the ray: not (0.05 m bound)
    out = _ray_case([[1.05, 0.0, 0.0], [2.05, 0.07, 0.0]], [4.0, 0.0, 0.0])
    assert out[(10, 0, 0)] == 1 and out[(20, 0, 0)] == 0
    # within the margin before the return (max(3 voxels, 3% of the range) = 0.3 m): spared
    out = _ray_case([[3.85, 0.0, 0.0]], [4.0, 0.0, 0.0])
    assert out[(38, 0, 0)] == 0
    # every = 3 over 7 rays: rays 0, 3, 6 are traced, each adding 3
    out = _ray_case([[1.05, 0.0, 0.0]], [4.0, 0.0, 0.0], every=3, n_rays=7)
    assert out[(10, 0, 0)] == 9
    # tagged returns are not traced
    m = O.IntegrationMap(1 << 12)
    _fill(m, [[1.05, 0.0, 0.0]], 0.1)
    O.carve(m, np.array([[4.0, 0, 0]], np.float32), np.array([1 << 8], np.uint32), np.zeros(1, np.int32),
            O.identity_poses(), 0.1, every=1)
    assert m.miss.sum() == 0


def test_carve_spares_a_floor_at_grazing_incidence(orc):
    """A floor patch the ray skims: its voxels form a surface, the ray does not cross the plane within 0.6 voxel
    of any centroid, so nothing is pierced; the same voxels crossed head-on are."""
    v = 0.1
    floor = [[(i + 0.5) * v, (j + 0.5) * v, -0.95] for i in range(5, 40) for j in range(-3, 3)]
    m = O.IntegrationMap(1 << 14)
    _fill(m, floor, v)
    grazing = np.array([[4.5, 0.05, -0.9]], np.float32)  # from the origin down to just above the floor's far end
    O.carve(m, grazing, np.zeros(1, np.uint32), np.zeros(1, np.int32), O.identity_poses(), v, every=1)
    assert m.miss.sum() == 0
    steep = np.array([[2.05, 0.05, -3.0]], np.float32)  # through the floor, nearly head-on
    O.carve(m, steep, np.zeros(1, np.uint32), np.zeros(1, np.int32), O.identity_poses(), v, every=1)
    assert m.miss.sum() >= 1

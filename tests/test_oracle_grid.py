"""The prior-map grid and coarse scoring, pinned without a GPU (mapgrid.py, localize.py:55-123, 426-528):
the host preparation sorts points by the cells they fall in; the oracle's seeds are the cell's point nearest
the centre, its jump-flooded seeds never beat and nearly always equal the exact nearest map point (a k-d tree),
every cell within half a metre of the map has one, the distance field is their rounded distance in cm; the
nearest-point query is exact within a cell and close beyond; the soft count meets hand-computed sums on a
hand-made distance field at headings 0 and 90 degrees (a transposed rotation lands elsewhere), gates
positions inside surfaces or outside the grid, and agrees with the hypothesis-list kernel; the host's local
maxima and suppression match their definitions."""

import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle as O


def _map(seed=3, spacing=0.02):
    """A floor, two walls and a tilted panel in a 2 x 2 x 1.5 m corner, points jittered off a 2 cm lattice."""
    rng = np.random.default_rng(seed)
    u = np.arange(0.0, 2.0, spacing)
    v = np.arange(0.0, 1.5, spacing)
    F = np.stack(np.meshgrid(u, u, indexing="ij"), -1).reshape(-1, 2)
    W = np.stack(np.meshgrid(u, v, indexing="ij"), -1).reshape(-1, 2)
    pts = [np.c_[F, np.zeros(len(F))], np.c_[np.zeros(len(W)), W], np.c_[W[:, 0], np.full(len(W), 2.0), W[:, 1]]]
    a, b = np.meshgrid(np.arange(0, 0.6, spacing), np.arange(0, 0.4, spacing), indexing="ij")
    pts.append(np.c_[1.0 + a.ravel(), 0.8 + 0.5 * a.ravel(), 0.3 + b.ravel()])
    xyz = np.concatenate(pts) + rng.normal(0, 0.004, (sum(len(p) for p in pts), 3))
    return xyz.astype(np.float32)


@pytest.fixture(scope="module")
def built(lw):
    O.lib()
    prep = lw.mapgrid.PriorGrid.prepare({"xyz": _map()})
    nn, df = O.grid_build(prep["xyz"], prep["origin"], prep["cell"], prep["dims"])
    return prep, nn, df


def _centres(prep):
    nx, ny, nz = (int(v) for v in prep["dims"])
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1).astype(np.float64)
    return prep["origin"].astype(np.float64) + (c + 0.5) * prep["cell"]


def test_prepare_sorts_by_cell(lw):
    prep = lw.mapgrid.PriorGrid.prepare({"xyz": _map()})
    xyz, start, dims = prep["xyz"], prep["start"], prep["dims"]
    u = (xyz.astype(np.float64) - prep["origin"]) / prep["cell"]
    ijk = np.floor(u).astype(np.int64)
    key = (ijk[:, 2] * dims[1] + ijk[:, 1]) * dims[0] + ijk[:, 0]
    clear = np.all(np.abs(u - np.round(u)) > 1e-4, axis=1)  # not on a cell boundary (float32 may round there)
    assert clear.mean() > 0.99 and start[-1] == len(xyz) and start[0] == 0
    assert np.all(np.diff(key[clear]) >= 0)
    # each point lies in the slice of its own cell
    assert np.array_equal((np.searchsorted(start, np.arange(len(xyz)), side="right") - 1)[clear], key[clear])
    assert np.all(prep["origin"] <= xyz.min(0) - 0.59) and prep["cell"] == 0.05


def test_seeds_band_and_field(built):
    prep, nn, df = built
    xyz = prep["xyz"].astype(np.float64)
    ctr = _centres(prep)
    exact, _ = cKDTree(xyz).query(ctr)
    has = nn >= 0
    seed_d = np.full(len(nn), np.inf)
    seed_d[has] = np.linalg.norm(ctr[has] - xyz[nn[has]], axis=1)
    # a seed is a real map point: never nearer than the nearest one, and the flood finds the nearest of the
    # cells' own seeds (points nearest their cell centres) for all but a few cells
    assert np.all(seed_d[has] >= exact[has] - 1e-6)
    nn0, _ = O.grid_build(prep["xyz"], prep["origin"], prep["cell"], prep["dims"], steps=())
    own = np.unique(nn0[nn0 >= 0])
    assert np.array_equal(np.unique(nn[has]), own)  # the flood only moves seeds
    nearest_seed, _ = cKDTree(xyz[own]).query(ctr)
    near = has & (exact < 0.4)
    assert np.mean(np.abs(seed_d[near] - nearest_seed[near]) < 1e-6) > 0.999
    assert np.max(seed_d[near] - nearest_seed[near]) < 0.01
    # so the field is within a couple of cm of the true distance (mapgrid.py:5-6: ~1 cm at 10 cm, ~2 cm beyond)
    assert np.percentile(seed_d[near] - exact[near], 95) < 0.025 and np.max(seed_d[near] - exact[near]) < 0.05
    # the band: every cell within 0.5 m of the map has a seed (steps 8 + 4 + 2 + 1 + 1 cells reach 0.8 m)
    assert np.all(has[exact < 0.5])
    # the distance field: the seed distance in cm, rounded, 255 beyond the band
    assert np.all(df[~has] == 255)
    cm = seed_d[has] * 100
    assert np.all(np.abs(df[has].astype(np.float64) - np.minimum(cm, 255)) <= 0.5 + 1e-3)
    # an occupied cell's seed is its own point nearest the centre (distances quantised to 1 mm)
    start = prep["start"]
    occ = np.flatnonzero(np.diff(start) > 0)[:2000]
    for c in occ:
        d = np.linalg.norm(xyz[start[c]:start[c + 1]] - ctr[c], axis=1)
        if start[c] <= nn[c] < start[c + 1]:
            assert seed_d[c] <= d.min() + 1e-3 + 1e-6
    inside = np.mean([start[c] <= nn[c] < start[c + 1] for c in occ])
    assert inside > 0.99  # a point on a cell boundary may be counted in the neighbour by the float32 floor


def test_nearest_point_exact_within_a_cell(built):
    prep, nn, df = built
    xyz = prep["xyz"]
    tree = cKDTree(xyz.astype(np.float64))
    rng = np.random.default_rng(4)
    args = (prep["origin"], prep["cell"], prep["dims"], prep["start"], nn, xyz)
    for scale, tol in ((0.02, 1e-6), (0.2, 0.03)):
        q = xyz[rng.integers(0, len(xyz), 300)] + rng.normal(0, scale, (300, 3)).astype(np.float32)
        for w in q:
            k = O.grid_nearest(w, *args)
            d_true, _ = tree.query(w.astype(np.float64))
            if d_true < 0.045:  # within one cell: the exact nearest (up to float32 ties)
                assert k >= 0 and np.linalg.norm(w - xyz[k]) <= d_true + tol + 1e-6
            elif d_true < 0.5:  # through the seeds: close
                assert k >= 0 and np.linalg.norm(w - xyz[k]) <= d_true + 0.03
    # distances through T: a pure translation equals querying the moved points; far points read 1e3
    T = np.eye(4, dtype=np.float32)
    T[:3, 3] = (0.3, -0.2, 0.1)
    p = xyz[:200] - T[:3, 3]
    d = O.grid_distances(p, T, *args)
    assert np.allclose(d, [np.linalg.norm(w - xyz[max(O.grid_nearest(w, *args), 0)]) for w in p + T[:3, 3]], atol=1e-6)
    assert np.all(O.grid_distances(np.array([[50.0, 50.0, 50.0]], np.float32), np.eye(4), *args) == 1e3)


def _hand_field():
    """A 40 x 40 x 10 grid of 5 cm cells at the origin, 255 everywhere but a few cells."""
    dims = np.array([40, 40, 10], np.int32)
    df = np.full(40 * 40 * 10, 255, np.uint8)

    def at(ix, iy, iz):
        return (iz * 40 + iy) * 40 + ix

    for ijk, v in (((22, 20, 5), 0), ((20, 24, 6), 5), ((14, 20, 5), 30), ((20, 22, 5), 10), ((20, 14, 5), 0),
                   ((30, 30, 5), 3)):
        df[at(*ijk)] = v
    return dims, df


def test_soft_count_closed_form(orc):
    dims, df = _hand_field()
    origin, cell = np.zeros(3, np.float32), 0.05
    pts = np.array([[0.1, 0.0, 0.0], [0.0, 0.2, 0.05], [-0.3, 0.0, 0.0]], np.float32)
    c = (np.array([20, 20, 5]) + 0.5) * 0.05
    pos = np.array([c, (np.array([30, 30, 5]) + 0.5) * 0.05, [5.0, 0.5, 0.2]], np.float32)
    sc = O.score_grid(pts, pos, 4, 0.0, math.pi / 2, origin, cell, dims, df, 25.0, 10.0).reshape(3, 4)
    # heading 0: cells (22,20,5) d=0, (20,24,6) d=5, (14,20,5) d=30 >= tau; 90 deg: (20,22,5) d=10,
    # (16,20,6) empty, (20,14,5) d=0; 180 deg: (18,20,5), (20,16,6), (26,20,5) all empty
    assert np.isclose(sc[0, 0], 1 + (1 - 25 / 625), rtol=1e-6) and np.isclose(sc[0, 1], (1 - 100 / 625) + 1, rtol=1e-6)
    assert sc[0, 2] == 0.0
    assert np.all(sc[1] == -1.0) and np.all(sc[2] == -1.0)  # inside a surface (3 cm < 10 cm), outside the grid
    # the list kernel gives the same sums for the same hypotheses
    hyp = np.array([[*c, 0.0], [*c, np.float32(math.pi / 2)], [*c, np.float32(math.pi)]], np.float32)
    assert np.array_equal(O.score_list(pts, hyp, origin, cell, dims, df, 25.0), sc[0, :3])


def test_local_maxima_and_nms(lw):
    CS = lw.localize.CoarseSearch
    rng = np.random.default_rng(7)
    shape = (2, 6, 5, 12)
    sc = rng.random(int(np.prod(shape))).astype(np.float32)
    got = set(CS.local_maxima(sc, shape, 10_000).tolist())
    S = sc.reshape(shape)
    want = set()
    for iz, ix, iy, ia in np.ndindex(*shape):
        v = S[iz, ix, iy, ia]
        if all(S[iz, ix + ox, iy + oy, (ia + oa) % shape[3]] <= v for ox in (-1, 0, 1) for oy in (-1, 0, 1)
               for oa in (-1, 0, 1) if 0 <= ix + ox < shape[1] and 0 <= iy + oy < shape[2]):
            want.add(int(np.ravel_multi_index((iz, ix, iy, ia), shape)))
    assert got == want
    hyps = np.array([[0, 0, 0, 0.0], [0.05, 0, 0, 0.01], [0.5, 0, 0, 0.0], [0.02, 0, 0, math.pi],
                     [0, 0, 0, 2 * math.pi - 0.01]])
    scores = np.array([3.0, 2.0, 1.0, 1.5, 2.5])
    assert CS.nms(hyps, scores, 16, 0.1, 1.5).tolist() == [0, 3, 2]
    assert CS.nms(hyps, scores, 2, 0.1, 1.5).tolist() == [0, 3]


@pytest.fixture(scope="module")
def room(lw):
    """The 8 x 6 x 3 m box room sampled every 4 cm with its face normals, prepared and built by the oracle."""
    S = lw.synth.box_room(1)
    xyz, nrm = lw.synth.sample_surfaces(S, 0.04, 3, normals=True)
    prep = lw.mapgrid.PriorGrid.prepare({"xyz": xyz, "normal": nrm})
    nn, _ = O.grid_build(prep["xyz"], prep["origin"], prep["cell"], prep["dims"])
    return S, prep, nn


def _room_scan(lw, S, n=4000, noise=0.0):
    T = lw.synth.pose_matrix(np.array([2.5, 3.0, 1.4]), 0.3, 0.02, -0.01)
    xyz, _, _ = lw.synth.scan(S, lw.synth.Rosette(5), lambda t: T, 0.0, n, 9, noise=noise)
    return T, xyz.astype(np.float32)


def _perturbed(lw, T, dp=(0.06, -0.05, 0.03), deg=(2.0, -1.0, 1.5)):
    d = np.r_[dp, np.radians(deg)]
    return lw.localize.twist_apply(T, d)


def test_plane_icp_drives_to_the_true_pose(lw, room):
    """The oracle's accumulations, solved as the reference's host loop does (localize.py:531-553), move a pose
    8 cm and 3 degrees off back onto a noise-free scan's true pose: a wrong sign or index in the Jacobian, the
    precision or the right-hand side would not converge there."""
    S, prep, nn = room
    T_true, pts = _room_scan(lw, S)
    T = _perturbed(lw, T_true)
    tri = np.triu_indices(6)
    for dmax, iters in ((0.40, 8), (0.20, 8), (0.10, 8), (0.05, 10)):
        for _ in range(iters):
            v = O.plane_icp(pts, [T], prep, nn, dmax * dmax, (dmax * 0.5) ** 2)[0]
            assert v[27] >= 30
            H = np.zeros((6, 6))
            H[tri] = v[:21]
            H += np.triu(H, 1).T
            assert np.all(np.linalg.eigvalsh(H) > -1e-9)  # a sum of weighted J^T M J: positive semidefinite
            delta = np.linalg.solve(H + np.diag(1e-3 * np.diag(H) + 1e-9), -v[21:27])
            T = lw.localize.twist_apply(T, delta)
            if max(np.linalg.norm(delta[:3]), 5 * np.linalg.norm(delta[3:])) < 2e-4:
                break
    d = np.linalg.inv(T_true) @ T
    assert np.linalg.norm(d[:3, 3]) < 3e-3
    assert np.degrees(np.arccos(np.clip((np.trace(d[:3, :3]) - 1) / 2, -1, 1))) < 0.05


def test_plane_eval_counts_match_kdtree(lw, room):
    """k_eval_batch's counts: within 3 and 5 cm the nearest point is exact (one cell), so the counts equal a k-d
    tree's up to float32 rounding at the radius; within 10 cm the seeds may be a little farther, never nearer."""
    S, prep, nn = room
    T_true, pts = _room_scan(lw, S, noise=0.03)
    c = O.plane_eval(pts, [T_true], prep, nn)[0]
    w = pts.astype(np.float64) @ T_true[:3, :3].T + T_true[:3, 3]
    d, _ = cKDTree(prep["xyz"].astype(np.float64)).query(w)
    want = [int((d < r).sum()) for r in O.EVAL_RADII]
    assert abs(int(c[0]) - want[0]) <= 2 and abs(int(c[1]) - want[1]) <= 2
    assert want[2] * 0.97 <= c[2] <= want[2] + 2


def test_prepare_refuses_grids_the_kernels_cannot_take(lw):
    """A map 4 km long in y at 5 cm has more rows than the kernels index: a clear error."""
    xyz = np.array([[0.0, 0.0, 0.0], [1.0, 4000.0, 1.0]], np.float32)
    with pytest.raises(ValueError, match="rows"):
        lw.mapgrid.PriorGrid.prepare({"xyz": xyz})
    with pytest.raises(ValueError, match="points"):
        lw.mapgrid.PriorGrid.prepare({"xyz": np.zeros((1 << 23, 3), np.float32)})

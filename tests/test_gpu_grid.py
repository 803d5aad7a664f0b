"""The prior-map grid and coarse scoring on the device (lw_grid_build, lw_grid_distances, lw_score_grid,
lw_score_list) against the oracle: the seeds and the distance field identical cell for cell, distances and
scores identical bit for bit (every hypothesis's sum runs in point order in float32, as the oracle's), the
hand-made field's closed-form sums, and on the 60 x 40 m multi-room floor (5.1M map points, 105M cells) the
coarse and medium stages keep the true pose among the candidates they hand to the ICP, and the batched ICP on
those candidates returns it as its best."""

import math

import numpy as np
import pytest
import torch

import oracle as O
import test_oracle_grid as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(lw, gpu_ok):
    prep = lw.mapgrid.PriorGrid.prepare({"xyz": G._map()})
    grid = lw.PriorGrid(prep, "cuda:0")
    grid.build()
    return prep, grid


def test_build_matches_oracle(small):
    prep, grid = small
    nn, df = O.grid_build(prep["xyz"], prep["origin"], prep["cell"], prep["dims"])
    assert np.array_equal(grid.nn.cpu().numpy(), nn)
    assert np.array_equal(grid.df.cpu().numpy(), df)
    assert np.array_equal(grid.xyz.cpu().numpy(), prep["xyz"]) and np.array_equal(grid.start.cpu().numpy(), prep["start"])


@pytest.mark.parametrize("steps", [(32, 16, 5, 3, 1), (7, 2, 1)])
def test_build_with_other_step_schedules(lw, small, steps):
    """Jump-flood schedules beyond the default: steps above the LDS slab's 16 (the direct kernel), odd steps
    whose lattices leave partial tiles in every direction; the seeds and the field equal the oracle's."""
    prep, _ = small
    grid = lw.PriorGrid(prep, "cuda:0")
    grid.build(steps=steps)
    nn, df = O.grid_build(prep["xyz"], prep["origin"], prep["cell"], prep["dims"], steps=steps)
    assert np.array_equal(grid.nn.cpu().numpy(), nn)
    assert np.array_equal(grid.df.cpu().numpy(), df)


def test_distances_match_oracle(small):
    prep, grid = small
    rng = np.random.default_rng(8)
    pts = (prep["xyz"][rng.integers(0, grid.n, 20000)] + rng.normal(0, 0.15, (20000, 3))).astype(np.float32)
    pts[:100] += 40.0  # beyond the band and the grid
    a = math.radians(17.0)
    T = np.array([[math.cos(a), -math.sin(a), 0, 0.1], [math.sin(a), math.cos(a), 0, -0.05], [0, 0, 1, 0.02],
                  [0, 0, 0, 1]], np.float32)
    d_pts = torch.from_numpy(pts).cuda()
    out = torch.empty(len(pts), dtype=torch.float32, device="cuda")
    grid.distances(d_pts, len(pts), T, out)
    nn = grid.nn.cpu().numpy()
    want = O.grid_distances(pts, T, prep["origin"], prep["cell"], prep["dims"], prep["start"], nn, prep["xyz"])
    got = out.cpu().numpy()
    assert np.array_equal(got, want) and np.all(got[:100] == 1e3)


def test_scores_match_oracle(lw, small):
    prep, grid = small
    cs = lw.CoarseSearch(grid, "cuda:0")
    rng = np.random.default_rng(9)
    a = rng.uniform(0, 2 * math.pi)
    pts = (prep["xyz"][rng.integers(0, grid.n, 700)] - [1.1, 0.9, 0.6]) @ lw.localize.rotz(a)  # a scan from inside
    heights = [0.45, 0.6]
    pos, shape, dyaw, sc = cs.coarse(pts, heights)
    want = O.score_grid(pts, pos, shape[3], 0.0, dyaw, prep["origin"], prep["cell"], prep["dims"],
                        grid.df.cpu().numpy(), cs.tau_coarse, cs.free_cm)
    assert sc.shape == want.shape and np.array_equal(sc, want)
    assert (sc == -1).any() and (sc > 0).any()
    hyps = np.c_[rng.uniform(0.0, 2.0, (5000, 3)), rng.uniform(-4, 4, 5000)].astype(np.float32)
    got = cs.score_list(pts, hyps, cs.tau_mid)
    assert np.array_equal(got, O.score_list(pts, hyps, prep["origin"], prep["cell"], prep["dims"],
                                            grid.df.cpu().numpy(), cs.tau_mid))


def test_closed_form_on_device(lw, gpu_ok):
    import ctypes

    dims, df = G._hand_field()
    geo = np.array([0, 0, 0, 0.05, 20.0], np.float32)
    pts = torch.tensor([[0.1, 0.0, 0.0], [0.0, 0.2, 0.05], [-0.3, 0.0, 0.0]], device="cuda")
    c = (np.array([20, 20, 5]) + 0.5) * 0.05
    pos = torch.tensor(np.array([c, (np.array([30, 30, 5]) + 0.5) * 0.05, [5.0, 0.5, 0.2]], np.float32), device="cuda")
    d_df = torch.from_numpy(df).cuda()
    out = torch.empty(12, dtype=torch.float32, device="cuda")
    L, A = lw._abi.lib(), lw._abi
    A.check(L.lw_score_grid(pts.data_ptr(), 3, pos.data_ptr(), 3, 4, 0.0, math.pi / 2, A.ptr(geo, A._f32p),
                            A.ptr(dims, A._i32p), d_df.data_ptr(), 25.0, 10.0, None, out.data_ptr(),
                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "lw_score_grid")
    sc = out.cpu().numpy().reshape(3, 4)
    assert np.isclose(sc[0, 0], 1 + (1 - 25 / 625), rtol=1e-6) and np.isclose(sc[0, 1], (1 - 100 / 625) + 1, rtol=1e-6)
    assert sc[0, 2] == 0.0 and np.all(sc[1:] == -1.0)


def _thin(pts, voxel):
    _, first = np.unique(np.floor(pts / voxel).astype(np.int64), axis=0, return_index=True)
    return pts[np.sort(first)]


def test_search_keeps_the_true_pose(lw, gpu_ok):
    """The localize() flow's coarse and medium stages (localize.py:618-666) for a level sensor at the known
    height, on the multi-room floor: scans from seeded poses in three rooms."""
    S = lw.synth.multi_room_floor(31)
    mp = lw.synth.sample_surfaces(S, 0.04, 32, prune=False)
    grid = lw.PriorGrid({"xyz": mp}, "cuda:0")
    grid.build()
    cs = lw.CoarseSearch(grid, "cuda:0")
    rng = np.random.default_rng(33)
    ranks, scans, cands, truth = [], [], [], []
    for q, (x, y, yaw) in enumerate(((6.3, 5.1, 0.7), (30.2, 20.4, 2.9), (53.5, 33.0, -1.8))):
        T = lw.synth.pose_matrix(np.array([x, y, 1.4]), yaw, 0.0, 0.0)
        xyz, _, _ = lw.synth.scan(S, lw.synth.Rosette(40 + q), lambda t, T=T: T, 0.0, 20000, 50 + q)
        pts = _thin(xyz.astype(np.float64), 0.05)
        spread = _thin(pts, 0.15)
        sub_c = spread[rng.choice(len(spread), min(cs.n_coarse_pts, len(spread)), replace=False)]
        sub_m = pts[rng.choice(len(pts), min(cs.n_mid_pts, len(pts)), replace=False)]
        Ts, mid, coarse = cs.search(sub_c, sub_m, [1.4])
        assert len(Ts)
        hit = [i for i, Tc in enumerate(Ts) if np.linalg.norm(Tc[:3, 3] - T[:3, 3]) < 0.15 and
               abs((math.atan2(Tc[1, 0], Tc[0, 0]) - yaw + math.pi) % (2 * math.pi) - math.pi) < math.radians(3)]
        print(f"query {q}: {len(pts)} points, coarse kernel {cs.last_kernel_ms:.2f} ms, true pose at rank "
              f"{hit[0] if hit else None} of {len(Ts)}, best medium score {mid[0]:.3f}")
        ranks.append(hit[0] if hit else None)
        scans.append(xyz)
        cands.append(Ts[:16] if len(Ts) >= 16 else np.concatenate([Ts, np.repeat(Ts[-1:], 16 - len(Ts), 0)]))
        truth.append(T)
    assert all(r is not None for r in ranks)
    # then the batched ICP on the candidates, as the reference's search does: its best score is the truth
    loc = lw.loc.BatchLocalizer(mp, voxel=0.1, reg_slots=1 << 22, device="cuda:0")
    Tf, S, _, _ = (a.cpu().numpy() for a in loc.run(scans, np.stack(cands)))
    for q in range(3):
        Tb = Tf[q, int(np.argmax(S[q]))]
        assert np.linalg.norm(Tb[:3, 3] - truth[q][:3, 3]) < 0.1, q


def test_edge_cases(lw, small):
    """No scan points: free positions score 0, gated ones -1; no hypotheses or distances: nothing happens."""
    prep, grid = small
    cs = lw.CoarseSearch(grid, "cuda:0")
    pos, shape, dyaw, sc = cs.coarse(np.zeros((0, 3)), [0.6])
    want = O.score_grid(np.zeros((0, 3), np.float32), pos, shape[3], 0.0, dyaw, prep["origin"], prep["cell"],
                        prep["dims"], grid.df.cpu().numpy(), cs.tau_coarse, cs.free_cm)
    assert np.array_equal(sc, want) and set(np.unique(sc).tolist()) <= {-1.0, 0.0}
    assert len(cs.score_list(np.zeros((5, 3)), np.zeros((0, 4)), cs.tau_mid)) == 0
    out = torch.full((4,), 7.0, device="cuda")
    grid.distances(torch.zeros((4, 3), device="cuda"), 0, np.eye(4), out)
    assert torch.all(out == 7.0)


@pytest.fixture(scope="module")
def room_grid(lw, gpu_ok):
    S = lw.synth.box_room(1)
    xyz, nrm = lw.synth.sample_surfaces(S, 0.04, 3, normals=True)
    prep = lw.mapgrid.PriorGrid.prepare({"xyz": xyz, "normal": nrm})
    grid = lw.PriorGrid(prep, "cuda:0")
    grid.build()
    return S, prep, grid


def test_plane_icp_and_eval_match_oracle(lw, room_grid):
    """One accumulation for 16 poses (lw_plane_icp) against the oracle's serial sums: counts exact, every
    float64 sum within 1e-9 of the row's scale (the float32 terms are the same; only the order of the float64
    additions differs); the fit counts (lw_plane_eval) exact."""
    S, prep, grid = room_grid
    cs = lw.CoarseSearch(grid, "cuda:0")
    T_true, pts = G._room_scan(lw, S, noise=0.02)
    rng = np.random.default_rng(17)
    Ts = [G._perturbed(lw, T_true, rng.normal(0, 0.05, 3), rng.normal(0, 2.0, 3)) for _ in range(15)] + [T_true]
    n = cs.upload_scan(pts)
    nn = grid.nn.cpu().numpy()
    for dmax in (0.40, 0.05):
        got = cs.accumulate(Ts, n, dmax)
        want = O.plane_icp(pts, Ts, prep, nn, dmax * dmax, (dmax * 0.5) ** 2)
        assert np.array_equal(got[:, 27], want[:, 27]) and np.all(want[:, 27] > 0)
        scale = np.abs(want).max(axis=1, keepdims=True)
        assert np.all(np.abs(got - want) <= 1e-9 * scale)
    fits = cs.evaluate(Ts, n)
    want = O.plane_eval(pts, Ts, prep, nn) / n
    assert np.array_equal(fits, want)

"""lw_plane_icp_noalloc (the allocation-free form of lw_plane_icp: its partial rows in the library's fixed device
array) against lw_plane_icp with a caller workspace, through the C ABI: two calls in a row, the second with fewer
poses and fewer points, each equal bit for bit (the same kernels in the same order); a pose count above the fixed
array's size is refused with LW_ERANGE before anything is launched."""

import ctypes

import numpy as np
import pytest
import torch

import test_oracle_grid as G

pytestmark = pytest.mark.gpu


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _call(lw, grid, pts, Ts, dmax, work=None):
    L = lw._abi.lib()
    k = len(Ts)
    d_pts = torch.from_numpy(np.ascontiguousarray(pts, np.float32)).cuda()
    d_T = torch.from_numpy(np.ascontiguousarray(np.array(Ts, np.float32).reshape(k, 16))).cuda()
    out = torch.full((k, 64), float("nan"), dtype=torch.float64, device="cuda:0")
    args = (_vp(d_pts), len(pts), _vp(d_T), k, *grid._geo_args(), _vp(grid.start), _vp(grid.nn), _vp(grid.xyz),
            _vp(grid.nrm), _vp(grid.planar), float(dmax * dmax), float((dmax * 0.5) ** 2), 0.05)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if work is None:
        rc = L.lw_plane_icp_noalloc(*args, _vp(out), s)
    else:
        rc = L.lw_plane_icp(*args, _vp(work), work.numel(), _vp(out), s)
    torch.cuda.synchronize()
    return rc, out.cpu().numpy()


def test_noalloc_equals_the_workspace_form(lw, gpu_ok):
    S = lw.synth.box_room(1)
    xyz, nrm = lw.synth.sample_surfaces(S, 0.04, 3, normals=True)
    grid = lw.PriorGrid(lw.mapgrid.PriorGrid.prepare({"xyz": xyz, "normal": nrm}), "cuda:0")
    grid.build()
    T_true, pts = G._room_scan(lw, S, n=6000, noise=0.02)
    rng = np.random.default_rng(23)
    L = lw._abi.lib()
    # the first call: 40 poses, the whole scan; the second: 7 poses, a third of it (what the first left in the
    # scratch must not leak into the second's rows)
    for k, n, dmax in ((40, len(pts), 0.40), (7, len(pts) // 3, 0.10)):
        Ts = [G._perturbed(lw, T_true, rng.normal(0, 0.05, 3), rng.normal(0, 2.0, 3)) for _ in range(k - 1)] + [T_true]
        work = torch.empty(int(L.lw_plane_work_bytes(k)), dtype=torch.uint8, device="cuda:0")
        rc_w, want = _call(lw, grid, pts[:n], Ts, dmax, work)
        rc_s, got = _call(lw, grid, pts[:n], Ts, dmax)
        print(f"noalloc: {k} poses x {n} points: rc {rc_s}, matched {got[:, 27].astype(int).tolist()[:4]}..., "
              f"equal {np.array_equal(got, want)}", flush=True)
        assert rc_w == 0 and rc_s == 0
        assert np.all(want[:, 27] > 0) and not np.isnan(got).any()
        assert np.array_equal(got, want)
    # beyond the fixed array: refused, nothing written
    big = [T_true] * (lw._abi.LW_PLANE_NOALLOC_MAX + 1)
    rc, out = _call(lw, grid, pts[:100], big, 0.4)
    assert rc == lw._abi.LW_ERANGE and np.isnan(out).all()
    with pytest.raises(lw._abi.LwError, match="fixed scratch"):
        lw._abi.check(rc, "lw_plane_icp_noalloc")

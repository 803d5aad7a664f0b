"""Throughput of the scan-to-map odometry hot path on one MI355X.

Workload (one "step" = one odometry frame): the plain odometry sequence -- 600 frames (60 s at 10 Hz) of a
seeded handheld walk through a 20 x 15 x 3 m office, 10,000 rose-curve rosette points per frame with
per-point timestamps (deskewed against the constant-velocity model), 0.2 m registration voxels, a 2^22-slot
map, up to 20 Gauss-Newton iterations per registration. All frames are generated on the host and uploaded
to HBM before the timed region; the timed region pushes frames from HBM through Odometry.push exactly as a
caller would (frame cutting, deskew, thinning, registration, keyframe inserts, one readback per
registration).

Printed: one JSON line with frames/s (value), points/s, drift against the ground truth, the roofline of the
dominant kernel (the persistent registration kernel: algorithmic bytes from its counters, duration from HIP
events on its stream), the CPU baseline (the oracle's serial restatement on a bounded sample of the same
sequence, one host thread) and, unless switched off, legs for the integration map (--no-imap), the prior-map
grid and coarse search (--no-grid), LVXR replay (--no-lvxr), the odometry worker (--no-worker), the batched global localisation (--no-loc) and the learned prior
(--no-prior), each with its own roofline and CPU baseline.
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "ICP frames/s @10k pts/frame, % HBM roofline; global-loc hypotheses/s"
SEED = 21
LOC_SEED = 31
LOC_VOXEL = 0.1


def accumulation_bytes(prof) -> float:
    """Algorithmic bytes of accumulation passes: per point-iteration the point (12 B) and one 8 B key per
    neighbour looked up, 16 B per occupied neighbour hit (mean + count) and 36 B per matched point (the chosen
    voxel's second moments)."""
    pt_iters, lookups, hits, matched = prof[2], prof[3], prof[4], prof[5]
    return 12.0 * pt_iters + 8.0 * lookups + 16.0 * hits + 36.0 * matched


def measured_traffic(kernel: str):
    """HBM bytes per launch of a kernel, from the committed hardware-counter summary (None if not profiled)."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            e = json.load(f).get(kernel)
        return None if e is None else float(e["fetch_bytes"] + e["write_bytes"])
    except (OSError, ValueError, KeyError):
        return None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def drift(seq, traj):
    T0 = seq.pose(traj[0][1])
    err = []
    dist = 0.0
    for (k, tm, T), nxt in zip(traj, traj[1:] + [None]):
        gt = np.linalg.inv(T0) @ seq.pose(tm)
        est = np.linalg.inv(traj[0][2]) @ T
        err.append(float(np.linalg.norm(est[:3, 3] - gt[:3, 3])))
        if nxt is not None:
            dist += float(np.linalg.norm(seq.pose(nxt[1])[:3, 3] - seq.pose(tm)[:3, 3]))
    err = np.array(err)
    return {"rmse_m": float(np.sqrt(np.mean(err ** 2))), "final_m": float(err[-1]), "max_m": float(err.max()),
            "distance_m": dist}


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    return oracle


def cpu_baseline(frames, seq, warmup: int, n_frames: int):
    """The oracle (serial C restatement of the kernels + the reference's numpy host step) on n_frames of the same
    sequence after the same untimed warm-up, one host thread."""
    oracle = _oracle()
    lw = importlib.import_module("livox-warp_amd")
    od = oracle.OracleOdometry(lw.PoseTable(), reg_slots=1 << 22, frame_cap=1 << 15)
    od.reg_voxel = 0.2
    od.max_iter = 20
    od.reset(seq.pose(0.05))
    for x, a, t in frames[:warmup]:
        od.push(x, a, t, len(t), float(t[0]), float(t[-1]))
    f0 = len(od.traj)
    t0 = time.perf_counter()
    for x, a, t in frames[warmup:warmup + n_frames]:
        od.push(x, a, t, len(t), float(t[0]), float(t[-1]))
    el = time.perf_counter() - t0
    done = len(od.traj) - f0
    return {"value": round(done / el, 3), "unit": "frames/s", "cores": 1, "kind": "port",
            "sample": f"{done} frames of the same sequence after the same {warmup}-frame warm-up; oracle serial C "
                      f"kernels + the reference's numpy host step (reference-semantics CPU restatement)",
            "seconds": round(el, 3), "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}


def _pose_err(A, B):
    d = np.linalg.inv(A) @ B
    ang = float(np.degrees(np.arccos(np.clip((np.trace(d[:3, :3]) - 1) / 2, -1, 1))))
    return float(np.linalg.norm(d[:3, 3])), ang


def localisation(args, torch, lw, dev):
    """Batched global localisation: the 60 x 40 x 4 m floor at 20 mm (about 20M points) voxelised at 0.1 m,
    args.loc_queries scans of 5k-30k points, args.loc_hyps hypotheses each (xy uniform over the floor, yaw
    uniform, z / roll / pitch at truth), 10 Gauss-Newton iterations per hypothesis, all in one launch."""
    t_gen = time.perf_counter()
    W = lw.synth.loc_workload(seed=LOC_SEED, n_queries=args.loc_queries)
    rng = np.random.default_rng(LOC_SEED + 7)
    H = args.loc_hyps
    hyps = np.stack([lw.loc.hypotheses(rng, H, W.T_true[q], (0.5, 0.5), (59.5, 39.5))
                     for q in range(len(W.scans))])
    t_gen = time.perf_counter() - t_gen
    X = torch.from_numpy(W.map_xyz).to(dev)
    torch.cuda.synchronize()
    loc = lw.loc.BatchLocalizer(X, voxel=LOC_VOXEL, reg_slots=1 << args.loc_slots, chunk=1 << 22, device=dev)
    # map build, timed on its own: clear, then the bulk insert of every map point (in index order)
    loc.eng.sync()
    t0 = time.perf_counter()
    loc.eng.reset()
    with torch.cuda.stream(loc.stream):
        loc.eng.insert(X, None, np.eye(4), LOC_VOXEL, max_range=1e30)
    loc.eng.sync()
    t_map = time.perf_counter() - t0
    n_map = int(X.shape[0])
    V = loc.map_stats()[0]
    map_bytes = 16.0 * n_map + 60.0 * V  # xyz + attr per point, one record written per new voxel
    # tile autotuner on a slice of the workload, then the timed launches (graph-replayed)
    tile, table, spilled = loc.autotune([W.scans[0]], hyps[:1, :256])
    staged = loc._stage(W.scans, hyps)
    L = loc.eng.lib
    ctx = loc.eng.ctx
    loc.run(W.scans, hyps, tile=tile, use_graph=True, staged=staged)
    loc.stream.synchronize()
    lw._abi.check(L.lw_prof_enable(ctx, 1), "lw_prof_enable")
    lw._abi.check(L.lw_prof_reset(ctx), "lw_prof_reset")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.loc_reps):
        out = loc.run(W.scans, hyps, tile=tile, use_graph=True, staged=staged)
    loc.stream.synchronize()
    el = time.perf_counter() - t0
    prof = np.zeros(8)
    lw._abi.check(L.lw_prof_read(ctx, lw._abi.ptr(prof, lw._abi._f64p)), "lw_prof_read")
    lw._abi.check(L.lw_prof_enable(ctx, 0), "lw_prof_enable")
    n_units = len(W.scans) * H
    T, S, C, I = (x.cpu().numpy() for x in out)
    best = [_pose_err(W.T_true[q], T[q, int(np.argmax(S[q]))]) for q in range(len(W.scans))]
    closest = [min(_pose_err(W.T_true[q], T[q, h])[0] for h in range(H)) for q in range(len(W.scans))]
    launch_s = prof[1] * 1e-3 / max(prof[0], 1.0)
    bytes_per_launch = accumulation_bytes(prof) / max(prof[0], 1.0)
    achieved = bytes_per_launch / launch_s / 1e9 if launch_s > 0 else 0.0
    res = {
        "value": round(n_units * args.loc_reps / el, 1), "unit": "hypotheses/s",
        "workload": f"{len(W.scans)} queries x {H} hypotheses, scans {[len(s) for s in W.scans]} points, map "
                    f"{n_map} points at 20 mm -> {V} voxels of 0.1 m in 2^{args.loc_slots} slots, 10 GN iterations at reach 1",
        "seed": LOC_SEED, "tile": {"block": tile[0], "workgroups_per_cu": tile[1], "chunk": tile[2]}, "autotune_ms": table,
        "autotune_skipped_spilling": [f"{b}x{g}x{c}" for b, g, c in spilled],
        "launch_ms": round(launch_s * 1e3, 3), "graph": True,
        "top1_recovered": sum(1 for e in best if e[0] < 0.3 and e[1] < 3.0), "queries": len(W.scans),
        "top1_error": [[round(e[0], 3), round(e[1], 2)] for e in best],
        "closest_final_m": [round(c, 3) for c in closest],
        "mean_iters": round(float(I.mean()), 2), "point_iterations": int(prof[2]),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic("k_loc"),
                     "kernel": f"k_loc<{tile[0]}, {tile[2]}>",
                     "bytes_per_launch": round(bytes_per_launch, 1)},
        "map_build": {"points": n_map, "voxels": V, "seconds": round(t_map, 4),
                      "points_per_s": round(n_map / t_map, 1), "algorithmic_GBps": round(map_bytes / t_map / 1e9, 2)},
        "gen_seconds": round(t_gen, 2),
    }
    res["bulk_voxelisation"] = bulk_voxelisation(torch, lw, dev, X)
    try:
        res["global_search"] = global_search(torch, lw, dev, W, loc)
    finally:
        loc.close()  # the map's context and its captured graph, before the interpreter's teardown
    if args.cpu_frames > 0:
        res["cpu_baseline"] = loc_cpu_baseline(W, hyps, args.loc_cpu_hyps)
    return res


def bulk_voxelisation(torch, lw, dev, X, voxel: float = 0.05, slots_log2: int = 25, reps: int = 3):
    """prior_map.voxel_downsample's device work on the floor's map points (already on the device): the insert
    into a cleared 2^25-slot table and the slot-order compaction, timed with HIP events on the table's stream;
    algorithmic bytes: 12 per point read, 40 per voxel record (key, count, sums) and 32 per voxel emitted."""
    st = torch.cuda.Stream(device=dev)
    st.wait_stream(torch.cuda.current_stream(dev))
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ins, emi = [], []
    with lw.VoxTable(slots_log2, dev, st) as t:
        for _ in range(reps):
            t.clear()
            e[0].record(st)
            t.insert(X, voxel)
            e[1].record(st)
            out = t.emit(voxel)
            e[2].record(st)
            st.synchronize()
            ins.append(e[0].elapsed_time(e[1]) * 1e-3)
            emi.append(e[1].elapsed_time(e[2]) * 1e-3)
        V, dropped = t.stats()
        del out
    n = int(X.shape[0])
    t_ins, t_emi = min(ins), min(emi)
    achieved = (12.0 * n + 40.0 * V) / t_ins / 1e9
    return {"points": n, "voxel_m": voxel, "slots": 1 << slots_log2, "voxels": V, "dropped": dropped,
            "insert_ms": round(t_ins * 1e3, 3), "emit_ms": round(t_emi * 1e3, 3),
            "points_per_s": round(n / t_ins, 1),
            "emit_algorithmic_GBps": round(((1 << slots_log2) / 8.0 + 40.0 * V + 32.0 * V) / t_emi / 1e9, 2),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic("k_vox_insert"),
                         "kernel": "k_vox_insert (one launch over every map point)",
                         "bytes_per_launch": 12.0 * n + 40.0 * V}}


def _thin(pts: np.ndarray, voxel: float) -> np.ndarray:
    """The first point of every voxel, in input order."""
    _, first = np.unique(np.floor(pts / voxel).astype(np.int64), axis=0, return_index=True)
    return pts[np.sort(first)]


def global_search(torch, lw, dev, W, loc):
    """The reference's global search order (localize.py:603-700) on the same queries: the distance-field
    scoring first (a PriorGrid of the same floor at 4 cm with its face normals, the coarse and medium stages at
    the known sensor height of a level scan), then each query's 16 candidates refined two ways: the batched
    ICP (top-1 = its best score) and the reference's own point-to-plane ICP on the grid (top-1 = the best
    3 cm fit)."""
    t0 = time.perf_counter()
    mxyz, mnrm = lw.synth.sample_surfaces(W.scene, GRID_SPACING, LOC_SEED + 1, prune=False, normals=True)
    grid = lw.PriorGrid({"xyz": mxyz, "normal": mnrm}, dev)
    grid.build()
    cs = lw.CoarseSearch(grid, dev)
    t_grid = time.perf_counter() - t0
    rng = np.random.default_rng(LOC_SEED + 11)
    cands, t_search, first, ref_best, t_ref = [], 0.0, [], [], 0.0
    for q, scan in enumerate(W.scans):
        pts = _thin(np.asarray(scan, np.float64), 0.05)
        spread = _thin(pts, 0.15)
        sub_c = spread[rng.choice(len(spread), min(cs.n_coarse_pts, len(spread)), replace=False)]
        sub_m = pts[rng.choice(len(pts), min(cs.n_mid_pts, len(pts)), replace=False)]
        t1 = time.perf_counter()
        Ts, _, _ = cs.search(sub_c, sub_m, [1.4])
        t_search += time.perf_counter() - t1
        Ts = np.concatenate([Ts, np.repeat(Ts[-1:], max(0, cs.n_icp - len(Ts)), 0)])[:cs.n_icp]
        first.append(_pose_err(W.T_true[q], Ts[0]))
        cands.append(Ts)
        # the reference's own refinement of the candidates: point-to-plane ICP on the grid, then the 3 cm fit
        t1 = time.perf_counter()
        n = cs.upload_scan(pts)
        Tr = cs.icp(list(Ts), n)
        fits = cs.evaluate(Tr, n)
        t_ref += time.perf_counter() - t1
        ref_best.append(_pose_err(W.T_true[q], Tr[int(np.argmax(fits[:, 0]))]))
    hyps = np.stack(cands)
    loc.stream.synchronize()
    t1 = time.perf_counter()
    out = loc.run(W.scans, hyps)
    loc.stream.synchronize()
    t_icp = time.perf_counter() - t1
    T, S = out[0].cpu().numpy(), out[1].cpu().numpy()
    best = [_pose_err(W.T_true[q], T[q, int(np.argmax(S[q]))]) for q in range(len(W.scans))]
    return {"grid_build_s": round(t_grid, 3), "search_s_per_query": round(t_search / len(W.scans), 4),
            "icp_s": round(t_icp, 4), "candidates": int(hyps.shape[1]),
            "coarse_first_error": [[round(e[0], 3), round(e[1], 2)] for e in first],
            "top1_recovered": sum(1 for e in best if e[0] < 0.3 and e[1] < 3.0), "queries": len(W.scans),
            "top1_error": [[round(e[0], 3), round(e[1], 2)] for e in best],
            "plane_icp": {"s_per_query": round(t_ref / len(W.scans), 4),
                          "top1_recovered": sum(1 for e in ref_best if e[0] < 0.3 and e[1] < 3.0),
                          "top1_error": [[round(e[0], 3), round(e[1], 2)] for e in ref_best]}}


def loc_cpu_baseline(W, hyps, n: int):
    """The oracle's loc_register (serial) on n hypotheses spread over the queries, extrapolated linearly
    (hypotheses are independent)."""
    oracle = _oracle()
    om = oracle.OracleMap(1 << 22)
    om.insert(W.map_xyz, np.zeros(len(W.map_xyz), np.uint32), np.eye(4), LOC_VOXEL, max_range=1e30)
    Q = len(W.scans)
    t0 = time.perf_counter()
    for j in range(n):
        q = j % Q
        oracle.loc_register(om, W.scans[q], hyps[q, j // Q], LOC_VOXEL, iters=10)
    el = time.perf_counter() - t0
    return {"value": round(n / el, 3), "unit": "hypotheses/s", "cores": 1, "kind": "port",
            "sample": f"{n} hypotheses ({n // Q} per query) through the oracle's serial loc_register, "
                      f"extrapolated linearly", "seconds": round(el, 3)}


MLP_FLOPS_PER_ROW = 2.0 * (152 * 256 + 256 * 256 + 256 * 256 + 256 * 6)  # forward multiply-adds x 2


IMAP_VOXEL = 0.05  # the integration map's voxel in the reference's odometry checks (check_odometry.py:27)
IMAP_IN_BYTES = 24  # per point: xyz 12, attribute 4, time 4, frame id 4
IMAP_REC_BYTES = 64  # one record read and written per kept point


def integration_map(args, torch, lw, dev, X, A, Tt, seq, n_frames: int, per: int):
    """The consumer of the frames (lw.IntegrationMap: the reference Pipeline's map_insert and carve): every
    generated frame inserted through its ground-truth pose into 0.05 m voxels, then carved (every 2nd return),
    each pass timed with events on the map's stream; the oracle's serial insert and carve on a sample beside
    it."""
    fdt = 0.1
    F = torch.floor(Tt / fdt).to(torch.int32)  # frame ids, as Pipeline.stage computes them
    ids = np.unique(F.cpu().numpy())
    pt = lw.PoseTable()
    poses = {}
    for k in ids:
        poses[int(k)] = seq.pose((int(k) + 0.5) * fdt).astype(np.float32)
        pt.set_pose(int(k), poses[int(k)])
    mp = lw.IntegrationMap(pt, dev, map_capacity=1 << 23)
    view = lw.imap.View(min_range=0.3, max_range=60.0)
    sx = [X[k].contiguous() for k in range(n_frames)]
    sa, st, sf = [A[k] for k in range(n_frames)], [Tt[k] for k in range(n_frames)], [F[k].contiguous() for k in range(n_frames)]
    mp.map_insert(sx[0], sa[0], st[0], sf[0], per, view, IMAP_VOXEL)  # warm-up (and the pose upload)
    mp.carve(sx[0], sa[0], sf[0], per, IMAP_VOXEL, every=2)
    mp.clear_map()
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    e[0].record(mp.stream)
    for k in range(n_frames):
        mp.map_insert(sx[k], sa[k], st[k], sf[k], per, view, IMAP_VOXEL)
    e[1].record(mp.stream)
    for k in range(n_frames):
        mp.carve(sx[k], sa[k], sf[k], per, IMAP_VOXEL, every=2)
    e[2].record(mp.stream)
    torch.cuda.synchronize()
    ins_ms, carve_ms = e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])
    occ, dropped = mp.map_stats()
    r = mp.records()
    o = r["key"] >= 0
    points = n_frames * per
    ins_gbs = points * (IMAP_IN_BYTES + IMAP_REC_BYTES) / (ins_ms * 1e-3) / 1e9
    # the oracle on a sample: the first frames, insert then carve
    oracle = _oracle()
    om = oracle.IntegrationMap(1 << 23)
    P = oracle.identity_poses()
    for k, T in poses.items():
        P[k & 0xFFFF] = T
    xs, as_, ts, fs = X.cpu().numpy(), A.cpu().numpy().view(np.uint32), Tt.cpu().numpy(), F.cpu().numpy()
    t0 = time.perf_counter()
    done = 0
    while done < n_frames and time.perf_counter() - t0 < 10.0:
        oracle.map_insert(om, xs[done], as_[done], ts[done], fs[done], P, IMAP_VOXEL, min_range=0.3, max_range=60.0)
        oracle.carve(om, xs[done], as_[done], fs[done], P, IMAP_VOXEL, every=2)
        done += 1
    cpu_s = time.perf_counter() - t0
    return {
        "workload": f"{n_frames} frames x {per} points of the odometry sequence through their ground-truth poses, "
                    f"{IMAP_VOXEL} m voxels in 2^23 slots, then carving every 2nd return",
        "insert_points_per_s": round(points / (ins_ms * 1e-3), 1),
        "carve_rays_per_s": round(points / 2 / (carve_ms * 1e-3), 1),
        "insert_ms_per_frame": round(ins_ms / n_frames, 4), "carve_ms_per_frame": round(carve_ms / n_frames, 4),
        "voxels": occ, "dropped": dropped, "carved_voxels": int((r["miss"][o] > 2 * r["cnt"][o]).sum()),
        "roofline": {"bound": "hbm", "achieved": round(ins_gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ins_gbs / HBM_PEAK_GBS, 6), "traffic": measured_traffic("k_imap_insert"),
                     "kernel": "k_imap_insert",
                     "bytes_per_point": IMAP_IN_BYTES + IMAP_REC_BYTES},
        "cpu_baseline": {"value": round(done / cpu_s, 3), "unit": "frames/s (insert + carve)", "cores": 1,
                         "kind": "port", "sample": f"the first {done} frames through the oracle's serial map insert "
                                                   f"and carve", "seconds": round(cpu_s, 3)},
        "frames_per_s": round(n_frames / ((ins_ms + carve_ms) * 1e-3), 1),
    }


GRID_SPACING = 0.04  # the prior map: the multi-room floor's surfaces every 4 cm (5.1M points)
GRID_HEIGHTS = tuple(round(1.0 + 0.1 * k, 2) for k in range(8))  # sensor heights tried (n_heights = 8)
GRID_QUERY = (30.2, 20.4, 2.9)  # the scan's true x, y (m) and heading (rad); sensor 1.4 m up, level
# VALU lane-instructions per (hypothesis, point) lookup in k_score_grid's inner loop: 411 for the 16 lookups
# of one unrolled iteration (4 positions x 4 headings) in the gfx950 code this repository builds
GRID_VALU_PER_LOOKUP = 411 / 16
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # 256 CUs x 64 lanes x 2.4 GHz, one VALU lane-op a cycle each


def prior_grid(args, torch, lw, dev):
    """The prior map's grid (lw.PriorGrid, the reference's mapgrid.py) built from the multi-room floor, and
    the coarse search over it (lw.CoarseSearch, localize.py's k_score_grid, local maxima, the medium grid and
    suppression) for one level scan at 8 sensor heights; the oracle's serial k_score_grid on a sample of the
    same hypotheses beside it (and compared with the device's scores for them)."""
    import math

    S = lw.synth.multi_room_floor(31)
    mp = lw.synth.sample_surfaces(S, GRID_SPACING, 32, prune=False)
    t0 = time.perf_counter()
    prep = lw.PriorGrid.prepare({"xyz": mp})
    prep_s = time.perf_counter() - t0
    grid = lw.PriorGrid(prep, dev)
    grid.build()  # first build: module load and allocator warm-up
    grid.build()
    x, y, yaw = GRID_QUERY
    T = lw.synth.pose_matrix(np.array([x, y, 1.4]), yaw, 0.0, 0.0)
    xyz, _, _ = lw.synth.scan(S, lw.synth.Rosette(41), lambda t: T, 0.0, 20000, 51)
    pts = xyz.astype(np.float64)
    _, first = np.unique(np.floor(pts / 0.05).astype(np.int64), axis=0, return_index=True)
    pts = pts[np.sort(first)]
    _, first = np.unique(np.floor(pts / 0.15).astype(np.int64), axis=0, return_index=True)
    spread = pts[np.sort(first)]
    rng = np.random.default_rng(0)
    cs = lw.CoarseSearch(grid, dev)
    sub_c = spread[rng.choice(len(spread), min(cs.n_coarse_pts, len(spread)), replace=False)]
    sub_m = pts[rng.choice(len(pts), min(cs.n_mid_pts, len(pts)), replace=False)]
    cs.coarse(sub_c, GRID_HEIGHTS[:1])  # warm-up
    t0 = time.perf_counter()
    pos, shape, dyaw, sc = cs.coarse(sub_c, GRID_HEIGHTS)
    coarse_s = time.perf_counter() - t0
    kernel_ms = cs.last_kernel_ms
    hyps = int(np.prod(shape))
    active = int((sc >= 0).sum())  # gated positions (inside a surface, outside the grid) score -1 unscored
    lookups = active * len(sub_c)
    t0 = time.perf_counter()
    Ts, mid, _ = cs.search(sub_c, sub_m, GRID_HEIGHTS)
    search_s = time.perf_counter() - t0
    err = [(float(np.linalg.norm(Tc[:3, 3] - T[:3, 3])),
            abs(math.degrees((math.atan2(Tc[1, 0], Tc[0, 0]) - yaw + math.pi) % (2 * math.pi) - math.pi))) for Tc in Ts]
    rank = next((i for i, (dt, da) in enumerate(err) if dt < 0.15 and da < 3.0), None)
    # the oracle on a sample: whole positions (every heading) from the middle of the list, about 10 s
    oracle = _oracle()
    n_yaw = shape[3]
    mid_pos = len(pos) // 2
    df_host = grid.df.cpu().numpy()
    done, exact, cpu_active = 0, True, 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 10.0 and mid_pos + done < len(pos):
        a, b = mid_pos + done, min(len(pos), mid_pos + done + 64)
        want = oracle.score_grid(sub_c, pos[a:b], n_yaw, 0.0, dyaw, prep["origin"], prep["cell"], prep["dims"],
                                 df_host, cs.tau_coarse, cs.free_cm)
        exact &= bool(np.array_equal(want, sc[a * n_yaw:b * n_yaw]))
        cpu_active += int((want >= 0).sum())
        done = b - mid_pos
    cpu_s = time.perf_counter() - t0
    cpu_lookups = cpu_active * len(sub_c)
    rate = lookups / (kernel_ms * 1e-3)
    valu = rate * GRID_VALU_PER_LOOKUP / 1e12
    return {
        "workload": f"multi-room floor 60 x 40 m at {GRID_SPACING} m ({grid.n} map points, {grid.ncell} cells of 5 cm), "
                    f"one level scan ({len(pts)} points after 5 cm thinning, {len(sub_c)} coarse, {len(sub_m)} medium): "
                    f"{shape[1]} x {shape[2]} positions at 0.25 m x {len(GRID_HEIGHTS)} heights x {n_yaw} headings",
        "prepare_s": round(prep_s, 3), "build_s": round(grid.build_s, 3),
        "build_kernels_ms": round(grid.build_kernel_ms, 3),
        "hypotheses": hyps, "scored_hypotheses": active, "coarse_kernel_ms": round(kernel_ms, 3),
        "coarse_s": round(coarse_s, 3),
        "lookups_per_s": round(rate, 1), "hypotheses_per_s": round(hyps / (kernel_ms * 1e-3), 1),
        "search_s": round(search_s, 3), "true_pose_rank": rank, "candidates": len(Ts),
        "best_medium_score": round(float(mid[0]), 4) if len(mid) else None,
        "roofline": {"bound": "valu", "achieved": round(valu, 3), "peak": round(VALU_PEAK_TOPS, 2),
                     "unit": "T lane-instr/s", "frac": round(valu / VALU_PEAK_TOPS, 4),
                     "traffic": measured_traffic("k_score_grid"),
                     "kernel": "k_score_grid", "valu_per_lookup": round(GRID_VALU_PER_LOOKUP, 3),
                     "df_bytes_per_lookup": 1},
        "cpu_baseline": {"value": round(cpu_lookups / max(cpu_s, 1e-9), 1), "unit": "lookups/s", "cores": 1,
                         "kind": "port", "sample": f"{done} positions x {n_yaw} headings x {len(sub_c)} "
                                                   f"points through the oracle's serial k_score_grid "
                                                   f"({cpu_active} hypotheses scored, the rest gated)",
                         "seconds": round(cpu_s, 3), "scores_identical": exact},
    }


LVXR_PER_PACKET = 96  # extended-cartesian records a packet (data type 2, 14 bytes each)
LVXR_OUT_BYTES = 8 + 12 + 1 + 1 + 1 + 12  # per firing written: time, xyz, reflectivity, tag, valid, direction


def _write_lvxr(path, xyz, attr, t):
    """The points as an LVXR2 recording (dev type 1, standard profile) of extended-cartesian packets: millimetre
    coordinates, reflectivity and tag, 96 records a packet stamped (ns) with its first record's time."""
    n = len(t) // LVXR_PER_PACKET * LVXR_PER_PACKET
    rec = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("refl", "u1"), ("tag", "u1")])
    pk = np.dtype([("host", "<u8"), ("len", "<u2"), ("head", "u1", (10,)), ("ts", "<u8"),
                   ("rec", rec, (LVXR_PER_PACKET,))])
    a = np.zeros(n // LVXR_PER_PACKET, pk)
    ts = np.round(t[:n:LVXR_PER_PACKET].astype(np.float64) * 1e9).astype(np.uint64)
    a["host"], a["ts"], a["len"] = ts, ts, 18 + 14 * LVXR_PER_PACKET
    a["head"][:, 0], a["head"][:, 2], a["head"][:, 9] = 5, 1, 2  # version, device id, data type
    mm = np.round(xyz[:n].astype(np.float64) * 1000).astype("<i4").reshape(-1, LVXR_PER_PACKET, 3)
    a["rec"]["x"], a["rec"]["y"], a["rec"]["z"] = mm[..., 0], mm[..., 1], mm[..., 2]
    a["rec"]["refl"] = (attr[:n] & 255).reshape(-1, LVXR_PER_PACKET)
    a["rec"]["tag"] = ((attr[:n] >> 8) & 255).reshape(-1, LVXR_PER_PACKET)
    with open(path, "wb") as f:
        f.write(b"LVXR2\n" + bytes([1, 0]) + a.tobytes())
    return n


def _scaled_traffic(kernel: str, units: int):
    """HBM bytes of a launch of `units` work items, from a profiled launch of another size whose traffic is
    linear in them (each item read and written once): the profile's bytes per item times units."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            e = json.load(f).get(kernel)
        if e is None or not e.get("units"):
            return None
        return round(float(e["fetch_bytes"] + e["write_bytes"]) / e["units"] * units, 1)
    except (OSError, ValueError, KeyError):
        return None


def lvxr_replay(args, torch, lw, dev, frames):
    """LVXR replay (lw.lvxr.read_firings, the reference's lvxr.read_firings): the odometry sequence's points
    written as a recording, read back with the decode on the device (kernel timed with HIP events), and the
    oracle's serial reader on the first packets beside it, compared firing for firing."""
    import shutil
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lvxr_oracle

    xyz = np.concatenate([f[0] for f in frames])
    attr = np.concatenate([f[1] for f in frames]).astype(np.uint32)
    t = np.concatenate([f[2] for f in frames])
    tmp = tempfile.mkdtemp()
    try:
        path = os.path.join(tmp, "walk.lvxr")
        n = _write_lvxr(path, xyz, attr, t)
        lw.lvxr.read_firings(path, dev)  # warm-up
        tm = {}
        t0 = time.perf_counter()
        f = lw.lvxr.read_firings(path, dev, timing=tm)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # the oracle on a sample: the first packets, about 10 s of serial Python
        sample = os.path.join(tmp, "sample.lvxr")
        k = 64
        while True:
            m = _write_lvxr(sample, xyz[:k * LVXR_PER_PACKET], attr[:k * LVXR_PER_PACKET], t[:k * LVXR_PER_PACKET])
            data = open(sample, "rb").read()
            t1 = time.perf_counter()
            ref = lvxr_oracle.read(data)
            spent = time.perf_counter() - t1
            if spent > 3.0 or k * 2 * LVXR_PER_PACKET > n:
                break
            k *= 2
        fs = lw.lvxr.read_firings(sample, dev)
        same = bool(np.array_equal(fs.t.cpu().numpy(), ref[2]) and np.array_equal(fs.xyz.cpu().numpy(), ref[3])
                    and np.array_equal(fs.valid.cpu().numpy(), ref[6]))
    finally:
        shutil.rmtree(tmp)
    dec_s = tm["decode_ms"] * 1e-3
    gbs = tm["firings"] * (14 + LVXR_OUT_BYTES) / dec_s / 1e9
    return {
        "workload": f"{len(frames)} office frames ({n} points) as an LVXR2 recording of {tm['packets']} "
                    f"extended-cartesian packets ({tm['file_bytes'] / 1e6:.1f} MB)",
        "firings_per_s": round(tm["firings"] / dec_s, 1), "decode_ms": round(tm["decode_ms"], 4),
        "host_framing_s": round(tm["host_s"], 4), "read_firings_s": round(wall, 4),
        "firings_per_s_end_to_end": round(tm["firings"] / wall, 1),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 6), "traffic": _scaled_traffic("k_lvxr_decode", tm["firings"]),
                     "kernel": "k_lvxr_decode",
                     "bytes_per_firing": 14 + LVXR_OUT_BYTES},
        "cpu_baseline": {"value": round(m / spent, 1), "unit": "firings/s", "cores": 1, "kind": "port",
                         "sample": f"the first {k} packets ({m} firings) through the oracle's serial reader",
                         "seconds": round(spent, 3), "identical": same},
    }


def odom_worker(args, torch, lw, dev, X, A, Tt, spans, seq, n_frames: int, per: int):
    """The threaded path (lw.OdomWorker, the reference's slam_worker.OdomWorker): the same frames submitted
    one 10k-point batch at a time from the caller's stream (copied into staging slots in, each deskewed frame
    copied into an output slot on the odometry's stream out), polled and released as they come; frames/s from the first submit until the worker is idle and
    the last frame is polled, against the direct Odometry's rate on the same frames."""
    w = lw.worker.OdomWorker(lw.PoseTable(), dev, max_batches=n_frames + 1, out_slots=64, frame_cap=1 << 15,
                             reg_slots=1 << 22)
    try:
        w.configure(reg_voxel=0.2, max_iter=20)
        w.reset(seq.pose(0.05))
        torch.cuda.synchronize()
        got = 0
        t0 = time.perf_counter()
        for k in range(n_frames):
            w.submit(X[k], A[k], Tt[k], per, spans[k][0], spans[k][1])
            got += len(w.poll())
            w.release_consumed()
        deadline = time.perf_counter() + 120.0
        while (not w.idle() or w.status()["processed"] < n_frames) and time.perf_counter() < deadline:
            got += len(w.poll())
            w.release_consumed()
            time.sleep(0.001)  # a caller polling at 1 kHz
        time.sleep(0.01)
        got += len(w.poll())
        torch.cuda.current_stream().synchronize()
        wall = time.perf_counter() - t0
        w.release_consumed()
        st = w.status()
    finally:
        w.close()
    return {"workload": f"{n_frames} office frames x {per} points, one batch a frame, through the worker (an input "
                        f"stream and the odometry's stream)", "frames": got, "frames_per_s": round(got / wall, 1), "drops": st["drops"],
            "busy_ms": st["busy_ms"]}


def learned_prior(args, torch, lw, dev):
    """The learned pose-increment prior (shipped weights, trained on 200 x 600 frames of the plain GPU
    odometry) against the constant-velocity guess on held-out office sequences, and the training step's speed
    on the matrix cores."""
    params, norm = lw.prior.load(os.path.join(ROOT, "livox-warp_amd", "prior_weights.npz"))
    rows = lw.prior.evaluate(lw, params, norm, range(5000, 5000 + args.prior_seqs), args.prior_frames, dev)
    res = lw.prior.summarise(rows)
    # the learned prior's targets: odometry frames/s at least 1.3x the constant-velocity guess's, and at least
    # 30% fewer Gauss-Newton steps
    res["targets"] = {"speedup": 1.3, "gn_reduction": 0.30, "speedup_met": bool(res["speedup"] >= 1.3),
                      "gn_reduction_met": bool(res["gn_reduction"] >= 0.30)}
    res["workload"] = (f"{args.prior_seqs} held-out office walks (seeds 5000..) x {args.prior_frames} frames, "
                       f"10k points/frame, both guesses on the same frames")
    res["training_step"] = training_step(torch, lw, dev, 4096)
    res["training_step_16k"] = training_step(torch, lw, dev, 16384)
    return res


def training_step(torch, lw, dev, B: int):
    """The prior's training step at batch B (forward, loss, backward with 2x the forward's products, Adam),
    replayed from its captured graph, HIP events over 50 steps; the same steps launch by launch beside it."""
    net = lw.prior.MLP(batch_cap=B, device=dev, stream=torch.cuda.Stream(device=dev))
    net.set_params(lw.prior.init_params(np.random.default_rng(0)))
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    X = torch.randn((B, 152), generator=g, device=dev)
    Y = torch.randn((B, 6), generator=g, device=dev)
    net.stream.wait_stream(torch.cuda.current_stream(dev))
    for _ in range(3):
        net.train_step(X, Y, 1e-4)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    steps = 50
    e0.record(net.stream)
    for _ in range(steps):
        net.train_step(X, Y, 1e-4)
    e1.record(net.stream)
    torch.cuda.synchronize()
    step_s = e0.elapsed_time(e1) * 1e-3 / steps
    e0.record(net.stream)
    for _ in range(steps):  # the same step launch by launch, for comparison
        net.train_step(X, Y, 1e-4, graph=False)
    e1.record(net.stream)
    torch.cuda.synchronize()
    eager_s = e0.elapsed_time(e1) * 1e-3 / steps
    net.close()
    tflops = 3.0 * MLP_FLOPS_PER_ROW * B / step_s / 1e12
    res = {"batch": B, "ms": round(step_s * 1e3, 4), "samples_per_s": round(B / step_s, 1),
           "eager_ms": round(eager_s * 1e3, 4),
           "roofline": {"bound": "mfma", "achieved": round(tflops, 3), "peak": 2500.0, "unit": "TFLOP/s",
                        "frac": round(tflops / 2500.0, 6),
                        "kernel": "k_mlp_fwd + k_gemm_pair + k_gemm + k_grad_sum (bf16 16x16x32 MFMA, fp32 "
                                  "accumulate), the step replayed from its captured hipGraph"}}
    return res


def multi_sequence(args, torch, lw, dev, seq, single_fps: float):
    """Independent sequences stepped together (BatchOdometry): args.batch_seqs maps, each fed one of
    args.batch_variants noise realisations of the same walk (frames_device seeds 100..), 30 untimed frames, then
    args.batch_frames timed frames; summed frames/s, the multiple of the single-stream headline, and the batched
    registration kernel's algorithmic bytes over its HIP-event launch time."""
    S, V, F = args.batch_seqs, args.batch_variants, SEQ_PREFIX + args.batch_frames
    var = [seq.frames_device(F, dev, seed=100 + v) for v in range(V)]
    X = torch.stack([var[s % V][0] for s in range(S)]).transpose(0, 1).contiguous()
    A = torch.stack([var[s % V][1] for s in range(S)]).transpose(0, 1).contiguous()
    T = torch.stack([var[s % V][2] for s in range(S)]).transpose(0, 1).contiguous()
    del var
    bo = lw.BatchOdometry(S, dev, reg_slots=1 << 18, frame_cap=1 << 15)
    try:
        bo.reg_voxel, bo.max_iter, bo.acc_mode = 0.2, 20, args.acc
        bo.reset(np.stack([seq.pose(0.05)] * S))
        for k in range(SEQ_PREFIX):
            bo.step(X[k], A[k], T[k], k)
        torch.cuda.synchronize()
        bo.prof_enable(True)
        bo.prof_reset()
        regs0 = bo.totals["registrations"]
        t0 = time.perf_counter()
        for k in range(SEQ_PREFIX, F):
            bo.step(X[k], A[k], T[k], k)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        pr = bo.prof_read()
        regs = bo.totals["registrations"] - regs0
        drifts = [drift(seq, bo.traj[s])["final_m"] for s in range(0, S, S // 4)]
    finally:
        bo.close()
        del X, A, T
        torch.cuda.empty_cache()
    fps = S * args.batch_frames / el
    launch_s = pr[1] * 1e-3 / max(pr[0], 1.0)
    bpl = accumulation_bytes(pr) / max(pr[0], 1.0)
    achieved = bpl / launch_s / 1e9 if launch_s > 0 else 0.0
    return {"sequences": S, "noise_variants": V, "frames_timed_per_sequence": args.batch_frames,
            "value": round(fps, 1), "unit": "frames/s (summed over sequences)",
            "x_single_stream": round(fps / single_fps, 2), "ms_per_frame_step": round(el * 1e3 / args.batch_frames, 3),
            "registrations": int(regs), "final_drift_m_sampled": [round(d, 4) for d in drifts],
            "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic("k_register_seq<0>"),
                         "kernel": "k_register_seq<0> (one launch per chain step: every sequence's whole "
                                   "registration, one workgroup each)",
                         "launches": int(pr[0]), "avg_launch_us": round(launch_s * 1e6, 2),
                         "bytes_per_launch": round(bpl, 1), "time_share": round(pr[1] * 1e-3 / el, 4)}}


def copy_probe(torch, dev, mib: int):
    """A plain device-to-device copy of known size, for calibrating hardware byte counters against."""
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()


SEQ_FRAMES = 600  # the plain odometry sequence: 60 s at 10 Hz
SEQ_PREFIX = 30  # frames of each pass kept out of the timed region: the 8 warm-up inserts, the walk's 1.5 s hold
PROF_STRIDE = 4  # the headline times the registration launches of every 4th bracketed frame
C4_OFFSET = (1000.0, -750.0, 30.0)  # the precision study's translation of the scene and the walk
C1_SEED = 11
C1_RUNS = 20


def odometry_pass(odo, push, reset_pose, n_frames: int, timed: bool, L=None, lw=None):
    """One pass of the odometry over the sequence on a fresh map: the first SEQ_PREFIX frames untimed, then the
    rest, bracketed by a stream and device synchronisation on both sides (returns its seconds). When L is given,
    the registration counters and HIP events are on for every PROF_STRIDE-th bracketed frame (its launches,
    timed where they run in the bracket; timing every launch costs about 4% of the frame rate)."""
    import torch

    odo.reset(reset_pose)
    for k in range(SEQ_PREFIX):
        push(k)
    odo.synchronize()
    torch.cuda.synchronize()
    sample = L is not None and timed
    t0 = time.perf_counter()
    for k in range(SEQ_PREFIX, n_frames):
        on = sample and (k - SEQ_PREFIX) % PROF_STRIDE == 0
        if on:
            lw._abi.check(L.lw_prof_enable(odo._ctx, 1), "lw_prof_enable")
        push(k)
        if on:
            lw._abi.check(L.lw_prof_enable(odo._ctx, 0), "lw_prof_enable")
    odo.synchronize()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def sampled_frames(n_frames: int) -> int:
    """Bracketed frames of one pass that odometry_pass times."""
    return (n_frames - SEQ_PREFIX + PROF_STRIDE - 1) // PROF_STRIDE


def precision_forms(odo, push, seq, n_frames: int):
    """C4: the sequence with the scene and the walk translated together by C4_OFFSET -- the sensor-frame points
    are the same, so the frames are reused and only the start pose moves 1250 m from the origin -- one pass per
    accumulation form: frames/s over the bracketed frames, and drift against the translated ground truth."""
    out = {}
    T0 = seq.pose(0.05)
    T0[:3, 3] += np.asarray(C4_OFFSET)
    shifted = _Shifted(seq, C4_OFFSET)
    mode0 = odo.acc_mode
    for mode in ("fp64", "kahan32", "fp32"):
        odo.acc_mode = mode
        regs0 = odo.totals["registrations"]
        el = odometry_pass(odo, push, T0, n_frames, timed=False)
        timed = len(odo.traj) - SEQ_PREFIX
        out[mode] = {"frames_per_s": round(timed / el, 2), "frames_timed": timed,
                     "registrations": odo.totals["registrations"] - regs0,
                     "drift": drift(shifted, odo.traj)}
    odo.acc_mode = mode0
    return {"workload": f"the {n_frames}-frame sequence with scene and walk translated by {C4_OFFSET} m (start "
                        f"1250 m from the origin), {n_frames - SEQ_PREFIX} bracketed frames a form",
            "forms": out}


class _Shifted:
    """A sequence's ground truth translated by an offset."""

    def __init__(self, seq, offset):
        self.seq, self.off = seq, np.asarray(offset, float)

    def pose(self, t):
        T = self.seq.pose(t)
        T[..., :3, 3] += self.off
        return T


def single_registration(args, torch, lw, dev):
    """C1: one scan-to-map registration of the box room (600k map points, 0.25 m voxels, 10k-point scan, 20
    iterations with the reference's 125-then-27-voxel schedule): median device register() time over C1_RUNS
    calls (thinning and the whole Gauss-Newton loop, one readback each), and the oracle's median beside it on
    the same inputs (its serial thinning and register, one host thread)."""
    w = lw.synth.single_scan(seed=C1_SEED)
    voxel = 0.25
    eng = lw.Engine(1 << 21, 1 << 15, device=dev)
    try:
        eng.insert(w.map_xyz, None, np.eye(4), voxel, max_range=1e30)
        m = eng.load_frame(w.scan_xyz, w.scan_attr, w.scan_t)
        eng.deskew(m, 0.05, 10.0, np.zeros(6))
        p = eng.params(reg_voxel=voxel, max_iter=20, prior_info=0.0)  # the start is a guess, not a prediction
        gpu = []
        for _ in range(C1_RUNS + 2):
            eng.sync()
            t0 = time.perf_counter()
            r = eng.register(m, p, w.T_init)
            gpu.append(time.perf_counter() - t0)
        gpu = gpu[2:]
        T_gpu = np.array(r.T).reshape(4, 4)
    finally:
        eng.close()
    res = {"workload": "box room 8x6x3 m, 600k map points, 0.25 m voxels, 10k-point rosette scan, <= 20 GN "
                       "iterations (125 voxels on the first 3), no motion prior", "seed": C1_SEED,
           "gpu_median_ms": round(float(np.median(gpu)) * 1e3, 4), "runs": C1_RUNS, "iters": int(r.iters),
           "n_ds": int(r.n_ds), "error_vs_truth": [round(v, 5) for v in _pose_err(w.T_true, T_gpu)]}
    if args.cpu_frames > 0:
        oracle = _oracle()
        om = oracle.OracleMap(1 << 21)
        om.insert(w.map_xyz, np.zeros(len(w.map_xyz), np.uint32), np.eye(4), voxel, max_range=1e30)
        P = oracle.Params(reg_voxel=voxel, max_iter=20, prior_info=0.0)
        cpu, ro = [], None
        for _ in range(args.c1_cpu_runs):
            t0 = time.perf_counter()
            ds = oracle.ds_insert(w.scan_xyz, w.scan_attr, 2.0 / voxel, 0.3, 60.0, 15, 1 << 15)
            ro = oracle.register(om, ds, w.T_init, P)
            cpu.append(time.perf_counter() - t0)
        dt, da = oracle.pose_error(T_gpu, ro.T)
        res["cpu_baseline"] = {"value": round(float(np.median(cpu)) * 1e3, 3), "unit": "ms per register() (median)",
                               "cores": 1, "kind": "port",
                               "sample": f"{args.c1_cpu_runs} registrations of the same scan on the same map "
                                         "through the oracle's serial thinning and register",
                               "gpu_vs_oracle": [round(dt * 1e3, 4), round(da, 5)]}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3,
                    help="timed passes over the whole sequence (each: a fresh map, its first 30 frames untimed)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes before them")
    ap.add_argument("--cpu-frames", type=int, default=150, help="frames of the CPU baseline sample (0: skip)")
    ap.add_argument("--c1-cpu-runs", type=int, default=5, help="oracle registrations of the C1 CPU row")
    ap.add_argument("--acc", default="fp64", choices=["fp64", "kahan32", "fp32"])
    ap.add_argument("--no-c4", action="store_true", help="skip the precision-form passes")
    ap.add_argument("--no-c1", action="store_true", help="skip the single-registration leg")
    ap.add_argument("--no-loc", action="store_true", help="skip the batched localisation leg")
    ap.add_argument("--no-imap", action="store_true", help="skip the integration-map leg")
    ap.add_argument("--no-grid", action="store_true", help="skip the prior-grid / coarse-search leg")
    ap.add_argument("--no-lvxr", action="store_true", help="skip the LVXR replay leg")
    ap.add_argument("--no-worker", action="store_true", help="skip the odometry-worker leg")
    ap.add_argument("--batch-seqs", type=int, default=256, help="sequences of the multi-sequence leg (0: skip)")
    ap.add_argument("--batch-variants", type=int, default=16, help="distinct noise realisations among them")
    ap.add_argument("--batch-frames", type=int, default=100, help="timed frames per sequence")
    ap.add_argument("--loc-queries", type=int, default=4)
    ap.add_argument("--loc-hyps", type=int, default=4096)
    ap.add_argument("--loc-reps", type=int, default=3)
    ap.add_argument("--loc-cpu-hyps", type=int, default=64)
    ap.add_argument("--loc-slots", type=int, default=24, help="log2 of the localisation map's hash slots")
    ap.add_argument("--copy-probe", type=int, default=0, help="MiB of a calibration copy at the end (0: none)")
    ap.add_argument("--no-prior", action="store_true", help="skip the learned-prior leg")
    ap.add_argument("--prior-seqs", type=int, default=4)
    ap.add_argument("--prior-frames", type=int, default=600)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the odometry is a sequential "
              f"per-frame pipeline; independent sequences and hypothesis batches would shard one per GPU)",
              file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be at least 1 and --warmup at least 0", file=sys.stderr)
        sys.exit(2)

    import torch

    lw = importlib.import_module("livox-warp_amd")
    n_frames = SEQ_FRAMES  # every step is a pass over the whole sequence, whatever --steps says
    seq = lw.synth.Sequence(seed=SEED, seconds=60.0)
    t_gen = time.perf_counter()
    frames = list(seq.frames(n_frames))
    t_gen = time.perf_counter() - t_gen
    per = len(frames[0][2])
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    X = torch.from_numpy(np.stack([f[0] for f in frames])).to(dev)
    A = torch.from_numpy(np.stack([f[1] for f in frames]).view(np.int32)).to(dev)
    Tt = torch.from_numpy(np.stack([f[2] for f in frames])).to(dev)
    spans = [(float(f[2][0]), float(f[2][-1])) for f in frames]
    torch.cuda.synchronize()

    t_run = time.perf_counter()

    def leg(name, fn):
        """One leg, with a progress line on stderr (stdout carries only the result line)."""
        t0 = time.perf_counter()
        r = fn()
        print(f"[bench] {name}: {time.perf_counter() - t0:.1f} s (at {time.perf_counter() - t_run:.1f} s)",
              file=sys.stderr, flush=True)
        return r

    odo = lw.Odometry(lw.PoseTable(), dev, reg_slots=1 << 22, frame_cap=1 << 15)
    try:
        out = leg("odometry", lambda: odometry_headline(args, torch, lw, dev, odo, X, A, Tt, spans, seq, n_frames,
                                                          per, t_gen))
    finally:
        odo.close()
    if args.batch_seqs > 0:
        out["multi_sequence"] = leg("multi_sequence", lambda: multi_sequence(args, torch, lw, dev, seq, out["value"]))
    if not args.no_c1:
        out["single_registration"] = leg("single_registration", lambda: single_registration(args, torch, lw, dev))
    if not args.no_imap:
        out["integration_map"] = leg("integration_map", lambda: integration_map(args, torch, lw, dev, X, A, Tt, seq,
                                                                                n_frames, per))
    if not args.no_grid:
        out["prior_grid"] = leg("prior_grid", lambda: prior_grid(args, torch, lw, dev))
    if not args.no_lvxr:
        out["lvxr_replay"] = leg("lvxr_replay", lambda: lvxr_replay(args, torch, lw, dev, frames))
    if not args.no_worker:
        out["odom_worker"] = leg("odom_worker", lambda: odom_worker(args, torch, lw, dev, X, A, Tt, spans, seq,
                                                                    n_frames, per))
    if not args.no_loc:
        out["localisation"] = leg("localisation", lambda: localisation(args, torch, lw, dev))
    if not args.no_prior:
        out["learned_prior"] = leg("learned_prior", lambda: learned_prior(args, torch, lw, dev))
    if args.cpu_frames > 0:
        out["cpu_baseline"] = leg("cpu_baseline", lambda: cpu_baseline(frames, seq, SEQ_PREFIX, args.cpu_frames))
    else:
        out["cpu_baseline"] = None
    if args.copy_probe > 0:
        copy_probe(torch, dev, args.copy_probe)
    torch.cuda.synchronize()
    print(json.dumps(out), flush=True)


def odometry_headline(args, torch, lw, dev, odo, X, A, Tt, spans, seq, n_frames: int, per: int, t_gen: float):
    """The headline: args.warmup untimed passes, then args.steps timed passes of the odometry over the whole
    sequence (each on a fresh map, its first SEQ_PREFIX frames outside the bracket), frames/s over the bracketed
    frames of every timed pass; the registration kernel's counters and HIP-event durations over every
    PROF_STRIDE-th of the same frames."""
    odo.reg_voxel = 0.2
    odo.max_iter = 20
    odo.acc_mode = args.acc
    L = odo._lib

    def push(k):
        odo.push(X[k], A[k], Tt[k], per, spans[k][0], spans[k][1], _no_frame)

    # the warm-up passes run with the device phase stamps on (where a pass's time goes)
    phases = np.zeros(24)
    if args.warmup > 0:
        lw._abi.check(L.lw_prof_enable(odo._ctx, 2), "lw_prof_enable")
        lw._abi.check(L.lw_prof_reset(odo._ctx), "lw_prof_reset")
        for _ in range(args.warmup):
            odometry_pass(odo, push, seq.pose(0.05), n_frames, timed=False)
        lw._abi.check(L.lw_prof_phases(odo._ctx, lw._abi.ptr(phases, lw._abi._f64p)), "lw_prof_phases")
    lw._abi.check(L.lw_prof_enable(odo._ctx, 1), "lw_prof_enable")
    lw._abi.check(L.lw_prof_reset(odo._ctx), "lw_prof_reset")
    lw._abi.check(L.lw_prof_enable(odo._ctx, 0), "lw_prof_enable")
    el, timed, regs, skipped, keyframes, times = 0.0, 0, 0, 0, 0, []
    for _ in range(args.steps):
        regs0 = odo.totals["registrations"]
        e = odometry_pass(odo, push, seq.pose(0.05), n_frames, timed=True, L=L, lw=lw)
        times.append(e)
        el += e
        timed += len(odo.traj) - SEQ_PREFIX
        regs += odo.totals["registrations"] - regs0
        skipped += odo.stats["skipped"]
        keyframes += odo.stats["keyframes"]
    prof = np.zeros(8)
    lw._abi.check(L.lw_prof_read(odo._ctx, lw._abi.ptr(prof, lw._abi._f64p)), "lw_prof_read")
    sampled = args.steps * sampled_frames(n_frames)
    fps = timed / el
    launch_s = prof[1] * 1e-3 / max(prof[0], 1.0)
    bytes_per_launch = accumulation_bytes(prof) / max(prof[0], 1.0)
    achieved = bytes_per_launch / launch_s / 1e9 if launch_s > 0 else 0.0
    dr = drift(seq, odo.traj)  # the last pass, all 600 frames
    out = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 residuals / f64 accumulation" if args.acc == "fp64" else f"f32 residuals / {args.acc}",
        "data": "synthetic",
        "config": {"workload": f"plain odometry sequence: {n_frames} frames x 10k rose-curve rosette points, office "
                               "20x15x3 m, 0.2 m voxels, 2^22-slot map, <=20 GN iterations; a step is one pass "
                               f"over the whole sequence on a fresh map, its first {SEQ_PREFIX} frames (8 warm-up "
                               "inserts, the walk's 1.5 s hold) outside the timed bracket",
                   "seed": SEED, "points_per_frame": per, "frames_generated": n_frames,
                   "frames_per_step": n_frames - SEQ_PREFIX, "frames_timed": timed, "acc_mode": args.acc},
        "points_per_s": round(fps * per, 1),
        "frames_registered": timed - skipped,
        "registrations": regs,
        "keyframes_per_pass": keyframes // args.steps,
        "step_ms": [round(t * 1e3, 2) for t in times],
        "drift": dr,
        "map_voxels": odo.stats["map_voxels"],
        "registration": {"launches": int(prof[0]), "passes": int(prof[6]),
                         "frames_sampled": sampled, "sample_stride": PROF_STRIDE,
                         "avg_launch_us": round(launch_s * 1e6, 3), "point_iterations": int(prof[2]),
                         "lookups": int(prof[3]), "hits": int(prof[4]), "matched": int(prof[5]),
                         "time_share": round(launch_s * prof[0] / max(sampled, 1) * timed / el, 4),
                         "phase_passes": args.warmup,
                         "us_per_pass": {"search": round(phases[0] / max(phases[3], 1), 2),
                                         "reduce": round(phases[1] / max(phases[3], 1), 2),
                                         "step": round(phases[2] / max(phases[3], 1), 2),
                                         "launch": round(prof[1] * 1e3 / max(prof[6], 1), 2),
                                         "searcher_wake": round(phases[4] / max(phases[7], 1), 2),
                                         "searcher_search": round(phases[5] / max(phases[7], 1), 2),
                                         "searcher_row_flag": round(phases[6] / max(phases[7], 1), 2),
                                         "searcher_neighbours": round(phases[15] / max(phases[7], 1), 2),
                                         "searcher_terms": round(phases[16] / max(phases[7], 1), 2),
                                         "searcher_row_sum": round(phases[17] / max(phases[7], 1), 2),
                                         "searcher_neighbours_narrow": round(phases[22] / max(phases[23], 1), 2),
                                         "searcher_neighbours_wide": round((phases[15] - phases[22]) /
                                                                           max(phases[7] - phases[23], 1), 2),
                                         "all_searchers_wake": round(phases[18] / max(phases[20], 1), 2),
                                         "all_searchers_flag": round(phases[19] / max(phases[20], 1), 2),
                                         "slowest_searcher_row": round(phases[21], 2)},
                         "steps_on_wave": int(phases[8]), "steps_one_lane": int(phases[9]),
                         "us_per_wave_step": {k: round(phases[10 + i] / max(phases[14], 1), 3) for i, k in
                                              enumerate(("setup", "factor", "substitute", "update"))}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic("k_register<0>"),
                     "kernel": "k_register<0> (one launch per registration: every accumulation pass and step)",
                     "bytes_per_launch": round(bytes_per_launch, 1)},
        "gen_seconds": round(t_gen, 2),
    }
    if not args.no_c4:
        out["precision_forms"] = precision_forms(odo, push, seq, n_frames)
    return out


def _no_frame(*_):
    pass


if __name__ == "__main__":
    main()

"""Helpers for the GPU tests that check whole odometry runs against the oracle registration by registration:
a recorder that watches an Odometry on its default (chain) path and keeps what each registration ran on -- its
thinned points, its start pose and the device's map as it stood (read back after every insert) -- and a threaded
replay of those registrations through the oracle, plus the oracle's own serial build of the same inserts."""

import ctypes
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

REPLICAS = 8


def err(A, B):
    d = np.linalg.inv(A) @ B
    return float(np.linalg.norm(A[:3, 3] - B[:3, 3])), float(
        np.degrees(np.arccos(np.clip((np.trace(d[:3, :3]) - 1) / 2, -1, 1))))


def gt_errors(seq, traj):
    T0 = seq.pose(traj[0][1])
    et, er, dist = [], [], 0.0
    for (k, tm, T), nxt in zip(traj, traj[1:] + [None]):
        gt = np.linalg.inv(T0) @ seq.pose(tm)
        est = np.linalg.inv(traj[0][2]) @ T
        et.append(float(np.linalg.norm(est[:3, 3] - gt[:3, 3])))
        d = est[:3, :3].T @ gt[:3, :3]
        er.append(float(np.degrees(np.arccos(np.clip((np.trace(d) - 1) / 2, -1, 1)))))
        if nxt is not None:
            dist += float(np.linalg.norm(seq.pose(nxt[1])[:3, 3] - seq.pose(tm)[:3, 3]))
    return np.array(et), np.array(er), dist


def odometry(lw, reg_slots=1 << 22):
    odo = lw.Odometry(lw.PoseTable(), "cuda:0", reg_slots=reg_slots, frame_cap=1 << 15)
    odo.reg_voxel = 0.2
    odo.max_iter = 20
    return odo


def read_map(odo):
    """The device map in the reference's per-slot layout (lw_map_read)."""
    cap = odo.reg_cap
    keys, rsum = np.empty(cap, np.int64), np.empty((cap, 3), np.float32)
    router, rcnt = np.empty((cap, 3, 3), np.float32), np.empty(cap, np.int32)
    P = ctypes.POINTER
    assert odo._lib.lw_map_read(odo._ctx, keys.ctypes.data_as(P(ctypes.c_int64)), rsum.ctypes.data_as(P(ctypes.c_float)),
                                router.ctypes.data_as(P(ctypes.c_float)), rcnt.ctypes.data_as(P(ctypes.c_int32))) == 0
    return keys, rsum, router, rcnt


def params(orc, odo):
    return orc.Params(reg_voxel=odo.reg_voxel, max_iter=odo.max_iter, min_corr=odo.min_corr,
                      min_range=odo.min_range, max_range=odo.max_range, noise=odo.noise, ret_mask=odo.ret_mask,
                      alpha=odo.alpha, damping=odo.damping, degen_thr=odo.degen_thr, prior_info=odo.prior_info)


class Recorder:
    """Watches an Odometry on the chain path: the inputs of each chained frame (prediction, first twist) and,
    after the frame, the device data the oracle needs -- raw and deskewed points, attributes, times, the thinned
    set of each registration -- plus every map insert (points and pose)."""

    def __init__(self, lw, odo):
        self.lw, self.odo = lw, odo
        self.L, self.ctx = odo._lib, odo._ctx
        self.events = []  # ("insert", xyz, attr, T) | ("reg", ds, T_start, T_gpu, frame)
        self.thin_checks = []  # (deskewed xyz, attr, thinned set) of the frame's last registration
        self.deskew_checks = []  # (raw, t, t_mid, xi, deskewed)
        self.frames = 0
        self.chained = None
        inner = odo._frame_chain

        def chain(t_mid):
            out = inner(t_mid)
            fo, fi = odo._fout, odo._fin
            self.chained = dict(T_pred=np.array(fi.T_pred).reshape(4, 4), xi0=np.array(fi.xi, np.float32),
                                t_mid=float(fi.t_mid), regs=[(np.array(r.T).reshape(4, 4), int(r.n_ds), bool(r.ok))
                                                             for r in fo.reg[: int(fo.registrations)]],
                                skipped=bool(fo.skipped), keyframe=bool(fo.keyframe), xi=np.array(fo.xi, np.float32))
            return out

        odo._frame_chain = chain
        self.keyframes = 0

    def _thin(self, m, xi, t_mid):
        o = self.odo
        x = np.ascontiguousarray(xi, np.float32)
        assert self.L.lw_deskew(self.ctx, m, t_mid, 1.0 / o.frame_dt, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
        n = (ctypes.c_int32 * 2)()
        assert self.L.lw_thin(self.ctx, m, 2.0 / o.reg_voxel, o.min_range, o.max_range, int(o.ret_mask), n) == 0
        return self._read_thin(int(n[0]))

    def _read_thin(self, n):
        ds = np.empty((n, 3), np.float32)
        assert self.L.lw_thin_read(self.ctx, ds.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n) == 0
        return ds

    def on_frame(self, k, T, fx, fa, ft, ff, m):
        o = self.odo
        self.frames += 1
        inserted = o.stats["keyframes"] > self.keyframes
        self.keyframes = o.stats["keyframes"]
        c, self.chained = self.chained, None
        torch.cuda.synchronize()
        xyz = fx[:m].cpu().numpy()
        attr = fa[:m].cpu().numpy().view(np.uint32)
        self._registrations(c, m, k, xyz, attr)
        if inserted:  # the frame's insert follows its registrations
            self.events.append(("insert", xyz, attr, T.copy()))
            self.events.append(("map",) + self.read_voxels())

    def read_voxels(self):
        """The device map's voxels in claim order (lw_map_read_voxels): keys, rsum, router, rcnt."""
        cnt = (ctypes.c_int32 * 1)()
        assert self.L.lw_map_read_voxels(self.ctx, 0, None, None, None, None, cnt) == 0
        V = int(cnt[0])
        keys, rsum = np.empty(V, np.int64), np.empty((V, 3), np.float32)
        router, rcnt = np.empty((V, 3, 3), np.float32), np.empty(V, np.int32)
        P = ctypes.POINTER
        assert self.L.lw_map_read_voxels(self.ctx, V, keys.ctypes.data_as(P(ctypes.c_int64)),
                                         rsum.ctypes.data_as(P(ctypes.c_float)), router.ctypes.data_as(P(ctypes.c_float)),
                                         rcnt.ctypes.data_as(P(ctypes.c_int32)), cnt) == 0
        return keys, rsum, router, rcnt

    def _registrations(self, c, m, k, xyz, attr):
        o = self.odo
        if c is None or c["skipped"] or not c["regs"]:
            return
        raw, t = o.f_raw[:m].cpu().numpy(), o.f_t[:m].cpu().numpy()
        last = self._read_thin(c["regs"][-1][1])
        self.thin_checks.append((xyz, attr, last))
        self.deskew_checks.append((raw, t, c["t_mid"], c["xi"], xyz))
        sets = [last]
        if len(c["regs"]) == 2:
            sets = [self._thin(m, c["xi0"], c["t_mid"]), last]
            # the same deskew and thinning with the final twist give the chain's last set back exactly
            assert np.array_equal(self._thin(m, c["xi"], c["t_mid"]), last)
        start = c["T_pred"]
        for (Tg, n_ds, ok), ds in zip(c["regs"], sets):
            assert len(ds) == n_ds
            self.events.append(("reg", ds, start.copy(), Tg, k))
            start = Tg


class SnapshotMap:
    """An oracle map that takes the device's map as it stands: new voxels are placed by the reference's linear probe
    from their home slots (voxels keep their claim order, so only the newly claimed ones move in), and every
    voxel's moments are copied in."""

    def __init__(self, orc, slots: int):
        self.orc = orc
        self.om = orc.OracleMap(slots)
        self.slot_of = np.zeros(0, np.int64)

    def load(self, keys, rsum, router, rcnt):
        om, old = self.om, len(self.slot_of)
        mask = om.cap - 1
        homes = self.orc.slot_hash_np(keys[old:]).astype(np.int64) & mask
        new = np.empty(len(keys) - old, np.int64)
        for j, (key, h) in enumerate(zip(keys[old:], homes)):
            s = int(h)
            while om.keys[s] != -1:
                s = (s + 1) & mask
            om.keys[s] = key
            new[j] = s
        self.slot_of = np.concatenate([self.slot_of, new])
        om.rsum[self.slot_of] = rsum
        om.router[self.slot_of] = router
        om.rcnt[self.slot_of] = rcnt


def replay(orc, events, P, slots):
    """Every registration of `events` through the oracle on the device's own map at that point (replicas in
    threads: the C accumulation releases the interpreter lock), and the oracle's serial build of the same
    inserts (the map-build comparison). Returns ({event index: (oracle result, (m, deg) from the device's)},
    the serial build)."""
    regs = [i for i, e in enumerate(events) if e[0] == "reg"]
    out = {}

    def worker(r):
        if r == REPLICAS:  # the serial build of the device's inserts
            om = orc.OracleMap(slots)
            for e in events:
                if e[0] == "insert":
                    om.insert(e[1], e[2], e[3], P.reg_voxel, P.max_range, P.ret_mask)
            return om
        sm = SnapshotMap(orc, 1 << 20)
        mine = set(regs[r::REPLICAS])
        for i, e in enumerate(events):
            if e[0] == "map":
                sm.load(*e[1:])
            elif i in mine:
                res = orc.register(sm.om, e[1], e[2], P)
                out[i] = (res, err(e[3], res.T))
        return None

    with ThreadPoolExecutor(REPLICAS + 1) as ex:
        maps = list(ex.map(worker, range(REPLICAS + 1)))
    return out, maps[REPLICAS]


def read_voxels(L, ctx):
    """A context's map voxels in claim order (lw_map_read_voxels): keys, rsum, router, rcnt."""
    cnt = (ctypes.c_int32 * 1)()
    assert L.lw_map_read_voxels(ctx, 0, None, None, None, None, cnt) == 0
    V = int(cnt[0])
    keys, rsum = np.empty(V, np.int64), np.empty((V, 3), np.float32)
    router, rcnt = np.empty((V, 3, 3), np.float32), np.empty(V, np.int32)
    P = ctypes.POINTER
    assert L.lw_map_read_voxels(ctx, V, keys.ctypes.data_as(P(ctypes.c_int64)), rsum.ctypes.data_as(P(ctypes.c_float)),
                                router.ctypes.data_as(P(ctypes.c_float)), rcnt.ctypes.data_as(P(ctypes.c_int32)),
                                cnt) == 0
    return keys, rsum, router, rcnt


class BatchRecorder:
    """Watches chosen sequences of a BatchOdometry: after each step, what each of their registrations ran on
    (the thinned points read from the sequence's context; the first registration's set rebuilt by the context's
    own deskew and thinning from the frame's raw points), the start poses, the device's map after every insert,
    and the frame's thinning and deskew for the oracle checks. events[s] has the form replay() takes."""

    def __init__(self, bo, seqs):
        self.bo, self.seqs = bo, list(seqs)
        self.events = {s: [] for s in self.seqs}
        self.thin_checks = []  # (deskewed xyz, attr, thinned set of the last registration)
        self.deskew_checks = []  # (raw, t, t_mid, xi, deskewed)

    def _thin(self, s, m, xi, t_mid):
        bo, L, ctx = self.bo, self.bo._lib, self.bo.context(s)
        x = np.ascontiguousarray(xi, np.float32)
        assert L.lw_deskew(ctx, m, t_mid, 1.0 / bo.frame_dt, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
        n = (ctypes.c_int32 * 2)()
        assert L.lw_thin(ctx, m, 2.0 / bo.reg_voxel, bo.min_range, bo.max_range, int(bo.ret_mask), n) == 0
        return self._read_thin(s, int(n[0]))

    def _read_thin(self, s, n):
        ds = np.empty((n, 3), np.float32)
        assert self.bo._lib.lw_thin_read(self.bo.context(s), ds.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n) == 0
        return ds

    def after_step(self, X, A, Tt, k):
        bo = self.bo
        out, vin = bo.last, bo._vin
        torch.cuda.synchronize()
        for s in self.seqs:
            L, ctx = bo._lib, bo.context(s)
            m = int(out["m"][s])
            nreg = int(out["registrations"][s])
            if nreg > 0:
                raw, attr, t = X[s, :m], A[s, :m], Tt[s, :m]
                f_raw, f_xyz, f_attr, f_t, _ = bo._bufs[s]
                xyz = f_xyz[:m].cpu().numpy()
                a = attr.cpu().numpy().view(np.uint32)
                last = self._read_thin(s, int(out["reg"]["n_ds"][s, nreg - 1]))
                t_mid, xi0, xi = float(vin["t_mid"][s]), vin["xi"][s].copy(), out["xi"][s].copy()
                self.thin_checks.append((xyz, a, last))
                self.deskew_checks.append((raw.cpu().numpy(), t.cpu().numpy(), t_mid, xi, xyz))
                sets = [last]
                if nreg == 2:
                    # the frame into the context's own buffers, then its first deskew and thinning again
                    f_raw[:m].copy_(raw)
                    f_attr[:m].copy_(attr)
                    f_t[:m].copy_(t)
                    torch.cuda.synchronize()
                    sets = [self._thin(s, m, xi0, t_mid), last]
                    assert np.array_equal(self._thin(s, m, xi, t_mid), last)
                start = vin["T_pred"][s].reshape(4, 4).copy()
                for r, ds in enumerate(sets):
                    Tg = out["reg"]["T"][s, r].reshape(4, 4).copy()
                    assert len(ds) == int(out["reg"]["n_ds"][s, r])
                    self.events[s].append(("reg", ds, start, Tg, k, float(out["reg"]["cond"][s, r])))
                    start = Tg
            if out["keyframe"][s]:
                self.events[s].append(("map",) + read_voxels(L, ctx))

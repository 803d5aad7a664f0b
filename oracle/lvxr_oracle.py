"""Oracle for LVXR recordings -- TEST INFRASTRUCTURE ONLY (tests call it as the checker).

A serial restatement of the reference's firing-level reader (python/livox_warp/lvxr.py:76-260) on the format of
crates/livox/src/record.rs:1-5 and point.rs:105-150: datagram by datagram with struct, firing by firing, in
Python; float32 where the reference's numpy is float32. Firings come back time-ordered, ties in the order the
reference concatenates them (by data type, then datagram length, then file order).
"""

from __future__ import annotations

import math
import struct

import numpy as np

SIZES = {0: 13, 1: 9, 2: 14, 3: 10, 4: 28, 5: 16, 7: 42, 8: 22}
RETURNS = {0: 1, 1: 1, 2: 1, 3: 1, 4: 2, 5: 2, 7: 3, 8: 3}
SPHERICAL = {1, 3, 5, 8}
f32 = np.float32
D2R = f32(np.float32(math.pi) / np.float32(180.0))


def _sph(depth, theta, phi):
    th = f32(f32(f32(theta) * f32(0.01)) * D2R)
    ph = f32(f32(f32(phi) * f32(0.01)) * D2R)
    unit = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], np.float32)
    return unit, (unit * f32(f32(depth) * f32(1e-3))).astype(np.float32)


def _record(rec: bytes, dt: int):
    """[(xyz, refl, tag)] per return, and the direction of a spherical record (else None)."""
    if dt in (0, 2):
        x, y, z, r = struct.unpack_from("<iiiB", rec)
        return [(np.array([x, y, z], np.float32) * f32(1e-3), r, rec[13] if dt == 2 else 0)], None
    if dt in (1, 3):
        depth, th, ph, r = struct.unpack_from("<IHHB", rec)
        u, p = _sph(depth, th, ph)
        return [(p, r, rec[9] if dt == 3 else 0)], u
    if dt in (4, 7):
        out = []
        for k in range(RETURNS[dt]):
            x, y, z, r, g = struct.unpack_from("<iiiBB", rec, 14 * k)
            out.append((np.array([x, y, z], np.float32) * f32(1e-3), r, g))
        return out, None
    th, ph = struct.unpack_from("<HH", rec)
    out = []
    for k in range(RETURNS[dt]):
        depth, r, g = struct.unpack_from("<IBB", rec, 4 + 6 * k)
        u, p = _sph(depth, th, ph)
        out.append((p, r, g))
    return out, u


def time_base(stamps, spans):
    """Packet times in ns on the live path's time base (net.rs Sink, lvxr.py:113-133): a stamp below the base
    starts a new base; a packet more than half its span late by a second or more moves the base by the gap."""
    out = []
    base = expected = None
    for ts, span in zip(stamps, spans):
        if base is None or ts < base:
            base, expected = ts, None
        if expected is not None and span > 0 and ts > expected + span // 2:
            if ts - expected >= 1_000_000_000:
                base += ts - expected
        expected = ts + span
        out.append(ts - base)
    return out


def read(data: bytes):
    """(dev_type, profile, t (n,) float64, xyz (n, R, 3), refl, tag, valid, direction (n, 3))."""
    if data[:6] == b"LVXR1\n":
        dev, profile, off = data[6], 0, 7
    elif data[:6] == b"LVXR2\n":
        dev, profile, off = data[6], data[7], 8
    else:
        raise ValueError("not an LVXR recording")
    dgrams = []
    while off + 10 <= len(data):
        (ln,) = struct.unpack_from("<H", data, off + 8)
        if off + 10 + ln > len(data):
            break
        dgrams.append(data[off + 10: off + 10 + ln])
        off += 10 + ln
    interval = 4167 if dev in (2, 3, 7) else 10000
    dual = profile == 1
    pk = []  # the point packets: (index, datagram, data type, records, t ns)
    for i, d in enumerate(dgrams):
        if len(d) < 18 or d[9] not in SIZES:
            continue
        dt = d[9]
        if d[8] == 3:
            hour = d[13]
            (us,) = struct.unpack_from("<I", d, 14)
            ts = (hour * 3_600_000_000 + us) * 1000
        else:
            (ts,) = struct.unpack_from("<Q", d, 10)
        pk.append([i, d, dt, (len(d) - 18) // SIZES[dt], ts])
    for p, tb in zip(pk, time_base([p[4] for p in pk], [(p[3] // 2 if dual else p[3]) * interval for p in pk])):
        p.append(tb)
    keep = [p for p in pk if len(p[1]) - 18 == p[3] * SIZES[p[2]]
            and (not dual or (p[2] in (0, 1) and p[3] == 100))]
    if not keep:
        raise ValueError("no point packets")
    R = max([RETURNS[p[2]] for p in keep] + [2 if dual else 1])
    rows = []  # (t, concat key, xyz, refl, tag, direction)
    for p in keep:
        i, d, dt, nrec, _, tpk = p
        size = SIZES[dt]
        nf = nrec // 2 if dual else nrec
        for j in range(nf):
            if dual:
                r0, u0 = _record(d[18 + 2 * j * size: 18 + (2 * j + 1) * size], dt)
                r1, _ = _record(d[18 + (2 * j + 1) * size: 18 + (2 * j + 2) * size], dt)
                rets, u = [r0[0], r1[0]], u0
            else:
                rets, u = _record(d[18 + j * size: 18 + (j + 1) * size], dt)
            rows.append(((tpk + j * interval) * 1e-9, (dt, len(d), i, j), rets, u))
    rows.sort(key=lambda r: (r[0], r[1]))
    n = len(rows)
    t = np.array([r[0] for r in rows], np.float64)
    xyz = np.zeros((n, R, 3), np.float32)
    refl = np.zeros((n, R), np.uint8)
    tag = np.zeros((n, R), np.uint8)
    direction = np.full((n, 3), np.nan, np.float32)
    for a, (_, _, rets, u) in enumerate(rows):
        for k, (p, rr, g) in enumerate(rets):
            xyz[a, k], refl[a, k], tag[a, k] = p, rr, g
        if u is not None:
            direction[a] = u
    valid = np.any(xyz != 0.0, axis=2)
    for a in range(n):
        if np.isnan(direction[a, 0]) and valid[a].any():
            v = xyz[a, int(np.argmax(valid[a]))]
            nv = np.sqrt(f32(f32(v[0] * v[0]) + f32(v[1] * v[1])) + f32(v[2] * v[2]), dtype=np.float32)
            direction[a] = v / nv
    return dev, profile, t, xyz, refl, tag, valid, direction


def points(t, xyz, refl, tag, valid, clean=True, min_range=0.3):
    """(xyz, attr, t, firing) of the returns, firing-major (lvxr.py:57-73)."""
    out = []
    for a in range(len(t)):
        for k in range(xyz.shape[1]):
            if not valid[a, k]:
                continue
            if clean and (tag[a, k] != 0 or not np.linalg.norm(xyz[a, k]) > min_range):
                continue
            out.append((xyz[a, k], int(refl[a, k]) | (int(tag[a, k]) << 8) | (k << 16), t[a], a))
    if not out:
        return np.zeros((0, 3), np.float32), np.zeros(0, np.int32), np.zeros(0), np.zeros(0, np.int64)
    return (np.array([o[0] for o in out], np.float32), np.array([o[1] for o in out], np.int32),
            np.array([o[2] for o in out]), np.array([o[3] for o in out], np.int64))

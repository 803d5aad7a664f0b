"""The LVXR oracle pinned without a GPU: records written with known values come back exactly (cartesian
millimetres, reflectivity, tags, firing times on the packet clock), spherical records land on closed-form
directions, the time base resets when the clock goes back and closes forward jumps of a second or more but
keeps shorter gaps (net.rs Sink), GPS stamps fold to ns of the day, and the reader's rejections hold
(truncated last datagram, short or malformed datagrams, IMU packets, dual packets that are not 100 plain
records). Also: the product's host half (framing and time base, livox-warp_amd/lvxr.py) agrees with it."""

import struct

import numpy as np
import pytest

import lvxr_oracle as X
import lvxr_write as W


def _read(tmp_path, dev, profile, dgrams, **kw):
    p = tmp_path / "r.lvxr"
    W.write(p, dev, profile, dgrams, **kw)
    return X.read(p.read_bytes()), p


def test_cartesian_values_and_times(tmp_path):
    recs = [(1000, -2000, 3500, 7), (0, 0, 0, 9), (-1, 2, -3, 255)]
    body = b"".join(struct.pack("<iiiB", *r) for r in recs)
    (dev, prof, t, xyz, refl, tag, valid, d), _ = _read(tmp_path, 1, 0, [W.header(0, 5_000_000) + body])
    assert (dev, prof) == (1, 0) and xyz.shape == (3, 1, 3)
    assert np.array_equal(t, (np.arange(3) * 10_000) * 1e-9)  # the base is the first packet's stamp
    assert np.array_equal(xyz[:, 0], np.array([r[:3] for r in recs], np.float32) * np.float32(1e-3))
    assert list(refl[:, 0]) == [7, 9, 255] and not tag.any()
    assert list(valid[:, 0]) == [True, False, True]
    assert np.isnan(d[1]).all() and abs(np.linalg.norm(d[0]) - 1) < 1e-6  # a firing with no return: no direction


def test_spherical_closed_forms(tmp_path):
    recs = [(2000, 0, 0, 1, 0), (3000, 9000, 0, 2, 1), (4000, 9000, 9000, 3, 2), (0, 4500, 18000, 4, 0)]
    body = b"".join(struct.pack("<IHHBB", *r) for r in recs)
    (_, _, _, xyz, refl, tag, valid, d), _ = _read(tmp_path, 2, 0, [W.header(3, 0) + body])
    assert np.allclose(d[0], [0, 0, 1], atol=1e-6) and np.allclose(xyz[0, 0], [0, 0, 2], atol=1e-6)
    assert np.allclose(d[1], [1, 0, 0], atol=1e-6) and np.allclose(xyz[1, 0], [3, 0, 0], atol=1e-6)
    assert np.allclose(d[2], [0, 1, 0], atol=1e-6) and np.allclose(xyz[2, 0], [0, 4, 0], atol=1e-6)
    h = np.sqrt(0.5)
    assert np.allclose(d[3], [-h, 0, h], atol=1e-6) and not valid[3, 0]  # zero depth: the direction, no return
    assert list(tag[:, 0]) == [0, 1, 2, 0] and list(refl[:, 0]) == [1, 2, 3, 4]


def test_time_base(tmp_path):
    span = 100_000
    st = [1_000_000_000 + k * span for k in range(3)]        # back to back
    st += [st[-1] + span + 500_000_000]                       # 0.5 s late: the gap stays
    st += [st[-1] + span + 4_000_000_000]                     # 4 s late: the base moves over it
    st += [st[-1] + span]
    st += [200_000_000, 200_000_000 + span]                   # below the base: a new base
    got = X.time_base(st, [span] * len(st))
    assert got[:3] == [0, span, 2 * span]
    assert got[3] == 3 * span + 500_000_000
    assert got[4] == got[3] + span and got[5] == got[4] + span
    assert got[6:] == [0, span]
    # a packet a little late (under half a span) or early stays where its stamp says
    assert X.time_base([0, span + span // 3, 2 * span - 10], [span] * 3) == [0, span + span // 3, 2 * span - 10]
    # the reader puts the firings of each packet on its base time
    rng = np.random.default_rng(0)
    dg = W.stream(rng, 0, 2, 10, 5_000_000_000, 10_000) + W.stream(rng, 0, 1, 10, 9_000_000_000, 10_000)
    (_, _, t, *_), _ = _read(tmp_path, 1, 0, dg)
    assert np.array_equal(t, np.arange(30) * 10_000 * 1e-9)  # the jump is closed: one continuous run


def test_gps_and_rejections(tmp_path):
    rng = np.random.default_rng(1)
    good = W.stream(rng, 2, 3, 8, 3_600_000_000_000 * 5 + 123_456_000, 4_167, ts_type=3)  # device 3: 240k/s
    bad_len = [W.header(2, 0, 3) + W.records(rng, 2, 8)[:-3]]      # not a whole number of records
    short = [bytes(10)]
    imu = [W.header(6, 0) + W.records(rng, 6, 1)]
    dg = good[:2] + bad_len + short + imu + good[2:]
    (_, _, t, xyz, *_), _ = _read(tmp_path, 3, 0, dg, truncate=5)   # the last datagram is cut
    assert len(t) == 2 * 8  # the third good packet was the truncated one
    # GPS stamps carry microseconds: packet k starts at its stamp truncated to the microsecond
    want = [((k * 8 * 4_167) // 1000 * 1000 + j * 4_167) * 1e-9 for k in range(2) for j in range(8)]
    assert np.array_equal(t, np.array(want))
    # LVXR1 (no profile byte) reads the same
    p1 = tmp_path / "v1.lvxr"
    W.write(p1, 3, None, good)
    assert len(X.read(p1.read_bytes())[2]) == 3 * 8


def test_dual_profile(tmp_path):
    rng = np.random.default_rng(2)
    ok = [W.header(0, 0) + W.records(rng, 0, 100)]
    wrong = [W.header(0, 500_000) + W.records(rng, 0, 98), W.header(2, 700_000) + W.records(rng, 2, 100)]
    (_, prof, t, xyz, refl, tag, valid, d), _ = _read(tmp_path, 1, 1, ok + wrong)
    assert prof == 1 and xyz.shape == (50, 2, 3) and len(t) == 50
    raw = ok[0][18:]
    second = struct.unpack_from("<iiiB", raw, 13)  # firing 0's second return is record 1
    assert np.array_equal(xyz[0, 1], np.array(second[:3], np.float32) * np.float32(1e-3))


def test_host_half_of_the_product(lw, tmp_path):
    """The product's framing and time base (its host half) against the oracle on a recording with resets,
    jumps, malformed packets and a truncated tail."""
    rng = np.random.default_rng(3)
    dg = (W.stream(rng, 0, 4, 20, 2_000_000_000, 10_000) + [W.header(0, 0) + W.records(rng, 0, 5)[:-2]] +
          W.stream(rng, 0, 3, 20, 7_000_000_000, 10_000) + W.stream(rng, 0, 3, 20, 100_000, 10_000, jitter=2000))
    p = tmp_path / "h.lvxr"
    W.write(p, 1, 0, dg, truncate=7)
    data = p.read_bytes()
    dev, prof, offs, lens = lw.lvxr._framing(data)
    assert (dev, prof) == (1, 0) and len(offs) == len(dg) - 1
    buf = np.frombuffer(data, np.uint8)
    ts = lw.lvxr._stamps(buf, offs)
    rec = (lens - 18) // 13
    base = lw.lvxr._clock(ts, rec * 10_000)
    t_or = X.read(data)[2]
    keep = (lens - 18) == rec * 13
    want = np.sort(np.concatenate([(b + np.arange(r) * 10_000) * 1e-9 for b, r in zip(base[keep], rec[keep])]))
    assert np.array_equal(want, t_or)


def test_dual_known_answer(tmp_path):
    """check_mid40_dual.py:33-41's expectations on the oracle."""
    p = tmp_path / "dual.lvxr"
    W.dual_known_answer(p)
    dev, prof, t, xyz, refl, tag, valid, d = X.read(p.read_bytes())
    assert prof == 1 and xyz.shape == (100, 2, 3)
    px, pa, pt, pf = X.points(t, xyz, refl, tag, valid, clean=False)
    assert len(pt) == 200
    assert np.array_equal((pa >> 16) & 255, np.tile([0, 1], 100))
    assert np.array_equal(pt[::2], pt[1::2]) and np.isclose(np.float32(pt[100]), np.float32(0.0005))
    assert np.array_equal(px[:, 0], (2000 + np.arange(200) % 100 * 100).astype(np.float32) * np.float32(1e-3))
    assert np.all(pa & 255 == 200)


def test_native_time_base_matches_oracle(lw):
    """lw_lvxr_clock (the product's native walk) against the oracle's time base on random stamp sequences with
    steady packets, jitter, resets backwards, jumps forward of more and less than a second, and empty spans."""
    rng = np.random.default_rng(12)
    for trial in range(40):
        n = int(rng.integers(1, 400))
        spans = rng.choice([0, 480_000, 960_000, 1_000_000], n).astype(np.int64)
        steps = spans + rng.integers(-3000, 3000, n)
        ev = rng.random(n)
        steps[ev < 0.03] = -rng.integers(1, 5_000_000_000, int((ev < 0.03).sum()))  # the clock goes back
        steps[(ev > 0.97)] += rng.integers(900_000_000, 3_000_000_000, int((ev > 0.97).sum()))  # pauses
        ts = np.maximum(np.cumsum(steps) + 10_000_000_000, 0).astype(np.uint64)
        got = lw.lvxr._clock(ts, spans)
        want = X.time_base([int(t) for t in ts], [int(s) for s in spans])
        assert np.array_equal(got, np.array(want, np.int64)), trial

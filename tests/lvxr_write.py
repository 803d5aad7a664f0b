"""Writes synthetic LVXR recordings for the reader's tests: the file layout of crates/livox/src/record.rs:1-5
(`LVXR2\\n`, dev_type, profile, then `host_ns:u64 len:u16 datagram`) with data packets laid out as point.rs
parses them (18-byte header, then records of one data type)."""

import struct

import numpy as np

SIZES = {0: 13, 1: 9, 2: 14, 3: 10, 4: 28, 5: 16, 6: 24, 7: 42, 8: 22}
RETURNS = {0: 1, 1: 1, 2: 1, 3: 1, 4: 2, 5: 2, 7: 3, 8: 3}


def header(dtype: int, ts_ns: int, ts_type: int = 0) -> bytes:
    """version, slot, id, rsvd, status u32, timestamp type, data type, timestamp (8 bytes)."""
    if ts_type == 3:  # GPS: year, month, day, hour, then microseconds within the hour
        hour, us = divmod(ts_ns // 1000, 3_600_000_000)
        ts = bytes([24, 1, 1, hour]) + struct.pack("<I", us)
    else:
        ts = struct.pack("<Q", ts_ns)
    return bytes([5, 1, 0, 0]) + struct.pack("<I", 0) + bytes([ts_type, dtype]) + ts


def records(rng, dtype: int, n: int, empty: float = 0.1) -> bytes:
    """n random records of a data type; about `empty` of the returns report nothing (zero xyz / depth)."""
    out = bytearray()
    for _ in range(n):
        if dtype in (0, 2):
            xyz = rng.integers(-30000, 30000, 3) * (rng.random() > empty)
            out += struct.pack("<iiiB", *xyz.tolist(), int(rng.integers(0, 256)))
            if dtype == 2:
                out += bytes([int(rng.integers(0, 4))])
        elif dtype in (1, 3):
            depth = int(rng.integers(100, 50000)) * (rng.random() > empty)
            out += struct.pack("<IHHB", depth, int(rng.integers(0, 18000)), int(rng.integers(0, 36000)),
                               int(rng.integers(0, 256)))
            if dtype == 3:
                out += bytes([int(rng.integers(0, 4))])
        elif dtype in (4, 7):
            for _ in range(RETURNS[dtype]):
                xyz = rng.integers(-30000, 30000, 3) * (rng.random() > empty)
                out += struct.pack("<iiiBB", *xyz.tolist(), int(rng.integers(0, 256)), int(rng.integers(0, 4)))
        elif dtype == 6:  # IMU: ignored by the reader
            out += bytes(24)
        else:
            out += struct.pack("<HH", int(rng.integers(0, 18000)), int(rng.integers(0, 36000)))
            for _ in range(RETURNS[dtype]):
                depth = int(rng.integers(100, 50000)) * (rng.random() > empty)
                out += struct.pack("<IBB", depth, int(rng.integers(0, 256)), int(rng.integers(0, 4)))
    return bytes(out)


def write(path, dev_type: int, profile: int | None, datagrams, truncate: int = 0):
    """datagrams: the data port's payloads in order. profile None writes a legacy LVXR1 file. truncate: bytes
    cut off the end (a recorder killed mid-write)."""
    body = bytearray(b"LVXR1\n" + bytes([dev_type]) if profile is None else b"LVXR2\n" + bytes([dev_type, profile]))
    for k, d in enumerate(datagrams):
        body += struct.pack("<QH", 1000 * k, len(d)) + d
    if truncate:
        body = body[:-truncate]
    with open(path, "wb") as f:
        f.write(bytes(body))


def stream(rng, dtype: int, packets: int, per: int, t0: int, interval: int, ts_type: int = 0, jitter: int = 0):
    """A run of packets of one data type, each `per` records, timed back to back from t0 ns."""
    out = []
    t = t0
    for _ in range(packets):
        out.append(header(dtype, t + (int(rng.integers(-jitter, jitter + 1)) if jitter else 0), ts_type) +
                   records(rng, dtype, per))
        t += per * interval
    return out


def dual_known_answer(path):
    """The reference's own dual-profile case (tests/check_mid40_dual.py:14-22): a Mid-40 dual recording of two
    packets 0.5 ms apart, each 100 plain records at x = 2.0 + 0.1 slot m, reflectivity 200. Expected: 200
    returns alternating return 0 and 1, both returns of a firing at one time, the second packet at 0.5 ms."""
    with open(path, "wb") as f:
        f.write(b"LVXR2\n\x01\x01")
        for packet in range(2):
            data = bytes([5, 0, 1, 0, 0, 0, 0, 0, 0, 0]) + struct.pack("<Q", packet * 500_000)
            data += b"".join(struct.pack("<iiiB", 2000 + slot * 100, 0, 0, 200) for slot in range(100))
            f.write(struct.pack("<QH", packet * 500_000, len(data)) + data)

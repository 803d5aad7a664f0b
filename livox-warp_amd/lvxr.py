"""LVXR recordings replayed through the device: the reference's firing-level reader (python/livox_warp/lvxr.py)
with the records decoded on the GPU (lw_lvxr_decode).

File layout (crates/livox/src/record.rs:1-5): `LVXR2\\n`, dev_type:u8, profile:u8 (LVXR1: no profile byte), then
`host_ns:u64 len:u16 datagram[len]` records. A point datagram is an 18-byte data header (version, slot, id,
rsvd, status u32, timestamp type, data type, timestamp u64) and records of one data type (cartesian or
spherical, one to three returns; the Mid-40 dual profile sends plain records in pairs, two returns of one
firing). The host walks the framing, computes each packet's time on the live path's time base (net.rs Sink:
reset when the clock goes back, shifted over forward jumps of a second or more; both walks native, in
liblwamd.so's host functions) and chooses the packets the reference keeps; the whole file goes to the device once and one thread per firing decodes it. Firings come
back as device tensors, in time order, as the reference's Firings holds them.
"""

from __future__ import annotations

import ctypes
import os
import time
from dataclasses import dataclass, replace

import numpy as np

from . import _abi

HEADER_LEN = 18
CLOCK_JUMP_NS = 1_000_000_000
PROFILE_STANDARD, PROFILE_MID40_DUAL = 0, 1
MID40_DUAL_RECORDS = 100
# data type -> (bytes per record, returns per record, spherical)
LAYOUTS = {0: (13, 1, False), 1: (9, 1, True), 2: (14, 1, False), 3: (10, 1, True), 4: (28, 2, False),
           5: (16, 2, True), 7: (42, 3, False), 8: (22, 3, True)}
FAST_DEVICES = {2, 3, 7}  # 240k samples/s; the rest 100k


def _torch():
    import torch

    return torch


@dataclass
class Firings:
    """Every firing of a recording, in time order (device tensors). R is the most returns any firing carries."""

    t: object  # (n,) float64 s since the time base
    xyz: object  # (n, R, 3) float32 m, sensor frame; zero where a return slot is empty
    refl: object  # (n, R) uint8
    tag: object  # (n, R) uint8
    valid: object  # (n, R) bool
    direction: object  # (n, 3) float32 unit firing direction, NaN where unknown
    dev_type: int
    profile: int
    interval: float  # s between firings

    @property
    def returns(self) -> int:
        return self.xyz.shape[1]

    @property
    def misses(self):
        return ~self.valid.any(dim=1)

    def head(self, seconds: float) -> Firings:
        k = self.t < self.t[0] + seconds
        return replace(self, t=self.t[k], xyz=self.xyz[k], refl=self.refl[k], tag=self.tag[k], valid=self.valid[k],
                       direction=self.direction[k])

    def points(self, clean: bool = True, min_range: float = 0.3):
        """The returns as flat device tensors (xyz, attr, t float64, firing index), in time order; attr packs
        reflectivity, tag << 8 and return << 16. clean: only untagged returns beyond min_range."""
        torch = _torch()
        f, r = torch.nonzero(self.valid, as_tuple=True)  # firing-major, as the firings are time-ordered
        xyz = self.xyz[f, r]
        tag = self.tag[f, r].to(torch.int32)
        attr = self.refl[f, r].to(torch.int32) | (tag << 8) | (r.to(torch.int32) << 16)
        if clean:
            keep = (tag == 0) & (torch.linalg.vector_norm(xyz, dim=1) > min_range)
            f, xyz, attr = f[keep], xyz[keep], attr[keep]
        return xyz.contiguous(), attr.contiguous(), self.t[f].contiguous(), f


def _framing(data):
    """(dev_type, profile, datagram offsets, lengths): the records of the file (bytes or a uint8 array), a
    truncated last one dropped (the walk is native: lw_lvxr_index)."""
    buf = np.ascontiguousarray(np.frombuffer(data, np.uint8) if isinstance(data, (bytes, bytearray)) else data)
    head = buf[:8].tobytes()
    if head[:6] == b"LVXR1\n" and len(head) >= 7:
        dev, profile, off = head[6], PROFILE_STANDARD, 7
    elif head[:6] == b"LVXR2\n" and len(head) >= 8:
        dev, profile, off = head[6], head[7], 8
    else:
        raise ValueError("not an LVXR recording")
    L = _abi.lib()
    count = ctypes.c_int64(0)
    _abi.check(L.lw_lvxr_index(buf.ctypes.data, len(buf), off, None, None, 0, ctypes.byref(count)), "lw_lvxr_index")
    offs = np.empty(count.value, np.int64)
    lens = np.empty(count.value, np.int32)
    if count.value:
        _abi.check(L.lw_lvxr_index(buf.ctypes.data, len(buf), off, offs.ctypes.data, lens.ctypes.data, count.value,
                                   ctypes.byref(count)), "lw_lvxr_index")
    return int(dev), int(profile), offs, lens.astype(np.int64)


def _stamps(buf: np.ndarray, offs: np.ndarray) -> np.ndarray:
    """Each packet's header timestamp in ns (type 3, GPS: hour and microseconds of the hour)."""
    le = np.array([1 << (8 * j) for j in range(8)], np.uint64)
    raw = buf[offs[:, None] + 10 + np.arange(8)].astype(np.uint64)
    ts = (raw * le).sum(axis=1, dtype=np.uint64)
    gps = buf[offs + 8] == 3
    if gps.any():
        hour = buf[offs[gps] + 13].astype(np.uint64)
        us = (buf[offs[gps][:, None] + 14 + np.arange(4)].astype(np.uint64) * le[:4]).sum(axis=1, dtype=np.uint64)
        ts[gps] = (hour * np.uint64(3_600_000_000) + us) * np.uint64(1000)
    return ts


def _clock(ts: np.ndarray, spans: np.ndarray) -> np.ndarray:
    """ns since the time base per packet: a packet earlier than the base starts a new base; one more than half
    a packet late by a second or more moves the base by the gap, so time stays continuous (lw_lvxr_clock)."""
    ts = np.ascontiguousarray(ts, np.uint64)
    spans = np.ascontiguousarray(spans, np.int64)
    out = np.empty(len(ts), np.int64)
    _abi.check(_abi.lib().lw_lvxr_clock(ts.ctypes.data, spans.ctypes.data, len(ts), CLOCK_JUMP_NS, out.ctypes.data),
               "lw_lvxr_clock")
    return out


def read_firings(path: str | os.PathLike, device=None, stream=None, timing: dict | None = None) -> Firings:
    """Every firing of an LVXR recording, decoded on `device` (the reference's read_firings). timing: a dict
    that receives the host framing time (s), the decode kernel's time (ms, HIP events) and the byte counts."""
    torch = _torch()
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    t0 = time.perf_counter()
    with open(path, "rb") as f:
        buf = np.fromfile(f, np.uint8)  # writable: it goes to the device without another copy
    dev_type, profile, offs, lens = _framing(buf)
    interval = 4_167 if dev_type in FAST_DEVICES else 10_000
    dual = profile == PROFILE_MID40_DUAL
    ok = lens >= HEADER_LEN
    offs, lens = offs[ok], lens[ok]
    dtypes = buf[offs + 9].astype(np.int64)
    point = np.isin(dtypes, list(LAYOUTS))
    offs, lens, dtypes = offs[point], lens[point], dtypes[point]
    sizes = np.array([LAYOUTS[int(d)][0] for d in dtypes], np.int64)
    records = (lens - HEADER_LEN) // np.maximum(sizes, 1)
    # every point packet moves the clock, the malformed ones too, before any is rejected
    fpp = records // 2 if dual else records
    t_pk = _clock(_stamps(buf, offs), fpp * interval) if len(offs) else np.zeros(0, np.int64)
    whole = (lens - HEADER_LEN) == records * sizes
    if dual:
        whole &= np.isin(dtypes, [0, 1]) & (records == MID40_DUAL_RECORDS)
    offs, lens, dtypes, t_pk, fpp = offs[whole], lens[whole], dtypes[whole], t_pk[whole], fpp[whole]
    if not len(offs):
        raise ValueError(f"{path}: no point packets")
    R = max([LAYOUTS[int(d)][1] for d in np.unique(dtypes)] + [2 if dual else 1])
    # rows in the reference's concatenation order (by data type, then datagram length, then file order), so a
    # stable sort by time breaks ties as it does
    order = np.lexsort((np.arange(len(offs)), lens, dtypes))
    first = np.zeros(len(offs), np.int64)
    first[order] = np.concatenate([[0], np.cumsum(fpp[order])[:-1]])
    n = int(fpp.sum())
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    # every device buffer is allocated and uploaded on the decode's stream (after what the caller's current
    # stream has queued), so the caching allocator ties each to the stream that reads it
    s.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(s):
        bytes_dev = torch.from_numpy(buf).to(dev)
        tab = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in
               (offs, dtypes.astype(np.int32), fpp.astype(np.int32), t_pk, first)]
        t = torch.empty(n, dtype=torch.float64, device=dev)
        xyz = torch.empty((n, R, 3), dtype=torch.float32, device=dev)
        refl = torch.empty((n, R), dtype=torch.uint8, device=dev)
        tag = torch.empty((n, R), dtype=torch.uint8, device=dev)
        valid = torch.empty((n, R), dtype=torch.uint8, device=dev)
        direction = torch.empty((n, 3), dtype=torch.float32, device=dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)] if timing is not None else None
    if ev:
        timing["host_s"] = time.perf_counter() - t0
        ev[0].record(s)
    _abi.check(_abi.lib().lw_lvxr_decode(bytes_dev.data_ptr(), *[a.data_ptr() for a in tab], len(offs),
                                         1 if dual else 0, interval, R, t.data_ptr(), xyz.data_ptr(), refl.data_ptr(),
                                         tag.data_ptr(), valid.data_ptr(), direction.data_ptr(),
                                         ctypes.c_void_p(s.cuda_stream)), "lw_lvxr_decode")
    if ev:
        ev[1].record(s)
        ev[1].synchronize()
        timing.update(decode_ms=ev[0].elapsed_time(ev[1]), firings=n, returns=R, packets=len(offs),
                      record_bytes=int((lens - HEADER_LEN).sum()), file_bytes=len(buf))
    with torch.cuda.stream(s):
        idx = torch.sort(t, stable=True).indices
        out = Firings(t=t[idx], xyz=xyz[idx], refl=refl[idx], tag=tag[idx], valid=valid[idx].bool(),
                      direction=direction[idx], dev_type=dev_type, profile=profile, interval=interval * 1e-9)
    if s != torch.cuda.current_stream(dev):
        # the caller reads the result on its current stream: that stream waits for the decode and the sort
        torch.cuda.current_stream(dev).wait_stream(s)
        for a in (out.t, out.xyz, out.refl, out.tag, out.valid, out.direction):
            a.record_stream(torch.cuda.current_stream(dev))
    return out

"""The odometry on its own thread and private stream: the reference's OdomWorker contract
(python/livox_warp/slam_worker.py:100-338).

    caller thread (its own stream)                      odometry thread (private stream "odom")
    staged batch ──own()──> snapshot on "odom-in" ────> Odometry.push: append, deskew, register, insert
    FrameResult snapshots <──own()── on "odom-out" <─── the deskewed frame (on_frame)
    pose table (host, locked) <──────────────────────── set_pose(k, T, twist)

- three streams, created here on the calling thread, at the default priority;
- a batch crosses into the worker as an owned device copy (the caller's staging buffers are rewritten by its
  next batch), and a finished frame crosses back as an owned copy from a fixed pool; a full queue or an
  exhausted pool drops the item and counts it (own-or-drop). Poses are small host data and always get through;
- ordering across streams is by events: a copy waits for its producer's stream, the producer's next writes
  wait for the copy to have read, and the consumer's kernels wait for the copy (`FrameResult.wait`);
- results carry a generation: reset() and restart_clock() bump it, drop the batches still queued for the old
  one, and anything produced for an older generation is dropped on arrival.

The reference builds this on an external helper (warp_worker.WarpWorker: own, pools, release_source_on); here
the same roles are torch streams and events (`_Lane`, `_Pool`, `_Snap`).
"""

from __future__ import annotations

import queue
import sys
import threading
import time
import traceback
from dataclasses import dataclass, field

import numpy as np

from .odom import Odometry


def _torch():
    import torch

    return torch


class _Pool:
    """Fixed device buffers of one shape and dtype, handed out one at a time (None when all are out)."""

    def __init__(self, shape, dtype, slots: int, device):
        torch = _torch()
        self.bufs = [torch.empty(shape, dtype=dtype, device=device) for _ in range(slots)]
        self._free = list(range(slots))
        self._lock = threading.Lock()

    def take(self):
        with self._lock:
            return self._free.pop() if self._free else None

    def give(self, slot: int):
        with self._lock:
            self._free.append(slot)

    @property
    def free_slots(self) -> int:
        with self._lock:
            return len(self._free)


class _Snap:
    """An owned copy: its tensor (the first n rows), the event its copy completes with, and its pool slot."""

    def __init__(self, array, event, pool: _Pool | None = None, slot: int | None = None):
        self.array = array
        self.event = event
        self._pool, self._slot = pool, slot

    def wait(self, stream):
        """The consumer side: work queued on `stream` from now on runs after the copy."""
        stream.wait_event(self.event)

    def release_source_on(self, stream):
        """The producer side: work queued on `stream` from now on (the next writes to the source) runs after
        the copy has read it."""
        stream.wait_event(self.event)

    def release(self):
        if self._pool is not None and self._slot is not None:
            self._pool.give(self._slot)
        self._pool = self._slot = None
        self.array = None


class _Lane:
    """A private stream that makes owned copies (WarpWorker's role in the reference)."""

    def __init__(self, name: str, device):
        torch = _torch()
        self.name = name
        self.device = device
        self.stream = torch.cuda.Stream(device=device)  # default priority
        self.copies = 0

    def own_many(self, arrays, produced_on, pools=None):
        """Owned copies of several device tensors at once (one wait on the producer, one completion event they
        share): a list of snaps, or None when a pool is exhausted (every slot taken for them is given back)."""
        torch = _torch()
        pools = list(pools) if pools is not None else [None] * len(arrays)
        dsts, slots = [], []
        for a, pool in zip(arrays, pools):
            if pool is None:
                with torch.cuda.stream(self.stream):
                    dsts.append(torch.empty_like(a))
                slots.append(None)
                continue
            slot = pool.take()
            if slot is None:
                for p, sl in zip(pools, slots):
                    if sl is not None:
                        p.give(sl)
                return None
            dsts.append(pool.bufs[slot][: a.shape[0]])
            slots.append(slot)
        self.stream.wait_stream(produced_on)
        with torch.cuda.stream(self.stream):
            for d, a in zip(dsts, arrays):
                d.copy_(a, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.stream)
        for a in arrays:
            a.record_stream(self.stream)  # each source stays allocated until the copy has read it
        self.copies += len(arrays)
        return [_Snap(d, ev, p, sl) for d, p, sl in zip(dsts, pools, slots)]

    def status(self) -> dict:
        return {"name": self.name, "copies": self.copies}


def _events(snaps):
    """The distinct completion events of a group of snaps (copies made together share one)."""
    out = []
    for sn in snaps or ():
        if all(sn.event is not e.event for e in out):
            out.append(sn)
    return out


@dataclass
class FrameResult:
    gen: int
    k: int  # frame id
    T: np.ndarray  # sensor-to-world at mid-frame
    xi: np.ndarray  # body-frame motion over the frame (rho, theta)
    m: int  # points in the frame
    snaps: list | None  # owned (xyz, attr, t, frame) of the deskewed frame, or None if dropped
    stats: dict = field(default_factory=dict)

    def arrays(self):
        return [s.array for s in self.snaps]

    def release(self):
        for s in self.snaps or ():
            s.release()
        self.snaps = None


# items of the input queue, caller thread -> odometry thread


@dataclass
class _Batch:
    gen: int  # the caller-side generation it was submitted under
    epoch: int  # the caller's pose epoch, written with every pose of this batch
    snaps: list  # owned (xyz, attr, t) copies of the staged points
    n: int
    t_min: float
    t_max: float


@dataclass
class _Reset:
    gen: int
    T0: np.ndarray


@dataclass
class _Restart:
    gen: int


@dataclass
class _Config:
    params: dict


class _Stop:
    """Ends the odometry thread."""


class OdomWorker:
    """An Odometry on its own thread: the caller submits staged batches and polls FrameResults."""

    def __init__(self, pipe, device=None, max_batches: int = 120, out_slots: int = 8, frame_cap: int = 1 << 18,
                 reg_slots: int = 1 << 21, caller_stream=None):
        torch = _torch()
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.pipe = pipe
        # the caller's stream: where its staged batches are produced and where it reads results
        self.render_stream = caller_stream if caller_stream is not None else torch.cuda.current_stream(self.device)
        self.w_in, self.w, self.w_out = (_Lane(n, self.device) for n in ("odom-in", "odom", "odom-out"))
        self.odom = Odometry(pipe, self.device, reg_slots=reg_slots, frame_cap=frame_cap, stream=self.w.stream)
        o = self.odom
        self.pools = [_Pool(shape, dt, out_slots, self.device) for shape, dt in
                      (((o.frame_cap, 3), torch.float32), ((o.frame_cap,), torch.int32),
                       ((o.frame_cap,), torch.float32), ((o.frame_cap,), torch.int32))]
        torch.cuda.synchronize(self.device)
        self.max_batches = max_batches
        self.q_in: queue.Queue = queue.Queue()
        self.q_out: queue.Queue = queue.Queue()
        self.gen = 0  # caller-side generation
        self._worker_gen = 0  # the generation the odometry thread is working on (written by that thread only)
        self._pending = 0  # batches queued and not yet processed
        self._lock = threading.Lock()
        self.drops = {"input_full": 0, "input_own": 0, "output_pool": 0, "stale": 0, "error": 0}
        self.processed = 0  # batches the odometry thread finished
        self.frames = 0  # frame results the caller polled
        self.last_error = ""
        self.busy_ms = 0.0  # moving average of worker time per batch
        self._consumed: list[FrameResult] = []
        self._stop = False
        self._thread = threading.Thread(target=self._run, name="livox-odometry", daemon=True)
        self._thread.start()

    # ---- caller thread -----------------------------------------------------------------------------------

    def submit(self, sx, sa, st, n: int, t_min: float, t_max: float, epoch: int = 0) -> bool:
        """Hand one staged batch (device tensors: points, attributes, times) to the worker as owned copies.
        False: dropped (counted)."""
        if n <= 0:
            return True
        with self._lock:
            if self._pending >= self.max_batches:
                self.drops["input_full"] += 1
                return False
        snaps = self.w_in.own_many([sx[:n], sa[:n], st[:n]], produced_on=self.render_stream)
        if snaps is None:
            with self._lock:
                self.drops["input_own"] += 1
            return False
        snaps[0].release_source_on(self.render_stream)  # the caller's next writes wait for the copies
        with self._lock:
            self._pending += 1
        self.q_in.put(_Batch(self.gen, epoch, snaps, n, t_min, t_max))
        return True

    def _new_generation(self):
        """Bump the generation and drop every batch still queued for the old one (control items keep their
        order)."""
        self.gen += 1
        keep = []
        while True:
            try:
                it = self.q_in.get_nowait()
            except queue.Empty:
                break
            if isinstance(it, _Batch):
                with self._lock:
                    self._pending -= 1
                    self.drops["stale"] += 1
            else:
                keep.append(it)
        for it in keep:
            self.q_in.put(it)
        self._drain_out()

    def reset(self, T0: np.ndarray):
        """Forget map and trajectory; the next frame is posed at T0. The old generation's batches and results
        are discarded."""
        self._new_generation()
        self.q_in.put(_Reset(self.gen, np.array(T0, dtype=np.float64)))

    def restart_clock(self):
        """The live clock restarted: keep the map and pose, restart frame counting at rest."""
        self._new_generation()
        self.q_in.put(_Restart(self.gen))

    def configure(self, **params):
        """Set Odometry tuning attributes between batches; a name that is not a scalar attribute raises
        AttributeError here, before anything is queued."""
        for name in params:
            if not isinstance(getattr(self.odom, name, None), (int, float, str)):
                raise AttributeError(f"Odometry has no tuning attribute {name!r}")
        self.q_in.put(_Config(dict(params)))

    def poll(self) -> list[FrameResult]:
        """Finished frames of the current generation, their copies ordered before the caller's stream's next
        work; call release_consumed() once the caller's stream has passed the work that reads them."""
        out = []
        while True:
            try:
                r = self.q_out.get_nowait()
            except queue.Empty:
                break
            if r.gen != self.gen:
                r.release()
                with self._lock:
                    self.drops["stale"] += 1
                continue
            for s in _events(r.snaps):
                s.wait(self.render_stream)
            self.frames += 1
            self._consumed.append(r)
            out.append(r)
        return out

    def release_consumed(self):
        for r in self._consumed:
            r.release()
        self._consumed.clear()

    def _drain_out(self):
        while True:
            try:
                self.q_out.get_nowait().release()
            except queue.Empty:
                break

    def status(self) -> dict:
        with self._lock:
            return {"pending": self._pending, "processed": self.processed, "frames": self.frames,
                    "drops": dict(self.drops), "busy_ms": round(self.busy_ms, 2), "last_error": self.last_error,
                    "pool_free": min(p.free_slots for p in self.pools),
                    "workers": [w.status() for w in (self.w_in, self.w, self.w_out)]}

    def idle(self) -> bool:
        with self._lock:
            return self._pending == 0 and self.q_in.empty()

    def close(self):
        if self._stop:
            return
        self._stop = True
        self._new_generation()  # the stop does not wait behind queued batches
        self.q_in.put(_Stop())
        self._thread.join(timeout=10.0)
        self._drain_out()
        self.release_consumed()
        if self._thread.is_alive():
            print("OdomWorker.close: the odometry thread did not stop", file=sys.stderr)

    # ---- odometry thread ---------------------------------------------------------------------------------

    def _run(self):
        torch = _torch()
        torch.cuda.set_device(self.device)
        while True:
            item = self.q_in.get()
            if isinstance(item, _Stop):
                break
            try:
                if isinstance(item, _Reset):
                    self._worker_gen = item.gen
                    self.odom.reset(item.T0)
                    self.odom.synchronize()
                elif isinstance(item, _Restart):
                    self._worker_gen = item.gen
                    self.odom.restart_clock()
                elif isinstance(item, _Config):
                    for name, value in item.params.items():
                        setattr(self.odom, name, value)
                elif isinstance(item, _Batch):
                    self._batch(item)
            except Exception as e:  # noqa: BLE001 - one bad batch must not end the worker
                with self._lock:
                    self.drops["error"] += 1
                    self.last_error = f"{type(e).__name__}: {e}"
                traceback.print_exc()

    def _batch(self, item: _Batch):
        try:
            if item.gen != self._worker_gen:
                with self._lock:
                    self.drops["stale"] += 1
                return
            t0 = time.perf_counter()
            s = self.w.stream
            for sn in _events(item.snaps):
                sn.wait(s)  # the odometry's kernels wait for the input copies
            sx, sa, st = (sn.array for sn in item.snaps)
            self.odom.epoch = item.epoch
            self.odom.push(sx, sa, st, item.n, item.t_min, item.t_max, self._on_frame)
            for sn in item.snaps:
                # the allocator keeps each copy until the odometry's work queued so far has read it, so the
                # thread moves on without waiting for the stream
                sn.array.record_stream(s)
                sn.release()
            dt = (time.perf_counter() - t0) * 1e3
            with self._lock:
                self.processed += 1
                self.busy_ms = dt if self.processed == 1 else 0.9 * self.busy_ms + 0.1 * dt
        finally:
            with self._lock:
                self._pending -= 1

    def _on_frame(self, k, T, fx, fa, ft, ff, m):
        """On the odometry thread when a frame's pose is final: copy its points out for the caller."""
        snaps = self.w_out.own_many([fx[:m], fa[:m], ft[:m], ff[:m]], produced_on=self.w.stream, pools=self.pools)
        if snaps is None:
            with self._lock:
                self.drops["output_pool"] += 1
        else:
            snaps[0].release_source_on(self.w.stream)  # the next frame's writes to fx..ff wait for the copies
        o = self.odom
        self.q_out.put(FrameResult(self._worker_gen, k, np.array(T, dtype=np.float64),
                                   np.array(o.xi, dtype=np.float32), m, snaps, dict(o.stats)))

"""The odometry on its own thread: the reference's OdomWorker contract (python/livox_warp/slam_worker.py:100-338:
submit / poll / release_consumed / reset / restart_clock / configure / status / close, FrameResult, generations,
own-or-drop counters), built on two HIP streams and fixed device slots.

    caller thread, caller stream                 odometry thread, stream "odom"
    submit(): copy into a staging slot  ──────>  Odometry.push (append, deskew, register, insert) reads the slot;
              on stream "odom-in"                the slot is free again once "odom" has passed that work
    poll(): FrameResults; the caller    <──────  on_frame: the frame's points copied into an output slot on
            stream waits for the copies          "odom" itself, right after the kernels that wrote them

Ownership without waits on the host:
  * a staging slot is written on "odom-in" only after "odom" has finished the push that read its previous batch
    (the slot keeps that push's completion event); an input copy waits for the caller's stream, and the caller's
    next writes to its own buffer wait for the copy;
  * the frame buffers the Odometry lends to on_frame are rewritten by the next frame's kernels on the same
    stream, so copies queued on that stream between the two need no event to be ordered against either; one
    event after them lets the caller's stream wait for the copies; a released output slot keeps an event on the
    caller's stream, which the next copy into it waits for.
Slots are allocated once (staging slots the first time each is used, up to max_batches); a batch larger than a
staging slot, a full queue, and a frame that finds no free output slot are dropped and counted.
"""

from __future__ import annotations

import collections
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


# ---- device slots -----------------------------------------------------------------------------------------------


class _Staging:
    """One staged batch's device copy: xyz (cap, 3) float32, attributes (cap,) int32, times (cap,) float32."""

    def __init__(self, cap: int, device):
        torch = _torch()
        self.xyz = torch.empty((cap, 3), dtype=torch.float32, device=device)
        self.attr = torch.empty(cap, dtype=torch.int32, device=device)
        self.t = torch.empty(cap, dtype=torch.float32, device=device)
        self.cap = cap
        self.filled = torch.cuda.Event()  # the copy into the slot is done (recorded on "odom-in")
        self.consumed = None  # the push that read the slot is done (recorded on "odom"); None: never used


class _Outputs:
    """A frame's owned copy: xyz, attributes, times and frame ids of up to cap points, and the events that order
    its reuse."""

    def __init__(self, cap: int, device):
        torch = _torch()
        self.arrays = (torch.empty((cap, 3), dtype=torch.float32, device=device),
                       torch.empty(cap, dtype=torch.int32, device=device),
                       torch.empty(cap, dtype=torch.float32, device=device),
                       torch.empty(cap, dtype=torch.int32, device=device))
        self.copied = torch.cuda.Event()  # recorded on "odom" after the copies
        self.read = None  # recorded on the caller's stream when the caller released it


class _Owned:
    """One array of a FrameResult: the first m rows of an output slot's array."""

    def __init__(self, array, result):
        self.array = array
        self._result = result

    def wait(self, stream):
        """Work queued on `stream` from now on runs after the copy."""
        stream.wait_event(self._result._slot.copied)

    def release(self):
        self._result.release()


@dataclass
class FrameResult:
    gen: int
    k: int  # frame id
    T: np.ndarray  # sensor-to-world at mid-frame
    xi: np.ndarray  # body-frame motion over the frame (rho, theta)
    m: int  # points in the frame
    snaps: list | None  # owned (xyz, attr, t, frame) of the deskewed frame, or None if dropped
    stats: dict = field(default_factory=dict)
    _slot: object = None
    _home: object = None

    def arrays(self):
        return [s.array for s in self.snaps]

    def release(self):
        """Give the output slot back; the next copy into it waits for what the caller's stream queued so far."""
        slot, home = self._slot, self._home
        self._slot = self._home = None
        self.snaps = None
        if slot is not None:
            home._give_back(slot)


# ---- the worker -------------------------------------------------------------------------------------------------

_BATCH, _RESET, _RESTART, _CONFIG, _STOP = range(5)


class OdomWorker:
    """An Odometry on its own thread: the caller submits staged batches and polls FrameResults."""

    def __init__(self, pipe, device=None, max_batches: int = 120, out_slots: int = 8, frame_cap: int = 1 << 18,
                 reg_slots: int = 1 << 21, caller_stream=None, batch_cap: int | None = None):
        torch = _torch()
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.pipe = pipe
        # the caller's stream: where its staged batches are produced and where it reads results
        self.render_stream = caller_stream if caller_stream is not None else torch.cuda.current_stream(self.device)
        self.in_stream = torch.cuda.Stream(device=self.device)  # default priority, as every stream here
        self.odom_stream = torch.cuda.Stream(device=self.device)
        self.odom = Odometry(pipe, self.device, reg_slots=reg_slots, frame_cap=frame_cap, stream=self.odom_stream)
        self.max_batches = max_batches
        self.batch_cap = int(batch_cap) if batch_cap is not None else int(self.odom.frame_cap)
        self._staging: list[_Staging] = []  # grown on demand up to max_batches
        self._staging_free: list[_Staging] = []
        self._outputs = [_Outputs(self.odom.frame_cap, self.device) for _ in range(out_slots)]
        self._outputs_free = list(self._outputs)
        torch.cuda.synchronize(self.device)
        self.gen = 0  # caller-side generation
        self._worker_gen = 0  # the generation the odometry thread runs (its thread writes it)
        self._cmds: collections.deque = collections.deque()
        self._results: collections.deque = collections.deque()
        self._cv = threading.Condition()
        self._lock = threading.Lock()  # counters and slot lists
        self._pending = 0
        self.drops = {"input_full": 0, "input_own": 0, "output_pool": 0, "stale": 0, "error": 0}
        self.processed = 0
        self.frames = 0
        self.copies = 0
        self.last_error = ""
        self.busy_ms = 0.0
        self._consumed: list[FrameResult] = []
        self._closed = False
        self._handlers = {_BATCH: self._do_batch, _RESET: self._do_reset, _RESTART: self._do_restart,
                          _CONFIG: self._do_config}
        self._thread = threading.Thread(target=self._loop, name="livox-odometry", daemon=True)
        self._thread.start()

    # ---- caller thread ------------------------------------------------------------------------------------------

    def _post(self, kind: int, payload, front: bool = False):
        with self._cv:
            (self._cmds.appendleft if front else self._cmds.append)((kind, payload))
            self._cv.notify()

    def _take_staging(self, n: int):
        with self._lock:
            if n > self.batch_cap:
                return None
            if self._staging_free:
                return self._staging_free.pop()
            if len(self._staging) < self.max_batches:
                s = _Staging(self.batch_cap, self.device)
                self._staging.append(s)
                return s
            return None

    def submit(self, sx, sa, st, n: int, t_min: float, t_max: float, epoch: int = 0) -> bool:
        """Copy one staged batch (device tensors: points, attributes, times) into a staging slot and queue it.
        False: dropped (counted)."""
        if n <= 0:
            return True
        with self._lock:
            if self._pending >= self.max_batches:
                self.drops["input_full"] += 1
                return False
        slot = self._take_staging(n)
        if slot is None:
            with self._lock:
                self.drops["input_own"] += 1
            return False
        torch = _torch()
        s = self.in_stream
        s.wait_stream(self.render_stream)  # after the caller's kernels that staged the batch
        if slot.consumed is not None:
            s.wait_event(slot.consumed)  # after the push that read this slot's previous batch
        with torch.cuda.stream(s):
            slot.xyz[:n].copy_(sx[:n], non_blocking=True)
            slot.attr[:n].copy_(sa[:n].view(torch.int32) if sa.dtype != torch.int32 else sa[:n], non_blocking=True)
            slot.t[:n].copy_(st[:n], non_blocking=True)
            slot.filled.record(s)
        self.render_stream.wait_event(slot.filled)  # the caller's next writes to its buffers come after the copy
        with self._lock:
            self._pending += 1
            self.copies += 3
        self._post(_BATCH, (self.gen, epoch, slot, n, float(t_min), float(t_max)))
        return True

    def _new_generation(self):
        """Bump the generation; queued batches of the old one are dropped (their slots freed), control commands
        keep their order, and results not yet polled are discarded."""
        self.gen += 1
        with self._cv:
            kept = collections.deque()
            for kind, payload in self._cmds:
                if kind == _BATCH:
                    self._recycle(payload[2])
                    with self._lock:
                        self._pending -= 1
                        self.drops["stale"] += 1
                else:
                    kept.append((kind, payload))
            self._cmds = kept
        self._discard_results()

    def _recycle(self, slot: _Staging):
        with self._lock:
            self._staging_free.append(slot)

    def reset(self, T0: np.ndarray):
        """Forget map and trajectory; the next frame is posed at T0. The old generation's batches and results
        are discarded."""
        self._new_generation()
        self._post(_RESET, (self.gen, np.array(T0, dtype=np.float64)))

    def restart_clock(self):
        """The live clock restarted: keep the map and pose, restart frame counting at rest."""
        self._new_generation()
        self._post(_RESTART, (self.gen,))

    def configure(self, **params):
        """Set Odometry tuning attributes between batches; a name that is not a scalar attribute raises
        AttributeError here, before anything is queued."""
        bad = [k for k in params if not isinstance(getattr(self.odom, k, None), (int, float, str))]
        if bad:
            raise AttributeError(f"Odometry has no tuning attribute {bad[0]!r}")
        self._post(_CONFIG, dict(params))

    def poll(self) -> list[FrameResult]:
        """Finished frames of the current generation; the caller's stream waits for their copies. Call
        release_consumed() once the caller's stream has queued the work that reads them."""
        out = []
        while True:
            with self._cv:
                if not self._results:
                    break
                r = self._results.popleft()
            if r.gen != self.gen:
                r.release()
                with self._lock:
                    self.drops["stale"] += 1
                continue
            if r._slot is not None:
                self.render_stream.wait_event(r._slot.copied)
            self.frames += 1
            self._consumed.append(r)
            out.append(r)
        return out

    def release_consumed(self):
        for r in self._consumed:
            r.release()
        self._consumed.clear()

    def _give_back(self, slot: _Outputs):
        """An output slot returns from the caller: its next copy waits for the caller's stream as it stands."""
        torch = _torch()
        ev = torch.cuda.Event()
        ev.record(self.render_stream)
        slot.read = ev
        with self._lock:
            self._outputs_free.append(slot)

    def _discard_results(self):
        while True:
            with self._cv:
                if not self._results:
                    return
                r = self._results.popleft()
            r.release()

    def status(self) -> dict:
        with self._lock:
            return {"pending": self._pending, "processed": self.processed, "frames": self.frames,
                    "drops": dict(self.drops), "busy_ms": round(self.busy_ms, 2), "last_error": self.last_error,
                    "pool_free": len(self._outputs_free), "staging_slots": len(self._staging),
                    "copies": self.copies}

    def idle(self) -> bool:
        with self._cv:
            queued = len(self._cmds)
        with self._lock:
            return self._pending == 0 and queued == 0

    def close(self):
        """Stop the thread (queued batches are dropped), then free the Odometry's context."""
        if self._closed:
            return
        self._closed = True
        self._new_generation()
        self._post(_STOP, None)
        self._thread.join(timeout=10.0)
        self._discard_results()
        self.release_consumed()
        if self._thread.is_alive():
            print("OdomWorker.close: the odometry thread did not stop", file=sys.stderr)
            return
        self.odom_stream.synchronize()
        self.odom.close()

    # ---- odometry thread ----------------------------------------------------------------------------------------

    def _loop(self):
        torch = _torch()
        torch.cuda.set_device(self.device)
        while True:
            with self._cv:
                while not self._cmds:
                    self._cv.wait()
                kind, payload = self._cmds.popleft()
            if kind == _STOP:
                return
            try:
                self._handlers[kind](payload)
            except Exception as e:  # noqa: BLE001 - one bad batch must not end the worker
                with self._lock:
                    self.drops["error"] += 1
                    self.last_error = f"{type(e).__name__}: {e}"
                traceback.print_exc()

    def _do_reset(self, payload):
        self._worker_gen, T0 = payload
        self.odom.reset(T0)

    def _do_restart(self, payload):
        (self._worker_gen,) = payload
        self.odom.restart_clock()

    def _do_config(self, params):
        for name, value in params.items():
            setattr(self.odom, name, value)

    def _do_batch(self, payload):
        gen, epoch, slot, n, t_min, t_max = payload
        torch = _torch()
        read = False
        try:
            if gen != self._worker_gen:
                with self._lock:
                    self.drops["stale"] += 1
                return
            t0 = time.perf_counter()
            s = self.odom_stream
            s.wait_event(slot.filled)
            read = True
            self.odom.epoch = epoch
            self.odom.push(slot.xyz, slot.attr, slot.t, n, t_min, t_max, self._on_frame)
            dt = (time.perf_counter() - t0) * 1e3
            with self._lock:
                self.processed += 1
                self.busy_ms = dt if self.processed == 1 else 0.9 * self.busy_ms + 0.1 * dt
        finally:
            if read:  # also after a failed push: whatever it queued may still read the slot
                ev = torch.cuda.Event()
                ev.record(self.odom_stream)  # the slot may be refilled once the stream has passed this point
                slot.consumed = ev
            self._recycle(slot)
            with self._lock:
                self._pending -= 1

    def _on_frame(self, k, T, fx, fa, ft, ff, m):
        """On the odometry thread when a frame's pose is final: copy its points into an output slot on the
        odometry stream, behind the kernels that wrote them and ahead of the next frame's."""
        torch = _torch()
        o = self.odom
        with self._lock:
            slot = self._outputs_free.pop() if self._outputs_free else None
            if slot is None:
                self.drops["output_pool"] += 1
        r = FrameResult(self._worker_gen, k, np.array(T, dtype=np.float64), np.array(o.xi, dtype=np.float32), m,
                        None, dict(o.stats))
        if slot is not None:
            s = self.odom_stream
            if slot.read is not None:
                s.wait_event(slot.read)
            with torch.cuda.stream(s):
                for dst, src in zip(slot.arrays, (fx, fa, ft, ff)):
                    dst[:m].copy_(src[:m], non_blocking=True)
                slot.copied.record(s)
            with self._lock:
                self.copies += 4
            r._slot, r._home = slot, self
            r.snaps = [_Owned(a[:m], r) for a in slot.arrays]
        with self._cv:
            self._results.append(r)

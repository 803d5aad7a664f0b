"""The learned pose-increment prior: a BF16 MLP on the matrix cores (lw_mlp_*) that predicts each frame's pose
increment from the last four increments and the frame's range / elevation histograms, in place of the
constant-velocity guess of Odometry.predict (python/livox_warp/odom.py:591-592, used at odom.py:666).

Plain ICP keeps the last word: the registration starts from the learned guess, and when the first pass's residual
at that start exceeds `fallback` times the running median of those residuals, the frame is registered again
from the constant-velocity guess (Odometry._estimate).

The network predicts the change of the increment: d_k = d_{k-1} + (y * sd_t + mu_t), with d the body twist
(translation, rotation vector) between consecutive frame poses, as Odometry._update_velocity forms it. Its
labels are the poses of the plain (constant-velocity) GPU odometry on synthetic sequences (`collect`), and it
is trained by the library's own forward / backward / Adam kernels (`fit`).
"""

from __future__ import annotations

import collections
import ctypes
import math

import numpy as np

from . import _abi

MOTION, IN, OUT = 24, 152, 6
R_MAX = 20.0                # range histogram over [0, 20) m
E_MAX = math.radians(20.0)  # elevation histogram over [-20, 20) degrees


def _torch():
    import torch

    return torch


def rotvec_of(R: np.ndarray) -> np.ndarray:
    """Rotation vector of a rotation matrix: angle atan2(|v| / 2, (tr R - 1) / 2) along v = vee(R - R^T)
    (accurate at small angles; the frame increments here are far from pi)."""
    v = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    s = 0.5 * float(np.sqrt(v @ v))
    c = 0.5 * (float(R[0, 0] + R[1, 1] + R[2, 2]) - 1.0)
    th = math.atan2(s, c)
    if th > 3.0:  # near pi: v loses the axis
        from scipy.spatial.transform import Rotation

        return Rotation.from_matrix(R).as_rotvec()
    return v * (0.5 if s < 1e-12 else th / (2.0 * s))


def rotmat_of(w: np.ndarray) -> np.ndarray:
    """Rodrigues: exp of the skew matrix of w."""
    th = float(np.sqrt(w @ w))
    K = np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    if th < 1e-8:
        return np.eye(3) + K + 0.5 * (K @ K)
    return np.eye(3) + (math.sin(th) / th) * K + ((1.0 - math.cos(th)) / (th * th)) * (K @ K)


def body_twist(d: np.ndarray) -> np.ndarray:
    """(translation, rotation vector) of a 4x4 increment, float64."""
    return np.concatenate([d[:3, 3], rotvec_of(d[:3, :3])])


def twist_matrix(xi: np.ndarray) -> np.ndarray:
    out = np.eye(4)
    out[:3, :3] = rotmat_of(np.asarray(xi[3:], np.float64))
    out[:3, 3] = xi[:3]
    return out


def init_params(rng, out_scale: float = 0.1) -> np.ndarray:
    """He-normal hidden weights, zero biases, a small last layer; the padded rows / columns stay zero."""
    P = _abi.lib().lw_mlp_param_count()
    p = np.zeros(P, np.float32)
    o = 0
    for rows, cols, real_rows, real_cols, fan_in, scale in ((160, 256, IN, 256, IN, 2.0), (256, 256, 256, 256, 256, 2.0),
                                                           (256, 256, 256, 256, 256, 2.0),
                                                           (256, 16, 256, OUT, 256, out_scale ** 2)):
        W = p[o:o + rows * cols].reshape(rows, cols)
        W[:real_rows, :real_cols] = rng.normal(0.0, math.sqrt(scale / fan_in), (real_rows, real_cols))
        o += rows * cols + cols  # weights, then the zero biases
    assert o == P
    return p


class MLP:
    """The network's device state (parameters, Adam moments, activations for up to batch_cap rows)."""

    def __init__(self, batch_cap: int = 4096, device=None, stream=None):
        torch = _torch()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        self.lib = _abi.lib()
        self.P = self.lib.lw_mlp_param_count()
        self.cap = batch_cap
        self.h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _abi.check(self.lib.lw_mlp_create(ctypes.byref(self.h), batch_cap, ctypes.c_void_p(self.stream.cuda_stream)),
                       "lw_mlp_create")
        self._y = torch.empty((batch_cap, OUT), dtype=torch.float32, device=self.device)
        self._loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        self.steps = 0

    def close(self):
        """Wait for the network's queued work and free its device state (idempotent)."""
        h = getattr(self, "h", None)
        self.h = None
        if h is not None and h.value:
            _abi.check(self.lib.lw_mlp_destroy(h), "lw_mlp_destroy")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        self.stream = stream
        _abi.check(self.lib.lw_mlp_set_stream(self.h, ctypes.c_void_p(stream.cuda_stream)), "lw_mlp_set_stream")

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (self.P,)
        _abi.check(self.lib.lw_mlp_set_params(self.h, _abi.ptr(p, _abi._f32p)), "lw_mlp_set_params")
        self.steps = 0

    def params(self) -> np.ndarray:
        out = np.empty(self.P, np.float32)
        _abi.check(self.lib.lw_mlp_get_params(self.h, _abi.ptr(out, _abi._f32p)), "lw_mlp_get_params")
        return out

    def grads(self) -> np.ndarray:
        out = np.empty(self.P, np.float32)
        _abi.check(self.lib.lw_mlp_get_grads(self.h, _abi.ptr(out, _abi._f32p)), "lw_mlp_get_grads")
        return out

    def forward(self, x):
        """x: (n, 152) float32 device tensor (rows contiguous). Returns a view (n, 6) of a reused buffer."""
        n = int(x.shape[0])
        assert x.dtype == _torch().float32 and x.stride(1) == 1 and 0 < n <= self.cap
        _abi.check(self.lib.lw_mlp_forward(self.h, x.data_ptr(), x.stride(0), n, self._y.data_ptr()), "lw_mlp_forward")
        return self._y[:n]

    def loss(self, y, t):
        n = int(y.shape[0])
        assert t.shape == (n, OUT) and t.is_contiguous()
        _abi.check(self.lib.lw_mlp_loss(self.h, y.data_ptr(), t.data_ptr(), n, self._loss.data_ptr()), "lw_mlp_loss")
        return self._loss

    def backward(self, n: int):
        _abi.check(self.lib.lw_mlp_backward(self.h, n), "lw_mlp_backward")

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        self.steps += 1
        _abi.check(self.lib.lw_mlp_adam(self.h, lr, beta1, beta2, eps, self.steps), "lw_mlp_adam")

    def train_step(self, x, t, lr: float, graph: bool = True, beta1: float = 0.9, beta2: float = 0.999,
                   eps: float = 1e-8):
        """Forward, loss, backward and Adam on (x, t) in one call (lw_mlp_train_step): with graph, replayed from
        a captured hipGraph while x, t and the row count stay the same (on a stream other than the null
        stream). Returns the loss (a one-element device tensor, reused)."""
        n = int(x.shape[0])
        assert x.dtype == _torch().float32 and x.stride(1) == 1 and 0 < n <= self.cap
        assert t.shape == (n, OUT) and t.is_contiguous()
        self.steps += 1
        _abi.check(self.lib.lw_mlp_train_step(self.h, x.data_ptr(), x.stride(0), t.data_ptr(), n, lr, beta1, beta2,
                                              eps, self.steps, self._loss.data_ptr(), 1 if graph else 0),
                   "lw_mlp_train_step")
        return self._loss


class Normalizer:
    """Scales of the motion inputs (per twist component, shared by the four lags) and of the targets."""

    def __init__(self, mu_d, sd_d, mu_t, sd_t):
        self.mu_d, self.sd_d = np.asarray(mu_d, np.float64), np.asarray(sd_d, np.float64)
        self.mu_t, self.sd_t = np.asarray(mu_t, np.float64), np.asarray(sd_t, np.float64)

    def motion(self, hist) -> np.ndarray:
        """hist: the last increments, most recent first (up to 4; missing lags read as zero motion)."""
        out = np.zeros(MOTION, np.float32)
        for j, d in enumerate(list(hist)[:4]):
            out[6 * j:6 * j + 6] = (np.asarray(d) - self.mu_d) / self.sd_d
        for j in range(len(hist), 4):
            out[6 * j:6 * j + 6] = -self.mu_d / self.sd_d
        return out

    def to_dict(self):
        return {"mu_d": self.mu_d, "sd_d": self.sd_d, "mu_t": self.mu_t, "sd_t": self.sd_t}


class LearnedPrior:
    """Odometry.prior: the learned guess of each frame's pose, with the residual test that hands a frame back to
    the constant-velocity guess."""

    def __init__(self, params: np.ndarray, norm: Normalizer, odo, fallback: float = 2.0, window: int = 51,
                 min_history: int = 10):
        torch = _torch()
        self.mlp = MLP(batch_cap=16, device=odo.device, stream=odo.stream)
        self.mlp.set_params(params)
        self.norm = norm
        self.fallback = fallback
        self.min_history = min_history
        self.hist = collections.deque(maxlen=4)   # increments, most recent first
        self.res = collections.deque(maxlen=window)
        self.feat = torch.zeros((1, IN), dtype=torch.float32, device=odo.device)
        self._fp = _abi.FramePrior()  # the frame chain's prior block, and numpy views onto it
        self._fp_T = np.ctypeslib.as_array(self._fp.T)
        self._fp_h = np.ctypeslib.as_array(self._fp.hist0)
        self._fp_sd = np.ctypeslib.as_array(self._fp.sd_t)
        self._fp_mu = np.ctypeslib.as_array(self._fp.mu_t)
        self.last = None
        self.guess_err = {"learned": [], "cv": []}

    def close(self):
        """Free the network's device state (idempotent; Odometry.close calls it)."""
        self.mlp.close()

    def guess(self, odo, m: int):
        if len(self.hist) < 4:
            return None
        mot = self.norm.motion(self.hist)
        _abi.check(odo._lib.lw_frame_features(odo._ctx, m, R_MAX, E_MAX, _abi.ptr(mot, _abi._f32p),
                                              self.feat.data_ptr()), "lw_frame_features")
        y = self.mlp.forward(self.feat)
        out = y.cpu().numpy()[0] if odo.stream == _torch().cuda.current_stream(odo.device) else self._read(odo, y)
        d = self.hist[0] + out.astype(np.float64) * self.norm.sd_t + self.norm.mu_t
        self.last = (d, body_twist(odo.delta))  # the learned and the constant-velocity increments
        return odo.T @ twist_matrix(d), d.astype(np.float32)

    @staticmethod
    def _read(odo, y):
        odo.synchronize()
        return y.cpu().numpy()[0]

    def chain_inputs(self, odo):
        """guess() for the device frame chain (lw_odom_frame_prior): the frame's inputs and the forward pass are
        queued on the odometry's stream (the frame's count stays on the device), and the chain forms the guess,
        and takes the residual test, itself. Returns the chain's prior block, or None while fewer than four
        increments are known (the frame then takes the constant-velocity chain)."""
        if len(self.hist) < 4:
            return None
        mot = self.norm.motion(self.hist)
        _abi.check(odo._lib.lw_frame_features(odo._ctx, -1, R_MAX, E_MAX, _abi.ptr(mot, _abi._f32p),
                                              self.feat.data_ptr()), "lw_frame_features")
        y = self.mlp.forward(self.feat)  # a reused buffer on the odometry's stream, read by the chain after it
        fp = self._fp
        fp.y = y.data_ptr()
        self._fp_T[:] = odo.T.ravel()
        self._fp_h[:] = self.hist[0]
        self._fp_sd[:] = self.norm.sd_t
        self._fp_mu[:] = self.norm.mu_t
        # accept()'s test, with the window as it stands before this frame's residual joins it
        fp.rms_gate = (self.fallback * float(np.median(self.res)) if len(self.res) >= self.min_history
                       else math.inf)
        return fp

    def chain_result(self, odo, fo):
        """The bookkeeping guess() and accept() do, from the chain's output fo."""
        if not fo.guess:
            return
        self.last = (np.array(fo.d, dtype=np.float64), body_twist(odo.delta))
        odo.totals["prior_used"] += 1
        self.accept(float(fo.reg[0].rms0))  # the first residual joins the window
        if fo.fallback:
            odo.totals["prior_fallbacks"] += 1

    def accept(self, rms0: float) -> bool:
        ok = len(self.res) < self.min_history or rms0 <= self.fallback * float(np.median(self.res))
        self.res.append(rms0)
        return ok

    def observe(self, T_prev: np.ndarray, T: np.ndarray):
        d = body_twist(np.linalg.inv(T_prev) @ T)
        if self.last is not None:  # how far each guess was from the registered increment
            for key, g in zip(("learned", "cv"), self.last):
                self.guess_err[key].append((float(np.linalg.norm(g[:3] - d[:3])), float(np.linalg.norm(g[3:] - d[3:]))))
            self.last = None
        self.hist.appendleft(d)


class Recorder:
    """Odometry.prior for collecting training data: computes each registered frame's inputs (on the device, as
    LearnedPrior does) from the raw motion history, and leaves the guess to the constant-velocity model."""

    def __init__(self, odo, capacity: int):
        torch = _torch()
        self.X = torch.zeros((capacity, IN), dtype=torch.float32, device=odo.device)
        self.frames = []   # for each row: the index of the frame in odo.traj it belongs to
        self.hist = collections.deque(maxlen=4)
        self.unit = Normalizer(np.zeros(6), np.ones(6), np.zeros(6), np.ones(6))
        self.odo = odo

    def guess(self, odo, m: int):
        if len(self.hist) == 4 and len(self.frames) < len(self.X):
            mot = self.unit.motion(self.hist)  # raw increments; scaled after collection
            row = self.X[len(self.frames)]
            _abi.check(odo._lib.lw_frame_features(odo._ctx, m, R_MAX, E_MAX, _abi.ptr(mot, _abi._f32p),
                                                  row.data_ptr()), "lw_frame_features")
            self.frames.append(len(odo.traj))  # this frame's index once it is finalised
        return None

    def accept(self, rms0: float) -> bool:
        return True

    def observe(self, T_prev, T):
        self.hist.appendleft(body_twist(np.linalg.inv(T_prev) @ T))


def run_sequence(lw, seq, n_frames: int, device, prior_factory=None, reg_slots: int = 1 << 22, max_iter: int = 20,
                 frames=None):
    """The plain odometry over n_frames of seq (frames generated on the device unless given); prior_factory(odo)
    optionally attaches a prior. Returns (odometry, seconds in the push loop)."""
    import time

    torch = _torch()
    X, A, Tt = frames if frames is not None else seq.frames_device(n_frames, device)
    odo = lw.Odometry(lw.PoseTable(), device, reg_slots=reg_slots, frame_cap=1 << 15)
    odo.reg_voxel = 0.2
    odo.max_iter = max_iter
    odo.reset(seq.pose(0.05))
    if prior_factory is not None:
        odo.prior = prior_factory(odo)
    per = X.shape[1]
    spans = Tt[:, [0, -1]].cpu().numpy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(n_frames):
        odo.push(X[k], A[k], Tt[k], per, float(spans[k, 0]), float(spans[k, 1]), lambda *_: None)
    odo.synchronize()
    return odo, time.perf_counter() - t0


def collect(lw, seeds, n_frames: int, device, seconds_margin: float = 1.0, progress=None):
    """Training rows from the plain GPU odometry on the sequences `seeds`: inputs (raw motion, histograms) and
    targets d_k - d_{k-1} from the odometry's own poses. Returns (X (N, 152), Y (N, 6)) device tensors."""
    torch = _torch()
    xs, ys = [], []
    for s in seeds:
        seq = lw.synth.Sequence(seed=int(s), seconds=n_frames * 0.1 + seconds_margin)
        rec = {}
        odo, _ = run_sequence(lw, seq, n_frames, device, prior_factory=lambda o: rec.setdefault("r", Recorder(o, n_frames)))
        r = rec["r"]
        traj = [T for _, _, T in odo.traj]
        odo.close()
        rows, targets = [], []
        for row, f in enumerate(r.frames):
            if f < 2 or f >= len(traj):
                continue
            d_k = body_twist(np.linalg.inv(traj[f - 1]) @ traj[f])
            d_p = body_twist(np.linalg.inv(traj[f - 2]) @ traj[f - 1])
            rows.append(row)
            targets.append(d_k - d_p)
        if rows:
            xs.append(r.X[torch.as_tensor(rows, device=r.X.device)])
            ys.append(torch.as_tensor(np.array(targets), dtype=torch.float64, device=r.X.device))
        if progress is not None:
            progress(f"sequence {s}: {len(rows)} rows")
    return torch.cat(xs), torch.cat(ys)


def make_normalizer(X, Y) -> Normalizer:
    d = X[:, :MOTION].double().view(-1, 4, 6).reshape(-1, 6)
    mu_d, sd_d = d.mean(0).cpu().numpy(), d.std(0).clamp_min(1e-6).cpu().numpy()
    mu_t, sd_t = Y.mean(0).cpu().numpy(), Y.std(0).clamp_min(1e-9).cpu().numpy()
    return Normalizer(mu_d, sd_d, mu_t, sd_t)


def normalise(X, Y, norm: Normalizer):
    torch = _torch()
    Xn = X.clone()
    mu = torch.as_tensor(np.tile(norm.mu_d, 4), dtype=torch.float32, device=X.device)
    sd = torch.as_tensor(np.tile(norm.sd_d, 4), dtype=torch.float32, device=X.device)
    Xn[:, :MOTION] = (X[:, :MOTION] - mu) / sd
    Yn = ((Y - torch.as_tensor(norm.mu_t, device=Y.device)) / torch.as_tensor(norm.sd_t, device=Y.device))
    return Xn.contiguous(), Yn.to(torch.float32).contiguous()


def fit(Xn, Yn, epochs: int = 30, batch: int = 1024, lr: float = 1e-3, seed: int = 0, params=None, device=None,
        val=None):
    """Train on (Xn, Yn) with the library's kernels (minibatch Adam, cosine-decayed rate, shuffled each epoch).
    With val = (Xv, Yv), the validation loss is measured after every epoch and the best epoch's parameters are
    returned. Returns (params, per-epoch training loss, per-epoch validation loss or [])."""
    torch = _torch()
    dev = device or Xn.device
    stream = torch.cuda.Stream(device=dev)  # the captured training step replays on a stream of its own
    mlp = MLP(batch_cap=batch, device=dev, stream=stream)
    rng = np.random.default_rng(seed)
    mlp.set_params(params if params is not None else init_params(rng))
    n = int(Xn.shape[0])
    hist, vhist = [], []
    best, best_params = float("inf"), None
    gen = torch.Generator(device=Xn.device)
    gen.manual_seed(seed)
    xb = torch.empty((batch, Xn.shape[1]), dtype=torch.float32, device=dev)  # fixed batch buffers: one graph
    tb = torch.empty((batch, Yn.shape[1]), dtype=torch.float32, device=dev)
    stream.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(stream):
        for ep in range(epochs):
            perm = torch.randperm(n, device=Xn.device, generator=gen)
            tot = torch.zeros(1, dtype=torch.float64, device=Xn.device)
            nb = 0
            decay = 0.5 * (1.0 + math.cos(math.pi * ep / max(epochs, 1)))
            for a in range(0, n - batch + 1, batch):
                idx = perm[a:a + batch]
                torch.index_select(Xn, 0, idx, out=xb)
                torch.index_select(Yn, 0, idx, out=tb)
                tot += mlp.train_step(xb, tb, lr * (0.1 + 0.9 * decay)).double()
                nb += 1
            hist.append(float(tot.item()) / max(nb, 1))
            if val is not None:
                v = evaluate_loss(mlp, *val)
                vhist.append(v)
                if v < best:
                    best, best_params = v, mlp.params()
        out = best_params if best_params is not None else mlp.params()
    mlp.close()
    return out, hist, vhist


def evaluate_loss(mlp: MLP, X, Y) -> float:
    """Mean squared error of the network over (X, Y), in chunks of the network's batch capacity."""
    tot, n = 0.0, int(X.shape[0])
    for a in range(0, n, mlp.cap):
        b = min(n, a + mlp.cap)
        tot += float(mlp.loss(mlp.forward(X[a:b].contiguous()), Y[a:b].contiguous()).item()) * (b - a)
    return tot / max(n, 1)


def split(Xn, Yn, fraction: float = 0.1, seed: int = 0):
    """A random (train, validation) split of the rows."""
    torch = _torch()
    g = torch.Generator(device=Xn.device)
    g.manual_seed(seed)
    perm = torch.randperm(int(Xn.shape[0]), device=Xn.device, generator=g)
    k = max(int(len(perm) * fraction), 1)
    v, t = perm[:k], perm[k:]
    return (Xn[t].contiguous(), Yn[t].contiguous()), (Xn[v].contiguous(), Yn[v].contiguous())


def save(path: str, params: np.ndarray, norm: Normalizer, meta: dict | None = None):
    np.savez_compressed(path, params=params, **norm.to_dict(), meta=np.array(repr(meta or {})))


def load(path: str):
    z = np.load(path, allow_pickle=False)
    return z["params"], Normalizer(z["mu_d"], z["sd_d"], z["mu_t"], z["sd_t"])


def drift(seq, traj) -> float:
    """RMS position error of the trajectory against the ground truth, both relative to their first pose."""
    T0 = seq.pose(traj[0][1])
    err = [np.linalg.norm((np.linalg.inv(traj[0][2]) @ T)[:3, 3] - (np.linalg.inv(T0) @ seq.pose(tm))[:3, 3])
           for _, tm, T in traj]
    return float(np.sqrt(np.mean(np.square(err))))


def evaluate(lw, params, norm: Normalizer, seeds, n_frames: int, device, fallback: float = 2.0, progress=None):
    """Each held-out sequence twice on the same device-generated frames: the constant-velocity guess, then the
    learned prior. Per sequence: frames/s, Gauss-Newton steps and registrations per frame, drift, fallbacks,
    and the mean error of both guesses against the registered increment (learned run)."""
    torch = _torch()
    rows = []
    for s in seeds:
        seq = lw.synth.Sequence(seed=int(s), seconds=n_frames * 0.1 + 1.0)
        frames = seq.frames_device(n_frames, device)
        torch.cuda.synchronize()
        out = {"seed": int(s)}
        for name, factory in (("cv", None), ("learned", lambda o: LearnedPrior(params, norm, o, fallback=fallback))):
            odo, el = run_sequence(lw, seq, n_frames, device, prior_factory=factory, frames=frames)
            n = len(odo.traj)
            tot = odo.totals
            r = {"fps": n / el, "gn_per_frame": tot["gn_iters"] / n, "reg_per_frame": tot["registrations"] / n,
                 "drift_m": drift(seq, odo.traj)}
            if factory is not None:
                ge = odo.prior.guess_err
                r["fallbacks"] = tot["prior_fallbacks"]
                r["used"] = tot["prior_used"]
                for key in ("learned", "cv"):
                    e = np.array(ge[key]) if ge[key] else np.zeros((1, 2))
                    r[f"guess_{key}_mm"] = float(e[:, 0].mean() * 1e3)
                    r[f"guess_{key}_mdeg"] = float(np.degrees(e[:, 1].mean()) * 1e3)
            odo.close()
            out[name] = r
        rows.append(out)
        if progress is not None:
            progress(f"held-out {s}: cv {out['cv']['fps']:.0f} fps {out['cv']['gn_per_frame']:.2f} GN/frame, learned "
                     f"{out['learned']['fps']:.0f} fps {out['learned']['gn_per_frame']:.2f} GN/frame")
    return rows


def summarise(rows) -> dict:
    def mean(k, key):
        return float(np.mean([r[k][key] for r in rows]))

    res = {}
    for k in ("cv", "learned"):
        res[k] = {key: round(mean(k, key), 4) for key in rows[0][k]}
    res["speedup"] = round(res["learned"]["fps"] / res["cv"]["fps"], 3)
    res["gn_reduction"] = round(1.0 - res["learned"]["gn_per_frame"] / res["cv"]["gn_per_frame"], 3)
    return res

"""What the trained network gains over the constant-velocity guess, on held-out office walks it never saw (the
shipped weights), and that the pipeline learns from the plain odometry's own poses when trained from scratch."""

import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_shipped_prior_beats_constant_velocity(lw, gpu_ok):
    params, norm = lw.prior.load(os.path.join(ROOT, "livox-warp_amd", "prior_weights.npz"))
    rows = lw.prior.evaluate(lw, params, norm, [5000, 5001], 300, "cuda:0",
                             progress=lambda m: print(m, flush=True))
    s = lw.prior.summarise(rows)
    print(s, flush=True)
    lp, cv = s["learned"], s["cv"]
    # the learned guess is closer to the registered pose than the constant-velocity guess
    assert lp["guess_learned_mm"] < 0.9 * lp["guess_cv_mm"]
    assert lp["guess_learned_mdeg"] < 0.7 * lp["guess_cv_mdeg"]
    # and ICP needs fewer Gauss-Newton steps from it
    assert lp["gn_per_frame"] < cv["gn_per_frame"]
    assert lp["fallbacks"] < 0.3 * lp["used"]


def test_training_from_odometry_poses(lw, gpu_ok):
    X, Y = lw.prior.collect(lw, range(1000, 1008), 300, "cuda:0", progress=lambda m: print(m, flush=True))
    norm = lw.prior.make_normalizer(X, Y)
    Xn, Yn = lw.prior.normalise(X, Y, norm)
    (Xt, Yt), (Xv, Yv) = lw.prior.split(Xn, Yn, fraction=0.2)
    params, hist, vhist = lw.prior.fit(Xt, Yt, epochs=30, batch=256, lr=1e-3, val=(Xv, Yv))
    print("train", [round(h, 3) for h in hist[::5]], "validation", [round(v, 3) for v in vhist[::5]], flush=True)
    # normalised targets: predicting their mean scores 1; the network must explain part of unseen rows
    assert hist[-1] < 0.6 * hist[0] and min(vhist) < 0.85


def test_prior_chain_decisions(lw, gpu_ok):
    """The learned prior in the device frame chain (lw_odom_frame_prior: the guess formed from the network's output,
    the first-residual test and the constant-velocity fallback taken on the device), frame by frame on a held-out
    walk, against the host's formulas on the same frame: the increment is hist0 + y sd_t + mu_t bit for bit (y read
    back from the network's buffer), the guess is taken exactly when the frame has min_points points, the fallback
    exactly when the first residual exceeds the gate (fallback x the median of the earlier first residuals), the
    re-deskew registration exactly when the registered twist leaves the one the points were deskewed with (the
    guess's, or the constant-velocity one after a fallback) by 5 mm / 2 mrad (a margin of 1e-5 set aside for the
    rotation-vector formulas' last bits). Then the same walk step by step (Odometry.chain off: each decision between
    device steps on the host): the same frames take the guess and about as many fall back. The trajectories are not
    compared: the map's float atomics make two runs of either path drift apart by centimetres."""
    params, norm = lw.prior.load(os.path.join(ROOT, "livox-warp_amd", "prior_weights.npz"))

    class Checked(lw.prior.LearnedPrior):
        def __init__(self, *a, **k):
            super().__init__(*a, **k)
            self.n = {"guess": 0, "fallback": 0, "redo_checked": 0, "redo": 0, "near": 0, "skipped": 0}

        def chain_inputs(self, odo):
            fp = super().chain_inputs(odo)
            if fp is not None:
                self.ctx = (fp.rms_gate, self.hist[0].copy(), odo.T_prev.copy(), odo.xi.copy(), odo.min_corr)
            return fp

        def chain_result(self, odo, fo):
            gate, h0, T_prev, xi_cv, min_corr = self.ctx
            assert bool(fo.guess) == (fo.m >= odo.min_points)
            if fo.guess:
                y = self.mlp._y[0].cpu().numpy().astype(np.float64)
                d = h0 + y * self.norm.sd_t + self.norm.mu_t
                assert np.array_equal(np.array(fo.d), d)
                self.n["guess"] += 1
                assert bool(fo.fallback) == (not fo.reg[0].rms0 <= gate)
                self.n["fallback"] += int(fo.fallback)
                fb = int(fo.fallback)
                r = fo.reg[fb]
                if r.n_ds >= min_corr and r.corr >= min_corr and not fo.skipped:
                    T1 = np.array(r.T).reshape(4, 4)
                    xn = lw.prior.body_twist(np.linalg.inv(T_prev) @ T1).astype(np.float32)
                    used = xi_cv.astype(np.float32) if fb else d.astype(np.float32)
                    e = xn - used
                    nr = float(np.sqrt(np.sum(e[3:] * e[3:], dtype=np.float32)))
                    nt = float(np.sqrt(np.sum(e[:3] * e[:3], dtype=np.float32)))
                    redo = fo.registrations - 1 - fb
                    if min(abs(nr - 2e-3), abs(nt - 5e-3)) > 1e-5:
                        assert redo == int(nr > 2e-3 or nt > 5e-3)
                        self.n["redo_checked"] += 1
                        self.n["redo"] += redo
                    else:
                        self.n["near"] += 1
            else:
                self.n["skipped"] += 1
            super().chain_result(odo, fo)

    seq = lw.synth.Sequence(seed=5002, seconds=25.0)
    n = 220
    frames = seq.frames_device(n, "cuda:0")
    runs = {}
    for chain in (True, False):
        def factory(odo, chain=chain):
            odo.chain = chain
            return (Checked if chain else lw.prior.LearnedPrior)(params, norm, odo)
        odo, secs = lw.prior.run_sequence(lw, seq, n, "cuda:0", prior_factory=factory, frames=frames)
        runs[chain] = (dict(odo.totals), getattr(odo.prior, "n", None))
        print(f"chain={chain}: {odo.totals}, {n / secs:.0f} frames/s", runs[chain][1], flush=True)
        odo.close()
    (ta, na), (tb, _) = runs[True], runs[False]
    assert na["guess"] == ta["prior_used"] > n // 2 and na["fallback"] == ta["prior_fallbacks"] > 0
    assert na["redo_checked"] > n // 2 and na["redo"] > 0
    assert ta["prior_used"] == tb["prior_used"]
    assert abs(ta["prior_fallbacks"] - tb["prior_fallbacks"]) <= 6

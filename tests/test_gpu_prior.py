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

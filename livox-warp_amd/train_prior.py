"""Train the learned pose-increment prior on the plain GPU odometry's trajectories, then compare it with the
constant-velocity guess on held-out sequences.

    python livox-warp_amd/train_prior.py --train 200 --frames 600 --heldout 20 --out livox-warp_amd/prior_weights.npz

Training sequences use seeds 1000.., held-out ones 5000.. (office walks of the odometry-sequence kind).
Prints one JSON line with the training losses and the held-out comparison.
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=200, help="training sequences")
    ap.add_argument("--frames", type=int, default=600, help="frames per sequence")
    ap.add_argument("--heldout", type=int, default=20)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=os.path.join(HERE, "prior_weights.npz"))
    args = ap.parse_args()
    import torch

    lw = importlib.import_module("livox-warp_amd")
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    def say(msg):
        print(msg, file=sys.stderr, flush=True)

    X, Y = lw.prior.collect(lw, range(1000, 1000 + args.train), args.frames, dev, progress=say)
    t_collect = time.perf_counter() - t0
    print(f"collected {tuple(X.shape)} rows in {t_collect:.1f} s", file=sys.stderr, flush=True)
    norm = lw.prior.make_normalizer(X, Y)
    Xn, Yn = lw.prior.normalise(X, Y, norm)
    # a random 5% of the rows watches generalisation; the best epoch's parameters are kept
    (Xt, Yt), (Xv, Yv) = lw.prior.split(Xn, Yn, fraction=0.05)
    t0 = time.perf_counter()
    params, hist, vhist = lw.prior.fit(Xt, Yt, epochs=args.epochs, batch=args.batch, lr=args.lr, device=dev,
                                       val=(Xv, Yv))
    t_fit = time.perf_counter() - t0
    val = min(vhist)
    say(f"trained in {t_fit:.1f} s: loss {hist[0]:.4f} -> {hist[-1]:.4f}, best validation {val:.4f} "
        f"(epoch {int(np.argmin(vhist))})")
    lw.prior.save(args.out, params, norm, {"train": args.train, "frames": args.frames, "epochs": args.epochs})
    rows = lw.prior.evaluate(lw, params, norm, range(5000, 5000 + args.heldout), args.frames, dev, progress=say)
    print(json.dumps({"rows": int(len(Xn)), "collect_s": round(t_collect, 1), "fit_s": round(t_fit, 2),
                      "loss_first": hist[0], "loss_last": hist[-1], "val_loss": val,
                      "heldout": lw.prior.summarise(rows), "per_sequence": rows}), flush=True)


if __name__ == "__main__":
    main()

"""Train the learned warm start on the GPU and write the model file.

    python -m splishsplash_for_cuda_amd.train_warmstart [--scenes 32] [--samples 2000000]
        [--steps 200] [--train-steps 6000] [--out splishsplash_for_cuda_amd/models/warmstart_mlp.npz]

1. Harvest (features, solver increments) pairs from `--scenes` seeded dam breaks
   (seed 1000 + k, 50k-200k particles, block aspect ratios in [0.5, 2]) run by
   the plain solver with the reference warm start.
2. Fit the input/output standardisation and train the network with the
   device's forward/backward kernels and Adam.
3. Compare iterations per step of a held-out scene (seed 99) with the
   reference, learned and no warm start at the reference's minimum
   iteration counts (divergence 3, density 2) and at (1, 1).
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import learned, scene


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=32)
    ap.add_argument("--samples", type=int, default=2_000_000)
    ap.add_argument("--steps", type=int, default=200, help="harvested steps per scene")
    ap.add_argument("--n-min", type=int, default=50_000)
    ap.add_argument("--n-max", type=int, default=200_000)
    ap.add_argument("--train-steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-n", type=int, default=100_000)
    ap.add_argument("--eval-steps", type=int, default=100)
    ap.add_argument("--out", default=learned.DEFAULT_MODEL)
    a = ap.parse_args(argv)

    t0 = time.time()
    x, t = learned.build_training_set(n_scenes=a.scenes, n_range=(a.n_min, a.n_max), total=a.samples,
                                      steps=a.steps, log=lambda m: print(m, flush=True))
    print("training set %s in %.1f s" % (x.shape, time.time() - t0), flush=True)
    print("feature mean", np.array2string(x.mean(0), precision=4), flush=True)
    print("feature std ", np.array2string(x.std(0), precision=4), flush=True)
    print("target mean %s std %s" % (t.mean(0), t.std(0)), flush=True)

    t0 = time.time()
    p, nrm, losses = learned.train(x, t, steps=a.train_steps, batch=a.batch, lr=a.lr)
    print("normalised loss first/last batch %.4f / %.4f in %.1f s" % (losses[0], losses[1], time.time() - t0),
          flush=True)
    learned.save(a.out, p, nrm, scenes=a.scenes, samples=len(x), harvested_steps=a.steps,
                 train_steps=a.train_steps, loss=np.asarray(losses))
    print("wrote", a.out, flush=True)

    sc = scene.seeded_dam_break(99, a.eval_n)
    report = {}
    for floors in ((3, 2), (1, 1)):
        r = learned.compare_iterations(sc, p, nrm, steps=a.eval_steps, warmup=10, min_iter_v=floors[0],
                                       min_iter=floors[1])
        key = "floors_%d_%d" % floors
        report[key] = {k: {kk: v[kk] for kk in ("iters_v", "iters", "err_v", "err")} for k, v in r.items()}
        print(key, json.dumps(report[key]), flush=True)
    return report


if __name__ == "__main__":
    main()

"""Trains the calibration surrogate on the GPU and writes its weights.

    python -m watergap2_amd.train_surrogate --out watergap2_amd/data/surrogate_weights.npz

The training set comes from the batched model itself: `--basins` synthetic basins
(log-uniform 50-5,000 cells, seeds disjoint from the evaluation basins), each with
`--candidates` (gamma, CFA) members, nine-year runs whose 30-year mean follows
from the three-year forcing cycle (surrogate.thirty_year_mean).  5 % of the
samples are held out for the validation error written with the weights.  Runs
with a non-finite discharge are dropped (the reference's land-area update can
turn a soil storage slightly negative when CFA > 1; the oracle does the same).
"""
import argparse
import sys
import time

import numpy as np

from . import surrogate as sg


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--basins", type=int, default=5000)
    ap.add_argument("--candidates", type=int, default=40)
    ap.add_argument("--seed", type=int, default=9001)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--data", default="", help="also write the training set (.npz) here")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)

    def log(msg):
        print("# " + msg, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    x, t = sg.generate_training_set(args.basins, args.candidates, seed=args.seed, log=log)
    gen_s = time.perf_counter() - t0
    if args.data:
        np.savez(args.data, x=x, t=t)
    finite = np.isfinite(t) & np.isfinite(x).all(axis=1)
    x, t = x[finite], t[finite]
    perm = np.random.default_rng(5).permutation(len(t))
    n_val = len(t) // 20
    val, tr = perm[:n_val], perm[n_val:]
    t1 = time.perf_counter()
    model, losses = sg.Surrogate.fit(x[tr], t[tr], epochs=(args.epochs, args.epochs // 2), log=log)
    train_s = time.perf_counter() - t1
    rmse = float(np.sqrt(np.mean((model.predict(x[val]) - t[val]) ** 2)))
    log(f"{len(t)} samples ({int((~finite).sum())} dropped), generation {gen_s:.0f} s, training {train_s:.1f} s, "
        f"validation RMSE of log runoff {rmse:.4f}")
    model.save(args.out, samples=len(t), dropped=int((~finite).sum()), basins=args.basins,
               candidates=args.candidates, seed=args.seed, epochs=len(losses), validation_rmse=rmse,
               generation_seconds=gen_s, training_seconds=train_s, residual_sd=float(np.std(
                   model.predict(x[val]) - t[val])))


if __name__ == "__main__":
    main()

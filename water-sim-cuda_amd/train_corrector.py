"""Trains the learned grid-velocity corrector on the 128^3 dam-break family
and evaluates it on held-out seeds (no reference counterpart).

    python water-sim-cuda_amd/train_corrector.py --out water-sim-cuda_amd/corrector_dambreak128.npz

Steps: training pairs from plain runs at DT = 5e-4 (seeds 0..N-1, pairs
spread over one simulated second), input normalisation from the pairs, Adam
training on the GPU with the library's kernels, then evaluation at 4x DT
against plain DT runs of held-out seeds.  Prints one progress line per step
and writes the weights (npz) and a JSON summary.
"""
from __future__ import annotations

import argparse
import importlib
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))


def log(msg: str) -> None:
    print(msg, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=32, help="training runs (seeds 0..N-1)")
    ap.add_argument("--samples", type=int, default=2_000_000, help="training pairs in all")
    ap.add_argument("--collections", type=int, default=20,
                    help="pair collections per seed, spread evenly over the simulated second")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init-seed", type=int, default=0, help="seed of learned.init_params")
    ap.add_argument("--eval-seeds", default="1000,1001")
    ap.add_argument("--sim-seconds", type=float, default=1.0)
    ap.add_argument("--out", default=str(HERE / "corrector_dambreak128.npz"))
    ap.add_argument("--json", default=None, help="summary file [next to --out]")
    args = ap.parse_args()

    m = importlib.import_module("water-sim-cuda_amd")
    learned = importlib.import_module("water-sim-cuda_amd.learned")

    t0 = time.time()
    # pairs at `collections` evenly spaced states of each one-second run, so
    # that every phase of the flow is in the training set
    per_sample = int(np.ceil(args.samples / args.seeds / args.collections))
    every = max(learned.RATIO, 2000 // args.collections)
    X, T = learned.generate_pairs(range(args.seeds), args.samples, every=every, per_sample=per_sample,
                                  substeps=2000, log=log)
    log(f"pairs: {len(X)} rows in {time.time() - t0:.1f}s; target rms {float(np.sqrt((T ** 2).mean())):.4e}")
    mean, inv, scale = learned.normalisation(X, T)
    corr = m.Corrector(learned.init_params(args.init_seed))
    corr.set_norm(mean, inv, scale)
    corr.set_clamp(learned.OUT_CLAMP)
    corr.set_deadzone(learned.DEADZONE)
    t1 = time.time()
    losses = learned.train(corr, X, T, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
    t_train = time.time() - t1
    log(f"train: {len(losses)} steps in {t_train:.1f}s, loss {float(losses[:50].mean()):.4f} -> "
        f"{float(losses[-50:].mean()):.4f}")
    del X, T
    meta = {"family": "dambreak128", "ratio": learned.RATIO, "train_seeds": args.seeds, "samples": args.samples,
            "init": f"learned.init_params(seed={args.init_seed}): He-uniform, output layer x0.1, zero biases",
            "collections_per_seed": args.collections, "deadzone": learned.DEADZONE, "clamp": learned.OUT_CLAMP,
            "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
            "loss_first": float(losses[:50].mean()), "loss_last": float(losses[-50:].mean()),
            "train_seconds": round(t_train, 2)}
    learned.save(corr, args.out, mean, inv, scale, meta)
    log(f"saved {args.out}")
    evals = [learned.evaluate(corr, int(s), sim_seconds=args.sim_seconds, log=log)
             for s in args.eval_seeds.split(",") if s]
    summary = dict(meta, evaluation=evals)
    path = Path(args.json) if args.json else Path(args.out).with_suffix(".json")
    path.write_text(json.dumps(summary, indent=1))
    log(f"summary {path}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

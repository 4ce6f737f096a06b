"""Trains the learned correction of the terminator (helix_amd.closure) on the GPU and writes the network.

Training data: `--sets` seeded parameter sets (closure.Params.draw(first_seed + i): 4 spins, lambda, gamma, T drawn
from their ranges, seeded couplings) propagated by the plain FP64 solver at tiers 0..deep (batched GPU form); along
each trajectory the response statistics of every virtual tier-(keep+1) ADO and eigen-element give the targets
(closure.correction_targets).  The network (2 x 256 tanh, BF16 weights, FP32 accumulation on BF16 MFMA) is
trained by the HIP kernels of include/helix_mlp.h (Adam).  Then, on held-out sets, the solver at tiers 0..keep
is run with plain truncation, the physics terminator and the learned correction against plain FP64 at tiers
0..deep, and the errors are stored with the network.

    python -m helix_amd.train_closure --sets 4096 --out helix_amd/data/closure_n16_L2.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def log(msg: str):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def held_out_errors(cl, mlp, seeds, keep: int, deep: int, steps: int, every: int = 20):
    import helix_amd as hx
    rows = []
    for seed in seeds:
        p = cl.Params.draw(seed)
        d = hx.Plan(p.model(deep, "fp64"))
        d.set_initial()
        t0 = time.perf_counter()
        ref, _ = d.step(steps, every, record_rho=True)
        t_deep = time.perf_counter() - t0
        d.close()
        s = cl.ClosureSolver(p, mlp, keep_tier=keep, precision="fp32")
        row = {"seed": seed, "t_deep_s": t_deep}
        for mode in ("plain", "physics", "learned"):
            s.use(mode)
            t0 = time.perf_counter()
            _, tr = s.run(steps, every)
            row[f"t_{mode}_s"] = time.perf_counter() - t0
            row[mode] = float(np.max(np.abs(tr - ref)))
        s.close()
        rows.append(row)
        log(f"held out {seed}: plain {row['plain']:.3e} physics {row['physics']:.3e} learned {row['learned']:.3e}")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--sets", type=int, default=4096)
    ap.add_argument("--first-seed", type=int, default=7000)
    ap.add_argument("--keep", type=int, default=2)
    ap.add_argument("--deep", type=int, default=6)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--chunk", type=int, default=1024)
    ap.add_argument("--per-set", type=int, default=2048)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--held-out", default="9000-9005")
    ap.add_argument("--out", default=os.path.join(ROOT, "helix_amd", "data", "closure_n16_L2.npz"))
    args = ap.parse_args(argv)

    from helix_amd import closure as cl
    sets = [cl.Params.draw(args.first_seed + i) for i in range(args.sets)]
    log(f"response statistics: {args.sets} sets, tiers 0..{args.deep}, {args.steps} steps of dt 0.01")
    t0 = time.perf_counter()
    terms, num, den = cl.response_statistics(sets, keep_tier=args.keep, deep_tier=args.deep, steps=args.steps,
                                             every=args.every, chunk=args.chunk,
                                             log=lambda m: log(m) if m.endswith(f"/{args.steps}") or "step 500/" in m else None)
    t_data = time.perf_counter() - t0
    X, Y = cl.correction_targets(terms, num, den, per_set=args.per_set)
    del num, den
    log(f"targets: {X.shape[0]} rows x {X.shape[1]} features ({t_data:.0f} s); mean square {np.mean(Y.astype(np.float64) ** 2):.3e}")
    mlp = cl.MLP(X.shape[1], 256, 256, 2, seed=11)
    t0 = time.perf_counter()
    losses = []
    for e0 in range(0, args.epochs, 4):  # a few epochs per call (Adam's state carries over), logging between them
        n = min(4, args.epochs - e0)
        losses.extend(float(v) for v in mlp.train(X, Y, epochs=n, batch=args.batch, lr=args.lr, seed=3 + e0))
        log(f"epoch {e0 + n}/{args.epochs}: loss {losses[-1]:.4e}")
    losses = np.array(losses, np.float32)
    t_train = time.perf_counter() - t0
    log(f"trained {args.epochs} epochs in {t_train:.0f} s: loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    a, b = (int(v) for v in args.held_out.split("-"))
    rows = held_out_errors(cl, mlp, range(a, b + 1), args.keep, args.deep, args.steps)
    summary = {k: float(np.mean([r[k] for r in rows])) for k in ("plain", "physics", "learned")}
    log("mean max|d rho_S| on held-out sets: " + json.dumps(summary))
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    cl.save_network(args.out, mlp, train_sets=args.sets, first_seed=args.first_seed, keep_tier=args.keep,
                    deep_tier=args.deep, steps=args.steps, every=args.every, rows=X.shape[0], epochs=args.epochs,
                    losses=losses, data_seconds=t_data, train_seconds=t_train,
                    held_out=json.dumps(rows), held_out_mean=json.dumps(summary))
    log(f"wrote {args.out}")
    mlp.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

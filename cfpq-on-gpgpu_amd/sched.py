"""Training pipeline of the learned tile scheduler.

The network (16 inputs -> 64 -> 64 -> 1 logit, bf16 parameters, fp32
arithmetic) predicts whether a (rule, C tile) work item of a fixpoint step
will add bits to its head tile.  Everything here runs through the C ABI:

  collect()   run the *plain* delta fixpoint on workloads with trace recording
              on; each scored (rule, tile) item of each step becomes a row
              (8 features, rule id, label = "the head tile gained a bit");
  train()     fit the network on the device with the engine's own forward /
              backward / Adam kernels;
  save/load   the bf16 parameters as a .npy array (uint16) next to this file.

The trained network only orders and defers work: a deferred tile is marked
dirty and later evaluated over the full operands, and convergence is declared
only after a step that evaluates every dirty tile and defers nothing adds no
bit.  Results are bit-exact whatever the network predicts.
"""
from __future__ import annotations

import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "sched_weights.npy")


def training_workloads(gen):
    """Delta-trace sources: the SG forest of 16,384 vertices, the dense-stress
    grammar at 4,096 vertices and the first 48 batch queries."""
    return [gen.config(1), gen.dense_stress(4096)] + gen.config(3)[:48]


def collect(pkg, cf, queries, capacity: int = 1 << 22, log=print):
    """Trace rows of the plain fixpoint on each query: (x, rule, y)."""
    xs, rs, ys = [], [], []
    with pkg.Engine(0) as eng:
        for q in queries:
            eng.trace_begin(capacity)
            gr = cf.Grammar(q.grammar, text=True)
            gr.intersection_with_graph(cf.Graph.from_query(q), eng)
            x, r, y, count = eng.trace_read(capacity)
            eng.trace_end()
            gr.free()
            xs.append(x)
            rs.append(r)
            ys.append(y)
            if count > capacity:
                log(f"[sched] {q.name}: {count} trace rows, kept {capacity}")
    return np.concatenate(xs), np.concatenate(rs), np.concatenate(ys)


def train(pkg, x, rule, y, epochs: int = 16, batch: int = 4096, lr: float = 2e-3, seed: int = 7, log=print):
    """Returns (bf16 weights, per-epoch training loss, seconds on the device)."""
    with pkg.Engine(0) as eng:
        eng.sched_init(seed)
        t0 = time.perf_counter()
        losses = eng.sched_train(x, rule, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
        secs = time.perf_counter() - t0
        w = eng.sched_get_weights()
    log(f"[sched] trained on {len(y)} rows ({y.mean():.3f} positive) in {secs:.2f} s; "
        f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    return w, losses, secs


def evaluate(pkg, w, x, rule, y) -> dict:
    """Accuracy, AUC and the share of active items among those below each
    probability threshold (what deferral at that threshold would skip)."""
    with pkg.Engine(0) as eng:
        logits, loss, _ = eng.sched_eval(w, x, rule, y, want_grad=False)
    p = 1.0 / (1.0 + np.exp(-logits.astype(np.float64)))
    pos, neg = p[y > 0.5], p[y <= 0.5]
    # AUC by ranks
    order = np.argsort(np.concatenate([pos, neg]), kind="stable")
    ranks = np.empty_like(order, dtype=np.float64)
    ranks[order] = np.arange(1, len(order) + 1)
    auc = (ranks[:len(pos)].sum() - len(pos) * (len(pos) + 1) / 2) / max(len(pos) * len(neg), 1)
    out = {"loss": float(loss), "accuracy": float(np.mean((p > 0.5) == (y > 0.5))), "auc": float(auc),
           "positive_rate": float(y.mean())}
    for th in (0.05, 0.1, 0.2, 0.3):
        sel = p < th
        out[f"deferred_share@{th}"] = float(sel.mean())
        out[f"active_among_deferred@{th}"] = float(y[sel].mean()) if sel.any() else 0.0
    return out


def save(w, path: str = WEIGHTS):
    np.save(path, np.asarray(w, np.uint16))


def load(path: str = WEIGHTS):
    return np.load(path) if os.path.exists(path) else None


def main(argv=None):
    """Collect traces, train on the device, evaluate on held-out rows, save."""
    import argparse
    import importlib
    import json
    import sys

    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--epochs", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args(argv)
    root = os.path.dirname(HERE)
    if root not in sys.path:
        sys.path.insert(0, root)
    pkg = importlib.import_module("cfpq-on-gpgpu_amd")
    gen = importlib.import_module("cfpq-on-gpgpu_amd.gen")
    cf = importlib.import_module("cfpq-on-gpgpu_amd.cfpq")
    say = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    x, r, y = collect(pkg, cf, training_workloads(gen), log=say)
    say(f"[sched] {len(y)} trace rows in {time.perf_counter() - t0:.1f} s")
    # hold out every 5th row for evaluation
    idx = np.arange(len(y))
    tr, te = idx % 5 != 0, idx % 5 == 0
    w, losses, secs = train(pkg, x[tr], r[tr], y[tr], epochs=args.epochs, lr=args.lr, seed=args.seed, log=say)
    ev = evaluate(pkg, w, x[te], r[te], y[te])
    save(w, args.out)
    print(json.dumps({"rows": int(len(y)), "train_rows": int(tr.sum()), "train_seconds": round(secs, 3),
                      "losses": [round(float(v), 5) for v in losses], "heldout": ev, "weights": args.out}))


if __name__ == "__main__":
    main()

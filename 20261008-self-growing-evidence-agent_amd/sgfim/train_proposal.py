"""Train and evaluate the proposal network that orders the anytime frontier.

  python -m sgfim.train_proposal --train 200 --eval 20 --out weights.npz --report report.json

1. For every training seed, a workload-4 stream (500,000 txns, 256 items) is mined
   exactly at the 2% floor with tracing on; every frontier record born during the
   search gives one row of 24 features and a label: the fraction of the record's
   children that its round confirms as MFIs (--label yield, the default; a soft BCE
   target), or 1 when that round confirms any MFI (--label direct), or also when the
   record is an ancestor of such a record (--label ancestor).
2. The rows (a per-stream sample) train the 24-256-1 network on the GPU with the
   library's own BF16-MFMA forward/backward kernels (minibatch Adam).
3. On held-out seeds the anytime search runs twice, support-ordered (plain) and
   score-ordered (learned), in rounds of --batch-nodes records; the report holds both
   coverage curves (fraction of the final MFI set confirmed vs wall time and vs
   support counts), time-to-95%, and whether the final sets are identical.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from . import autotune, gen, learned
from ._lib import ORDER_LEARNED
from .core import Context


def _gen(args):
    seed, n_txn = args
    T = gen.workload4(seed, n_txn)
    return T.offsets, T.items


def streams(seeds, n_txn, workers):
    if workers <= 1:
        for s in seeds:
            yield s, _gen((s, n_txn))
        return
    with ProcessPoolExecutor(max_workers=workers) as ex:
        for s, r in zip(seeds, ex.map(_gen, [(s, n_txn) for s in seeds])):
            yield s, r


def anytime_curve(ctx: Context, min_count: int, learned_order: bool):
    """Run one anytime search round by round: (per-round records, final MFI set).

    MFIs are taken as flat arrays each round (no per-itemset Python objects inside the
    timed loop); the final set is built once at the end for the identity check."""
    rec = []
    parts = []
    n = 0
    # no cyclic garbage collection inside the timed loop: a full collection over the
    # millions of objects an earlier result holds stalls a round by a quarter second
    gc.collect()
    gc.disable()
    try:
        ctx.anytime_begin(min_count, ORDER_LEARNED if learned_order else 0)
        while True:
            st = ctx.anytime_step(1)
            r = ctx.anytime_take(convert=False)
            if r.count:
                parts.append(r.arrays)
                n += r.count
            rec.append((st["elapsed_s"], st["counts"], n))
            if st["done"]:
                break
    finally:
        gc.enable()
    found = {}
    for it, off, sup in parts:
        for i in range(len(sup)):
            found[frozenset(it[off[i]:off[i + 1]].tolist())] = int(sup[i])
    return rec, found


def compare_orders(ctx: Context, min_count: int, plain_first: bool = True):
    """Plain and learned anytime runs on the loaded data: (plain record, plain set, learned
    record, learned set). The first search after a load is slower than the next (about 17%
    of a 0.12 s search on 500,000 x 256 items), so an untimed run goes first and
    plain_first says which order is timed first."""
    anytime_curve(ctx, min_count, False)
    if plain_first:
        rp, fp = anytime_curve(ctx, min_count, False)
        rl, fl = anytime_curve(ctx, min_count, True)
    else:
        rl, fl = anytime_curve(ctx, min_count, True)
        rp, fp = anytime_curve(ctx, min_count, False)
    return rp, fp, rl, fl


def t_at(rec, frac, total, col):
    for r in rec:
        if r[2] >= frac * total:
            return r[col]
    return rec[-1][col]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=200)
    ap.add_argument("--eval", type=int, default=20)
    ap.add_argument("--n-txn", type=int, default=500_000)
    ap.add_argument("--floor", type=float, default=0.02)
    ap.add_argument("--per-stream", type=int, default=20_000)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--batch-nodes", type=int, default=4096)
    ap.add_argument("--label", choices=("yield", "direct", "ancestor"), default="yield",
                    help="yield: fraction of the record's children its round confirms as MFIs; "
                         "direct: the record's own round confirms an MFI; ancestor: also its ancestors")
    ap.add_argument("--workers", type=int, default=min(16, len(os.sched_getaffinity(0))))
    ap.add_argument("--out", default=learned.WEIGHTS_PATH)
    ap.add_argument("--eval-only", action="store_true",
                    help="skip tracing and training; evaluate the weights at --out")
    ap.add_argument("--report", default="")
    args = ap.parse_args(argv)
    rng = np.random.default_rng(12345)
    ctx = Context(device=0, batch_nodes=args.batch_nodes)
    if args.eval_only:
        ctx.set_scorer(*learned.load_weights(args.out))
        report = {"weights": args.out, "batch_nodes": args.batch_nodes, "floor": args.floor, "n_txn": args.n_txn,
                  "eval": []}
        evaluate(ctx, args, report)
        ctx.close()
        return
    X, Y = [], []
    t0 = time.time()
    for s, (off, items) in streams(range(args.train), args.n_txn, args.workers):
        ctx.load_csr(off, items)
        thr = gen.min_count(args.floor, ctx.n_txn)
        feat, lab, res = learned.collect_trace(ctx, thr, mode=args.label)
        if feat.shape[0] > args.per_stream:
            idx = rng.choice(feat.shape[0], args.per_stream, replace=False)
            feat, lab = feat[idx], lab[idx]
        X.append(feat)
        Y.append(lab)
        print(f"trace seed {s}: {len(res)} MFIs, {feat.shape[0]} rows, positive {lab.mean():.3f} "
              f"({time.time() - t0:.0f} s)", flush=True)
    X = np.concatenate(X)
    Y = np.concatenate(Y)
    tr = learned.Trainer(seed=0)
    tr.set_data(X, Y)
    t1 = time.time()
    losses = tr.train(args.epochs, args.batch, args.lr, seed=0)
    p = tr.forward(X, probs=True)
    acc = float(((p > 0.5) == (Y > 0.5)).mean())
    print(f"trained on {X.shape[0]} rows in {time.time() - t1:.1f} s: losses {np.round(losses, 4).tolist()}, "
          f"train accuracy {acc:.3f}", flush=True)
    params = tr.params()
    learned.save_weights(params, args.out, {"rows": int(X.shape[0]), "epochs": args.epochs, "lr": args.lr,
                                            "train_seeds": args.train, "n_txn": args.n_txn, "floor": args.floor})
    ctx.set_scorer(*params)
    report = {"train_rows": int(X.shape[0]), "label": args.label, "positive": float(Y.mean()),
              "losses": [float(x) for x in losses], "train_accuracy": acc,
              "batch_nodes": args.batch_nodes, "floor": args.floor, "n_txn": args.n_txn, "eval": []}
    evaluate(ctx, args, report)
    ctx.close()
    tr.close()


def evaluate(ctx: Context, args, report: dict):
    """Plain vs learned anytime runs on the held-out seeds; fills and writes the report."""
    seeds = range(1_000_000, 1_000_000 + args.eval)
    warm = True
    for k, (s, (off, items)) in enumerate(streams(seeds, args.n_txn, args.workers)):
        ctx.load_csr(off, items)
        autotune.apply(ctx, 1)   # the counting tile tuned for this device and input shape
        thr = gen.min_count(args.floor, ctx.n_txn)
        if warm:  # first held-out stream: warm the captured rounds of both orders
            anytime_curve(ctx, thr, True)
            warm = False
        rp = fp = rl = fl = None   # release the previous stream's results before timing
        rp, fp, rl, fl = compare_orders(ctx, thr, plain_first=k % 2 == 0)
        tot = len(fp)
        e = {"seed": s, "mfis": tot, "identical": fp == fl, "plain_first": k % 2 == 0,
             "plain": {"t95_s": t_at(rp, 0.95, tot, 0), "counts95": t_at(rp, 0.95, tot, 1),
                       "total_s": rp[-1][0], "total_counts": rp[-1][1], "rounds": len(rp)},
             "learned": {"t95_s": t_at(rl, 0.95, tot, 0), "counts95": t_at(rl, 0.95, tot, 1),
                         "total_s": rl[-1][0], "total_counts": rl[-1][1], "rounds": len(rl)},
             "curve_plain": rp, "curve_learned": rl}
        report["eval"].append(e)
        print(f"eval seed {s}: {tot} MFIs identical={e['identical']} t95 plain {e['plain']['t95_s']:.4f} s "
              f"learned {e['learned']['t95_s']:.4f} s; counts95 plain {e['plain']['counts95']} learned "
              f"{e['learned']['counts95']}", flush=True)
    ev = report["eval"]
    if ev:
        report["summary"] = {
            "all_identical": all(e["identical"] for e in ev),
            "median_t95_ratio": float(np.median([e["learned"]["t95_s"] / e["plain"]["t95_s"] for e in ev])),
            "median_counts95_ratio": float(np.median([e["learned"]["counts95"] / max(e["plain"]["counts95"], 1)
                                                      for e in ev])),
        }
        for frac in (0.25, 0.5):   # learned / plain, time and counts, to a part of the family
            for col, name in ((0, "t"), (1, "counts")):
                report["summary"][f"median_{name}{int(frac * 100)}_ratio"] = float(np.median(
                    [t_at(e["curve_learned"], frac, e["mfis"], col) / max(t_at(e["curve_plain"], frac, e["mfis"], col),
                                                                          1e-12) for e in ev]))
        print(json.dumps(report["summary"]), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f)


if __name__ == "__main__":
    sys.exit(main())

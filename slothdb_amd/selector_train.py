"""Collects the learned selector's training data on the GPU and trains it there.

    python -m slothdb_amd.selector_train [--queries N] [--rows-hi R] [--out PATH]

For each synthetic query (rows log-uniform in [1e4, rows_hi], distinct keys
log-uniform in [1, rows], Zipf s uniform in [0, 1.5], key = one INTEGER
(sparse ids, or dense ranks from an offset), one BIGINT, two INTEGER or two
BIGINT columns, i.e. 4, 8 or 16 key bytes;
SELECT key, COUNT(*), SUM(v) GROUP BY key):

  * the AUTO plan's sample features (gbagg_plan_features),
  * the exact group count (every path returns it; they must agree),
  * the device time of each path DIRECT, LDS_HASH, GLOBAL_HASH, RADIX (best
    of 2; DIRECT only where the key range allows it, else never the label).

Labels: the group count as selector.target (a correction to the sample's own
estimate) and the fastest path. The model trains on 80% of
the queries with the GPU training kernel; the other 20% report the distinct
error (selector vs the sample's default estimate) and the plan quality (time of
the selector's plan and of the default plan over the fastest plan). Writes the
parameters and these figures to slothdb_amd/selector_weights.json.
"""
import argparse
import math
import sys
import time

import numpy as np

from . import datagen, gbagg as G, selector as S

KEYS = ("i32", "i32dense", "i64", "2xi32", "2xi64")


def _mix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return x ^ (x >> np.uint64(31))


def make_query(rows, distinct, s, key, seed):
    """(Query, host columns) of one synthetic query."""
    if key == "i32":
        k, v = datagen.query_i32(rows, distinct, s, seed)
        q = G.Query([G.INTEGER, G.BIGINT], [0], [(G.COUNT_STAR, None), (G.SUM, 1)], nullable=[False, False])
        return q, [G.Column(G.INTEGER, k), G.Column(G.BIGINT, v)]
    ranks = datagen.zipf_ranks(rows, distinct, s, seed=seed)
    ids = _mix64(ranks ^ np.uint64(seed & 0xFFFFFFFF))
    v = (ranks.astype(np.int64) % 1000) - 500
    if key == "i32dense":
        k = (ranks.astype(np.int64) - (seed % 100_000)).astype(np.int32)
        q = G.Query([G.INTEGER, G.BIGINT], [0], [(G.COUNT_STAR, None), (G.SUM, 1)], nullable=[False, False])
        return q, [G.Column(G.INTEGER, k), G.Column(G.BIGINT, v)]
    if key == "2xi64":
        ids2 = _mix64(ranks ^ np.uint64(0xA5A5A5A5 ^ (seed & 0xFFFF)))
        q = G.Query([G.BIGINT, G.BIGINT, G.BIGINT], [0, 1], [(G.COUNT_STAR, None), (G.SUM, 2)],
                    nullable=[False] * 3)
        return q, [G.Column(G.BIGINT, ids.view(np.int64)), G.Column(G.BIGINT, ids2.view(np.int64)),
                   G.Column(G.BIGINT, v)]
    if key == "i64":
        q = G.Query([G.BIGINT, G.BIGINT], [0], [(G.COUNT_STAR, None), (G.SUM, 1)], nullable=[False, False])
        return q, [G.Column(G.BIGINT, ids.view(np.int64)), G.Column(G.BIGINT, v)]
    a = (ids >> np.uint64(32)).astype(np.uint32).view(np.int32)
    b = (ids & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32)
    q = G.Query([G.INTEGER, G.INTEGER, G.BIGINT], [0, 1], [(G.COUNT_STAR, None), (G.SUM, 2)],
                nullable=[False] * 3)
    return q, [G.Column(G.INTEGER, a), G.Column(G.INTEGER, b), G.Column(G.BIGINT, v)]


def specs(n, rows_hi, seed):
    out = []
    for i in range(n):
        u = [datagen.rand(seed, 60 + j, i) / 2.0 ** 64 for j in range(4)]
        rows = int(round(math.exp(math.log(1e4) + u[0] * (math.log(rows_hi) - math.log(1e4)))))
        distinct = max(1, min(rows, int(round(math.exp(u[1] * math.log(rows))))))
        out.append((rows, distinct, 1.5 * u[2], KEYS[int(u[3] * len(KEYS)) % len(KEYS)],
                    (seed * 7919 + i) & 0xFFFFFFFFFFFF))
    return out


def run_plan(q, dcols, strategy, reps=2, want_features=False):
    q.strategy = strategy
    op = G.GpuHashAggregate(q, {"features": 1} if want_features else None)
    best, groups, feat = 1e30, None, None
    for _ in range(reps):
        op.Init()
        op.consume_device(dcols)
        op.finalize()
        info = op.plan_info()
        # a forced DIRECT over a too-wide key range runs another plan: not a DIRECT time
        if strategy == G.DIRECT and info["strategy"] != G.DIRECT:
            best = math.inf
        else:
            best = min(best, info["ms_total"])
        groups = info["groups"]
    if want_features:
        feat = S.features(op)
    used = op.plan_info()["strategy"]
    op.close()
    q.strategy = G.AUTO
    return best, groups, feat, used


def collect(n, rows_hi, seed, log=print, first=0, count=None, dump=None):
    X, Y, T, AUTO_T, DEF, G_ = [], [], [], [], [], []
    t0 = time.time()
    todo = specs(n, rows_hi, seed)[first:None if count is None else first + count]
    for i, (rows, distinct, s, key, qseed) in enumerate(todo, start=first):
        q, cols = make_query(rows, distinct, s, key, qseed)
        dcols = [G.DeviceColumn(c) for c in cols]
        t_auto, g_auto, feat, used = run_plan(q, dcols, G.AUTO, want_features=True)
        times, groups = [], []
        for st in S.PATHS:
            t, g, _, _ = run_plan(q, dcols, st)
            times.append(t)
            groups.append(g)
        if len(set(groups + [g_auto])) != 1:
            raise RuntimeError(f"query {i}: paths disagree on the group count {groups} {g_auto}")
        for c in dcols:
            c.free()
        X.append(feat)
        Y.append([float(S.target(g_auto, feat)), int(np.argmin(times))])
        G_.append(g_auto)
        T.append(times)
        AUTO_T.append(t_auto)
        DEF.append(float(S.default_estimate(feat)))
        if (i + 1) % 50 == 0:
            log(f"[selector] collected {i + 1}/{n} queries in {time.time() - t0:.0f}s", flush=True)
            if dump:  # progress survives an interrupted collection
                _save(dump, X, Y, T, AUTO_T, DEF, G_)
    out = (np.array(X, np.float32), np.array(Y, np.float32), np.array(T), np.array(AUTO_T), np.array(DEF),
           np.array(G_, np.float64))
    if dump:
        _save(dump, *out)
    return out


def _save(path, X, Y, T, AUTO_T, DEF, GT):
    np.savez(path, X=np.array(X, np.float32), Y=np.array(Y, np.float32), T=np.array(T), AUTO_T=np.array(AUTO_T),
             DEF=np.array(DEF), GT=np.array(GT, np.float64))


def _load(paths):
    parts = [np.load(p) for p in paths.split(",")]
    return tuple(np.concatenate([z[k] for z in parts]) for k in ("X", "Y", "T", "AUTO_T", "DEF", "GT"))


def evaluate(params, X, Y, T, AUTO_T, DEF, true):
    out = S.infer(params, X)
    pred = S.distinct_from_output(out[:, 0], X)
    err_sel = np.abs(np.log2(pred / np.maximum(true, 1)))
    err_def = np.abs(np.log2(DEF / np.maximum(true, 1)))
    path = np.argmax(out[:, 1:], axis=1)
    # the operator's rule: the selector's plan unless it is 4x off the default estimate
    used = (pred <= 4 * DEF) & (pred * 4 >= DEF)
    t_path = T[np.arange(len(T)), path]
    # DIRECT over a too-wide key range falls back to LDS_HASH in the operator
    t_path = np.where(np.isfinite(t_path), t_path, T[:, S.PATHS.index(G.LDS_HASH)])
    t_sel = np.where(used, t_path, AUTO_T)
    best = T.min(axis=1)
    return {
        "queries": int(len(X)),
        "distinct_abs_log2_err_selector": {"median": float(np.median(err_sel)), "p90": float(np.percentile(err_sel, 90))},
        "distinct_abs_log2_err_default": {"median": float(np.median(err_def)), "p90": float(np.percentile(err_def, 90))},
        "path_accuracy": float(np.mean(path == Y[:, 1])),
        "selector_used": float(np.mean(used)),
        "time_selector_over_best": float(t_sel.sum() / best.sum()),
        "time_default_over_best": float(AUTO_T.sum() / best.sum()),
        "time_radix_only_over_best": float(T[:, S.PATHS.index(G.RADIX)].sum() / best.sum()),
        "path_label_counts": [int(np.sum(Y[:, 1] == c)) for c in range(len(S.PATHS))],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--queries", type=int, default=1000)
    ap.add_argument("--rows-hi", type=float, default=1e7)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default=S.WEIGHTS)
    ap.add_argument("--dump", help="also save the collected data (.npz), every 50 queries")
    ap.add_argument("--first", type=int, default=0, help="collect queries first .. first + count - 1 only")
    ap.add_argument("--count", type=int, default=None)
    ap.add_argument("--collect-only", action="store_true")
    ap.add_argument("--data", help="train on data saved by --dump (comma-separated files) instead of collecting")
    args = ap.parse_args()
    if args.data:
        X, Y, T, AUTO_T, DEF, GT = _load(args.data)
    else:
        X, Y, T, AUTO_T, DEF, GT = collect(args.queries, args.rows_hi, args.seed, first=args.first,
                                           count=args.count, dump=args.dump)
        if args.collect_only:
            return 0
    n_train = int(0.8 * len(X))
    mu = X[:n_train].mean(axis=0)
    sigma = X[:n_train].std(axis=0) + 1e-3
    p0 = S.init_params(mu, sigma, seed=args.seed)
    t = time.time()
    params, loss = S.train(X[:n_train], Y[:n_train], p0, epochs=args.epochs, lr=args.lr)
    train_s = time.time() - t
    ev = evaluate(params, X[n_train:], Y[n_train:], T[n_train:], AUTO_T[n_train:], DEF[n_train:], GT[n_train:])
    ev_train = evaluate(params, X[:n_train], Y[:n_train], T[:n_train], AUTO_T[:n_train], DEF[:n_train], GT[:n_train])
    meta = {"queries": args.queries, "rows_hi": args.rows_hi, "epochs": args.epochs, "lr": args.lr,
            "seed": args.seed, "train_rows": n_train, "train_seconds": train_s,
            "loss_first": float(loss[0]), "loss_last": float(loss[-1]), "held_out": ev, "train": ev_train}
    S.save(params, args.out, meta)
    print(f"[selector] trained in {train_s:.1f}s, loss {loss[0]:.4f} -> {loss[-1]:.4f}", flush=True)
    print(f"[selector] held-out {ev}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Collect ADDS phase traces on seeded training graphs, train the delta
network on the device, and write the model (and a summary) for the engine.

    python -m mi355sssp.train_delta_mlp --graphs 2000 --out mi355sssp/data/delta_mlp.json

Collection can be split over runs (--first/--graphs write one .npz shard
each); --train-only trains on the shards given with --shards.  --heads-for
adds the first-phase records (learn.collect_heads) to a shard collected
without them.  Training graphs come from learn.training_specs(seed);
evaluation graphs use other seeds.

The light-width stage trains the network's fourth output into a shipped
model (the teacher) under pulls:

    python -m mi355sssp.train_delta_mlp --light-from mi355sssp/data/delta_mlp.json \
        --graphs 240 --large-grids 8 --out mi355sssp/data/delta_mlp.json
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    __package__ = "mi355sssp"

from . import learn as L  # noqa: E402


def log(msg: str) -> None:
    print(f"[train_delta_mlp] {msg}", flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=2000, help="training graphs in this shard")
    ap.add_argument("--first", type=int, default=0, help="index of the first spec of this shard")
    ap.add_argument("--total", type=int, default=2000, help="size of the full spec list")
    ap.add_argument("--seed", type=int, default=16001, help="training spec seed")
    ap.add_argument("--kind", choices=["rmat", "route", "grid"], default=None,
                    help="collect only the training specs of this graph kind")
    ap.add_argument("--large-grids", type=int, default=0,
                    help="collect this many 4e6-3e7 vertex grids instead (learn.large_grid_specs)")
    ap.add_argument("--max-rmat-scale", type=int, default=None,
                    help="leave out R-MAT specs above this scale (21: the training set of scales 16-21 "
                         "and grids of 1e5-2e6 vertices, evaluated on held-out scale 22 / 24M graphs)")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--no-heads", action="store_true",
                    help="collect without the two-phase schedule search (first-phase records)")
    ap.add_argument("--budget-s", type=float, default=None, help="stop collecting after this long")
    ap.add_argument("--shard-out", default=None, help="write the collected shard here (.npz)")
    ap.add_argument("--shards", nargs="*", default=[], help="train on these shards")
    ap.add_argument("--train-only", action="store_true")
    ap.add_argument("--heads-for", default=None,
                    help="search the first phase of every graph of this shard and write the shard "
                         "with its first-phase records to --shard-out")
    ap.add_argument("--head-weight", type=float, default=1.0,
                    help="weight of a graph's first-phase record relative to all of its "
                         "phase-boundary records together")
    ap.add_argument("--epochs", type=int, default=80)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=None, help="model JSON to write")
    ap.add_argument("--light-from", default=None,
                    help="teacher model: collect light-width data on --graphs training specs (and "
                         "--large-grids large grids) with pulls on, train its light-width output, "
                         "write --out (learn.collect_light, learn.fit_light)")
    ap.add_argument("--pulled", action="store_true",
                    help="with --light-from: on R-MAT graphs search static settings with pulls "
                         "(learn.collect_pulled) and train all four outputs there")
    ap.add_argument("--grid-static", action="store_true",
                    help="with --pulled: on grid graphs too, search static settings "
                         "(learn.collect_grid_static) and train all four outputs there")
    ap.add_argument("--rmat-teacher", action="store_true",
                    help="light stage: the records of R-MAT graphs target the teacher's own outputs "
                         "(its decisions kept there; the new targets apply to the other graphs)")
    ap.add_argument("--light-epochs", type=int, default=40)
    ap.add_argument("--light-lr", type=float, default=1e-3)
    ap.add_argument("--graph-weight", choices=["class", "time", "sqrt-time", "equal"], default="class",
                    help="weight of a graph's records: every size class (R-MAT scale, grid kind x "
                         "decade of n) the same, its plain-ADDS time (the time at stake), its square "
                         "root, or every graph the same")
    args = ap.parse_args()

    shards = list(args.shards)
    if args.light_from:
        return light_stage(args)
    if args.heads_for:
        z = np.load(args.heads_for, allow_pickle=False)
        metas = json.loads(str(z["meta"]))
        t0 = time.time()
        y = z["y"].copy()
        hx, hy, hidx, hm = L.collect_heads(metas, z["x"], y, z["w"], log=log, reps=args.reps,
                                           time_budget_s=args.budget_s)
        for i, m in zip(hidx, hm):
            metas[int(i)].update(m)
        gains = [metas[int(i)]["plain_ms"] / metas[int(i)]["head_ms"] for i in hidx]
        log(f"first phases of {len(hidx)} graphs in {time.time() - t0:.0f}s; plain/best two-phase "
            f"mean {np.mean(gains) if gains else 0:.3f}")
        np.savez_compressed(args.shard_out, x=z["x"], y=y, w=z["w"], hx=hx, hy=hy, hidx=hidx,
                            meta=np.array(json.dumps(metas)))
        return 0
    if not args.train_only:
        if args.large_grids > 0:
            specs = L.large_grid_specs(args.large_grids, args.seed)
        else:
            specs = L.training_specs(args.total, args.seed)[args.first:args.first + args.graphs]
        if args.kind:
            specs = [s for s in specs if s["kind"] == args.kind]
        if args.max_rmat_scale is not None:
            specs = [s for s in specs if s["kind"] != "rmat" or s["scale"] <= args.max_rmat_scale]
        t0 = time.time()
        x, y, w, metas, hx, hy, hidx = L.collect(specs, log=log, reps=args.reps, time_budget_s=args.budget_s,
                                                  heads=not args.no_heads)
        log(f"collected {len(metas)} graphs, {len(x)} samples in {time.time() - t0:.0f}s")
        if args.shard_out:
            np.savez_compressed(args.shard_out, x=x, y=y, w=w, hx=hx, hy=hy, hidx=hidx,
                                meta=np.array(json.dumps(metas)))
            shards.append(args.shard_out)
        else:
            shards = []
            data = [(x, y, w, metas, hx, hy, hidx)]
    if shards:
        data = []
        for p in shards:
            z = np.load(p, allow_pickle=False)
            empty = (np.zeros((0, L.N_IN), np.float32), np.zeros((0, L.N_OUT), np.float32), np.zeros(0, np.int64))
            h = (z["hx"], z["hy"], z["hidx"]) if "hx" in z.files else empty
            data.append((z["x"], z["y"], z["w"], json.loads(str(z["meta"]))) + h)
    x = np.concatenate([d[0] for d in data])
    y = np.concatenate([d[1] for d in data])
    metas = [m for d in data for m in d[3]]
    if args.out is None or len(x) == 0:
        return 0
    # every shard (or collection) gives its graphs equal weights; bring them to
    # one scale (1 per graph), then re-weight
    w = np.concatenate([d[2] / (d[2].sum() / max(len(d[3]), 1)) for d in data])
    if args.graph_weight == "class":
        w = L.class_weights(w, metas)
    elif args.graph_weight != "equal":
        w = L.time_weights(w, metas, 1.0 if args.graph_weight == "time" else 0.5)
    # first-phase records: each weighs head_weight x its graph's total
    counts = np.concatenate([L.record_counts(d[3], d[2]) for d in data])
    bounds = np.concatenate([[0], np.cumsum(counts)])
    per_graph = np.array([w[bounds[i]:bounds[i + 1]].sum() for i in range(len(metas))])
    base = np.cumsum([0] + [len(d[3]) for d in data])[:-1]
    hidx = np.concatenate([d[6] + b for d, b in zip(data, base)]).astype(np.int64)
    hx = np.concatenate([d[4] for d in data]).reshape(-1, L.N_IN)
    hy = np.concatenate([d[5] for d in data]).reshape(-1, L.N_OUT)
    hw = args.head_weight * per_graph[hidx] if len(hidx) else np.zeros(0)
    # a boundary record with at most one queued item has the first phase's
    # input (the end of a run): left out, its decision does not matter
    keep = x[:, 14] > 1.0 + 1e-6
    x = np.concatenate([x[keep], hx])
    y = np.concatenate([y[keep], hy])
    w = np.concatenate([w[keep], hw]).astype(np.float32)
    w = w * (len(w) / w.sum())
    log(f"{int(keep.sum())} phase-boundary records ({int((~keep).sum())} end-of-run records left out), "
        f"{len(hidx)} first-phase records")
    t0 = time.time()
    model, losses = L.fit(x, y, w, epochs=args.epochs, batch=args.batch, lr=args.lr)
    log(f"trained on {len(x)} samples from {len(metas)} graphs in {time.time() - t0:.1f}s: "
        f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    pred = model.forward(x)
    best_gain = [m["plain_ms"] / m["best_ms"] for m in metas]
    head_gain = [m["plain_ms"] / m["head_ms"] for m in metas if "head_ms" in m]
    summary = {
        "graphs": len(metas), "samples": int(len(x)), "epochs": args.epochs, "batch": args.batch,
        "lr": args.lr, "train_seed": args.seed, "graph_weight": args.graph_weight,
        "loss_first_epoch": float(losses[0]),
        "loss_last_epoch": float(losses[-1]),
        "mean_abs_err_log2_scale": float(np.abs(pred[:, 0] - y[:, 0]).mean()),
        "mean_abs_err_active": float(np.abs(pred[:, 1] - y[:, 1]).mean()),
        "mean_abs_err_log2_share": float(np.abs(pred[:, 2] - y[:, 2]).mean()),
        "oracle_static_gain_mean": float(np.mean(best_gain)),
        "oracle_static_gain_median": float(np.median(best_gain)),
        "oracle_two_phase_gain_mean": float(np.mean(head_gain)) if head_gain else None,
        "first_phase_records": int(len(hidx)), "head_weight": args.head_weight, "heads": list(L.HEADS),
        "worker_shares": list(L.SHARES), "settings": [list(s) for s in L.SETTINGS],
    }
    model.save(args.out, meta=summary)
    log(f"wrote {args.out}: {json.dumps(summary)}")
    return 0


def light_stage(args) -> int:
    """Collect light-width data under the teacher (or read --shards), train
    the fourth output, write --out and, with --shard-out, the data."""
    teacher = L.DeltaModel.load(args.light_from)
    if args.shards:
        parts = [np.load(p, allow_pickle=False) for p in args.shards]
        x = np.concatenate([z["x"] for z in parts])
        y = np.concatenate([z["y"] for z in parts])
        w = np.concatenate([z["w"] for z in parts])
        metas = [m for z in parts for m in json.loads(str(z["meta"]))]
    else:
        specs = L.training_specs(args.total, args.seed)[args.first:args.first + args.graphs]
        if args.max_rmat_scale is not None:
            specs = [s for s in specs if s["kind"] != "rmat" or s["scale"] <= args.max_rmat_scale]
        if args.large_grids > 0:
            specs += L.large_grid_specs(args.large_grids, args.seed)
        t0 = time.time()
        if args.pulled:
            x, y, w, metas = L.collect_pulled(specs, teacher, log=log, reps=args.reps, time_budget_s=args.budget_s,
                                              grid_static=args.grid_static)
        else:
            x, y, w, metas = L.collect_light(specs, teacher, log=log, reps=args.reps, time_budget_s=args.budget_s)
        log(f"light: {len(metas)} graphs, {len(x)} records in {time.time() - t0:.0f}s")
        if args.shard_out:
            np.savez_compressed(args.shard_out, x=x, y=y, w=w, meta=np.array(json.dumps(metas)))
    if args.out is None or len(x) == 0:
        return 0
    if args.rmat_teacher:
        kinds = np.repeat(np.array([m["kind"] for m in metas]), [m["records"] for m in metas])
        keep = kinds == "rmat"
        y = y.copy()
        if keep.any():
            y[keep] = teacher.forward(x[keep])
        log(f"light: {int(keep.sum())} R-MAT records target the teacher's outputs")
    if args.graph_weight == "class":
        w = L.class_weights(w, metas)
    w = (w * (len(w) / w.sum())).astype(np.float32)
    t0 = time.time()
    model, losses = L.fit_light(teacher, x, y, w, epochs=args.light_epochs, lr=args.light_lr)
    pred = model.forward(x)
    first = np.concatenate([[0], np.cumsum([m["records"] for m in metas])[:-1]]).astype(int)
    chosen = np.exp2(np.clip(pred[first, 3], -5.0, 0.0))
    summary = dict(teacher.meta or {})
    summary["light_stage"] = {
        "graphs": len(metas), "samples": int(len(x)), "epochs": args.light_epochs, "lr": args.light_lr,
        "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1]),
        "light_scales": list(L.LIGHT_SCALES),
        "targets": {str(k): int(sum(1 for m in metas if m["light"] == k))
                    for k in sorted(set(L.LIGHT_SCALES) | set(L.PULL_LIGHTS), reverse=True)},
        "mean_abs_err_log2_light_first_phase": float(np.abs(pred[first, 3] - y[first, 3]).mean()),
        "mean_abs_dev_teacher_outputs": float(np.abs(pred[:, :3] - y[:, :3]).mean()),
        "oracle_light_gain_mean": float(np.mean([m["light_costs"]["1.0"] / min(m["light_costs"].values())
                                                 for m in metas if "best" not in m] or [1.0])),
        "pulled_graphs": int(sum(1 for m in metas if "best" in m and m["kind"] == "rmat")),
        "grid_static_graphs": int(sum(1 for m in metas if "best" in m and m["kind"] != "rmat")),
        "grid_static_over_teacher_mean": float(np.mean([m["teacher_ms"] / m["best_ms"] for m in metas
                                                        if "best" in m and m["kind"] != "rmat"] or [1.0])),
        "pulled_static_over_teacher_mean": float(np.mean([m["teacher_ms"] / m["best_ms"] for m in metas
                                                          if "best" in m and m["kind"] == "rmat"] or [1.0])),
        "pulled_static_over_plain_mean": float(np.mean([m["plain_ms"] / m["best_ms"] for m in metas
                                                        if "best" in m and m["kind"] == "rmat"] or [1.0])),
        "network_light_choice_mean": float(chosen.mean()),
    }
    model.save(args.out, meta=summary)
    log(f"trained in {time.time() - t0:.1f}s, wrote {args.out}: {json.dumps(summary['light_stage'])}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

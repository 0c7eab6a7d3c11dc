"""Learned delta policy for ADDS: the model, phase traces, trace datasets and
training.  Every forward pass, gradient and optimiser step runs on the device
through the C ABI (mi355_mlp_*); this module only moves arrays.

The network maps the 32 features of a phase boundary to (log2 of the bucket
width relative to the heuristic delta0, number of buckets served at once,
log2 of the share of the persistent workgroups that claim work).  It is
trained by regression on traces that ADDS itself records: for each training
graph ADDS is timed under static settings (first the worker share, then the
width scale x active buckets grid at the best share), and every timed setting
is also run once with phase records.  Then two-phase schedules are timed (the
first phase, the source's bucket, at a width of HEADS; every later phase at
the best static setting or that setting at a width of TAIL_SCALES): on R-MAT
graphs a narrow first bucket (the ramp-up, where a wide bucket relaxes
vertices that are not settled yet) and wide later buckets beat every static
setting.  The fastest schedule gives two targets: its first phase for one
record with the input the engine gives the network before the first claim,
and its tail for every phase-boundary record of the graph.  The work done and
relaxations wasted of each static setting are kept in the metadata.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time

import numpy as np

from ._abi import (ADDS_TRACE_STRIDE, ALG_ADDS, MlpParams, MlpTrainOptions, SsspOptions, SsspStats,
                   check, engine)

N_IN, N_HID, N_OUT = 32, 64, 4
N_GRAPH = 14  # graph features: inputs 0..13
_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(_HERE, "data", "delta_mlp.json")

# static settings searched per training graph: the worker share (at width
# scale 1/2, 2 active buckets), then (width scale, active buckets) at the best
# share
SHARES = (1.0, 0.5, 0.25, 0.125)
SCALES = (0.25, 0.5, 1.0, 2.0, 4.0)
ACTIVE = (1, 2, 4)
SETTINGS = [(s, a) for a in ACTIVE for s in SCALES]
# two-phase schedules searched per training graph: first-phase width scales,
# and the tail widths tried besides the best static setting's
HEADS = (0.0625, 0.125, 0.25)
TAIL_SCALES = (1.0, 2.0, 4.0)


def _f32(a, shape=None) -> np.ndarray:
    out = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        out = out.reshape(shape)
    return out


class DeltaModel:
    """fp32 master weights of the 32 -> 64 -> 4 network plus the input
    normalisation (not trained)."""

    def __init__(self, w1, b1, w2, b2, mean=None, inv_std=None, feat_lo=None, feat_hi=None):
        self.w1 = _f32(w1, (N_HID, N_IN))
        self.b1 = _f32(b1, (N_HID,))
        self.w2 = _f32(w2, (N_OUT, N_HID))
        self.b2 = _f32(b2, (N_OUT,))
        self.mean = _f32(np.zeros(N_IN) if mean is None else mean, (N_IN,))
        self.inv_std = _f32(np.ones(N_IN) if inv_std is None else inv_std, (N_IN,))
        # training range of the graph features (inputs 0..13); lo == hi
        # (the default) leaves a feature unbounded
        self.feat_lo = _f32(np.zeros(N_GRAPH) if feat_lo is None else feat_lo, (N_GRAPH,))
        self.feat_hi = _f32(np.zeros(N_GRAPH) if feat_hi is None else feat_hi, (N_GRAPH,))

    @classmethod
    def init(cls, seed: int = 0) -> "DeltaModel":
        rng = np.random.default_rng(seed)
        w1 = rng.normal(0.0, math.sqrt(2.0 / N_IN), (N_HID, N_IN))
        w2 = rng.normal(0.0, 0.1 * math.sqrt(1.0 / N_HID), (N_OUT, N_HID))
        return cls(w1, np.zeros(N_HID), w2, np.array([0.0, 1.0, 0.0, 0.0]))

    # -- conversions --------------------------------------------------------
    def to_c(self) -> MlpParams:
        p = MlpParams()
        C.memmove(p.w1, self.w1.ctypes.data, self.w1.nbytes)
        C.memmove(p.b1, self.b1.ctypes.data, self.b1.nbytes)
        C.memmove(p.w2, self.w2.ctypes.data, self.w2.nbytes)
        C.memmove(p.b2, self.b2.ctypes.data, self.b2.nbytes)
        C.memmove(p.feat_mean, self.mean.ctypes.data, self.mean.nbytes)
        C.memmove(p.feat_inv_std, self.inv_std.ctypes.data, self.inv_std.nbytes)
        C.memmove(p.graph_feat_lo, self.feat_lo.ctypes.data, self.feat_lo.nbytes)
        C.memmove(p.graph_feat_hi, self.feat_hi.ctypes.data, self.feat_hi.nbytes)
        return p

    @classmethod
    def from_c(cls, p: MlpParams) -> "DeltaModel":
        return cls(np.ctypeslib.as_array(p.w1).copy(), np.ctypeslib.as_array(p.b1).copy(),
                   np.ctypeslib.as_array(p.w2).copy(), np.ctypeslib.as_array(p.b2).copy(),
                   np.ctypeslib.as_array(p.feat_mean).copy(),
                   np.ctypeslib.as_array(p.feat_inv_std).copy(),
                   np.ctypeslib.as_array(p.graph_feat_lo).copy(),
                   np.ctypeslib.as_array(p.graph_feat_hi).copy())

    def flat(self) -> np.ndarray:
        return np.concatenate([self.w1.ravel(), self.b1, self.w2.ravel(), self.b2])

    def save(self, path: str, meta: dict | None = None) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        doc = {"w1": self.w1.tolist(), "b1": self.b1.tolist(), "w2": self.w2.tolist(),
               "b2": self.b2.tolist(), "feat_mean": self.mean.tolist(),
               "feat_inv_std": self.inv_std.tolist(), "graph_feat_lo": self.feat_lo.tolist(),
               "graph_feat_hi": self.feat_hi.tolist(), "meta": meta or {}}
        with open(path, "w") as f:
            json.dump(doc, f)

    @classmethod
    def load(cls, path: str) -> "DeltaModel":
        with open(path) as f:
            d = json.load(f)
        w2, b2 = np.asarray(d["w2"], np.float32), np.asarray(d["b2"], np.float32)
        if w2.shape[0] == N_OUT - 1:
            # a model without the light-width output: the light edges follow
            # the first bucket width (output 0: log2 1)
            w2 = np.vstack([w2, np.zeros((1, N_HID), np.float32)])
            b2 = np.append(b2, np.float32(0.0))
        m = cls(d["w1"], d["b1"], w2, b2, d["feat_mean"], d["feat_inv_std"],
                d.get("graph_feat_lo"), d.get("graph_feat_hi"))
        m.meta = d.get("meta", {})
        return m

    # -- device compute -----------------------------------------------------
    def forward(self, x) -> np.ndarray:
        x = _f32(x, (-1, N_IN))
        out = np.empty((x.shape[0], N_OUT), np.float32)
        check(engine().mi355_mlp_forward(C.byref(self.to_c()), x.ctypes.data, x.shape[0],
                                         out.ctypes.data))
        return out

    def loss_grad(self, x, y, w=None):
        x = _f32(x, (-1, N_IN))
        y = _f32(y, (-1, N_OUT))
        wp = None if w is None else _f32(w, (-1,))
        loss = C.c_float()
        g = MlpParams()
        check(engine().mi355_mlp_loss_grad(C.byref(self.to_c()), x.ctypes.data, y.ctypes.data,
                                           None if wp is None else wp.ctypes.data, x.shape[0],
                                           C.byref(loss), C.byref(g)))
        return float(loss.value), DeltaModel.from_c(g)

    def train(self, x, y, w=None, epochs: int = 60, batch: int = 512, lr: float = 2e-3,
              weight_decay: float = 0.0, seed: int = 2026) -> np.ndarray:
        x = _f32(x, (-1, N_IN))
        y = _f32(y, (-1, N_OUT))
        wp = None if w is None else _f32(w, (-1,))
        o = MlpTrainOptions()
        engine().mi355_mlp_train_options_default(C.byref(o))
        o.epochs, o.batch, o.lr, o.weight_decay, o.seed = epochs, batch, lr, weight_decay, seed
        p = self.to_c()
        losses = np.zeros(epochs, np.float32)
        check(engine().mi355_mlp_train(C.byref(p), x.ctypes.data, y.ctypes.data,
                                       None if wp is None else wp.ctypes.data, x.shape[0],
                                       C.byref(o), losses.ctypes.data))
        t = DeltaModel.from_c(p)
        self.w1, self.b1, self.w2, self.b2 = t.w1, t.b1, t.w2, t.b2
        return losses

    def install(self) -> None:
        """Make this the model mi355_sssp uses when use_mlp is set."""
        check(engine().mi355_adds_mlp_set(C.byref(self.to_c())))


def clear_delta_model() -> None:
    check(engine().mi355_adds_mlp_clear())


def set_schedule(table) -> None:
    """Phase schedule for use_mlp=2: rows (width scale, active buckets, worker
    share); phase i runs with row min(i, len - 1).  None or [] clears it."""
    t = np.ascontiguousarray(np.zeros((0, 3)) if table is None else table, np.float32).reshape(-1, 3)
    check(engine().mi355_adds_schedule_set(t.ctypes.data if len(t) else None, len(t)))


def load_default_delta_model(install: bool = True) -> DeltaModel | None:
    if not os.path.exists(DEFAULT_MODEL):
        return None
    m = DeltaModel.load(DEFAULT_MODEL)
    if install:
        m.install()
    return m


# -- traces -----------------------------------------------------------------
def trace(dg, source: int, cap: int = 4096, **opts):
    """One ADDS run with phase records: (records [phases x 40], stats)."""
    o = dg.options(ALG_ADDS, **opts)
    buf = np.zeros((cap, ADDS_TRACE_STRIDE), np.float32)
    phases = C.c_int32()
    st = SsspStats()
    check(engine().mi355_adds_trace(dg.handle, int(source), C.byref(o), buf.ctypes.data, cap,
                                    C.byref(phases), C.byref(st), None))
    return buf[: min(phases.value, cap)].copy(), st


_default_workers = None


def default_workers() -> int:
    global _default_workers
    if _default_workers is None:
        w = C.c_int32()
        check(engine().mi355_adds_default_workers(C.byref(w)))
        _default_workers = int(w.value)
    return _default_workers


def workers_for(share: float) -> int:
    """Launch size of a static worker share (0 = the default launch)."""
    return 0 if share >= 1.0 else max(16, int(round(default_workers() * share)))


def timed_adds(dg, source: int, scale: float, active: int, delta0: float, reps: int = 2,
               use_mlp: bool = False, share: float = 1.0):
    """Best kernel time (ms) over reps of ADDS at width scale * delta0."""
    o = dg.options(ALG_ADDS, delta=0.0 if use_mlp else scale * delta0, active_buckets=active,
                   use_mlp=int(use_mlp), worker_blocks=workers_for(share))
    best, st_best = float("inf"), None
    for _ in range(reps):
        st = dg.sssp_labels(int(source), o)
        if st.kernel_ms < best:
            best, st_best = st.kernel_ms, st
    return best, st_best


# -- training graphs --------------------------------------------------------
def training_specs(count: int, seed: int):
    """Seeded mix of R-MAT (scale 16-21, every 20th at 22; fp32 weights) and perturbed grids of
    1e5-2e6 vertices (routing grids with lognormal integer weights, 4-neighbour
    grids with uniform integer weights)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        kind = ("rmat", "route", "grid")[i % 3]
        gseed = int(rng.integers(1, 2**62))
        if kind == "rmat":
            scale = int(rng.integers(16, 22))
            if i % 60 == 0:
                scale = 22  # every 20th R-MAT graph at scale 22 (4.2M vertices)
            specs.append({"kind": kind, "scale": scale, "seed": gseed})
        else:
            nv = math.exp(rng.uniform(math.log(1e5), math.log(2e6)))
            side = int(round(math.sqrt(nv)))
            specs.append({"kind": kind, "side": side, "seed": gseed})
    return specs


def large_grid_specs(count: int, seed: int):
    """Seeded routing and 4-neighbour grids of 4e6-3e7 vertices (the size of
    a full-chip routing graph): few of them, since each takes seconds per
    setting."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        gseed = int(rng.integers(1, 2**62))
        nv = math.exp(rng.uniform(math.log(4e6), math.log(3e7)))
        specs.append({"kind": ("route", "grid")[i % 2], "side": int(round(math.sqrt(nv))), "seed": gseed})
    return specs


def make_graph(spec):
    from . import graph as G

    if spec["kind"] == "rmat":
        return G.rmat(spec["scale"], 16, seed=spec["seed"])
    if spec["kind"] == "route":
        return G.routing_grid(spec["side"], seed=spec["seed"])
    return G.grid4(spec["side"], spec["side"], 1, 100, seed=spec["seed"])


def collect_graph(dg, source: int, reps: int = 2, max_records: int = 24):
    """Time every static setting on one graph, pick the fastest, and return
    trace records of a run at every timed setting, all with the fastest
    setting as their target.  (Records only from the fastest run would carry
    the target in the delta-ratio feature itself, and a network trained on
    them follows its own previous delta instead of correcting it.)"""
    _, st0 = trace(dg, source, cap=1)
    delta0 = st0.initial_delta
    m_r = int(st0.reached_edges)
    costs = {}

    def run(s, a, w):
        ms, st = timed_adds(dg, source, s, a, delta0, reps, share=w)
        # (kernel ms, edges relaxed, relaxations beyond one per reached edge)
        costs[(s, a, w)] = (ms, int(st.edges_relaxed), int(st.edges_relaxed) - m_r)

    for w in SHARES:
        run(0.5, 2, w)
    w_best = min((k for k in costs), key=lambda k: costs[k][0])[2]
    for (s, a) in SETTINGS:
        if (s, a, w_best) not in costs:
            run(s, a, w_best)
    if (1.0, 1, 1.0) not in costs:
        run(1.0, 1, 1.0)
    best = min(costs, key=lambda k: costs[k][0])
    # light edges at the first width in these searches (light-width output 0)
    target = np.array([math.log2(best[0]), float(best[1]), math.log2(best[2]), 0.0], np.float32)
    xs = []
    for (s, a, w) in costs:
        rec, _ = trace(dg, source, cap=4096, delta=s * delta0, active_buckets=a,
                       worker_blocks=workers_for(w))
        if len(rec) > max_records:
            rec = rec[np.linspace(0, len(rec) - 1, max_records).astype(int)]
        xs.append(rec[:, :N_IN])
    x = np.concatenate(xs) if xs else np.zeros((0, N_IN), np.float32)
    y = np.repeat(target[None, :], len(x), axis=0)
    meta = {"delta0": delta0, "best": list(best), "best_ms": costs[best][0],
            "plain_ms": costs[(1.0, 1, 1.0)][0],
            "costs": {f"{s}x{a}x{w}": list(v) for (s, a, w), v in costs.items()}}
    return x, y, meta


def first_phase_features(graph_feat) -> np.ndarray:
    """The 32 inputs the engine gives the network before the first claim
    (adds.hip, k_adds_first_phase): the 14 graph features, a window of one
    queued item (log2(1 + 1) = 1), the heuristic width (log2 ratio 0) and that
    one item in the first window slot (log2(1 + 1) = 1), other slots empty."""
    x = np.zeros(N_IN, np.float32)
    x[:N_GRAPH] = _f32(graph_feat, (N_GRAPH,))
    x[14], x[15], x[16] = 1.0, 0.0, 1.0
    return x


def collect_head(dg, source: int, graph_feat, static, delta0: float, reps: int = 2):
    """Search two-phase schedules on one graph: the first phase at width h x
    delta0 (1 bucket) for every h in HEADS, then a tail setting (scale,
    active, share) from the best static setting and the same setting at widths
    1, 2 and 4 x delta0 (a narrow first bucket leaves the later, shallower
    phases free to run wide); against the best static setting alone.  Returns
    (the first-phase record, its target, the tail target for the graph's
    phase-boundary records, metadata)."""
    ss, sa, sw = (float(static[0]), int(static[1]), float(static[2]))
    tails = [(ss, sa, sw)] + [(s, sa, sw) for s in TAIL_SCALES if s != ss]
    o_static = dg.options(ALG_ADDS, delta=ss * delta0, active_buckets=sa, worker_blocks=workers_for(sw))
    o_sched = dg.options(ALG_ADDS, use_mlp=2)

    def best_of(o):
        return min(dg.sssp_labels(int(source), o).kernel_ms for _ in range(reps))

    costs = {"static": best_of(o_static)}
    try:
        for h in HEADS:
            for t in tails:
                set_schedule([(h, 1, t[2]), t])
                costs[(h,) + t] = best_of(o_sched)
    finally:
        set_schedule(None)
    best = min(costs, key=costs.get)
    if best == "static":
        head = tail = (ss, sa, sw)
    else:
        head, tail = (best[0], 1, best[3]), best[1:]

    def target(s):
        return np.array([math.log2(s[0]), float(s[1]), math.log2(s[2]), 0.0], np.float32)

    meta = {"head": list(head), "tail": list(tail), "head_ms": costs[best], "static_ms": costs["static"],
            "head_costs": {"x".join(str(v) for v in k) if k != "static" else k: c for k, c in costs.items()}}
    return first_phase_features(graph_feat), target(head), target(tail), meta


def collect(specs, log=print, reps: int = 2, time_budget_s: float | None = None, heads: bool = True):
    """Dataset over training specs: phase-boundary features, targets,
    per-sample weights (each graph weighs the same in total), per-graph
    metadata, and (heads) one first-phase record and target per graph
    (hx, hy; hidx: its graph's index in metas)."""
    from .device import DeviceGraph

    X, Y, W, metas = [], [], [], []
    HX, HY, HI = [], [], []
    t0 = time.time()
    for i, spec in enumerate(specs):
        if time_budget_s is not None and time.time() - t0 > time_budget_s:
            log(f"collect: time budget reached after {i} graphs")
            break
        g = make_graph(spec)
        dg = DeviceGraph(g)
        src = int(g.pick_sources(1, seed=spec["seed"] & 0xFFFFFFFF)[0])
        x, y, meta = collect_graph(dg, src, reps=reps)
        if heads and len(x):
            hx, hy, ty, hmeta = collect_head(dg, src, x[0, :N_GRAPH], meta["best"], meta["delta0"], reps=reps)
            meta.update(hmeta)
            y[:] = ty  # the phase-boundary records' target: the schedule's tail
            HX.append(hx)
            HY.append(hy)
            HI.append(len(metas))
        meta.update(spec)
        meta["n"], meta["m"], meta["source"] = g.n, g.m, src
        meta["records"] = int(len(x))
        X.append(x)
        Y.append(y)
        W.append(np.full(len(x), 1.0 / max(len(x), 1), np.float32))
        metas.append(meta)
        if i % 10 == 0:
            hd = f" head={meta['head']} gain={meta['plain_ms'] / meta['head_ms']:.2f}" if "head" in meta else ""
            log(f"collect {i + 1}/{len(specs)} {spec['kind']} n={g.n} best={meta['best']} "
                f"gain={meta['plain_ms'] / meta['best_ms']:.2f}{hd} ({time.time() - t0:.0f}s)")
        del dg, g
    x = np.concatenate(X) if X else np.zeros((0, N_IN), np.float32)
    y = np.concatenate(Y) if Y else np.zeros((0, N_OUT), np.float32)
    w = np.concatenate(W) if W else np.zeros(0, np.float32)
    if len(w):
        w *= len(w) / w.sum()
    hx = np.array(HX, np.float32).reshape(-1, N_IN)
    hy = np.array(HY, np.float32).reshape(-1, N_OUT)
    return x, y, w, metas, hx, hy, np.array(HI, np.int64)


def collect_heads(metas, x, y, w, log=print, reps: int = 2, time_budget_s: float | None = None):
    """First-phase records for the graphs of an existing dataset (metas, its
    phase-boundary features x, targets y and weights w, in collect()'s order):
    each graph is built again from its spec and collect_head() runs from its
    best static setting; y is rewritten in place with each searched graph's
    tail target.  Returns (hx, hy, hidx, per-graph head metadata)."""
    from .device import DeviceGraph

    counts = record_counts(metas, w)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    HX, HY, HI, HM = [], [], [], []
    t0 = time.time()
    for i, meta in enumerate(metas):
        if time_budget_s is not None and time.time() - t0 > time_budget_s:
            log(f"heads: time budget reached after {i} graphs")
            break
        if counts[i] == 0:
            continue
        g = make_graph(meta)
        dg = DeviceGraph(g)
        hx, hy, ty, hmeta = collect_head(dg, int(meta["source"]), x[first[i], :N_GRAPH], meta["best"],
                                         float(meta["delta0"]), reps=reps)
        y[first[i]:first[i] + counts[i]] = ty
        HX.append(hx)
        HY.append(hy)
        HI.append(i)
        HM.append(hmeta)
        if i % 20 == 0:
            log(f"heads {i + 1}/{len(metas)} {meta['kind']} n={g.n} static={meta['best']} head={hmeta['head']} "
                f"tail={hmeta['tail']} {hmeta['static_ms']:.2f} -> {hmeta['head_ms']:.2f} ms "
                f"({time.time() - t0:.0f}s)")
        del dg, g
    return (np.array(HX, np.float32).reshape(-1, N_IN), np.array(HY, np.float32).reshape(-1, N_OUT),
            np.array(HI, np.int64), HM)


# light widths searched per graph (relative to the first bucket width), widest
# first; a search stops once a width runs at more than twice the best time
LIGHT_SCALES = (1.0, 0.5, 0.25, 0.125, 0.0625)


def collect_light(specs, teacher: "DeltaModel", log=print, reps: int = 2, max_records: int = 24,
                  time_budget_s: float | None = None):
    """Data for the network's light-width output under pulls.  For each graph
    the shipped configuration (the teacher network's widths, active buckets
    and worker share, buckets pulled when cheaper) is timed at every light
    width of LIGHT_SCALES; the fastest width is the graph's target for output
    3, and one traced run at that width gives the phase records, whose targets
    for outputs 0-2 are the teacher's own outputs (so the new network keeps
    the teacher's phase decisions).  The record the engine evaluates before
    the first claim (first_phase_features) is added with the same targets;
    output 3 is read only there.  Returns (x, y, w, metas): each graph's
    records weigh 1 in total, half of it on the first-phase record."""
    from .device import DeviceGraph

    X, Y, W, metas = [], [], [], []
    t0 = time.time()
    teacher.install()
    for i, spec in enumerate(specs):
        if time_budget_s is not None and time.time() - t0 > time_budget_s:
            log(f"light: time budget reached after {i} graphs")
            break
        g = make_graph(spec)
        dg = DeviceGraph(g)
        src = int(g.pick_sources(1, seed=spec["seed"] & 0xFFFFFFFF)[0])
        costs = {}
        for sc in LIGHT_SCALES:
            o = dg.options(ALG_ADDS, use_mlp=1, pull=1, light_scale=sc)
            costs[sc] = min(dg.sssp_labels(src, o).kernel_ms for _ in range(reps))
            if costs[sc] > 2.0 * min(costs.values()):
                break
        best = min(costs, key=costs.get)
        rec, st = trace(dg, src, use_mlp=1, pull=1, light_scale=best)
        if len(rec) > max_records:
            rec = rec[np.linspace(0, len(rec) - 1, max_records).round().astype(int)]
        x = np.ascontiguousarray(rec[:, :N_IN])
        x = np.vstack([first_phase_features(x[0, :N_GRAPH]), x])
        y = teacher.forward(x)
        y[:, 3] = math.log2(best)
        w = np.full(len(x), 0.5 / max(len(x) - 1, 1), np.float32)
        w[0] = 0.5
        meta = dict(spec, n=g.n, m=g.m, source=src, light=best, light_costs={str(k): v for k, v in costs.items()},
                    records=int(len(x)), policy=int(st.policy))
        X.append(x)
        Y.append(y)
        W.append(w)
        metas.append(meta)
        if i % 10 == 0:
            log(f"light {i + 1}/{len(specs)} {spec['kind']} n={g.n} best={best} "
                f"gain={costs[1.0] / costs[best]:.2f} ({time.time() - t0:.0f}s)")
        del dg, g
    x = np.concatenate(X) if X else np.zeros((0, N_IN), np.float32)
    y = np.concatenate(Y) if Y else np.zeros((0, N_OUT), np.float32)
    w = np.concatenate(W) if W else np.zeros(0, np.float32)
    if len(w):
        w *= len(w) / w.sum()
    return x, y, w, metas


# the pulled-configuration search on graphs with a giant bucket: width scale x
# light width at one bucket and half the workgroups, then the share and the
# active buckets at the best pair
PULL_WIDTHS = (0.125, 0.25, 0.5, 1.0)
PULL_LIGHTS = (0.25, 0.125, 0.0625)


# the static search on grid graphs (high diameter, no bucket ever pulled):
# width x active buckets at a quarter of the workgroups, then the share at the
# best pair, then the light width
GRID_WIDTHS = (1.0, 2.0, 3.0, 4.0)
GRID_ACTIVE = (1, 2, 4)
GRID_SHARES = (0.5, 0.25, 0.125)
GRID_LIGHTS = (1.0, 0.5)


def collect_grid_static(specs, teacher: "DeltaModel", log=print, reps: int = 2, max_records: int = 24,
                        time_budget_s: float | None = None):
    """Targets of all four outputs on grid graphs from the current kernel: on
    each graph, static settings (width x delta0, active buckets, worker share,
    light width; pulls on, as shipped) are timed (GRID_WIDTHS x GRID_ACTIVE at
    a quarter of the workgroups, then GRID_SHARES at the best pair, then
    GRID_LIGHTS), the fastest becomes the target of every record of one traced
    run of that setting and of the first-phase record.  Returns (x, y, w,
    metas) as collect_light."""
    from .device import DeviceGraph

    X, Y, W, metas = [], [], [], []
    t0 = time.time()
    teacher.install()
    for i, spec in enumerate(specs):
        if time_budget_s is not None and time.time() - t0 > time_budget_s:
            log(f"grid static: time budget reached after {i} graphs")
            break
        g = make_graph(spec)
        dg = DeviceGraph(g)
        src = int(g.pick_sources(1, seed=spec["seed"] & 0xFFFFFFFF)[0])
        delta0 = dg.sssp_labels(src, dg.options(ALG_ADDS)).initial_delta

        def run(sc, act, sh, lt):
            o = dg.options(ALG_ADDS, delta=sc * delta0, active_buckets=act, worker_blocks=workers_for(sh),
                           pull=1, light_scale=lt)
            return min(dg.sssp_labels(src, o).kernel_ms for _ in range(reps))

        costs = {(sc, act, 0.25, 1.0): run(sc, act, 0.25, 1.0) for sc in GRID_WIDTHS for act in GRID_ACTIVE}
        sc, act, _, _ = min(costs, key=costs.get)
        for sh in GRID_SHARES:
            if (sc, act, sh, 1.0) not in costs:
                costs[(sc, act, sh, 1.0)] = run(sc, act, sh, 1.0)
        sc, act, sh, _ = min(costs, key=costs.get)
        for lt in GRID_LIGHTS:
            if (sc, act, sh, lt) not in costs:
                costs[(sc, act, sh, lt)] = run(sc, act, sh, lt)
        best = min(costs, key=costs.get)
        o = dg.options(ALG_ADDS, use_mlp=1, pull=1)
        learned = min(dg.sssp_labels(src, o).kernel_ms for _ in range(reps))
        plain = min(dg.sssp_labels(src, dg.options(ALG_ADDS)).kernel_ms for _ in range(reps))
        rec, _ = trace(dg, src, delta=best[0] * delta0, active_buckets=best[1],
                       worker_blocks=workers_for(best[2]), pull=1, light_scale=best[3])
        if len(rec) > max_records:
            rec = rec[np.linspace(0, len(rec) - 1, max_records).round().astype(int)]
        x = np.ascontiguousarray(rec[:, :N_IN])
        x = np.vstack([first_phase_features(x[0, :N_GRAPH]), x])
        y = np.repeat(np.array([[math.log2(best[0]), float(best[1]), math.log2(best[2]), math.log2(best[3])]],
                               np.float32), len(x), axis=0)
        w = np.full(len(x), 0.5 / max(len(x) - 1, 1), np.float32)
        w[0] = 0.5
        meta = dict(spec, n=g.n, m=g.m, source=src, best=list(best), best_ms=costs[best],
                    teacher_ms=learned, plain_ms=plain, records=int(len(x)), light=best[3],
                    light_costs={"1.0": learned})
        X.append(x)
        Y.append(y)
        W.append(w)
        metas.append(meta)
        if i % 10 == 0 or g.n > 2_000_000:
            log(f"grid static {i + 1}/{len(specs)} {spec['kind']} n={g.n} best={best} {costs[best]:.3f} ms "
                f"(network {learned:.3f}, plain {plain:.3f}) ({time.time() - t0:.0f}s)")
        del dg, g
    x = np.concatenate(X) if X else np.zeros((0, N_IN), np.float32)
    y = np.concatenate(Y) if Y else np.zeros((0, N_OUT), np.float32)
    w = np.concatenate(W) if W else np.zeros(0, np.float32)
    if len(w):
        w *= len(w) / w.sum()
    return x, y, w, metas


def collect_pulled(specs, teacher: "DeltaModel", log=print, reps: int = 2, max_records: int = 24,
                   time_budget_s: float | None = None, grid_static: bool = False):
    """Targets for the shipped configuration (pulls on) where it differs from
    the push-only training: on each R-MAT graph, static settings (width x
    delta0, active buckets, worker share, light width) are timed with pulls
    (PULL_WIDTHS x PULL_LIGHTS at one bucket and half the workgroups, then
    shares {1, 1/2, 1/4} and 1 or 2 buckets at the best pair), the fastest
    becomes the target of all four outputs for every record of the graph (and
    the first-phase record), and the records come from one traced run of that
    setting.  Other graphs go through collect_light (the teacher's decisions
    kept, the light width searched), or with grid_static through
    collect_grid_static.  Returns (x, y, w, metas) as collect_light."""
    from .device import DeviceGraph

    rmat = [sp for sp in specs if sp["kind"] == "rmat"]
    other = [sp for sp in specs if sp["kind"] != "rmat"]
    X, Y, W, metas = [], [], [], []
    t0 = time.time()
    teacher.install()
    for i, spec in enumerate(rmat):
        if time_budget_s is not None and time.time() - t0 > time_budget_s:
            log(f"pulled: time budget reached after {i} graphs")
            break
        g = make_graph(spec)
        dg = DeviceGraph(g)
        src = int(g.pick_sources(1, seed=spec["seed"] & 0xFFFFFFFF)[0])
        delta0 = dg.sssp_labels(src, dg.options(ALG_ADDS)).initial_delta

        def run(sc, act, sh, lt):
            o = dg.options(ALG_ADDS, delta=sc * delta0, active_buckets=act, worker_blocks=workers_for(sh),
                           pull=1, light_scale=lt)
            return min(dg.sssp_labels(src, o).kernel_ms for _ in range(reps))

        costs = {(sc, 1, 0.5, lt): run(sc, 1, 0.5, lt) for sc in PULL_WIDTHS for lt in PULL_LIGHTS}
        sc, _, _, lt = min(costs, key=costs.get)
        for act in (1, 2):
            for sh in (1.0, 0.5, 0.25):
                if (sc, act, sh, lt) not in costs:
                    costs[(sc, act, sh, lt)] = run(sc, act, sh, lt)
        best = min(costs, key=costs.get)
        o = dg.options(ALG_ADDS, use_mlp=1, pull=1)
        learned = min(dg.sssp_labels(src, o).kernel_ms for _ in range(reps))
        plain = min(dg.sssp_labels(src, dg.options(ALG_ADDS)).kernel_ms for _ in range(reps))
        rec, _ = trace(dg, src, delta=best[0] * delta0, active_buckets=best[1],
                       worker_blocks=workers_for(best[2]), pull=1, light_scale=best[3])
        if len(rec) > max_records:
            rec = rec[np.linspace(0, len(rec) - 1, max_records).round().astype(int)]
        x = np.ascontiguousarray(rec[:, :N_IN])
        x = np.vstack([first_phase_features(x[0, :N_GRAPH]), x])
        y = np.repeat(np.array([[math.log2(best[0]), float(best[1]), math.log2(best[2]), math.log2(best[3])]],
                               np.float32), len(x), axis=0)
        w = np.full(len(x), 0.5 / max(len(x) - 1, 1), np.float32)
        w[0] = 0.5
        meta = dict(spec, n=g.n, m=g.m, source=src, best=list(best), best_ms=costs[best],
                    teacher_ms=learned, plain_ms=plain, records=int(len(x)), light=best[3],
                    light_costs={"1.0": learned})
        X.append(x)
        Y.append(y)
        W.append(w)
        metas.append(meta)
        if i % 10 == 0:
            log(f"pulled {i + 1}/{len(rmat)} scale {spec['scale']} best={best} {costs[best]:.3f} ms "
                f"(network {learned:.3f}, plain {plain:.3f}) ({time.time() - t0:.0f}s)")
        del dg, g
    collect_other = collect_grid_static if grid_static else collect_light
    xo, yo, wo, mo = collect_other(other, teacher, log=log, reps=reps, max_records=max_records,
                                   time_budget_s=None if time_budget_s is None else
                                   max(0.0, time_budget_s - (time.time() - t0)))
    x = np.concatenate(X + [xo]) if X else xo
    y = np.concatenate(Y + [yo]) if Y else yo
    w = np.concatenate(W + [wo]) if W else wo
    if len(w):
        w = w * (len(w) / w.sum())
    return x, y, w, metas + mo


def fit_light(teacher: "DeltaModel", x, y, w=None, epochs: int = 40, batch: int = 512, lr: float = 1e-3,
              seed: int = 2026):
    """The teacher's network with the light-width output trained in: starts
    from the teacher's weights (output 3 at the mean target) and its input
    normalisation and graph-feature range, then Adam on the device over all
    outputs (outputs 0-2 are the teacher's own, so their error starts at
    zero)."""
    m = DeltaModel(teacher.w1.copy(), teacher.b1.copy(), teacher.w2.copy(), teacher.b2.copy(),
                   teacher.mean.copy(), teacher.inv_std.copy(), teacher.feat_lo.copy(), teacher.feat_hi.copy())
    m.w2[3, :] = 0.0
    m.b2[3] = float(np.average(y[:, 3], weights=w)) if len(y) else 0.0
    losses = m.train(x, y, w, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return m, losses


def record_counts(metas, w) -> np.ndarray:
    """Phase-boundary records per graph of a dataset, in the order of metas:
    from the metadata (records) where collect() stored it, else from the
    weights (collect() gives each graph the same total c, spread evenly over
    its records, so a graph's count is c / w of its first record)."""
    w = np.asarray(w, np.float32)
    c = float(w.astype(np.float64).sum()) / max(len(metas), 1)
    out = np.zeros(len(metas), np.int64)
    i = 0
    for j, m in enumerate(metas):
        k = int(m["records"]) if "records" in m else (int(round(c / float(w[i]))) if i < len(w) else 0)
        out[j] = k
        i += k
    if i != len(w):
        raise ValueError("sample weights do not match the graphs' record counts")
    return out


def time_weights(w, metas, power: float = 1.0):
    """Per-sample weights that give each graph a total weight proportional to
    its plain-ADDS kernel time to the given power (power 1: the time a policy
    can save on it) instead of an equal share.  w gives every graph the same total weight c, spread
    evenly over its records, in the order of metas (as collect() writes it),
    so a graph's record count is c / w."""
    w = np.asarray(w, np.float32)
    out = np.empty_like(w)
    c = float(w.astype(np.float64).sum()) / max(len(metas), 1)
    i = 0
    for m in metas:
        k = int(round(c / float(w[i])))
        out[i:i + k] = w[i:i + k] * float(m["plain_ms"]) ** power
        i += k
    if i != len(w):
        raise ValueError("sample weights do not match the graphs' record counts")
    return out


def graph_class(meta) -> str:
    """Size class of a training graph: R-MAT by scale, grids by kind and
    decade of the vertex count."""
    if meta["kind"] == "rmat":
        return f"rmat{meta['scale']}"
    decade = int(math.floor(math.log10(max(int(meta["n"]), 1))))
    return f"{meta['kind']}{min(max(decade, 5), 7)}"  # 1e5-1e6, 1e6-1e7, above


def class_weights(w, metas, class_of=graph_class):
    """Per-sample weights that give every size class the same total weight
    (graphs of a class share it equally), from the equal-per-graph weights w
    in the order of metas."""
    w = np.asarray(w, np.float32)
    counts = {}
    for m in metas:
        counts[class_of(m)] = counts.get(class_of(m), 0) + 1
    out = np.empty_like(w)
    c = float(w.astype(np.float64).sum()) / max(len(metas), 1)
    i = 0
    for m in metas:
        k = int(m["records"]) if "records" in m else int(round(c / float(w[i])))
        out[i:i + k] = w[i:i + k] / counts[class_of(m)]
        i += k
    if i != len(w):
        raise ValueError("sample weights do not match the graphs' record counts")
    return out


def graph_feature_range(x):
    """Per-feature min / max of the graph features (inputs 0..13) over a
    dataset: the range the engine trusts the network in."""
    x = _f32(x, (-1, N_IN))
    return _f32(x[:, :N_GRAPH].min(axis=0)), _f32(x[:, :N_GRAPH].max(axis=0))


def fit(x, y, w=None, epochs: int = 60, batch: int = 512, lr: float = 2e-3, seed: int = 2026):
    """Normalisation and the graph-feature range from the data, then Adam on
    the device."""
    m = DeltaModel.init(seed)
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    m.mean = _f32(mean)
    m.inv_std = _f32(np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 0.0))
    m.feat_lo, m.feat_hi = graph_feature_range(x)
    losses = m.train(x, y, w, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return m, losses

"""Batched multi-source queries on grid graphs (routing nets): the query
workload generator and the ctypes front of mi355_sssp_batched."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._abi import BatchOptions, BatchStats, QueryBatch, W_F32, check, engine
from .device import DeviceBuffer


class Queries:
    """count queries: sources (concatenated, global vertex ids) with
    src_offsets, and rectangles boxes[q] = (row0, col0, row1, col1)."""

    def __init__(self, grid_cols: int, src_offsets, sources, boxes):
        self.grid_cols = int(grid_cols)
        self.src_offsets = np.ascontiguousarray(src_offsets, np.int32)
        self.sources = np.ascontiguousarray(sources, np.int32)
        self.boxes = np.ascontiguousarray(boxes, np.int32).reshape(-1, 4)
        self.count = len(self.boxes)

    def c(self) -> QueryBatch:
        return QueryBatch(self.grid_cols, self.count, self.src_offsets.ctypes.data,
                          self.sources.ctypes.data, self.boxes.ctypes.data)

    def query_sources(self, q: int) -> np.ndarray:
        return self.sources[self.src_offsets[q]:self.src_offsets[q + 1]]

    def areas(self) -> np.ndarray:
        b = self.boxes.astype(np.int64)
        return (b[:, 2] - b[:, 0]) * (b[:, 3] - b[:, 1])

    def subset(self, idx) -> "Queries":
        idx = np.asarray(idx)
        srcs = [self.query_sources(int(q)) for q in idx]
        off = np.concatenate([[0], np.cumsum([len(s) for s in srcs])])
        return Queries(self.grid_cols, off, np.concatenate(srcs), self.boxes[idx])


def routing_queries(rows: int, cols: int, count: int, seed: int, side_min: int = 50,
                    side_max: int = 2000, max_sources: int = 8, square: bool = True) -> Queries:
    """Net-like queries: a bounding box of side log-uniform in
    [side_min, side_max] (square unless square=False, clipped to the grid),
    placed uniformly; 1..max_sources sources uniform inside it."""
    rng = np.random.default_rng(seed)
    boxes = np.zeros((count, 4), np.int32)
    srcs, off = [], [0]
    for q in range(count):
        side = math.exp(rng.uniform(math.log(side_min), math.log(side_max)))
        other = side if square else math.exp(rng.uniform(math.log(side_min), math.log(side_max)))
        h = min(rows, int(round(side)))
        w = min(cols, int(round(other)))
        r0 = int(rng.integers(0, rows - h + 1))
        c0 = int(rng.integers(0, cols - w + 1))
        boxes[q] = (r0, c0, r0 + h, c0 + w)
        k = int(rng.integers(1, max_sources + 1))
        rr = rng.integers(r0, r0 + h, k)
        cc = rng.integers(c0, c0 + w, k)
        srcs.append((rr * cols + cc).astype(np.int32))
        off.append(off[-1] + k)
    return Queries(cols, np.array(off), np.concatenate(srcs), boxes)


def routing_sinks(qs: Queries, seed: int, max_sinks: int = 8):
    """Net-like sinks for the queries of `qs`: 1..max_sinks cells uniform in
    each query's rectangle, as (target offsets, target vertices) for run()."""
    rng = np.random.default_rng(seed)
    off, tv = [0], []
    cols = qs.grid_cols
    for q in range(qs.count):
        r0, c0, r1, c1 = (int(x) for x in qs.boxes[q])
        k = int(rng.integers(1, max_sinks + 1))
        rr = rng.integers(r0, r1, k)
        cc = rng.integers(c0, c1, k)
        tv.append((rr * cols + cc).astype(np.int32))
        off.append(off[-1] + k)
    return np.array(off, np.int32), (np.concatenate(tv) if tv else np.zeros(0, np.int32))


def block_slots(bh: int, bw: int) -> int:
    """Label slots the engine stores for a bh x bw rectangle: 4 x 4 tiles
    padded to whole tiles, or row-major (no padding) below 4 rows."""
    bh, bw = int(bh), int(bw)
    return bh * bw if bh < 4 else ((bh + 3) & ~3) * ((bw + 3) & ~3)


def storage_slots(qs: "Queries") -> int:
    """Label slots the engine allocates for a batch (tile padding included)."""
    b = qs.boxes.astype(np.int64)
    return int(sum(block_slots(h, w) for h, w in zip(b[:, 2] - b[:, 0], b[:, 3] - b[:, 1])))


def options(**kw) -> BatchOptions:
    o = BatchOptions()
    engine().mi355_batch_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def layout(dg, qs: Queries) -> np.ndarray:
    off = np.zeros(qs.count + 1, np.int64)
    check(engine().mi355_batch_layout(dg.handle, C.byref(qs.c()), off.ctypes.data))
    return off


def run(dg, qs: Queries, want_pred: bool = True, download: bool = True, targets=None, device_out: bool = False,
        **kw):
    """Solve all queries; returns (dist, pred, offsets, stats) with dist/pred
    concatenated per query (None when not downloaded; the device buffers
    themselves when device_out, for paths()).

    targets: None, or (target_offsets, target_vertices) with count + 1 offsets:
    query q stops once all of its targets (global vertex ids inside its
    rectangle) are settled; its outputs are then exact below the threshold it
    stopped at (its targets among them) and unreached above.  A query with no
    targets runs to the end (mi355_sssp_batched_targets)."""
    off = layout(dg, qs)
    total = int(off[-1])
    d_dist = DeviceBuffer(4 * total)
    d_pred = DeviceBuffer(4 * total) if want_pred else None
    st = BatchStats()
    if targets is None:
        check(engine().mi355_sssp_batched(dg.handle, C.byref(qs.c()), C.byref(options(**kw)), d_dist.ptr,
                                          d_pred.ptr if d_pred else None, C.byref(st), None))
    else:
        t_off = np.ascontiguousarray(targets[0], np.int32)
        t_v = np.ascontiguousarray(targets[1], np.int32)
        if t_off.shape != (qs.count + 1,):
            raise ValueError(f"target offsets: expected {qs.count + 1} entries, got {t_off.shape}")
        if int(t_off[-1]) != len(t_v):
            raise ValueError("target offsets do not match the target list")
        check(engine().mi355_sssp_batched_targets(dg.handle, C.byref(qs.c()), t_off.ctypes.data,
                                                  t_v.ctypes.data if len(t_v) else None, C.byref(options(**kw)),
                                                  d_dist.ptr, d_pred.ptr if d_pred else None, C.byref(st), None))
    if device_out:
        return d_dist, d_pred, off, st
    if not download:
        return None, None, off, st
    dist = d_dist.to_array(np.float32 if dg.weight_type == W_F32 else np.uint32, total)
    pred = d_pred.to_array(np.int32, total) if d_pred else None
    return dist, pred, off, st


def paths(dg, qs: Queries, off, d_dist, d_pred, query, target):
    """Parent chains of targets out of run(..., device_out=True)'s buffers,
    read on the device (mi355_batch_paths): target i is global vertex
    target[i] of query query[i].  Returns (path offsets [count + 1], target
    distances [count], path vertices, source first); an unreached target has
    an empty path."""
    query = np.ascontiguousarray(query, np.int32)
    target = np.ascontiguousarray(target, np.int32)
    if query.shape != target.shape:
        raise ValueError("query and target arrays differ in length")
    off = np.ascontiguousarray(off, np.int64)
    count = len(target)
    poff = np.zeros(count + 1, np.int64)
    tdist = np.zeros(count, np.float32 if dg.weight_type == W_F32 else np.uint32)
    args = (dg.handle, C.byref(qs.c()), off.ctypes.data, d_dist.ptr, d_pred.ptr, query.ctypes.data,
            target.ctypes.data, count, poff.ctypes.data, tdist.ctypes.data)
    check(engine().mi355_batch_paths(*args, None, 0, None))
    nodes = np.zeros(max(int(poff[-1]), 1), np.int32)
    check(engine().mi355_batch_paths(*args, nodes.ctypes.data, int(poff[-1]), None))
    return poff, tdist, nodes[:int(poff[-1])]


TUNE_CANDIDATES = 18


def autotune(dg, qs: Queries, sample_stride: int = 8, confirm: int = 0, **kw):
    """(best options, device ms of every candidate) on every sample_stride-th
    query.  confirm > 0: for each bucket width, the fastest candidate on the
    sample; the `confirm` fastest of those are timed once more on the whole
    batch and the fastest there is returned (a sample has the batch's round
    count but a fraction of its work, which favours wide buckets)."""
    best = options(**kw)
    t = np.zeros(TUNE_CANDIDATES, np.float64)
    check(engine().mi355_batch_autotune(dg.handle, C.byref(qs.c()), int(sample_stride), C.byref(best),
                                        t.ctypes.data, None))
    if confirm > 0:
        # candidate order of mi355_batch_autotune: order x width x workgroups per CU
        cands = [(order, sc, bpc) for order in (0, 1) for sc in (0.5, 1.0, 2.0) for bpc in (2, 4, 8)]
        per_width = [min((i for i in range(len(cands)) if cands[i][1] == sc), key=lambda i: t[i])
                     for sc in (0.5, 1.0, 2.0)]
        full = {}
        d_dist = DeviceBuffer(4 * int(layout(dg, qs)[-1]))
        for i in sorted(per_width, key=lambda i: t[i])[:confirm]:
            order, sc, bpc = cands[int(i)]
            o = options(**kw)
            o.order, o.delta_scale, o.blocks_per_cu = order, sc, bpc
            o.rounds_per_graph = best.rounds_per_graph
            st = BatchStats()
            check(engine().mi355_sssp_batched(dg.handle, C.byref(qs.c()), C.byref(o), d_dist.ptr, None,
                                              C.byref(st), None))
            full[int(i)] = (float(st.device_ms), o)
        best = min(full.values(), key=lambda x: x[0])[1]
    return best, t


def induced_subgraph(g, qs: Queries, q: int):
    """Host CSR of query q's rectangle (local ids, row-major) plus one extra
    vertex (the last) joined to every source by a zero-weight edge: plain
    single-source ADDS from it solves the multi-source query."""
    from .graph import from_csr

    r0, c0, r1, c1 = (int(x) for x in qs.boxes[q])
    bw = c1 - c0
    rows = np.arange(r0, r1)
    verts = (rows[:, None] * qs.grid_cols + np.arange(c0, c1)[None, :]).ravel()
    area = len(verts)
    rp = np.asarray(g.rowptr, np.int64)
    deg = rp[verts + 1] - rp[verts]
    eidx = np.repeat(rp[verts], deg) + (np.arange(deg.sum()) - np.repeat(np.cumsum(deg) - deg, deg))
    src_local = np.repeat(np.arange(area), deg)
    dst = np.asarray(g.col)[eidx].astype(np.int64)
    w = np.asarray(g.wbits)[eidx]
    dr, dc = dst // qs.grid_cols - r0, dst % qs.grid_cols - c0
    keep = (dr >= 0) & (dr < r1 - r0) & (dc >= 0) & (dc < bw)
    src_local, w = src_local[keep], w[keep]
    dst_local = (dr * bw + dc)[keep]
    s = qs.query_sources(q).astype(np.int64)
    s_local = np.unique((s // qs.grid_cols - r0) * bw + (s % qs.grid_cols - c0))
    src_all = np.concatenate([src_local, np.full(len(s_local), area)])
    dst_all = np.concatenate([dst_local, s_local])
    w_all = np.concatenate([w, np.zeros(len(s_local), w.dtype)])
    order = np.argsort(src_all, kind="stable")
    rowptr = np.zeros(area + 2, np.int64)
    np.add.at(rowptr, src_all + 1, 1)
    rowptr = np.cumsum(rowptr)
    wt = g.weight_type
    wv = w_all[order].view(np.float32) if wt == W_F32 else w_all[order]
    return from_csr(rowptr, dst_all[order], wv, wt), area

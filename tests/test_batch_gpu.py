"""Batched multi-source queries (routing nets) against the CPU oracle's
multi-source Dijkstra on each query's induced subgraph."""
import numpy as np
import pytest

import mi355sssp as M
from mi355sssp import batch as B
from oracle import orc

pytestmark = pytest.mark.gpu


def _check(g, qs, dist, pred, off, sample=None):
    w = g.weights
    qids = range(qs.count) if sample is None else sample
    for q in qids:
        box = tuple(int(x) for x in qs.boxes[q])
        srcs = qs.query_sources(q)
        want = orc.dijkstra_box(g.rowptr, g.col, w, qs.grid_cols, box, srcs)
        got = dist[off[q]:off[q + 1]]
        bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
        assert bad.size == 0, f"query {q} box {box}: {bad.size} distances differ"
        if pred is not None:
            wp = orc.canonical_pred_box(g.rowptr, g.col, w, qs.grid_cols, box, srcs, want)
            badp = np.flatnonzero(pred[off[q]:off[q + 1]] != wp)
            assert badp.size == 0, f"query {q}: {badp.size} predecessors differ"


def _f32_grid(rows, cols, seed, holes=0):
    """4-neighbour grid with fp32 weights, 8 % of the edges missing; holes > 0:
    no edge into a vertex v with v % holes == 1 (unreachable unless a source)."""
    rng = np.random.default_rng(seed)
    src, dst = [], []
    for r in range(rows):
        for c in range(cols):
            v = r * cols + c
            for dr, dc in ((-1, 0), (0, -1), (0, 1), (1, 0)):
                rr, cc = r + dr, c + dc
                if holes and (rr * cols + cc) % holes == 1:
                    continue
                if 0 <= rr < rows and 0 <= cc < cols and rng.random() > 0.08:
                    src.append(v)
                    dst.append(rr * cols + cc)
    w = ((rng.integers(1, 1 << 24, len(src)) * 2.0 ** -24) * 4).astype(np.float32)
    return M.from_edges(rows * cols, np.array(src), np.array(dst), w, M.W_F32)


@pytest.mark.parametrize("use_graph", [1, 0])
def test_routing_grid_queries_bit_exact(use_graph):
    side = 150
    g = M.routing_grid(side, seed=21)
    dg = M.DeviceGraph(g)
    qs = B.routing_queries(side, side, 48, seed=5, side_min=5, side_max=140)
    print(f"batch: {qs.count} queries, {int(qs.areas().sum())} slots, use_graph={use_graph}", flush=True)
    dist, pred, off, st = B.run(dg, qs, use_graph=use_graph, count_edges=1)
    assert st.label_slots == off[-1] == qs.areas().sum()
    assert st.rounds > 0 and st.edges_relaxed > 0
    _check(g, qs, dist, pred, off)
    # the same batch again on the cached state, other schedule parameters
    dist2, pred2, _, _ = B.run(dg, qs, use_graph=use_graph, rounds_per_graph=6, blocks_per_cu=1)
    assert dist2.tobytes() == dist.tobytes() and np.array_equal(pred2, pred)


@pytest.mark.parametrize("delta_scale", [0.1, 1.0, 8.0, -1.0])
def test_near_far_widths_and_bellman_ford_rounds_agree(delta_scale):
    """Any bucket width (and Bellman-Ford rounds, delta_scale < 0) gives the
    oracle's distances and parents; bucketed runs use the far pile, Bellman-
    Ford rounds never do, and narrow buckets relax fewer edges."""
    side = 130
    g = M.routing_grid(side, seed=31)
    dg = M.DeviceGraph(g)
    qs = B.routing_queries(side, side, 40, seed=12, side_min=8, side_max=130)
    print(f"batch: {qs.count} queries, delta_scale={delta_scale}", flush=True)
    dist, pred, off, st = B.run(dg, qs, delta_scale=delta_scale, count_edges=1)
    _check(g, qs, dist, pred, off)
    if delta_scale < 0:
        assert st.far_passes == 0
    else:
        assert st.far_passes > 0
    if delta_scale == 1.0:
        _, _, _, bf = B.run(dg, qs, delta_scale=-1.0, count_edges=1, download=False)
        assert st.edges_relaxed < bf.edges_relaxed, (st.edges_relaxed, bf.edges_relaxed)


def test_f32_grid_queries_bit_exact():
    g = _f32_grid(60, 70, seed=3)
    dg = M.DeviceGraph(g)
    qs = B.routing_queries(60, 70, 30, seed=9, side_min=3, side_max=60)
    dist, pred, off, _ = B.run(dg, qs)
    _check(g, qs, dist, pred, off)


def test_whole_grid_single_source_equals_sssp():
    side = 90
    g = M.routing_grid(side, seed=4)
    dg = M.DeviceGraph(g)
    src = 17 * side + 33
    qs = B.Queries(side, [0, 1], [src], [[0, 0, side, side]])
    dist, pred, off, _ = B.run(dg, qs)
    d1, p1, _ = dg.sssp(src, M.ALG_DELTA)
    assert dist.tobytes() == d1.tobytes() and np.array_equal(pred, p1)


def test_degenerate_queries():
    side = 40
    g = M.routing_grid(side, seed=8)
    dg = M.DeviceGraph(g)
    v = 5 * side + 6
    qs = B.Queries(side, [0, 1, 4, 6],
                   [v, v, v, v, 0, side * side - 1],  # 1x1 box; duplicate sources; corners
                   [[5, 6, 6, 7], [0, 0, side, side], [0, 0, side, side]])
    dist, pred, off, _ = B.run(dg, qs)
    assert dist[0] == 0 and pred[0] == -1
    _check(g, qs, dist, pred, off)


def test_invalid_batches_are_rejected():
    side = 20
    g = M.routing_grid(side, seed=2)
    dg = M.DeviceGraph(g)
    bad = [
        B.Queries(side, [0, 1], [0], [[1, 1, 5, 5]]),          # source outside its box
        B.Queries(side, [0, 1], [0], [[0, 0, side + 1, 5]]),   # box outside the grid
        B.Queries(side, [0, 0], [], [[0, 0, 5, 5]]),           # no source
        B.Queries(side + 1, [0, 1], [0], [[0, 0, 5, 5]]),      # grid width does not divide n
    ]
    for qs in bad:
        with pytest.raises(M.Mi355Error):
            B.run(dg, qs)


def test_queue_overflow_is_reported():
    side = 100
    g = M.routing_grid(side, seed=6)
    dg = M.DeviceGraph(g)
    qs = B.routing_queries(side, side, 16, seed=1, side_min=60, side_max=100)
    with pytest.raises(M.Mi355Error) as e:
        B.run(dg, qs, queue_capacity=64)
    assert "overflow" in str(e.value)
    # and the handle still works afterwards
    dist, pred, off, _ = B.run(dg, qs)
    _check(g, qs, dist, pred, off, sample=[0, 5])


def test_autotune_and_schedule_orders():
    side = 120
    g = M.routing_grid(side, seed=13)
    dg = M.DeviceGraph(g)
    qs = B.routing_queries(side, side, 40, seed=2, side_min=10, side_max=120)
    best, t = B.autotune(dg, qs, sample_stride=4)
    assert np.all(t > 0) and best.rounds_per_graph % 2 == 0 and best.blocks_per_cu in (2, 4, 8)
    assert best.delta_scale in (0.5, 1.0, 2.0)
    # confirmed on the whole batch: the fastest candidate of its width on the sample
    best3, t3 = B.autotune(dg, qs, sample_stride=4, confirm=3)
    cands = [(o, sc, b) for o in (0, 1) for sc in (0.5, 1.0, 2.0) for b in (2, 4, 8)]
    i = cands.index((best3.order, best3.delta_scale, best3.blocks_per_cu))
    assert t3[i] == min(t3[j] for j in range(len(cands)) if cands[j][1] == best3.delta_scale)
    ref, pref, off, _ = B.run(dg, qs)
    for order in (0, 1):
        d, p, _, _ = B.run(dg, qs, order=order, rounds_per_graph=best.rounds_per_graph,
                           blocks_per_cu=best.blocks_per_cu, delta_scale=best.delta_scale)
        assert d.tobytes() == ref.tobytes() and np.array_equal(p, pref)
    _check(g, qs, ref, pref, off, sample=range(0, 40, 7))


def test_one_query_at_a_time_adds_matches_the_batch():
    side = 110
    g = M.routing_grid(side, seed=17)
    dg = M.DeviceGraph(g)
    qs = B.routing_queries(side, side, 6, seed=4, side_min=20, side_max=110)
    dist, pred, off, _ = B.run(dg, qs)
    for q in range(qs.count):
        sub, area = B.induced_subgraph(g, qs, q)
        sdg = M.DeviceGraph(sub)
        d, p, _ = sdg.sssp(area, M.ALG_ADDS)
        assert d[:area].tobytes() == dist[off[q]:off[q + 1]].tobytes()
        r0, c0, r1, c1 = qs.boxes[q]
        bw = c1 - c0
        gp = np.where((p[:area] < 0) | (p[:area] == area), -1,
                      (r0 + p[:area] // bw) * side + c0 + p[:area] % bw)
        assert np.array_equal(gp, pred[off[q]:off[q + 1]])


@pytest.mark.parametrize("use_graph", [1, 0])
def test_short_rectangles_row_major_blocks(use_graph):
    """Rectangles of 1-3 rows are stored row-major (no tile padding); mixed
    with tiled ones in one batch they give the oracle's answers."""
    side = 120
    g = M.routing_grid(side, seed=41)
    dg = M.DeviceGraph(g)
    rng = np.random.default_rng(3)
    boxes, srcs, off = [], [], [0]
    for q in range(36):
        h = (1, 2, 3, 4, 5, 9)[q % 6]
        w = int(rng.integers(1, side + 1))
        r0, c0 = int(rng.integers(0, side - h + 1)), int(rng.integers(0, side - w + 1))
        boxes.append((r0, c0, r0 + h, c0 + w))
        k = int(rng.integers(1, 4))
        s = (rng.integers(r0, r0 + h, k) * side + rng.integers(c0, c0 + w, k)).astype(np.int32)
        srcs.append(s)
        off.append(off[-1] + k)
    qs = B.Queries(side, np.array(off), np.concatenate(srcs), np.array(boxes))
    assert B.storage_slots(qs) < sum(B.block_slots(max(4, b[2] - b[0]), b[3] - b[1]) for b in boxes)
    print(f"batch: {qs.count} queries, use_graph={use_graph}", flush=True)
    dist, pred, off, st = B.run(dg, qs, use_graph=use_graph)
    assert st.label_slots == off[-1] == qs.areas().sum()
    _check(g, qs, dist, pred, off)


def test_batch_on_a_degree_ordered_graph():
    """The batch reads the input numbering's grid geometry: on a graph handle
    renumbered by degree it runs on an input-order copy, same answers."""
    side = 64
    g = M.routing_grid(side, seed=43)
    qs = B.routing_queries(side, side, 12, seed=9, side_min=4, side_max=64)
    dg = M.DeviceGraph(g, order=M.ORDER_DEGREE)
    assert dg.info.vertex_order == M.ORDER_DEGREE
    dist, pred, off, _ = B.run(dg, qs)
    _check(g, qs, dist, pred, off)


@pytest.mark.parametrize("use_graph", [1, 0])
def test_queries_stop_once_their_targets_are_settled(use_graph):
    """mi355_sssp_batched_targets: every target's distance and parent equal the
    full run's; a query that stopped keeps exact outputs below its threshold and
    reads unreached above it (a cut in distance order); queries without
    targets equal the full run."""
    side = 140
    g = M.routing_grid(side, seed=41)
    dg = M.DeviceGraph(g)
    qs = B.routing_queries(side, side, 40, seed=17, side_min=20, side_max=140)
    rng = np.random.default_rng(3)
    t_off, t_v = [0], []
    for q in range(qs.count):
        r0, c0, r1, c1 = (int(x) for x in qs.boxes[q])
        k = 0 if q % 5 == 4 else int(rng.integers(1, 4))  # every fifth query has no targets
        for _ in range(k):
            # near a source: the query should stop long before its rectangle is done
            s = int(qs.query_sources(q)[0])
            r = min(max(s // side + int(rng.integers(-3, 4)), r0), r1 - 1)
            c = min(max(s % side + int(rng.integers(-3, 4)), c0), c1 - 1)
            t_v.append(r * side + c)
        t_off.append(len(t_v))
    full_d, full_p, off, _ = B.run(dg, qs, use_graph=use_graph)
    print(f"batch: {qs.count} queries, {len(t_v)} targets, use_graph={use_graph}", flush=True)
    d, p, off2, st = B.run(dg, qs, use_graph=use_graph, targets=(np.array(t_off), np.array(t_v)))
    assert np.array_equal(off, off2)
    with_targets = sum(1 for q in range(qs.count) if t_off[q + 1] > t_off[q])
    assert 0 < st.stopped_queries <= with_targets, (st.stopped_queries, with_targets)
    stopped_cells = 0
    for q in range(qs.count):
        r0, c0, r1, c1 = (int(x) for x in qs.boxes[q])
        a, b = off[q], off[q + 1]
        fd, fp, gd, gp = full_d[a:b], full_p[a:b], d[a:b], p[a:b]
        for v in t_v[t_off[q]:t_off[q + 1]]:
            i = (v // side - r0) * (c1 - c0) + v % side - c0
            assert gd[i].tobytes() == fd[i].tobytes() and gp[i] == fp[i], (q, v)
        same = (gd.view(np.uint32) == fd.view(np.uint32)) & (gp == fp)
        cut = ~same
        if t_off[q + 1] == t_off[q]:
            assert not cut.any(), f"query {q} without targets differs from the full run"
            continue
        # cells that differ read unreached (uint32 weights: 0xFFFFFFFF) and lie
        # at or beyond every kept distance
        unreached = gd == np.uint32(0xFFFFFFFF)
        assert np.all(unreached[cut]) and np.all(gp[cut] == -1), q
        if cut.any():
            stopped_cells += int(cut.sum())
            assert fd[cut].min() >= gd[~unreached].max(), q
    assert stopped_cells > 0


def test_invalid_targets_are_rejected():
    side = 30
    g = M.routing_grid(side, seed=4)
    dg = M.DeviceGraph(g)
    qs = B.Queries(side, [0, 1], [3 * side + 3], [[2, 2, 10, 10]])
    for t_off, t_v in (([0, 1], [0]),                  # outside the rectangle
                       ([0, 1], [side * side]),        # outside the graph
                       ([1, 1], [3 * side + 4])):      # offsets not starting at 0
        with pytest.raises((M.Mi355Error, ValueError)):
            B.run(dg, qs, targets=(np.array(t_off), np.array(t_v)))
    with pytest.raises(ValueError):                     # offsets and list disagree
        B.run(dg, qs, targets=(np.array([0, 2]), np.array([3 * side + 4])))
    # no targets at all: the query runs to the end
    d, p, off, st = B.run(dg, qs, targets=(np.array([0, 0]), np.array([], np.int32)))
    full, fp, _, _ = B.run(dg, qs)
    assert d.tobytes() == full.tobytes() and np.array_equal(p, fp) and st.stopped_queries == 0


def test_paths_read_on_the_device_match_the_parent_outputs():
    """mi355_batch_paths: every target's path is its parent chain in the
    downloaded outputs (source first), its distance the output distance, and
    an unreached target has an empty path; also on a stopped batch."""
    side = 120
    g = _f32_grid(side, side, seed=9, holes=13)  # every 13th vertex unreachable
    dg = M.DeviceGraph(g)
    qs = B.routing_queries(side, side, 24, seed=6, side_min=3, side_max=120)
    rng = np.random.default_rng(1)
    query, target = [], []
    for q in range(qs.count):
        r0, c0, r1, c1 = (int(x) for x in qs.boxes[q])
        for _ in range(5):
            query.append(q)
            target.append(int(rng.integers(r0, r1)) * side + int(rng.integers(c0, c1)))
    for stop in (False, True):
        tg = None
        if stop:
            t_off = np.searchsorted(np.array(query), np.arange(qs.count + 1))
            tg = (t_off, np.array(target))
        d_dist, d_pred, off, _ = B.run(dg, qs, targets=tg, device_out=True)
        dist = d_dist.to_array(np.float32, int(off[-1]))
        pred = d_pred.to_array(np.int32, int(off[-1]))
        print(f"paths: {len(target)} targets, stop={stop}", flush=True)
        poff, tdist, nodes = B.paths(dg, qs, off, d_dist, d_pred, query, target)
        assert poff[0] == 0 and len(poff) == len(target) + 1 and poff[-1] == len(nodes)
        unreached = 0
        for i, (q, t) in enumerate(zip(query, target)):
            r0, c0, r1, c1 = (int(x) for x in qs.boxes[q])
            loc = lambda v: off[q] + (v // side - r0) * (c1 - c0) + v % side - c0  # noqa: E731
            assert tdist[i].tobytes() == dist[loc(t)].tobytes()
            want = []
            if np.isfinite(dist[loc(t)]):
                v = t
                while True:
                    want.append(v)
                    if pred[loc(v)] < 0:
                        break
                    v = int(pred[loc(v)])
                want = want[::-1]
                assert want[0] in qs.query_sources(q)
            else:
                unreached += 1
            assert np.array_equal(nodes[poff[i]:poff[i + 1]], want), (q, t)
        assert unreached > 0
    with pytest.raises(M.Mi355Error):  # target outside its query's rectangle
        r0, c0, r1, c1 = (int(x) for x in qs.boxes[0])
        outside = (r1 % side) * side + c0 if r1 < side else (r0 - 1) * side + c0
        B.paths(dg, qs, off, d_dist, d_pred, [0], [outside])

"""GPU parity of the SSSP engine against the CPU oracle (bit-exact).

Every case compares the HIP path, called through the C ABI, with the oracle
on the same seeded inputs: distances bit for bit, predecessors bit for bit
against the canonical (smallest-index) parent, plus the predecessor
invariant dist[v] == dist[p] (+) w(p, v).
"""
import math

import numpy as np
import pytest

import mi355sssp as M
from oracle import orc
from golden import bf_cases

pytestmark = pytest.mark.gpu

ALGS = [M.ALG_BF, M.ALG_BF_JACOBI, M.ALG_DELTA]


def _oracle(g: M.HostGraph, src: int):
    w = g.weights
    dist = orc.dijkstra(g.rowptr, g.col, w, src)
    pred = orc.canonical_pred(g.rowptr, g.col, w, src, dist)
    return dist, pred


def _check(g: M.HostGraph, dg: M.DeviceGraph, src: int, alg: int, delta: float = 0.0):
    print(f"case n={g.n} m={g.m} src={src} alg={alg} delta={delta}", flush=True)
    dist, pred, st = dg.sssp(src, alg, delta=delta)
    want_d, want_p = _oracle(g, src)
    bad = np.flatnonzero(dist.view(np.uint32) != want_d.view(np.uint32))
    assert bad.size == 0, f"{bad.size} distances differ, first at {bad[:8].tolist()}: " \
                          f"{dist[bad[:4]].tolist()} vs {want_d[bad[:4]].tolist()}"
    assert orc.check_pred(g.rowptr, g.col, g.weights, src, dist, pred) == 0
    badp = np.flatnonzero(pred != want_p)
    assert badp.size == 0, f"{badp.size} predecessors differ, first at {badp[:8].tolist()}"
    mask = want_p >= 0
    mask[src] = True
    assert st.reached == int(mask.sum())
    assert st.reached_edges == int(np.diff(g.rowptr)[mask].sum())
    return st


@pytest.fixture(scope="module")
def grid256():
    g = M.grid4(256, 256, 1, 100, seed=2026)
    return g, M.DeviceGraph(g)


@pytest.mark.parametrize("alg", ALGS)
def test_grid256_u32(grid256, alg):
    g, dg = grid256
    assert g.m == 261120
    st = _check(g, dg, 0, alg)
    assert st.converged == 1


@pytest.mark.parametrize("alg", ALGS)
@pytest.mark.parametrize("seed", [1, 2])
def test_chain_random_f32(alg, seed):
    # the reference's verify generator shape: chain + 8 random out-edges,
    # weights uniform in [1, 16] (DS/testing/verify_bf_delta_vs_dijkstra.cpp:144-201)
    g = M.chain_random(2048, 8, 1.0, 16.0, M.W_F32, seed=seed)
    dg = M.DeviceGraph(g)
    for src in (0, 17, 2047):
        _check(g, dg, src, alg, delta=4.0)


@pytest.fixture(scope="module")
def rmat14():
    g = M.rmat(14, 16, seed=7)
    return g, M.DeviceGraph(g)


@pytest.mark.parametrize("alg", ALGS)
def test_rmat14_f32(rmat14, alg):
    g, dg = rmat14
    assert dg.info.large_vertices > 0  # exercises the hub chunk path
    for src in g.pick_sources(4, seed=11):
        _check(g, dg, int(src), alg)


@pytest.mark.parametrize("alg", ALGS)
def test_routing_grid_u32(alg):
    g = M.routing_grid(120, seed=5)
    dg = M.DeviceGraph(g)
    for src in g.pick_sources(3, seed=9):
        _check(g, dg, int(src), alg)


@pytest.mark.parametrize("delta", [0.05, 0.3, 2.0])
def test_rmat_delta_sweep(rmat14, delta):
    g, dg = rmat14
    src = int(g.pick_sources(1, seed=3)[0])
    _check(g, dg, src, M.ALG_DELTA, delta=delta)


def test_tiny_graph_known_answer():
    ka = bf_cases.known_answers()["tiny_gr"]
    rp = [0] + ka["row_ends"]
    g = M.from_csr(rp, ka["destinations"], np.array(ka["weights"], np.float32))
    dg = M.DeviceGraph(g)
    for alg in ALGS:
        dist, pred, _ = dg.sssp(0, alg, delta=1.0)
        want = np.array([np.inf if x is None else x for x in ka["expected"]], np.float32)
        assert dist.tobytes() == want.tobytes()
        assert list(pred) == [-1, 0, 1, 2, -1]


def test_rrg_path_known_answer():
    ka = bf_cases.known_answers()["rrg"]
    e = np.array(ka["edges"], dtype=object)
    g = M.from_edges(ka["n"], e[:, 0].astype(np.int32), e[:, 1].astype(np.int32),
                     e[:, 2].astype(np.float32))
    dg = M.DeviceGraph(g)
    for alg in ALGS:
        dist, pred, _ = dg.sssp(ka["source"], alg, delta=0.5)
        path = [ka["target"]]
        while path[-1] != ka["source"]:
            path.append(int(pred[path[-1]]))
        assert path[::-1] == ka["path"]
        # the reference's hand value 3.9 is the fp32 left fold along that path
        want = np.float32(np.float32(np.float32(0.7) + np.float32(1.1)) + np.float32(2.1))
        assert dist[ka["target"]] == want
        assert math.isclose(float(dist[ka["target"]]), ka["expected_target_approx"], rel_tol=1e-6)


def test_single_vertex_and_unreachable():
    g = M.from_csr([0, 0], [], np.zeros(0, np.float32))
    dg = M.DeviceGraph(g)
    for alg in ALGS:
        dist, pred, st = dg.sssp(0, alg, delta=1.0)
        assert dist.tolist() == [0.0] and pred.tolist() == [-1]
    g = M.from_edges(4, [0, 2], [1, 3], np.array([1.0, 1.0], np.float32))
    dg = M.DeviceGraph(g)
    for alg in ALGS:
        dist, pred, _ = dg.sssp(0, alg, delta=1.0)
        assert dist[0] == 0 and dist[1] == 1 and np.isinf(dist[2]) and np.isinf(dist[3])
        assert pred.tolist() == [-1, 0, -1, -1]


def test_source_out_of_range_and_negative_delta_rejected():
    g = M.chain_random(64, 2, 1.0, 2.0, M.W_F32, seed=1)
    dg = M.DeviceGraph(g)
    with pytest.raises(M.Mi355Error) as e:
        dg.sssp(64, M.ALG_BF)
    assert e.value.code == -2
    neg = M.from_edges(3, [0, 1], [1, 2], np.array([1.0, -1.0], np.float32))
    with pytest.raises(M.Mi355Error):
        M.DeviceGraph(neg)  # negative weights need allow_negative
    dneg = M.DeviceGraph(neg, allow_negative=True)
    with pytest.raises(M.Mi355Error):
        dneg.sssp(0, M.ALG_DELTA, delta=1.0)
    dist, _, _ = dneg.sssp(0, M.ALG_BF_JACOBI)
    assert dist.tolist() == [0.0, 1.0, 0.0]


# ---- reference entry points ---------------------------------------------------

def _in_graph(n, edges):
    e = np.array(edges, dtype=np.float64).reshape(-1, 3)
    return M.from_edges(n, e[:, 0].astype(np.int32), e[:, 1].astype(np.int32),
                        e[:, 2].astype(np.float32), orientation=M.IN, dedup_min=True)


@pytest.mark.parametrize("case", bf_cases.all_cases(), ids=lambda c: c[0])
def test_bf_csr_known_answers(case):
    name, n, src, edges, expected, _neg = case
    print(f"bf_csr case {name}", flush=True)
    g = _in_graph(n, edges)
    for ci in (1, 0, 3):
        dist, st, prog = M.bf_csr(g, src, -1, ci)
        want_d, want_it, want_conv = orc.bf_jacobi_in(g.rowptr, g.col, g.weights, src, -1, ci)
        assert dist.tobytes() == want_d.tobytes()
        assert (st.iterations_used, bool(st.converged)) == (want_it, want_conv)
        assert len(prog) == st.iterations_used
    if n > 1:
        exp = np.array(expected, np.float32)
        dist, st, _ = M.bf_csr(g, src, -1, 1)
        assert dist.tobytes() == exp.tobytes()


def test_bf_csr_path_graph_hits_iteration_cap():
    # A path of n vertices needs n-1 changing rounds; with the default cap of
    # n-1 iterations the no-change round never runs, so converged is false.
    n = 50
    g = _in_graph(n, [(i, i + 1, 1.0) for i in range(n - 1)])
    dist, st, prog = M.bf_csr(g, 0, -1, 1)
    assert dist.tolist() == [float(i) for i in range(n)]
    assert (st.iterations_used, st.converged) == (n - 1, 0)
    assert all(p[3] == 1 for p in prog)
    dist, st, prog = M.bf_csr(g, 0, -1, 7)
    assert (st.iterations_used, st.converged, st.checks_performed) == (n - 1, 0, 7)


@pytest.mark.parametrize("case", [c for c in bf_cases.all_cases() if not c[5]], ids=lambda c: c[0])
def test_delta_stepping_csr_known_answers(case):
    name, n, src, edges, expected, _neg = case
    g = _in_graph(n, edges)
    for delta in (0.5, 3.0, 1000.0):
        dist, it, conv, prog = M.delta_stepping_csr(g, src, delta)
        assert dist.tobytes() == np.array(expected, np.float32).tobytes()
        assert it == orc.ds_bucket_count(dist, delta)
        assert conv
        assert len(prog) == it and prog[-1][3] == 0


def test_delta_stepping_csr_iteration_cap():
    g = M.chain_random(512, 4, 1.0, 16.0, M.W_F32, seed=4)
    gi = g.transpose()
    dist, it, conv, _ = M.delta_stepping_csr(gi, 0, 2.0)
    full = orc.ds_bucket_count(dist, 2.0)
    assert it == full and conv
    assert dist.tobytes() == orc.dijkstra(g.rowptr, g.col, g.weights, 0).tobytes()
    _, it2, conv2, prog = M.delta_stepping_csr(gi, 0, 2.0, max_iters=3)
    assert (it2, conv2) == (3, False) and prog[-1][3] == 1
    with pytest.raises(M.Mi355Error):
        M.delta_stepping_csr(gi, 0, 0.0)
    with pytest.raises(M.Mi355Error) as e:
        M.delta_stepping_csr(gi, 512, 1.0)
    assert e.value.code == -2


def _weight_keys(wbits, wtype):
    b = np.asarray(wbits, np.uint32)
    if wtype != M.W_F32:
        return b.astype(np.uint64)
    return np.where(b & 0x80000000, ~b, b ^ 0x80000000).astype(np.uint64)


@pytest.mark.parametrize("wtype", [M.W_F32, M.W_U32])
def test_engine_rows_sorted_by_weight(wtype):
    """The engine keeps every row ordered by weight, ties in column order
    (delta-stepping reads light edges as a row prefix), from either input
    orientation, and SSSP results do not depend on it."""
    rng = np.random.default_rng(17)
    n, m = 300, 4000
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    if wtype == M.W_F32:
        w = rng.choice(np.array([-1.5, -0.0, 0.0, 0.25, 0.5, 2.0], np.float32), m)
    else:
        w = rng.choice(np.array([0, 1, 7, 7, 100], np.uint32), m)
    for orient in (M.OUT, M.IN):
        g = M.from_edges(n, s, d, w, wtype, orientation=orient)
        dg = M.DeviceGraph(g, allow_negative=True)
        rp, col, wb = dg.download()
        go = g if orient == M.OUT else g.transpose()
        assert rp.tolist() == go.rowptr.tolist()
        rows = np.repeat(np.arange(n, dtype=np.uint64), np.diff(go.rowptr))
        key_in = (rows << np.uint64(32)) | _weight_keys(go.wbits, wtype)
        key_out = (rows << np.uint64(32)) | _weight_keys(wb, wtype)
        assert np.all(key_out[1:] >= key_out[:-1])
        # same (column, weight) multiset in every row
        pin = np.lexsort((go.col, key_in))
        pout = np.lexsort((col, key_out))
        assert np.array_equal(key_in[pin], key_out[pout])
        assert np.array_equal(go.col[pin], col[pout])
        if orient == M.OUT:
            order = np.argsort(key_in, kind="stable")  # ties keep the input column order
            assert np.array_equal(col, go.col[order])


@pytest.mark.parametrize("alg", [M.ALG_BF, M.ALG_BF_JACOBI, M.ALG_DELTA, M.ALG_ADDS])
def test_vertex_order_gives_identical_results(alg):
    """A skewed graph is renumbered by degree internally (auto), or kept in
    input order on request: distances, canonical parents and counts are the
    same either way, in input ids."""
    g = M.rmat(13, 16, seed=99)
    srcs = g.pick_sources(3, seed=4)
    sts = {}
    for order in (M.ORDER_INPUT, M.ORDER_DEGREE, M.ORDER_AUTO):
        dg = M.DeviceGraph(g, order=order)
        assert dg.info.vertex_order == (M.ORDER_DEGREE if order == M.ORDER_AUTO else order)
        sts[order] = [(_check(g, dg, int(s), alg).iterations_used if alg in (M.ALG_BF_JACOBI, M.ALG_DELTA)
                       else 0) for s in srcs]
    assert sts[M.ORDER_INPUT] == sts[M.ORDER_DEGREE] == sts[M.ORDER_AUTO]


def test_degree_order_download_is_in_input_order():
    g = M.rmat(12, 16, seed=5)
    a = M.DeviceGraph(g, order=M.ORDER_INPUT).download()
    b = M.DeviceGraph(g, order=M.ORDER_DEGREE).download()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert np.array_equal(a[0], g.rowptr)


def test_hip_adds_warm_start_on_a_renumbered_graph():
    """initialize_distances = false on a graph renumbered by degree: the
    caller's distances (input order) seed the labels.  One vertex v starts
    below its true distance: it keeps that value and is never expanded, so
    the rest equals Dijkstra with v's out-edges removed."""
    g = M.rmat(13, 16, seed=12)
    src = int(g.pick_sources(1, seed=3)[0])
    true = orc.dijkstra(g.rowptr, g.col, g.weights, src)
    deg = np.diff(g.rowptr)
    v = int(np.flatnonzero((deg > 50) & np.isfinite(true) & (true > 0))[0])
    init = np.full(g.n, np.inf, np.float32)
    init[v] = np.float32(true[v] / 2)
    o = M.AddsOptions.defaults()
    o.initialize_distances = 0
    rc, dist, _ = M.hip_adds(g.rowptr, g.col, g.weights, src, M.W_F32, o, init)
    assert rc == 0
    keep = np.ones(g.m, bool)
    keep[g.rowptr[v]:g.rowptr[v + 1]] = False
    deg_kept = deg.copy()
    deg_kept[v] = 0
    rp = np.concatenate([[0], np.cumsum(deg_kept)])
    want = orc.dijkstra(rp, g.col[keep], g.weights[keep], src)
    want[v] = init[v]
    assert dist.tobytes() == want.tobytes()


@pytest.mark.parametrize("alg", [M.ALG_BF, M.ALG_BF_JACOBI, M.ALG_DELTA, M.ALG_ADDS])
def test_inf_mode_flt_max_marks_unreached(alg):
    """MI355_INF_FLT_MAX writes FLT_MAX (bits 0x7F7FFFFF) for an unreached
    vertex, the reference ADDS driver's convention; MI355_INF_IEEE writes +inf
    (0x7F800000).  Finite distances are the same bits either way."""
    rng = np.random.default_rng(41)
    n = 600
    src = rng.integers(0, 400, 3000)
    dst = rng.integers(0, 400, 3000)  # vertices 400.. have no in-edges: unreachable
    keep = src != dst
    w = rng.uniform(0.01, 1.0, keep.sum()).astype(np.float32)
    g = M.from_edges(n, src[keep], dst[keep], w, M.W_F32)
    dg = M.DeviceGraph(g)
    want = orc.dijkstra(g.rowptr, g.col, g.weights, 0)
    unreached = ~np.isfinite(want)
    assert unreached.sum() >= 200
    for mode, bits in ((M.INF_FLT_MAX, 0x7F7FFFFF), (M.INF_IEEE, 0x7F800000)):
        dist, pred, _ = dg.sssp(0, alg, inf_mode=mode)
        u = dist.view(np.uint32)
        assert np.all(u[unreached] == bits), (mode, np.unique(u[unreached])[:4])
        assert np.array_equal(u[~unreached], want.view(np.uint32)[~unreached])
        assert np.all(pred[unreached] == -1)


@pytest.mark.gpu
def test_round_log_matches_the_jacobi_frontiers():
    # mi355_sssp_round_log: the next BF_JACOBI run writes |F_0| .. |F_rounds|,
    # where F_k = {v : d_k[v] < d_(k-1)[v]} of the Jacobi iterates (F_0 = the
    # source); the oracle's iterates after k rounds give every size.
    print("round log: R-MAT scale 9, Jacobi BF", flush=True)
    g = M.rmat(9, 16, seed=77)
    dg = M.DeviceGraph(g)
    src = int(g.pick_sources(1, 5)[0])
    log = dg.round_log(4096)
    dist, _, st = dg.sssp(src, M.ALG_BF_JACOBI)
    rounds = st.rounds
    assert log[rounds + 1] == -1 and (log[: rounds + 1] >= 0).all()
    prev = np.full(g.n, np.inf, np.float32)
    prev[src] = 0.0
    want = [1]
    for k in range(1, rounds + 1):
        d_k, _, _ = orc.bf_jacobi_in(g.rowptr, g.col, g.weights, src, k, 1)
        want.append(int((d_k < prev).sum()))
        prev = d_k
    assert log[: rounds + 1].tolist() == want
    assert want[-1] == 0 and dist.tobytes() == prev.tobytes()
    # the log is for one run only: the next run leaves the array alone
    log2 = log.copy()
    dg.sssp(src, M.ALG_BF_JACOBI)
    assert (log == log2).all()


@pytest.mark.gpu
def test_in_place_bf_chains_inside_rounds():
    # In-place Bellman-Ford relaxes chains of improvements inside a round
    # (per-wave lists, a bounded number of passes, the rest to the next
    # round's frontier).  A bare path makes every chain hit the pass limit, a
    # grid overflows the lists into the frontier; both stay exact and need
    # far fewer frontier rounds than hops.  Negative weights without a cycle
    # and graphs with chunked rows (no chaining) are exact too.
    n = 5000
    path = M.from_edges(n, np.arange(n - 1, dtype=np.int32), np.arange(1, n, dtype=np.int32),
                        np.full(n - 1, 0.5, np.float32))
    dg = M.DeviceGraph(path)
    log = dg.round_log(2 * n)
    _check(path, dg, 0, M.ALG_BF)
    busy = int((log > 0).sum())
    assert busy < (n - 1) // 4, busy  # more than four hops per round on average
    g = M.grid4(300, 300, 1, 100, seed=12)
    dg = M.DeviceGraph(g)
    for src in (0, 300 * 150 + 150):
        _check(g, dg, src, M.ALG_BF)
    rng = np.random.default_rng(3)
    m = 30000
    s = rng.integers(0, 3000, m).astype(np.int32)
    d = rng.integers(0, 3000, m).astype(np.int32)
    # a potential-shifted weight keeps every cycle positive: w(u,v) = c + p(v) - p(u) >= 0
    p = rng.uniform(0, 5, 3000).astype(np.float64)
    w = (rng.uniform(0.5, 2.0, m) + p[d] - p[s]).astype(np.float32)
    keep = (s != d) & (w + p[s] - p[d] > 0.25)
    neg = M.from_edges(3000, s[keep], d[keep], w[keep])
    dneg = M.DeviceGraph(neg, allow_negative=True)
    dist, _, st = dneg.sssp(0, M.ALG_BF)
    # every label is some walk's fp32 left fold and the fixpoint is at most
    # every walk's (rounding is monotone), so any relaxation order reaches the
    # oracle's Jacobi fixpoint bit for bit
    t = neg.transpose()
    want, _, conv = orc.bf_jacobi_in(t.rowptr, t.col, t.weights, 0, -1, 1)
    assert conv and st.converged == 1
    assert dist.tobytes() == want.tobytes()
    hub = M.rmat(12, 16, seed=21)  # rows above 1024 edges: chunked, chaining off
    dh = M.DeviceGraph(hub)
    _check(hub, dh, int(hub.pick_sources(1, seed=2)[0]), M.ALG_BF)

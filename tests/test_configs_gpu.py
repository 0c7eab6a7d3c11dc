"""GPU parity at the sizes of the benchmark configurations (bit-exact).

  C0  256 x 256 4-neighbour grid, uint32 weights in [1, 100], source 0:
      Bellman-Ford (in place and Jacobi), delta-stepping and ADDS
  C1  R-MAT scale 22, edge factor 16, symmetrised, fp32 weights in (0, 1]:
      Bellman-Ford, Jacobi Bellman-Ford, delta-stepping and plain ADDS on 4
      seeded sources, learned-delta ADDS on 2
  C2  4900 x 4900 routing grid (10 % of edges deleted, 2 % diagonals,
      ceil(lognormal(3, 1)) weights): plain and learned ADDS on 2 sources
  C3  the 4096 box-limited multi-source queries on the C2 grid as one batch
      (HIP-graph rounds, autotuned options), 64 queries stratified by box area
      checked against the oracle's multi-source Dijkstra; and all 512 min-plus
      products of the batched GEMM (n log-uniform in [32, 512]) against the
      oracle's product
  C4  learned-delta ADDS on R-MAT scale 22 and the 4900 x 4900 grid with seeds
      that were not trained on

Distances are compared bit for bit with the oracle's Dijkstra and predecessors
with its canonical (smallest-index) parents.  The seeds and sizes are those
bench.py times.
"""
import ctypes as C
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import mi355sssp as M
from mi355sssp import batch as B
from mi355sssp import learn as L
from oracle import orc

pytestmark = pytest.mark.gpu

SEED = 2026          # the benchmark's graphs and sources
HELD_OUT = (31337, 31338)  # C4: R-MAT and routing-grid seeds not used in training
THREADS = min(16, os.cpu_count() or 1)
HAVE_MODEL = os.path.exists(L.DEFAULT_MODEL)


def _check(g, dg, src, alg, **kw):
    print(f"case n={g.n} m={g.m} src={src} alg={alg} {kw}", flush=True)
    dist, pred, st = dg.sssp(int(src), alg, **kw)
    want = orc.dijkstra(g.rowptr, g.col, g.weights, int(src))
    bad = np.flatnonzero(dist.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, f"{bad.size} distances differ, first at {bad[:8].tolist()}"
    wpred = orc.canonical_pred(g.rowptr, g.col, g.weights, int(src), want)
    badp = np.flatnonzero(pred != wpred)
    assert badp.size == 0, f"{badp.size} predecessors differ, first at {badp[:8].tolist()}"
    assert st.converged == 1
    reached = wpred >= 0
    reached[src] = True
    assert st.reached == int(reached.sum())
    return st


# ---- C0 ----------------------------------------------------------------------
@pytest.mark.parametrize("alg", [M.ALG_BF, M.ALG_BF_JACOBI, M.ALG_DELTA, M.ALG_ADDS])
def test_c0_grid256_u32(alg):
    g = M.grid4(256, 256, 1, 100, seed=SEED)
    assert g.n == 65536 and g.m == 261120
    _check(g, M.DeviceGraph(g), 0, alg)


# ---- C1 ----------------------------------------------------------------------
@pytest.fixture(scope="module")
def c1():
    g = M.rmat(22, 16, seed=SEED)
    assert g.n == 1 << 22 and g.m == 128308228
    return g, M.DeviceGraph(g), g.pick_sources(64, seed=SEED)


@pytest.mark.parametrize("alg", [M.ALG_BF, M.ALG_BF_JACOBI, M.ALG_DELTA, M.ALG_ADDS])
def test_c1_rmat22_four_sources(c1, alg):
    g, dg, srcs = c1
    for s in srcs[[0, 21, 42, 63]]:
        _check(g, dg, s, alg)


@pytest.mark.skipif(not HAVE_MODEL, reason="no trained delta model shipped")
def test_c1_rmat22_learned_adds(c1):
    g, dg, srcs = c1
    L.load_default_delta_model()
    try:
        for s in srcs[[1, 2]]:
            assert _check(g, dg, s, M.ALG_ADDS, use_mlp=1).policy == 2
            # the shipped configuration: finished buckets pulled when cheaper
            st = _check(g, dg, s, M.ALG_ADDS, use_mlp=1, pull=1)
            assert st.policy == 2 and st.pulled_phases > 0
        # plain ADDS with the pull rule, and every bucket pulled
        _check(g, dg, srcs[3], M.ALG_ADDS, pull=1)
        _check(g, dg, srcs[3], M.ALG_ADDS, use_mlp=1, pull=1, pull_margin=-60.0)
    finally:
        L.clear_delta_model()


@pytest.mark.skipif(not HAVE_MODEL, reason="no trained delta model shipped")
def test_c1_rmat22_many_runs_complete_and_stay_exact(c1):
    # A claim keeps its bucket's pushes in the workgroup between passes and
    # offers the rest at the end; nothing may reach the workgroup's local queue
    # after its last pass (an entry left there corrupted a later claim of the
    # same launch, which the engine reported as a bucket queue overflow in a
    # few runs out of dozens).  32 learned runs with pulls and 8 plain ones
    # must all finish; the next of each is checked bit for bit.
    g, dg, srcs = c1
    L.load_default_delta_model()
    try:
        o = dg.options(M.ALG_ADDS, use_mlp=1, pull=1)
        for s in srcs[4:36]:
            dg.sssp_labels(int(s), o)
        _check(g, dg, srcs[36], M.ALG_ADDS, use_mlp=1, pull=1)
    finally:
        L.clear_delta_model()
    o = dg.options(M.ALG_ADDS)
    for s in srcs[40:48]:
        dg.sssp_labels(int(s), o)
    _check(g, dg, srcs[48], M.ALG_ADDS)


# ---- C2 ----------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2():
    g = M.routing_grid(4900, seed=SEED)
    assert g.n == 4900 * 4900
    return g, M.DeviceGraph(g), g.pick_sources(16, seed=SEED)


def test_c2_routing_grid_plain_adds(c2):
    g, dg, srcs = c2
    for s in srcs[[0, 9]]:
        _check(g, dg, s, M.ALG_ADDS)


@pytest.mark.skipif(not HAVE_MODEL, reason="no trained delta model shipped")
def test_c2_routing_grid_learned_adds(c2):
    g, dg, srcs = c2
    L.load_default_delta_model()
    try:
        for s in srcs[[0, 9]]:
            _check(g, dg, s, M.ALG_ADDS, use_mlp=1)
        _check(g, dg, srcs[5], M.ALG_ADDS, use_mlp=1, pull=1)
    finally:
        L.clear_delta_model()


# ---- C3 ----------------------------------------------------------------------
def _d2h_slice(buf, dtype, lo, hi):
    out = np.empty(hi - lo, dtype)
    if out.nbytes:
        M._abi.check(M.engine().mi355_memcpy_d2h(out.ctypes.data, C.c_void_p(buf.ptr.value + 4 * lo),
                                                 out.nbytes, None))
    return out


def test_c3_routing_batch_4096_queries(c2):
    g, dg, _ = c2
    side = 4900
    qs = B.routing_queries(side, side, 4096, seed=SEED)
    best, _ = B.autotune(dg, qs, sample_stride=8)
    off = B.layout(dg, qs)
    total = int(off[-1])
    d_dist, d_pred = M.DeviceBuffer(4 * total), M.DeviceBuffer(4 * total)
    st = M._abi.BatchStats()
    M._abi.check(M.engine().mi355_sssp_batched(dg.handle, C.byref(qs.c()), C.byref(best), d_dist.ptr,
                                               d_pred.ptr, C.byref(st), None))
    print(f"batch: {qs.count} queries, {total} slots, {st.rounds} rounds, {st.graph_replays} graph "
          f"replays, {st.device_ms:.1f} ms", flush=True)
    assert st.graph_replays > 0  # the HIP-graph path ran
    # 64 queries stratified by area (the largest included)
    order = np.argsort(qs.areas(), kind="stable")
    sample = order[np.linspace(0, qs.count - 1, 64).astype(int)]

    def check(q):
        q = int(q)
        sq = qs.subset([q])
        box = tuple(int(x) for x in sq.boxes[0])
        want = orc.dijkstra_box(g.rowptr, g.col, g.weights, side, box, sq.sources)
        wpred = orc.canonical_pred_box(g.rowptr, g.col, g.weights, side, box, sq.sources, want)
        return q, want, wpred

    orc.lib()  # loaded before the worker threads call it
    with ThreadPoolExecutor(THREADS) as ex:
        for q, want, wpred in ex.map(check, sample):
            lo, hi = int(off[q]), int(off[q + 1])
            dist = _d2h_slice(d_dist, want.dtype, lo, hi)
            pred = _d2h_slice(d_pred, np.int32, lo, hi)
            assert dist.tobytes() == want.tobytes(), f"query {q}: distances differ"
            assert np.array_equal(pred, wpred), f"query {q}: predecessors differ"


def test_c3_minplus_gemm_512_products():
    rng = np.random.default_rng(SEED)
    ns = [int(round(math.exp(rng.uniform(math.log(32), math.log(512))))) for _ in range(512)]
    As, Bs = [], []
    for n in ns:
        As.append((rng.integers(1, 1 << 24, (n, n)) * 2.0 ** -24).astype(np.float32))
        Bs.append((rng.integers(1, 1 << 24, (n, n)) * 2.0 ** -24).astype(np.float32))
    got = M.kernels.minplus_gemm_batched(As, Bs)

    def check(i):
        n = ns[i]
        want = np.asarray(orc.minplus_gemm(As[i], Bs[i], n, n, n), np.float32).reshape(n, n)
        return i, got[i].tobytes() == want.tobytes()

    orc.lib()
    with ThreadPoolExecutor(THREADS) as ex:
        bad = [i for i, ok in ex.map(check, range(len(ns))) if not ok]
    assert not bad, f"{len(bad)} products differ, first {bad[:8]}"


# ---- C4 ----------------------------------------------------------------------
@pytest.mark.skipif(not HAVE_MODEL, reason="no trained delta model shipped")
@pytest.mark.parametrize("kind", ["rmat22", "grid4900"])
def test_c4_learned_adds_held_out_seeds(kind):
    if kind == "rmat22":
        g = M.rmat(22, 16, seed=HELD_OUT[0])
        srcs = g.pick_sources(2, seed=HELD_OUT[0])
    else:
        g = M.routing_grid(4900, seed=HELD_OUT[1])
        srcs = g.pick_sources(2, seed=HELD_OUT[1])
    dg = M.DeviceGraph(g)
    L.load_default_delta_model()
    try:
        for s in srcs:
            st = _check(g, dg, s, M.ALG_ADDS, use_mlp=1)
            assert st.policy == 2  # the network ran (the graph lies in its range)
            st = _check(g, dg, s, M.ALG_ADDS, use_mlp=1, pull=1)
            assert st.policy == 2
    finally:
        L.clear_delta_model()

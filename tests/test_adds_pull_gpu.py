"""GPU parity of ADDS with pulled buckets.

With pull = 1 a finished bucket's heavy edges are relaxed by the vertices they
could still improve (each reads the weight-sorted prefix of its in-edge row
that can lower or tie its label) instead of being pushed from the bucket's
vertices.  Distances must stay bit-exact and predecessors canonical whichever
buckets are pulled: symmetric graphs (their own rows serve as in-edge rows),
directed graphs with different weights per direction (a transpose is built),
zero weights, both vertex orders, the fixed, feedback and learned policies.
"""
import os

import numpy as np
import pytest

import mi355sssp as M
from mi355sssp import learn as L
from oracle import orc

pytestmark = pytest.mark.gpu

# pull every bucket that listed heavy edges (the margin makes the push work
# estimate exceed any pull estimate)
ALWAYS = -60.0


def _check(g, dg, src, **kw):
    print(f"adds pull n={g.n} m={g.m} src={src} {kw}", flush=True)
    dist, pred, st = dg.sssp(src, M.ALG_ADDS, **kw)
    want = orc.dijkstra(g.rowptr, g.col, g.weights, src)
    bad = np.flatnonzero(dist.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, f"{bad.size} distances differ, first {bad[:6].tolist()}"
    wp = orc.canonical_pred(g.rowptr, g.col, g.weights, src, want)
    badp = np.flatnonzero(pred != wp)
    assert badp.size == 0, f"{badp.size} predecessors differ, first {badp[:6].tolist()}"
    assert st.converged == 1
    return st


def _digraph(n, m, weight_type, seed, zero_share=0.0):
    """Random directed graph: every edge drawn independently (an edge and its
    reverse, where both exist, have unrelated weights), duplicates kept."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    if weight_type == M.W_F32:
        w = rng.uniform(0.0, 1.0, src.size).astype(np.float32)
        w[w == 0] = np.float32(0.5)
    else:
        w = rng.integers(1, 50, src.size).astype(np.uint32)
    if zero_share > 0:
        w[rng.random(src.size) < zero_share] = 0
    return M.from_edges(n, src, dst, w, weight_type)


@pytest.mark.parametrize("margin", [ALWAYS, 0.0])
@pytest.mark.parametrize("delta", [0.0, 0.02, 0.5])
def test_pull_rmat(margin, delta):
    g = M.rmat(14, 16, seed=7)
    dg = M.DeviceGraph(g)
    pulled = 0
    for src in g.pick_sources(3, seed=11):
        st = _check(g, dg, int(src), delta=delta, pull=1, pull_margin=margin)
        pulled += st.pulled_phases
    if margin == ALWAYS:
        assert pulled > 0


def test_pull_rmat_input_order_and_feedback_policy():
    g = M.rmat(13, 16, seed=12)
    for order in (M.ORDER_INPUT, M.ORDER_DEGREE):
        dg = M.DeviceGraph(g, order=order)
        for src in g.pick_sources(2, seed=3):
            st = _check(g, dg, int(src), pull=1, pull_margin=ALWAYS)
            assert st.pulled_phases > 0
            _check(g, dg, int(src), pull=1, pull_margin=ALWAYS, dynamic_delta=1)


def test_pull_grid_u32_and_routing_grid():
    g = M.grid4(96, 96, 1, 100, seed=2026)
    dg = M.DeviceGraph(g)
    st = _check(g, dg, 0, pull=1, pull_margin=ALWAYS, delta=20.0)
    assert st.pulled_phases > 0
    _check(g, dg, 5000, pull=1, pull_margin=0.0)
    g = M.routing_grid(120, seed=5)
    dg = M.DeviceGraph(g)
    for src in g.pick_sources(2, seed=9):
        _check(g, dg, int(src), pull=1, pull_margin=ALWAYS)


@pytest.mark.parametrize("weight_type", [M.W_F32, M.W_U32])
def test_pull_directed_graph_with_asymmetric_weights(weight_type):
    g = _digraph(4000, 48000, weight_type, seed=5 + weight_type)
    dg = M.DeviceGraph(g)
    for src in (0, 17, 3999):
        st = _check(g, dg, src, pull=1, pull_margin=ALWAYS)
    assert st.pulled_phases >= 0


def test_pull_zero_weights_u32():
    g = _digraph(3000, 30000, M.W_U32, seed=31, zero_share=0.3)
    dg = M.DeviceGraph(g)
    for src in (0, 1, 2):
        _check(g, dg, src, pull=1, pull_margin=ALWAYS, delta=3.0)


def test_pull_high_diameter_chain():
    g = M.chain_random(3000, 6, 1.0, 16.0, M.W_F32, seed=3)
    dg = M.DeviceGraph(g)
    _check(g, dg, 0, delta=2.0, pull=1, pull_margin=ALWAYS)


def test_pull_and_push_runs_share_a_handle():
    g = M.rmat(12, 16, seed=9)
    dg = M.DeviceGraph(g)
    srcs = [int(s) for s in g.pick_sources(4, seed=1)]
    for s in srcs:
        _check(g, dg, s, pull=1, pull_margin=ALWAYS)
        _check(g, dg, s)
    for alg in (M.ALG_DELTA, M.ALG_BF):
        dist, _, _ = dg.sssp(srcs[0], alg)
        assert dist.tobytes() == orc.dijkstra(g.rowptr, g.col, g.weights, srcs[0]).tobytes()


def test_pull_rejects_bad_options():
    g = M.grid4(16, 16, 1, 9, seed=1)
    dg = M.DeviceGraph(g)
    with pytest.raises(M.Mi355Error):
        dg.sssp(0, M.ALG_ADDS, pull=2)
    with pytest.raises(M.Mi355Error):
        dg.sssp(0, M.ALG_ADDS, pull=1, pull_margin=float("inf"))


@pytest.mark.skipif(not os.path.exists(L.DEFAULT_MODEL), reason="no trained delta model shipped")
def test_pull_under_the_learned_policy():
    L.load_default_delta_model()
    try:
        for g in (M.rmat(16, 16, seed=424243), M.routing_grid(300, seed=424242)):
            dg = M.DeviceGraph(g)
            for src in g.pick_sources(2, seed=77):
                _check(g, dg, int(src), use_mlp=1, pull=1, pull_margin=0.0)
                _check(g, dg, int(src), use_mlp=1, pull=1, pull_margin=ALWAYS)
    finally:
        L.clear_delta_model()


def test_device_light_split_follows_the_width_across_runs():
    # Under a device-side policy (here an explicit phase schedule) the light
    # split is chosen on the device: recounted when the width changes, kept
    # when it does not.  Alternating light widths on one handle exercises both
    # (for a symmetric graph the split serves the out-rows and the pull's
    # in-rows; for a directed one the transposed in-rows have their own).
    L.set_schedule([(1.0, 1, 1.0), (0.5, 2, 0.5)])
    try:
        for g in (_digraph(4000, 48000, M.W_F32, seed=11), M.rmat(14, 16, seed=3)):
            dg = M.DeviceGraph(g)
            srcs = [int(s) for s in g.pick_sources(2, seed=4)]
            for ls in (0.5, 0.25, 0.25, 1.0, 0.5):
                for s in srcs:
                    _check(g, dg, s, use_mlp=2, pull=1, pull_margin=ALWAYS, light_scale=ls)
    finally:
        L.set_schedule(None)

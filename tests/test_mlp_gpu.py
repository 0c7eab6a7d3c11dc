"""The delta network's HIP kernels against the CPU oracle, training on the
device, and the network inside the ADDS scheduler (bit-exact distances; the
delta it picks is the network's output)."""
import math
import os

import numpy as np
import pytest

import mi355sssp as M
from mi355sssp import learn as L
from oracle import mlp_ref as R
from oracle import orc

pytestmark = pytest.mark.gpu


def _model_and_data(seed, n=500):
    rng = np.random.default_rng(seed)
    m = L.DeltaModel.init(seed)
    m.w2 = rng.normal(0, 0.3, (L.N_OUT, 64)).astype(np.float32)
    m.b1 = rng.normal(0, 0.1, 64).astype(np.float32)
    m.mean = rng.normal(0, 1, 32).astype(np.float32)
    m.inv_std = rng.uniform(0.5, 2.0, 32).astype(np.float32)
    x = (rng.normal(0, 1, (n, 32)) / m.inv_std + m.mean).astype(np.float32)
    y = rng.normal(0, 1, (n, L.N_OUT)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return m, x, y, w


@pytest.mark.parametrize("seed", [1, 2])
def test_forward_matches_oracle(seed):
    m, x, _, _ = _model_and_data(seed)
    got = m.forward(x)
    want, _, _ = R.forward(m.flat(), m.mean, m.inv_std, x)
    assert np.allclose(got, want, rtol=1e-4, atol=1e-4 * np.abs(want).max())


@pytest.mark.parametrize("seed,weighted", [(3, False), (4, True)])
def test_gradient_matches_oracle(seed, weighted):
    m, x, y, w = _model_and_data(seed, n=777)
    loss, g = m.loss_grad(x, y, w if weighted else None)
    L_ref, g_ref = R.loss_grad(m.flat(), m.mean, m.inv_std, x, y, w if weighted else None)
    assert loss == pytest.approx(L_ref, rel=1e-4)
    got = g.flat()
    scale = np.abs(g_ref).max()
    bad = np.flatnonzero(~np.isclose(got, g_ref, rtol=1e-3, atol=1e-4 * scale))
    assert bad.size == 0, f"{bad.size} gradient entries differ, first {bad[:5].tolist()}"


def test_training_fits_a_teacher_and_is_deterministic():
    teacher, x, _, _ = _model_and_data(7, n=4096)
    y = R.forward(teacher.flat(), teacher.mean, teacher.inv_std, x)[0].astype(np.float32)
    a, la = L.fit(x, y, epochs=25, batch=256, lr=3e-3, seed=11)
    b, lb = L.fit(x, y, epochs=25, batch=256, lr=3e-3, seed=11)
    assert np.array_equal(a.flat(), b.flat()) and np.array_equal(la, lb)
    assert la[-1] < 0.1 * la[0], la
    final = R.loss(a.flat(), a.mean, a.inv_std, x, y)
    assert final == pytest.approx(float(la[-1]), rel=0.5)


def test_trace_records_features_and_policy():
    g = M.grid4(200, 200, 1, 100, seed=3)
    dg = M.DeviceGraph(g)
    rec, st = L.trace(dg, 0)
    assert len(rec) == st.iterations_used > 5
    # graph features (lanes 0..13) are constant over the run, phase features vary
    assert np.all(rec[:, :14] == rec[0, :14])
    assert np.all(rec[:, 15] == 0.0)  # log2(delta / delta0) under the plain policy
    assert np.all(rec[:, 33] == st.initial_delta) and np.all(rec[:, 34] == 1)
    assert np.all(rec[:, 37] == L.default_workers())  # every workgroup claims
    assert np.all(np.diff(rec[:, 35]) >= 0)  # edges relaxed so far


def _check_adds(g, dg, src, **kw):
    dist, pred, st = dg.sssp(src, M.ALG_ADDS, **kw)
    want = orc.dijkstra(g.rowptr, g.col, g.weights, src)
    assert dist.tobytes() == want.tobytes()
    assert np.array_equal(pred, orc.canonical_pred(g.rowptr, g.col, g.weights, src, want))
    return st


def test_adds_with_network_is_exact_and_follows_the_network():
    m, _, _, _ = _model_and_data(9)
    m.mean[:] = 0.0
    m.inv_std[:] = 0.25
    m.install()
    try:
        for g in (M.grid4(128, 128, 1, 100, seed=5), M.rmat(13, 16, seed=8)):
            dg = M.DeviceGraph(g)
            for src in g.pick_sources(2, seed=4):
                _check_adds(g, dg, int(src), use_mlp=1)
            rec, st = L.trace(dg, int(g.pick_sources(1, seed=4)[0]), use_mlp=1)
            out = m.forward(rec[:, :32])
            d0 = np.float32(st.initial_delta)
            scale = np.exp2(np.clip(out[:, 0], -4, 4)).astype(np.float32)
            want_delta = np.clip(d0 * scale, d0 / 64, d0 * 64)
            assert np.allclose(rec[:, 33], want_delta, rtol=2e-6)
            want_active = np.clip(np.floor(out[:, 1] + 0.5), 1, 8)
            assert np.array_equal(rec[:, 34], want_active)
            wk = np.float32(L.default_workers())
            share = np.exp2(np.clip(out[:, 2], -5, 0)).astype(np.float32)
            want_workers = np.minimum(np.floor(np.maximum(wk * share, np.float32(16))), wk)
            assert np.all(np.abs(rec[:, 37] - want_workers) <= 1)
            # the first phase's width is the network's output for the input the
            # training set uses for it (learn.first_phase_features); record 0
            # holds log2(that width / delta0) as its input 15
            first = m.forward(L.first_phase_features(rec[0, :14])[None, :])[0]
            want_first = np.clip(d0 * np.float32(np.exp2(np.clip(first[0], -4, 4))), d0 / 64, d0 * 64)
            assert rec[0, 15] == pytest.approx(float(np.log2(want_first / d0)), abs=1e-4)
    finally:
        L.clear_delta_model()


def test_adds_follows_a_phase_schedule():
    """use_mlp = 2: phase i runs with row min(i, len - 1) of the schedule
    (width, active buckets, share of claiming workgroups); distances and
    parents stay exact."""
    sched = [(0.125, 1, 0.5), (0.5, 2, 1.0), (2.0, 4, 0.25)]
    L.set_schedule(sched)
    try:
        for g in (M.rmat(14, 16, seed=12), M.grid4(150, 150, 1, 100, seed=13)):
            dg = M.DeviceGraph(g)
            for src in g.pick_sources(2, seed=5):
                st = _check_adds(g, dg, int(src), use_mlp=2)
                assert st.policy == 3
            rec, st = L.trace(dg, int(g.pick_sources(1, seed=5)[0]), use_mlp=2)
            assert len(rec) > 3
            d0 = np.float32(st.initial_delta)
            assert rec[0, 15] == pytest.approx(math.log2(sched[0][0]), abs=1e-5)
            row = np.minimum(np.arange(len(rec)) + 1, len(sched) - 1)  # record i sets phase i + 1
            want_delta = np.array([d0 * np.float32(sched[r][0]) for r in row], np.float32)
            assert np.allclose(rec[:, 33], want_delta, rtol=2e-6)
            assert np.array_equal(rec[:, 34], [sched[r][1] for r in row])
            wk = L.default_workers()
            want_workers = np.array([min(wk, max(int(wk * sched[r][2]), 16)) for r in row])
            assert np.all(np.abs(rec[:, 37] - want_workers) <= 1)
    finally:
        L.set_schedule(None)
    with pytest.raises(M.Mi355Error):  # no schedule set
        dg.sssp(0, M.ALG_ADDS, use_mlp=2)


def test_graph_outside_the_training_range_runs_the_fixed_delta():
    """The model's graph-feature range decides whether the network is applied:
    inside it the run reports policy 2 (learned), outside it the engine keeps
    the fixed heuristic delta (policy 0, same phases as a plain run)."""
    m, _, _, _ = _model_and_data(10)
    m.mean[:] = 0.0
    m.inv_std[:] = 0.25
    g = M.rmat(12, 16, seed=21)
    dg = M.DeviceGraph(g)
    src = int(g.pick_sources(1, seed=5)[0])
    rec, _ = L.trace(dg, src)
    f = rec[0, :L.N_GRAPH]
    try:
        m.feat_lo, m.feat_hi = f - 1.0, f + 1.0  # contains the graph
        m.install()
        st = _check_adds(g, dg, src, use_mlp=1)
        assert st.policy == 2
        m.feat_lo, m.feat_hi = f - 1.0, f + 1.0
        m.feat_hi[13] = f[13] - 0.5  # graph larger than every training graph
        m.feat_lo[13] = f[13] - 1.0
        m.install()
        st = _check_adds(g, dg, src, use_mlp=1)
        plain = _check_adds(g, dg, src)
        assert st.policy == 0 and plain.policy == 0
        assert st.initial_delta == plain.initial_delta
        rec2, st2 = L.trace(dg, src, use_mlp=1)  # every phase at the heuristic delta
        assert st2.policy == 0
        assert np.all(rec2[:, 33] == plain.initial_delta) and np.all(rec2[:, 34] == 1)
        assert np.all(rec2[:, 37] == L.default_workers())
        m.feat_lo[:] = 0.0  # lo == hi: unbounded
        m.feat_hi[:] = 0.0
        m.install()
        assert _check_adds(g, dg, src, use_mlp=1).policy == 2
    finally:
        L.clear_delta_model()


def test_use_mlp_without_a_model_is_rejected():
    L.clear_delta_model()
    g = M.grid4(16, 16, 1, 9, seed=1)
    dg = M.DeviceGraph(g)
    with pytest.raises(M.Mi355Error):
        dg.sssp(0, M.ALG_ADDS, use_mlp=1)


@pytest.mark.skipif(not os.path.exists(L.DEFAULT_MODEL), reason="no trained delta model shipped")
def test_trained_network_gain_on_held_out_graphs():
    """What the shipped learned configuration (the trained network's widths,
    active buckets and worker share, buckets pulled when the engine estimates
    it cheaper) gains over plain ADDS (fixed heuristic delta, every bucket
    pushed) on graphs whose seeds were not trained on; distances stay
    bit-exact.  It must be a real gain on every graph, not merely no loss."""
    model = L.load_default_delta_model()
    try:
        gains = []
        for name, g in (("routing grid 700^2", M.routing_grid(700, seed=424242)),
                        ("R-MAT scale 18", M.rmat(18, 16, seed=424243)),
                        ("R-MAT scale 20", M.rmat(20, 16, seed=424244))):
            dg = M.DeviceGraph(g)
            src = int(g.pick_sources(1, seed=77)[0])
            _check_adds(g, dg, src, use_mlp=1, pull=1)
            o_plain = dg.options(M.ALG_ADDS)
            o_learned = dg.options(M.ALG_ADDS, use_mlp=1, pull=1)
            dg.sssp_labels(src, o_plain)
            dg.sssp_labels(src, o_learned)
            plain = np.median([dg.sssp_labels(src, o_plain).kernel_ms for _ in range(5)])
            learned = np.median([dg.sssp_labels(src, o_learned).kernel_ms for _ in range(5)])
            gains.append(plain / learned)
            print(f"{name}: plain ADDS {plain:.2f} ms, learned {learned:.2f} ms, "
                  f"gain {plain / learned:.2f}x", flush=True)
        assert min(gains) >= 1.3, gains
    finally:
        L.clear_delta_model()

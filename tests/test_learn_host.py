"""Host side of the learned delta policy (no device): sample weighting,
size classes, the trusted feature range and model files."""
import os

import numpy as np
import pytest

from mi355sssp import learn as L

DATA = os.path.join(os.path.dirname(L.__file__), "data", "delta_mlp.json")


def _equal_per_graph(records, c=6.0):
    """Weights as collect() writes them: every graph the same total c, spread
    evenly over its records."""
    return np.concatenate([np.full(k, c / k, np.float32) for k in records])


def test_class_weights_give_every_class_the_same_total():
    metas = [dict(kind="rmat", scale=18, n=1 << 18), dict(kind="rmat", scale=18, n=1 << 18),
             dict(kind="grid", n=2_000_000), dict(kind="rmat", scale=20, n=1 << 20)]
    records = [3, 4, 2, 5]
    w = _equal_per_graph(records)
    out = L.class_weights(w, metas)
    bounds = np.cumsum([0] + records)
    per_graph = [float(out[bounds[i]:bounds[i + 1]].sum()) for i in range(len(metas))]
    # rmat18 holds two graphs: each gets half of the class total
    assert per_graph[0] == pytest.approx(per_graph[1]) == pytest.approx(per_graph[2] / 2)
    assert per_graph[2] == pytest.approx(per_graph[3])
    # within a graph the records stay equal
    assert np.allclose(out[0:3], out[0])
    with pytest.raises(ValueError):
        L.class_weights(w[:-1], metas)


def test_time_weights_scale_each_graph_by_its_plain_time():
    metas = [dict(plain_ms=1.0), dict(plain_ms=4.0), dict(plain_ms=9.0)]
    records = [2, 3, 4]
    w = _equal_per_graph(records)
    bounds = np.cumsum([0] + records)
    for power, want in ((1.0, [1, 4, 9]), (0.5, [1, 2, 3])):
        out = L.time_weights(w, metas, power)
        tot = np.array([out[bounds[i]:bounds[i + 1]].sum() for i in range(3)])
        assert tot / tot[0] == pytest.approx(want)


def test_record_counts_from_metadata_or_weights():
    records = [3, 4, 2, 5]
    w = _equal_per_graph(records)
    metas = [dict(kind="rmat") for _ in records]
    assert L.record_counts(metas, w).tolist() == records
    metas[1]["records"] = 4
    assert L.record_counts(metas, w).tolist() == records
    metas[1]["records"] = 5
    with pytest.raises(ValueError):
        L.record_counts(metas, w)


def test_first_phase_features():
    """The input the engine builds before the first claim: graph features,
    a window of one item in the first slot, the heuristic width."""
    g = np.arange(L.N_GRAPH, dtype=np.float32) * 0.5
    x = L.first_phase_features(g)
    assert x.shape == (L.N_IN,) and np.array_equal(x[:L.N_GRAPH], g)
    assert x[14] == np.log2(2.0) and x[15] == 0.0 and x[16] == np.log2(2.0)
    assert not x[17:].any()


def test_graph_class_and_feature_range():
    assert L.graph_class(dict(kind="rmat", scale=22, n=1 << 22)) == "rmat22"
    assert L.graph_class(dict(kind="grid", n=50_000)) == "grid5"        # clamped up
    assert L.graph_class(dict(kind="route", n=24_010_000)) == "route7"
    assert L.graph_class(dict(kind="grid", n=10 ** 9)) == "grid7"       # clamped down
    x = np.random.default_rng(0).normal(size=(50, L.N_IN)).astype(np.float32)
    lo, hi = L.graph_feature_range(x)
    assert lo.shape == hi.shape == (L.N_GRAPH,)
    assert np.array_equal(lo, x[:, :L.N_GRAPH].min(axis=0)) and np.array_equal(hi, x[:, :L.N_GRAPH].max(axis=0))


def test_model_file_round_trip_and_shipped_model(tmp_path):
    m = L.DeltaModel.init(seed=3)
    m.feat_lo[:] = np.arange(L.N_GRAPH, dtype=np.float32)
    m.feat_hi[:] = m.feat_lo + 1
    p = str(tmp_path / "m.json")
    m.save(p, meta={"note": "test"})
    r = L.DeltaModel.load(p)
    for a in ("w1", "b1", "w2", "b2", "mean", "inv_std", "feat_lo", "feat_hi"):
        assert np.array_equal(getattr(m, a), getattr(r, a)), a
    assert r.meta == {"note": "test"}
    c = L.DeltaModel.from_c(m.to_c())  # the C struct carries every field
    assert np.array_equal(c.flat(), m.flat()) and np.array_equal(c.feat_hi, m.feat_hi)
    shipped = L.DeltaModel.load(DATA)
    assert shipped.w1.shape == (L.N_HID, L.N_IN) and shipped.w2.shape == (L.N_OUT, L.N_HID)
    assert np.all(np.isfinite(shipped.flat())) and np.all(shipped.inv_std > 0)
    assert np.all(shipped.feat_hi >= shipped.feat_lo) and np.any(shipped.feat_hi > shipped.feat_lo)

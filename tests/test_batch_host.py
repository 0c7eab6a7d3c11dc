"""Host side of the batched queries (no device): the query generator, the
Queries container and the induced-subgraph builder that the one-at-a-time
comparison runs plain ADDS on."""
import numpy as np

import mi355sssp as M
from mi355sssp import batch as B
from oracle import orc


def test_routing_queries_shape_and_determinism():
    rows, cols = 300, 400
    qs = B.routing_queries(rows, cols, 200, seed=4, side_min=5, side_max=500, max_sources=6)
    again = B.routing_queries(rows, cols, 200, seed=4, side_min=5, side_max=500, max_sources=6)
    assert np.array_equal(qs.boxes, again.boxes) and np.array_equal(qs.sources, again.sources)
    b = qs.boxes.astype(np.int64)
    assert np.all(b[:, 0] >= 0) and np.all(b[:, 2] <= rows) and np.all(b[:, 1] >= 0) and np.all(b[:, 3] <= cols)
    assert np.all(b[:, 2] > b[:, 0]) and np.all(b[:, 3] > b[:, 1])
    assert np.array_equal(qs.areas(), (b[:, 2] - b[:, 0]) * (b[:, 3] - b[:, 1]))
    # square boxes unless clipped by the grid's height or width
    h, w = b[:, 2] - b[:, 0], b[:, 3] - b[:, 1]
    assert np.all((h == w) | (h == rows) | (w == cols))
    for q in range(qs.count):
        s = qs.query_sources(q).astype(np.int64)
        assert 1 <= len(s) <= 6
        r, c = s // cols, s % cols
        assert np.all((r >= b[q, 0]) & (r < b[q, 2]) & (c >= b[q, 1]) & (c < b[q, 3]))
    sub = qs.subset([5, 0, 199])
    assert sub.count == 3 and sub.grid_cols == cols
    for i, q in enumerate([5, 0, 199]):
        assert np.array_equal(sub.boxes[i], qs.boxes[q])
        assert np.array_equal(sub.query_sources(i), qs.query_sources(q))


def test_induced_subgraph_solves_the_query():
    """Single-source Dijkstra from the super source on the induced subgraph
    equals the oracle's multi-source Dijkstra on the rectangle."""
    side = 90
    g = M.routing_grid(side, seed=21)
    out = g if g.orientation == M.OUT else g.transpose()
    qs = B.routing_queries(side, side, 12, seed=3, side_min=4, side_max=70)
    for q in range(qs.count):
        sub, area = B.induced_subgraph(out, qs, q)
        assert sub.n == area + 1
        d_sub = orc.dijkstra(sub.rowptr, sub.col, sub.weights, area)
        want = orc.dijkstra_box(out.rowptr, out.col, out.weights, side, tuple(qs.boxes[q]), qs.query_sources(q))
        assert d_sub[:area].tobytes() == want.tobytes(), q

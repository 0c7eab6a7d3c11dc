"""Known-answer graphs of the reference's dense Bellman-Ford test.

Loads the transcribed cases of tests/golden/known_answers.json and rebuilds
the four generated cases from the construction rules of
HIP_kernel/bellman_ford/tests/test_correctness.cpp:246-371 (grid 10x10,
complete 10, banded 1024, star 10000) together with the expected distances
that file states for them.  Every case is returned as
(name, n, source, edges[(u, v, w)], expected[float or inf], has_negative).
"""
from __future__ import annotations

import json
import math
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
INF = math.inf


def _load():
    with open(os.path.join(_HERE, "known_answers.json")) as f:
        return json.load(f)


def _flat_to_edges(n, flat):
    # flat[v*n+u] = w(u->v)
    out = []
    for v in range(n):
        for u in range(n):
            w = flat[v * n + u]
            if w is not None:
                out.append((u, v, float(w)))
    return out


def grid_10x10():
    # test_correctness.cpp:246-275: right edges weight 1, down edges weight 2,
    # expected distance 2*row + col.
    w, h = 10, 10
    edges, exp = [], []
    for r in range(h):
        for c in range(w):
            node = r * w + c
            exp.append(float(2 * r + c))
            if c + 1 < w:
                edges.append((node, node + 1, 1.0))
            if r + 1 < h:
                edges.append((node, node + w, 2.0))
    return ("sparse_grid_10x10_directed_right_down", w * h, 0, edges, exp)


def complete_10():
    # :277-305: from<to weight to-from, else 100+from-to; expected = node.
    n = 10
    edges = [(a, b, float(b - a) if a < b else float(100 + a - b))
             for a in range(n) for b in range(n) if a != b]
    return ("dense_complete_10_positive_cycles", n, 0, edges, [float(i) for i in range(n)])


def banded_1024():
    # :307-343: 0->node weight node, and for offsets in [-16, 48] \ {0}:
    # from -> (from+offset) mod n with weight (to - from) + (|offset| % 5 + 1).
    n, back, fwd = 1024, 16, 48
    edges = [(0, v, float(v)) for v in range(1, n)]
    for a in range(n):
        for off in range(-back, fwd + 1):
            if off == 0:
                continue
            b = (a + off + n) % n
            slack = float(abs(off) % 5 + 1)
            edges.append((a, b, float(b - a) + slack))
    return ("medium_connected_1024_local_density", n, 0, edges, [float(i) for i in range(n)])


def star_10000():
    # :345-371: 0->node weight (node % 97) + 1; plus 0->5000 (1) and
    # 5000->9999 (-5), so dist[5000] = 1 and dist[9999] = -4.
    n = 10000
    edges, exp = [], [0.0] * n
    for v in range(1, n):
        wv = float(v % 97 + 1)
        edges.append((0, v, wv))
        exp[v] = wv
    edges.append((0, 5000, 1.0))
    exp[5000] = 1.0
    edges.append((5000, 9999, -5.0))
    exp[9999] = -4.0
    return ("large_sparse_star_10000_with_two_hop_improvement", n, 0, edges, exp)


def all_cases(include_large: bool = True):
    data = _load()
    out = []
    for c in data["bf_cases"]:
        n = c["n"]
        edges = ([tuple(e) for e in c["edges"]] if "edges" in c else _flat_to_edges(n, c["flat"]))
        exp = [INF if x is None else float(x) for x in c["expected"]]
        out.append((c["name"], n, c["source"], edges, exp))
    out += [grid_10x10(), complete_10(), banded_1024()]
    if include_large:
        out.append(star_10000())
    return [(name, n, s, e, x, any(w < 0 for _, _, w in e)) for name, n, s, e, x in out]


def known_answers():
    return _load()

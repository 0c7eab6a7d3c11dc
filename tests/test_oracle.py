"""Pins the CPU oracle before it is trusted as the parity reference.

Each check is independent of the oracle's own code: the reference's known
answers (tests/golden), closed forms (hop counts, bucket counts), brute
force over all simple paths on tiny graphs with numpy fp32 arithmetic,
scipy's Dijkstra on integer weights, and cross-checks between oracle
functions that compute the same quantity by different algorithms.  A few
tests plant a mistake in an oracle output and require the checker to notice.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra as sp_dijkstra

from golden import bf_cases
from oracle import orc

INF = np.float32(np.inf)


def _csr_from_edges(n, edges, incoming=False, wdtype=np.float32):
    rows = [[] for _ in range(n)]
    for u, v, w in edges:
        (rows[v] if incoming else rows[u]).append((u if incoming else v, w))
    rp, col, ww = [0], [], []
    for r in rows:
        for c, w in sorted(r, key=lambda t: t[0]):
            col.append(c)
            ww.append(w)
        rp.append(len(col))
    return np.array(rp, np.int64), np.array(col, np.int32), np.array(ww, wdtype)


def _brute_force(n, edges, src):
    """min over all simple paths of the fp32 left-fold sum (numpy float32)."""
    adj = [[] for _ in range(n)]
    for u, v, w in edges:
        adj[u].append((v, np.float32(w)))
    best = [INF] * n
    best[src] = np.float32(0)

    def dfs(u, d, seen):
        for v, w in adj[u]:
            if v in seen:
                continue
            nd = np.float32(d + w)
            if nd < best[v]:
                best[v] = nd
            dfs(v, nd, seen | {v})

    dfs(src, np.float32(0), {src})
    return np.array(best, np.float32)


@pytest.mark.parametrize("case", [c for c in bf_cases.all_cases(include_large=True) if not c[5]],
                         ids=lambda c: c[0])
def test_dijkstra_known_answers(case):
    name, n, src, edges, expected, _ = case
    rp, col, w = _csr_from_edges(n, edges)
    d = orc.dijkstra(rp, col, w, src)
    assert d.tobytes() == np.array(expected, np.float32).tobytes()


@pytest.mark.parametrize("seed", range(25))
def test_dijkstra_matches_brute_force_fp32(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 7))
    edges = []
    for u in range(n):
        for v in range(n):
            if u != v and rng.random() < 0.5:
                # fp32 weights with awkward rounding: mixes of 2^-24-ish and ~1
                w = np.float32(rng.choice([rng.random(), rng.random() * 1e-7, 1 + rng.random()]))
                edges.append((u, v, w))
    rp, col, w = _csr_from_edges(n, edges)
    for src in range(n):
        d = orc.dijkstra(rp, col, w, src)
        assert d.tobytes() == _brute_force(n, edges, src).tobytes()


@pytest.mark.parametrize("seed", range(5))
def test_dijkstra_u32_matches_scipy(seed):
    rng = np.random.default_rng(100 + seed)
    n, m = 300, 2000
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    w = rng.integers(1, 1000, m)
    keep = u != v
    edges = list(zip(u[keep], v[keep], w[keep]))
    rp, col, ww = _csr_from_edges(n, edges, wdtype=np.uint32)
    d = orc.dijkstra(rp, col, ww, 0)
    A = sp.csr_matrix((ww.astype(np.float64), col, rp), shape=(n, n))
    want = sp_dijkstra(A, indices=0)
    reach = np.isfinite(want)
    assert np.array_equal(d[reach].astype(np.float64), want[reach])
    assert np.all(d[~reach] == 0xFFFFFFFF)


def test_dijkstra_u32_saturates_to_infinity():
    # JAS/hip_adds_sssp.cpp:205-208: a wrapped sum saturates to INF, which
    # never improves a label.
    rp, col, w = _csr_from_edges(3, [(0, 1, 0xFFFFFFF0), (1, 2, 0x20)], wdtype=np.uint32)
    d = orc.dijkstra(rp, col, w, 0)
    assert d.tolist() == [0, 0xFFFFFFF0, 0xFFFFFFFF]


def test_canonical_pred_and_planted_mistake():
    rng = np.random.default_rng(3)
    n, m = 200, 1500
    u, v = rng.integers(0, n, m), rng.integers(0, n, m)
    w = rng.integers(1, 5, m).astype(np.float32)  # many ties
    keep = u != v
    rp, col, ww = _csr_from_edges(n, list(zip(u[keep], v[keep], w[keep])))
    d = orc.dijkstra(rp, col, ww, 0)
    p = orc.canonical_pred(rp, col, ww, 0, d)
    assert orc.check_pred(rp, col, ww, 0, d, p) == 0
    # canonical = smallest valid parent, found independently by a scan
    for t in np.flatnonzero(np.isfinite(d))[:50]:
        if t == 0:
            continue
        parents = [s for s in range(n) for e in range(rp[s], rp[s + 1])
                   if col[e] == t and np.float32(d[s] + ww[e]) == d[t]]
        assert p[t] == min(parents)
    bad = p.copy()
    victim = int(np.flatnonzero(p > 0)[0])
    bad[victim] = (p[victim] + 1) % n
    while any(col[e] == victim and np.float32(d[bad[victim]] + ww[e]) == d[victim]
              for e in range(rp[bad[victim]], rp[bad[victim] + 1])):
        bad[victim] = (bad[victim] + 1) % n
    assert orc.check_pred(rp, col, ww, 0, d, bad) == 1
    dd = d.copy()
    dd[victim] = np.nextafter(dd[victim], np.float32(np.inf))
    assert orc.check_pred(rp, col, ww, 0, dd, p) >= 1


@pytest.mark.parametrize("case", bf_cases.all_cases(include_large=False), ids=lambda c: c[0])
def test_bf_jacobi_and_dense_known_answers(case):
    name, n, src, edges, expected, _ = case
    rp, col, w = _csr_from_edges(n, edges, incoming=True)
    exp = np.array(expected, np.float32)
    d, it, conv = orc.bf_jacobi_in(rp, col, w, src)
    if n > 1:
        assert d.tobytes() == exp.tobytes()
    A = np.full(n * n, np.inf, np.float32)
    for a, b, ww in edges:
        A[b * n + a] = min(A[b * n + a], np.float32(ww))
    dd, it2, conv2 = orc.bf_dense(A, n, src)
    assert dd.tobytes() == d.tobytes()
    assert (it, conv) == (it2, conv2)  # dense and sparse Jacobi are the same iteration


def test_bf_jacobi_iteration_closed_forms():
    # directed 10x10 grid (right 1, down 2): hop count to (r,c) is r+c, so 18
    # changing rounds and a 19th round that changes nothing.
    _, n, src, edges, exp = bf_cases.grid_10x10()
    rp, col, w = _csr_from_edges(n, edges, incoming=True)
    assert orc.bf_jacobi_in(rp, col, w, src, -1, 1)[1:] == (19, True)
    assert orc.bf_jacobi_in(rp, col, w, src, -1, 0)[1:] == (99, False)   # no checks: n-1
    assert orc.bf_jacobi_in(rp, col, w, src, -1, 4)[1:] == (20, True)    # next multiple of 4
    assert orc.bf_jacobi_in(rp, col, w, src, 10, 1)[1:] == (10, False)
    # a path of 50 vertices: 49 changing rounds fill the default cap of n-1
    n = 50
    rp, col, w = _csr_from_edges(n, [(i, i + 1, 1.0) for i in range(n - 1)], incoming=True)
    d, it, conv = orc.bf_jacobi_in(rp, col, w, 0)
    assert (it, conv) == (49, False) and d.tolist() == list(range(n))
    d, it, conv = orc.bf_jacobi_in(rp, col, w, 0, max_iters=0)
    assert (it, conv) == (0, False) and d[0] == 0 and np.isinf(d[1:]).all()


def test_gauss_seidel_closed_forms():
    # edges u -> v with u < v only: one ordered sweep is exact.
    rng = np.random.default_rng(9)
    n = 40
    A = np.full(n * n, np.inf, np.float32)
    edges = []
    for v in range(n):
        for u in range(v):
            if rng.random() < 0.2:
                ww = np.float32(rng.random() * 10)
                A[v * n + u] = ww
                edges.append((u, v, ww))
    rp, col, w = _csr_from_edges(n, edges)
    d, it = orc.gauss_seidel(A, n, 0, 1)
    assert it == 1
    assert d.tobytes() == orc.dijkstra(rp, col, w, 0).tobytes()
    # reversed order (u > v only) needs one sweep per hop: Jacobi-like
    Ar = np.full(n * n, np.inf, np.float32)
    for v in range(n - 1):
        Ar[v * n + v + 1] = 1.0  # edge (v+1) -> v
    d1, _ = orc.gauss_seidel(Ar, n, n - 1, 1)
    assert np.isfinite(d1).sum() == 2
    dn, _ = orc.gauss_seidel(Ar, n, n - 1, n - 1)
    assert dn.tolist() == [float(n - 1 - i) for i in range(n)]


@pytest.mark.parametrize("case", bf_cases.all_cases(include_large=False), ids=lambda c: c[0])
def test_gauss_seidel_known_answers(case):
    name, n, src, edges, expected, _ = case
    A = np.full(n * n, np.inf, np.float32)
    for a, b, ww in edges:
        A[b * n + a] = min(A[b * n + a], np.float32(ww))
    d, it = orc.gauss_seidel(A, n, src)
    assert it == max(n - 1, 0)
    if n > 1:
        assert d.tobytes() == np.array(expected, np.float32).tobytes()


def test_ds_bucket_count_closed_form():
    d = np.array([0.0, 0.5, 0.99, 1.0, 2.5, np.inf], np.float32)
    assert orc.ds_bucket_count(d, 1.0) == 3          # buckets {0, 1, 2}
    assert orc.ds_bucket_count(d, 1000.0) == 1
    assert orc.ds_bucket_count(d, 0.1) == 5          # 0, 5, 9, 10, 25
    # floorf(d * (1/delta)) with 1/delta rounded to fp32: 0.3 * fl(1/0.1)
    assert orc.ds_bucket_count(np.array([0.0, 0.3], np.float32), 0.1) == 2


@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 5, 3), (64, 64, 64), (33, 70, 129)])
def test_minplus_gemm_matches_broadcast(shape):
    M, N, K = shape
    rng = np.random.default_rng(M * 1000 + N)
    A = rng.random((M, K), dtype=np.float32)
    B = rng.random((K, N), dtype=np.float32)
    A[rng.random((M, K)) < 0.1] = np.inf
    C = orc.minplus_gemm(A.ravel(), B.ravel(), M, N, K).reshape(M, N)
    want = np.min(A[:, :, None] + B[None, :, :], axis=1)
    assert C.tobytes() == want.astype(np.float32).tobytes()


def test_minplus_gemm_leading_dimensions():
    rng = np.random.default_rng(5)
    M, N, K, lda, ldb, ldc = 5, 4, 3, 7, 6, 9
    A = rng.random(M * lda, dtype=np.float32)
    B = rng.random(K * ldb, dtype=np.float32)
    C = orc.minplus_gemm(A, B, M, N, K, lda, ldb, ldc).reshape(M, ldc)[:, :N]
    want = np.min(A.reshape(M, lda)[:, :K, None] + B.reshape(K, ldb)[None, :, :N], axis=1)
    assert C.tobytes() == want.tobytes()


def test_spgemm_matches_dense_product():
    rng = np.random.default_rng(11)
    M, K, N = 60, 50, 40

    def rand_csr(r, c, p):
        dense = np.where(rng.random((r, c)) < p, rng.random((r, c)).astype(np.float32), np.inf)
        rp, col, val = [0], [], []
        for i in range(r):
            for j in range(c):
                if np.isfinite(dense[i, j]):
                    col.append(j)
                    val.append(dense[i, j])
            rp.append(len(col))
        return dense.astype(np.float32), np.array(rp), np.array(col, np.int32), np.array(val, np.float32)

    Ad, arp, acol, aval = rand_csr(M, K, 0.1)
    Bd, brp, bcol, bval = rand_csr(K, N, 0.1)
    rp, col, val = orc.minplus_spgemm(arp, acol, aval, K, brp, bcol, bval, N)
    dense = np.min(Ad[:, :, None] + Bd[None, :, :], axis=1)
    got = np.full((M, N), np.inf, np.float32)
    for i in range(M):
        cols = col[rp[i]:rp[i + 1]]
        assert np.all(np.diff(cols) > 0)
        got[i, cols] = val[rp[i]:rp[i + 1]]
    assert got.tobytes() == dense.tobytes()


def test_spgemm_skips_inf_and_nan_entries():
    # MP/src/minplus_sparse_hip.cpp:93-99: explicit +INF / NaN inputs ignored
    arp = np.array([0, 3]); acol = np.array([0, 1, 2], np.int32)
    aval = np.array([1.0, np.inf, np.nan], np.float32)
    brp = np.array([0, 1, 2, 3]); bcol = np.array([0, 1, 2], np.int32)
    bval = np.array([2.0, 5.0, 7.0], np.float32)
    rp, col, val = orc.minplus_spgemm(arp, acol, aval, 3, brp, bcol, bval, 3)
    assert rp.tolist() == [0, 1] and col.tolist() == [0] and val.tolist() == [3.0]


def test_transpose_matches_scipy():
    rng = np.random.default_rng(2)
    A = sp.random(50, 30, density=0.2, format="csr", random_state=3, dtype=np.float32)
    A.sort_indices()
    trp, tcol, tw = orc.transpose(A.indptr, A.indices, A.data.view(np.uint32), 30)
    T = A.T.tocsr()
    T.sort_indices()
    assert np.array_equal(trp, T.indptr) and np.array_equal(tcol, T.indices)
    assert np.array_equal(tw.view(np.float32), T.data)
    rp2, c2, w2 = orc.transpose(trp, tcol, tw, 50)
    assert np.array_equal(rp2, A.indptr) and np.array_equal(c2, A.indices)


def _grid_graph(rows, cols, seed, wtype):
    """4-neighbour grid as an outgoing CSR (v = r * cols + c)."""
    rng = np.random.default_rng(seed)
    src, dst = [], []
    for r in range(rows):
        for c in range(cols):
            v = r * cols + c
            for dr, dc in ((0, 1), (1, 0), (0, -1), (-1, 0)):
                rr, cc = r + dr, c + dc
                if 0 <= rr < rows and 0 <= cc < cols:
                    src.append(v)
                    dst.append(rr * cols + cc)
    src, dst = np.array(src), np.array(dst)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    rowptr = np.zeros(rows * cols + 1, np.int64)
    np.add.at(rowptr, src + 1, 1)
    rowptr = np.cumsum(rowptr)
    if wtype == "u32":
        w = rng.integers(1, 20, len(src)).astype(np.uint32)
    else:
        w = (rng.integers(1, 1 << 24, len(src)) * 2.0 ** -24).astype(np.float32)
    return rowptr, dst.astype(np.int32), w


@pytest.mark.parametrize("wtype", ["u32", "f32"])
def test_box_dijkstra_closed_forms(wtype):
    rows, cols = 17, 23
    rp, col, w = _grid_graph(rows, cols, 5, wtype)
    # the whole grid as the box and one source: plain Dijkstra
    s = 5 * cols + 7
    full = orc.dijkstra_box(rp, col, w, cols, (0, 0, rows, cols), [s])
    assert full.tobytes() == orc.dijkstra(rp, col, w, s).tobytes()
    # several sources: the element-wise minimum of the single-source runs
    box = (3, 4, 14, 19)
    srcs = [5 * cols + 7, 12 * cols + 15, 3 * cols + 4]
    got = orc.dijkstra_box(rp, col, w, cols, box, srcs)
    sub_rows, sub_cols = box[2] - box[0], box[3] - box[1]
    # the induced subgraph built explicitly, relabelled to local ids
    keep = {}
    for r in range(box[0], box[2]):
        for c in range(box[1], box[3]):
            keep[r * cols + c] = (r - box[0]) * sub_cols + (c - box[1])
    s_src, s_dst, s_w = [], [], []
    for u, lu in keep.items():
        for e in range(rp[u], rp[u + 1]):
            if int(col[e]) in keep:
                s_src.append(lu)
                s_dst.append(keep[int(col[e])])
                s_w.append(w[e])
    order = np.lexsort((s_dst, s_src))
    s_src, s_dst, s_w = np.array(s_src)[order], np.array(s_dst)[order], np.array(s_w)[order]
    srp = np.zeros(sub_rows * sub_cols + 1, np.int64)
    np.add.at(srp, s_src + 1, 1)
    srp = np.cumsum(srp)
    per = [orc.dijkstra(srp, s_dst.astype(np.int32), s_w.astype(w.dtype), keep[s]) for s in srcs]
    assert got.tobytes() == np.minimum.reduce(per).tobytes()
    # predecessors: sources have none; every other reached vertex has an in-box
    # parent that gives its distance exactly, the smallest such global id
    pred = orc.canonical_pred_box(rp, col, w, cols, box, srcs, got)
    for s in srcs:
        assert pred[keep[s]] == -1
    inv = {lv: v for v, lv in keep.items()}
    for lv in range(len(got)):
        if lv in (keep[s] for s in srcs):
            continue
        v = inv[lv]
        cands = []
        for u, lu in keep.items():
            for e in range(rp[u], rp[u + 1]):
                if col[e] == v:
                    du = got[lu]
                    nd = (np.float32(du) + w[e]) if wtype == "f32" else min(int(du) + int(w[e]), 0xFFFFFFFF)
                    if nd == got[lv]:
                        cands.append(u)
        assert pred[lv] == (min(cands) if cands else -1)

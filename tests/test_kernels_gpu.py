"""GPU parity of the min-plus products, dense Bellman-Ford / Gauss-Seidel and
the reference-family SSSP entry points (ONE, hip_adds) against the oracle."""
import numpy as np
import pytest

import mi355sssp as M
from golden import bf_cases
from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 5, 3), (64, 64, 64), (33, 70, 129),
                                   (200, 130, 257), (512, 512, 512)])
def test_minplus_gemm_bit_exact(shape):
    M_, N, K = shape
    print(f"gemm {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    A = ((rng.integers(0, 1 << 24, (M_, K)) + 1) / 2 ** 24).astype(np.float32)
    B = ((rng.integers(0, 1 << 24, (K, N)) + 1) / 2 ** 24).astype(np.float32)
    A[rng.random((M_, K)) < 0.05] = np.inf
    C = M.minplus_gemm(A, B)
    want = orc.minplus_gemm(A.ravel(), B.ravel(), M_, N, K).reshape(M_, N)
    assert C.tobytes() == want.tobytes()


def test_minplus_gemm_rejects_bad_arguments():
    A = M.DeviceBuffer(64)
    lib = M.engine()
    assert lib.mi355_minplus_gemm_f32(A.ptr, A.ptr, A.ptr, 2, 2, 2, 1, 2, 2, None) == -1  # lda < K
    assert lib.mi355_minplus_gemm_f32(None, A.ptr, A.ptr, 2, 2, 2, 2, 2, 2, None) == -1
    assert lib.mi355_minplus_gemm_f32(A.ptr, A.ptr, A.ptr, 0, 2, 2, 2, 2, 2, None) == -1


def test_minplus_gemm_batched_and_autotune():
    rng = np.random.default_rng(3)
    sizes = [32, 45, 100, 257, 64, 512, 33]
    As = [rng.random((n, n), dtype=np.float32) for n in sizes]
    Bs = [rng.random((n, n), dtype=np.float32) for n in sizes]
    outs = M.minplus_gemm_batched(As, Bs)
    for a, b, c in zip(As, Bs, outs):
        n = a.shape[0]
        assert c.tobytes() == orc.minplus_gemm(a.ravel(), b.ravel(), n, n, n).reshape(n, n).tobytes()
    names, times = M.minplus_autotune([64, 256])
    assert times.shape == (2, len(names)) and np.all(times > 0)
    table = M.minplus_tuning()
    assert table[6] >= 0 and table[8] >= 0  # classes of 64 and 256 were tuned
    outs2 = M.minplus_gemm_batched(As, Bs)  # tuned choices give the same bits
    assert all(x.tobytes() == y.tobytes() for x, y in zip(outs, outs2))


def test_minplus_gemm_batched_ragged_classes():
    """160 products from two size classes in shuffled order, ragged M, N and
    K (K not a multiple of the tile's k-step), once on the default tile and
    once with each class forced onto another tile shape."""
    rng = np.random.default_rng(11)
    shapes = [(int(rng.integers(257, 301)), int(rng.integers(1, 41)), int(rng.integers(257, 301))) for _ in range(120)]
    shapes += [(int(rng.integers(65, 129)), int(rng.integers(1, 41)), int(rng.integers(65, 129))) for _ in range(40)]
    order = rng.permutation(len(shapes))
    shapes = [shapes[i] for i in order]
    As = [rng.random((m, k), dtype=np.float32) for m, k, _ in shapes]
    Bs = [rng.random((k, n), dtype=np.float32) for _, k, n in shapes]
    want = [orc.minplus_gemm(a.ravel(), b.ravel(), m, n, k).reshape(m, n).tobytes()
            for (m, k, n), a, b in zip(shapes, As, Bs)]
    saved = M.minplus_tuning()
    try:
        for table in (None, {7: 0, 9: 4}, {7: 3, 9: 1}, {7: 2, 9: 5}):
            t = [-1] * 32
            for cls, tile in (table or {}).items():
                t[cls] = tile
            M.minplus_tuning(t)
            outs = M.minplus_gemm_batched(As, Bs)
            assert [c.tobytes() for c in outs] == want, table
    finally:
        M.minplus_tuning(saved)


def test_minplus_gemm_batched_repeated_calls():
    """The same products twice (the uploaded plan is reused) with new data in
    the same buffers in between, then a different, smaller batch right after
    (the plan is rebuilt), then the first batch again."""
    import ctypes as C
    rng = np.random.default_rng(17)
    eng = M.engine()
    sizes = [40, 300, 77, 129, 512, 64]

    def batch(ns):
        mats = []
        for n in ns:
            a, b = rng.random((n, n), dtype=np.float32), rng.random((n, n), dtype=np.float32)
            mats.append([n, a, b, M.DeviceBuffer.from_array(a), M.DeviceBuffer.from_array(b),
                         M.DeviceBuffer(4 * n * n)])
        return mats

    def run(mats):
        cnt = len(mats)
        ptr = lambda j: (C.c_void_p * cnt)(*[m[j].ptr.value for m in mats])  # noqa: E731
        d = (C.c_int * cnt)(*[m[0] for m in mats])
        M._abi.check(eng.mi355_minplus_gemm_f32_batched(cnt, ptr(3), ptr(4), ptr(5), d, d, d, d, d, d, None))
        M._abi.check(eng.mi355_stream_sync(None))
        for n, a, b, _, _, dc in mats:
            want = orc.minplus_gemm(a.ravel(), b.ravel(), n, n, n)
            assert dc.to_array(np.float32, n * n).tobytes() == want.tobytes(), n

    first = batch(sizes)
    run(first)
    for m in first:  # new values, same buffers and sizes
        m[1] = rng.random((m[0], m[0]), dtype=np.float32)
        M._abi.check(eng.mi355_memcpy_h2d(m[3].ptr, m[1].ctypes.data, m[1].nbytes, None))
    run(first)
    run(batch([33, 90]))
    run(first)


def _rand_csr(rng, rows, cols, p):
    dense = np.where(rng.random((rows, cols)) < p, rng.random((rows, cols)).astype(np.float32), np.inf)
    rp, col, val = [0], [], []
    for i in range(rows):
        for j in range(cols):
            if np.isfinite(dense[i, j]):
                col.append(j)
                val.append(dense[i, j])
        rp.append(len(col))
    return np.array(rp), np.array(col, np.int32), np.array(val, np.float32)


@pytest.mark.parametrize("hc,bs", [(512, 64), (256, 64), (1024, 128), (4096, 256)])
def test_spgemm_matches_oracle(hc, bs):
    rng = np.random.default_rng(hc + bs)
    arp, acol, aval = _rand_csr(rng, 120, 90, 0.08)
    brp, bcol, bval = _rand_csr(rng, 90, 70, 0.1)
    rp, col, val, _ = M.spgemm(arp, acol, aval, 90, brp, bcol, bval, 70, hc, bs)
    wrp, wcol, wval = orc.minplus_spgemm(arp, acol, aval, 90, brp, bcol, bval, 70)
    assert np.array_equal(rp, wrp)
    for i in range(120):
        o = np.argsort(col[rp[i]:rp[i + 1]])
        assert np.array_equal(col[rp[i]:rp[i + 1]][o], wcol[wrp[i]:wrp[i + 1]])
        assert val[rp[i]:rp[i + 1]][o].tobytes() == wval[wrp[i]:wrp[i + 1]].tobytes()


def test_spgemm_empty_overflow_and_duplicates():
    # empty product: zero rowptr, null colind / values (hpp:51-57)
    z = np.zeros(4, np.int64)
    rp, col, val, nulls = M.spgemm(z, np.zeros(0, np.int32), np.zeros(0, np.float32), 3,
                                   z, np.zeros(0, np.int32), np.zeros(0, np.float32), 3)
    assert rp.tolist() == [0, 0, 0, 0] and nulls == (True, True)
    # one row with 300 distinct columns overflows a 256-slot table
    arp = np.array([0, 1]); acol = np.array([0], np.int32); aval = np.array([1.0], np.float32)
    brp = np.array([0, 300]); bcol = np.arange(300, dtype=np.int32); bval = np.ones(300, np.float32)
    with pytest.raises(M.Mi355Error) as e:
        M.spgemm(arp, acol, aval, 1, brp, bcol, bval, 300, 256, 64)
    assert e.value.code == -6
    rp, col, val, _ = M.spgemm(arp, acol, aval, 1, brp, bcol, bval, 300, 512, 64)
    assert rp.tolist() == [0, 300] and sorted(col.tolist()) == list(range(300))
    # duplicate columns reduce by min; +INF / NaN entries are not stored
    arp = np.array([0, 2]); acol = np.array([0, 1], np.int32); aval = np.array([1.0, 0.5], np.float32)
    brp = np.array([0, 2, 4]); bcol = np.array([0, 1, 0, 1], np.int32)
    bval = np.array([3.0, np.inf, 2.0, np.nan], np.float32)
    rp, col, val, _ = M.spgemm(arp, acol, aval, 2, brp, bcol, bval, 2)
    assert rp.tolist() == [0, 1] and col.tolist() == [0] and val.tolist() == [2.5]


def _dense(n, edges):
    A = np.full(n * n, np.inf, np.float32)
    for a, b, w in edges:
        A[b * n + a] = min(A[b * n + a], np.float32(w))
    return A



def _sparse_vec(rng, rows, density, dup):
    """Random sparse n x 1 vector: row counts 0, 1 or (dup) 2; some entries
    NaN or +inf (both ignored), the rest in (0, 8]."""
    cnt = (rng.random(rows) < density).astype(np.int64)
    if dup:
        cnt += (rng.random(rows) < 0.2) & (cnt > 0)
    rp = np.concatenate([[0], np.cumsum(cnt)])
    val = (rng.integers(1, 1 << 24, int(rp[-1])) * 2.0 ** -21).astype(np.float32)
    odd = rng.random(len(val))
    val[odd < 0.03] = np.nan
    val[(odd >= 0.03) & (odd < 0.06)] = np.inf
    return rp, val


def _dense_of(rp, val):
    rows = len(rp) - 1
    d = np.full(rows, np.inf, np.float32)
    r = np.repeat(np.arange(rows), np.diff(rp))
    ok = ~np.isnan(val)
    np.minimum.at(d, r[ok], val[ok])
    return d


@pytest.mark.parametrize("rows,dup", [(1, False), (777, False), (50000, True)])
def test_sparse_vector_min_merge_matches_elementwise_min(rows, dup):
    """The bf_csr_detail merge (mi355_sparse_vector_min_merge): per row the
    smaller entry (absent, NaN = +inf), +inf results dropped, every output row
    one entry in column 0; changed when some entry of b beats a's."""
    rng = np.random.default_rng(rows)
    arp, aval = _sparse_vec(rng, rows, 0.5, dup)
    brp, bval = _sparse_vec(rng, rows, 0.5, dup)
    print(f"merge rows={rows} a_nnz={len(aval)} b_nnz={len(bval)}", flush=True)
    rp, col, val, changed = M.kernels.sparse_vector_min_merge(arp, aval, brp, bval)
    da, db = _dense_of(arp, aval), _dense_of(brp, bval)
    want = np.minimum(da, db)
    keep = np.isfinite(want)
    assert np.array_equal(np.diff(rp), keep.astype(np.int64))
    assert np.all(col == 0)
    assert val.tobytes() == want[keep].tobytes()
    assert changed == bool(np.any(db < da))
    # b never below a: unchanged, and merging a vector with itself as well
    _, _, _, ch2 = M.kernels.sparse_vector_min_merge(arp, aval, arp, aval)
    assert not ch2


def test_sparse_vector_min_merge_empty_vectors():
    z = np.zeros(11, np.int64)
    rp, col, val, changed = M.kernels.sparse_vector_min_merge(z, np.zeros(0, np.float32), z,
                                                             np.zeros(0, np.float32))
    assert np.all(rp == 0) and len(val) == 0 and not changed

@pytest.mark.parametrize("case", bf_cases.all_cases(include_large=True), ids=lambda c: c[0])
def test_dense_bellman_ford_known_answers(case):
    name, n, src, edges, expected, _ = case
    A = _dense(n, edges)
    dist, it, conv = M.bf_dense(A, n, src, on_device=(n % 2 == 0))
    wd, wit, wconv = orc.bf_dense(A, n, src)
    assert dist.tobytes() == wd.tobytes()
    assert (it, conv) == (wit, wconv)
    if n > 1:
        assert dist.tobytes() == np.array(expected, np.float32).tobytes()


@pytest.mark.parametrize("case", bf_cases.all_cases(include_large=False), ids=lambda c: c[0])
def test_gauss_seidel_known_answers(case):
    name, n, src, edges, expected, _ = case
    A = _dense(n, edges)
    for iters in (-1, 1, 2):
        dist, it = M.gauss_seidel(A, n, src, iters)
        wd, wit = orc.gauss_seidel(A, n, src, iters)
        assert dist.tobytes() == wd.tobytes() and it == wit


def _csr_out(n, edges):
    g = M.from_edges(n, [e[0] for e in edges], [e[1] for e in edges],
                     np.array([e[2] for e in edges], np.float32))
    return g


@pytest.mark.parametrize("case", bf_cases.all_cases(include_large=False), ids=lambda c: c[0])
def test_one_bellman_ford_known_answers(case):
    name, n, src, edges, expected, neg = case
    if not edges:
        g = M.from_csr(np.zeros(n + 1, np.int64), [], np.zeros(0, np.float32))
    else:
        g = _csr_out(n, edges)
    dist, r = M.one_sssp(M.ALG_BF, g.rowptr, g.col, g.weights, src)
    assert dist.tobytes() == np.array(expected, np.float32).tobytes()
    assert r.has_negative_cycle == 0


def test_one_bellman_ford_detects_negative_cycle():
    edges = [(0, 1, 1.0), (1, 2, -2.0), (2, 1, 1.0), (2, 3, 1.0)]
    g = _csr_out(4, edges)
    dist, r = M.one_sssp(M.ALG_BF, g.rowptr, g.col, g.weights, 0)
    assert r.has_negative_cycle == 1 and r.converged == 0


@pytest.mark.parametrize("seed", [1, 2])
def test_one_adds_matches_oracle(seed):
    g = M.chain_random(4096, 8, 1.0, 20.0, M.W_F32, seed=seed)  # ONE/test_sssp_hip.cpp:26-66 shape
    for src in (0, 100):
        dist, r = M.one_sssp(M.ALG_ADDS, g.rowptr, g.col, g.weights, src)
        want = orc.dijkstra(g.rowptr, g.col, g.weights, src)
        assert dist.tobytes() == want.tobytes()
        assert r.converged == 1 and r.final_delta > 0


@pytest.mark.parametrize("wt", [M.W_F32, M.W_U32])
def test_hip_adds_matches_oracle(wt):
    g = M.rmat(12, 8, seed=4) if wt == M.W_F32 else M.routing_grid(64, seed=4)
    for src in g.pick_sources(2, seed=5):
        rc, dist, st = M.hip_adds(g.rowptr, g.col, g.wbits.view(np.float32) if wt == M.W_F32 else g.wbits,
                                  int(src), wt)
        assert rc == 0
        want = orc.dijkstra(g.rowptr, g.col, g.weights, int(src))
        assert dist.tobytes() == want.tobytes()
        assert st.edge_relaxations > 0
    rc, _, _ = M.hip_adds(g.rowptr, g.col, g.wbits.view(np.float32) if wt == M.W_F32 else g.wbits,
                          g.n, wt)
    assert rc == 1  # Status::InvalidArgument


def test_hip_adds_warm_start_keeps_existing_distances():
    # initialize_distances = false: labels start from the caller's buffer and
    # only the source is expanded (JAS/hip_adds_sssp.hpp:64-66)
    g = M.from_edges(4, [0, 1, 2], [1, 2, 3], np.array([5.0, 5.0, 5.0], np.float32))
    o = M.AddsOptions.defaults()
    o.initialize_distances = 0
    init = np.array([np.inf, 1.0, np.inf, np.inf], np.float32)
    rc, dist, _ = M.hip_adds(g.rowptr, g.col, g.weights, 0, M.W_F32, o, init)
    assert rc == 0
    assert dist.tolist() == [0.0, 1.0, np.inf, np.inf]


@pytest.mark.parametrize("nbytes,off", [(0, 0), (1, 0), (15, 0), (16, 0), (1 << 20, 0), ((1 << 20) + 7, 0),
                                        (4099, 3)])
def test_device_copy(nbytes, off):
    """mi355_memcpy_d2d (the streaming copy kernel) copies every byte,
    including unaligned ranges and tails, and nothing beyond."""
    import ctypes as C
    from mi355sssp._abi import check, engine

    rng = np.random.default_rng(nbytes + off)
    src = rng.integers(0, 256, nbytes + 64, dtype=np.uint8)
    a = M.DeviceBuffer.from_array(src)
    b = M.DeviceBuffer.from_array(np.full(nbytes + 64, 0xA5, np.uint8))
    check(engine().mi355_memcpy_d2d(C.c_void_p(b.ptr.value + off), C.c_void_p(a.ptr.value + off),
                                    nbytes, None))
    check(engine().mi355_stream_sync(None))
    got = b.to_array(np.uint8, nbytes + 64)
    assert np.array_equal(got[off:off + nbytes], src[off:off + nbytes])
    assert np.all(got[:off] == 0xA5) and np.all(got[off + nbytes:] == 0xA5)


def test_minplus_signed_zero_ties():
    """The one documented divergence from the reference's k-ordered select:
    where -0.0 (from -0.0 + -0.0) ties with +0.0 the zero's sign may differ;
    the value, and every other entry, is exact."""
    rng = np.random.default_rng(5)
    n = 64
    A = (rng.integers(1, 9, (n, n)) * 0.25).astype(np.float32)
    B = (rng.integers(1, 9, (n, n)) * 0.25).astype(np.float32)
    A[:8, :] = 0.0
    B[:, :8] = 0.0
    A[:8, ::2] = -0.0
    B[::2, :8] = -0.0
    got = M.minplus_gemm(A, B)
    want = np.asarray(orc.minplus_gemm(A, B, n, n, n), np.float32).reshape(n, n)
    assert np.array_equal(got, want)  # equal values (+0.0 == -0.0)
    zero = want == 0.0
    assert zero[:8, :8].all() and zero.sum() == 64
    assert got[~zero].tobytes() == want[~zero].tobytes()


@pytest.mark.gpu
def test_minplus_batch_tuner_keeps_results_exact():
    # the batch-level tuner only reorders which tile variant computes each
    # size class; every choice it installs must give bit-exact products
    from mi355sssp import kernels as K
    import ctypes as C
    rng = np.random.default_rng(8)
    ns = [int(x) for x in rng.integers(20, 300, 24)]
    mats = []
    for n in ns:
        a = rng.random((n, n), dtype=np.float32) + np.float32(0.01)
        b = rng.random((n, n), dtype=np.float32) + np.float32(0.01)
        mats.append((n, a, b, M.DeviceBuffer.from_array(a.ravel()), M.DeviceBuffer.from_array(b.ravel()),
                     M.DeviceBuffer(4 * n * n)))
    cnt = len(mats)
    pa = (C.c_void_p * cnt)(*[m[3].ptr.value for m in mats])
    pb = (C.c_void_p * cnt)(*[m[4].ptr.value for m in mats])
    pc = (C.c_void_p * cnt)(*[m[5].ptr.value for m in mats])
    dims = (C.c_int * cnt)(*ns)
    eng = M.engine()

    def run():
        M._abi.check(eng.mi355_minplus_gemm_f32_batched(cnt, pa, pb, pc, dims, dims, dims, dims, dims, dims,
                                                        None))
        eng.mi355_stream_sync(None)

    try:
        table, best_ms, static_ms, names = K.minplus_tune_batch(run, ns, reps=1, rounds=1)
        assert best_ms <= static_ms and len(names) == eng.mi355_minplus_tile_count()
        run()
        for n, a, b, _, _, dc in mats:
            want = orc.minplus_gemm(a.ravel(), b.ravel(), n, n, n)
            got = dc.to_array(np.float32, n * n)
            assert got.tobytes() == np.ascontiguousarray(want, np.float32).ravel().tobytes(), n
    finally:
        K.minplus_tuning([-1] * 32)

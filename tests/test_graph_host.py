"""Host graph library: generators, hashes, file formats (CPU only)."""
import hashlib
import json
import os
import struct
import subprocess

import numpy as np
import pytest

import mi355sssp as M
from golden import bf_cases

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF_BIN = os.path.join(REPO, "oracle", "_ref")


def _hash_bytes(g):
    h = hashlib.sha256()
    h.update(g.rowptr.tobytes())
    h.update(g.col.tobytes())
    h.update(g.wbits.tobytes())
    return h.hexdigest()


def test_sha256_matches_hashlib():
    g = M.grid4(16, 9, 1, 100, seed=3)
    assert g.sha256() == _hash_bytes(g)
    empty = M.from_csr([0, 0], [], np.zeros(0, np.float32))
    assert empty.sha256() == _hash_bytes(empty)


def test_generator_hashes_are_pinned():
    with open(os.path.join(HERE, "golden", "graph_hashes.json")) as f:
        want = json.load(f)
    got = {
        "grid4_256x256_w1_100_seed2026": M.grid4(256, 256, 1, 100, seed=2026).sha256(),
        "rmat_s12_ef16_seed7": M.rmat(12, 16, seed=7).sha256(),
        "routing_grid_120_seed5": M.routing_grid(120, seed=5).sha256(),
        "chain_random_2048_8_f32_seed1": M.chain_random(2048, 8, 1.0, 16.0, M.W_F32, seed=1).sha256(),
    }
    assert got == want


def test_grid4_shape_and_weights():
    g = M.grid4(256, 256, 1, 100, seed=2026)
    assert (g.n, g.m) == (65536, 261120)  # 4n - 4*256 directed edges
    w = g.weights
    assert w.dtype == np.uint32 and w.min() == 1 and w.max() == 100
    counts = np.bincount(w, minlength=101)[1:]
    assert counts.min() > 0.8 * g.m / 100 and counts.max() < 1.2 * g.m / 100
    deg = np.diff(g.rowptr)
    assert deg.min() == 2 and deg.max() == 4
    assert np.all(np.diff(g.col[g.rowptr[300]:g.rowptr[301]]) > 0)


def test_rmat_is_simple_symmetric_and_weighted_in_unit_interval():
    g = M.rmat(12, 16, seed=7)
    src = np.repeat(np.arange(g.n), np.diff(g.rowptr))
    assert not np.any(src == g.col)  # no self loops
    key = src.astype(np.int64) * g.n + g.col
    assert np.all(np.diff(key) > 0)  # rows sorted, no duplicates
    w = g.weights
    assert w.min() > 0 and w.max() <= 1
    assert np.all((w * 2 ** 24) == np.round(w * 2 ** 24))
    rkey = g.col.astype(np.int64) * g.n + src
    order = np.argsort(rkey)
    assert np.array_equal(rkey[order], key)  # symmetric
    assert np.array_equal(w[order], w)       # shared weight per undirected pair
    # duplicates of R-MAT's skewed draws are removed, so m < 2 * 16 * n
    assert g.m % 2 == 0 and 0.5 * 2 * 16 * g.n < g.m < 2 * 16 * g.n


def test_rmat_is_deterministic_and_seeded():
    a, b, c = M.rmat(10, 8, seed=1), M.rmat(10, 8, seed=1), M.rmat(10, 8, seed=2)
    assert a.sha256() == b.sha256() != c.sha256()


def test_routing_grid_properties():
    g = M.routing_grid(120, seed=5)
    deg = np.diff(g.rowptr)
    assert deg.max() <= 8
    assert 3.4 < g.m / g.n < 3.8  # 4*0.9 + 4*0.02 in the interior
    src = np.repeat(np.arange(g.n), deg)
    r, c = src // 120, src % 120
    vr, vc = g.col // 120, g.col % 120
    assert np.all(np.maximum(np.abs(r - vr), np.abs(c - vc)) == 1)
    key = src.astype(np.int64) * g.n + g.col
    rkey = g.col.astype(np.int64) * g.n + src
    order = np.argsort(rkey)
    assert np.array_equal(rkey[order], key) and np.array_equal(g.weights[order], g.weights)
    assert g.weights.min() >= 1


def test_pick_sources_have_out_edges_and_are_distinct():
    g = M.rmat(10, 4, seed=3)
    s = g.pick_sources(32, seed=4)
    assert len(set(s.tolist())) == 32
    assert np.all(np.diff(g.rowptr)[s] > 0)


def test_from_edges_dedup_and_orientation():
    g = M.from_edges(3, [0, 0, 1, 0], [1, 1, 2, 2], np.array([5.0, 2.0, 1.0, 9.0], np.float32),
                     orientation=M.IN, dedup_min=True)
    assert g.rowptr.tolist() == [0, 0, 1, 3]
    assert g.col.tolist() == [0, 0, 1] and g.weights.tolist() == [2.0, 9.0, 1.0]
    t = g.transpose()
    assert t.orientation == M.OUT and t.rowptr.tolist() == [0, 2, 3, 3]
    assert t.col.tolist() == [1, 2, 2] and t.weights.tolist() == [2.0, 9.0, 1.0]


def test_ripscsr1_roundtrip(tmp_path):
    g = M.chain_random(300, 3, 0.5, 2.0, M.W_F32, seed=8).transpose()
    p = str(tmp_path / "g.csrbin")
    g.write_ripscsr1(p)
    h = M.read_ripscsr1(p)
    assert h.orientation == M.IN and h.sha256() == g.sha256()
    raw = open(p, "rb").read()
    assert raw[:8] == b"RIPSCSR1"
    ver, orient, rows, cols = struct.unpack_from("<4Q", raw, 8)
    assert (ver, orient, rows, cols) == (1, 1, 300, 300)


def test_ripscsr1_reader_rejects_corrupt_files(tmp_path):
    """Column indices, row offsets and array counts are checked (counts against
    the file size, before anything is allocated)."""
    g = M.chain_random(50, 3, 0.5, 2.0, M.W_F32, seed=9).transpose()
    p = tmp_path / "g.csrbin"
    g.write_ripscsr1(str(p))
    good = p.read_bytes()
    rows, nnz = 50, g.m
    col_at = 8 + 80 + 8 * (rows + 1)
    cases = {}
    b = bytearray(good)
    struct.pack_into("<i", b, col_at + 4 * 3, rows)  # column == cols
    cases["column"] = b
    b = bytearray(good)
    struct.pack_into("<i", b, col_at, -1)
    cases["negative column"] = b
    b = bytearray(good)
    _, second = struct.unpack_from("<2q", b, 8 + 80 + 8 * 10)
    struct.pack_into("<q", b, 8 + 80 + 8 * 10, second + 1)  # rowptr[10] > rowptr[11]
    cases["rowptr"] = b
    b = bytearray(good)
    for k, v in ((2, 1 << 40), (6, 1 << 40), (7, (1 << 40) + 1), (8, 1 << 40), (9, 1 << 40)):
        struct.pack_into("<Q", b, 8 + 8 * k, v)  # counts far beyond the file
    cases["huge counts"] = b
    cases["short"] = bytearray(good[:-4])
    b = bytearray(good)
    # rows = 2^64 - 1 makes rows + 1 wrap to 0: with an empty rowptr count the
    # count checks alone would pass
    for k, v in ((2, (1 << 64) - 1), (7, 0)):
        struct.pack_into("<Q", b, 8 + 8 * k, v)
    cases["rows wrap"] = b
    b = bytearray(good)
    for k, v in ((2, 1 << 31), (7, (1 << 31) + 1)):
        struct.pack_into("<Q", b, 8 + 8 * k, v)
    cases["rows beyond 32-bit ids"] = b
    for name, data in cases.items():
        q = tmp_path / "bad.csrbin"
        q.write_bytes(bytes(data))
        with pytest.raises(M.Mi355Error):
            M.read_ripscsr1(str(q))
    assert nnz > 3


def test_galois_gr_fixtures(tmp_path):
    ka = bf_cases.known_answers()
    t = ka["tiny_gr"]
    p = tmp_path / "tiny.gr"
    with open(p, "wb") as f:  # layout of the reference's make_tiny_gr.py
        f.write(struct.pack("<QQQQ", 1, 4, t["num_nodes"], len(t["destinations"])))
        f.write(struct.pack(f"<{t['num_nodes']}Q", *t["row_ends"]))
        f.write(struct.pack(f"<{len(t['destinations'])}I", *t["destinations"]))
        if len(t["destinations"]) % 2:
            f.write(struct.pack("<I", 0))
        f.write(struct.pack(f"<{len(t['weights'])}f", *t["weights"]))
    g = M.read_galois_gr(str(p))
    assert g.rowptr.tolist() == [0] + t["row_ends"]
    assert g.col.tolist() == t["destinations"] and g.weights.tolist() == t["weights"]
    q = tmp_path / "back.gr"
    g.write_galois_gr(str(q))
    assert open(q, "rb").read() == open(p, "rb").read()
    u = ka["unweighted_gr"]
    p2 = tmp_path / "unweighted.gr"
    with open(p2, "wb") as f:
        f.write(struct.pack("<QQQQ", 1, 0, u["num_nodes"], len(u["destinations"])))
        f.write(struct.pack(f"<{u['num_nodes']}Q", *u["row_ends"]))
        f.write(struct.pack(f"<{len(u['destinations'])}I", *u["destinations"]))
    g2 = M.read_galois_gr(str(p2))
    assert g2.weights.tolist() == [1.0, 1.0, 1.0]
    mf = ka["malformed_gr"]
    p3 = tmp_path / "malformed.gr"
    with open(p3, "wb") as f:
        f.write(struct.pack("<QQQQ", 1, 4, mf["num_nodes"], len(mf["destinations"])))
        f.write(struct.pack(f"<{mf['num_nodes']}Q", *mf["row_ends"]))
        f.write(struct.pack(f"<{len(mf['destinations'])}I", *mf["destinations"]))
        f.write(struct.pack(f"<{len(mf['weights'])}f", *mf["weights"]))
    with pytest.raises(M.Mi355Error):
        M.read_galois_gr(str(p3))


DIMACS = """c tiny test graph
p sp 5 8
a 1 2 3.5
a 1 3 1
a 2 3 0.25
a 3 2 4
a 1 2 2.5
a 4 5 7
a 5 4 1e-3
a 3 1 10
"""


def test_dimacs_reader(tmp_path):
    p = tmp_path / "t.gr"
    p.write_text(DIMACS)
    g = M.read_dimacs(str(p))
    assert g.orientation == M.IN and g.n == 5
    assert (g.declared_edges, g.loaded_edges) == (8, 8)
    # incoming rows, columns ascending, duplicate 1->2 reduced to 2.5
    assert g.rowptr.tolist() == [0, 1, 3, 5, 6, 7]
    assert g.col.tolist() == [2, 0, 2, 0, 1, 4, 3]
    assert g.weights.tolist() == np.array([10, 2.5, 4, 1, 0.25, 1e-3, 7], np.float32).tolist()


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_BIN, "gr_to_csr")),
                    reason="reference converter not built (needs the reference tree)")
def test_ripscsr1_writer_matches_reference_converter(tmp_path):
    p = tmp_path / "t.gr"
    p.write_text(DIMACS)
    out_ref = tmp_path / "ref.csrbin"
    subprocess.run([os.path.join(REF_BIN, "gr_to_csr"), str(p), str(out_ref)], cwd=tmp_path,
                   check=True, capture_output=True)
    g = M.read_dimacs(str(p))
    mine = tmp_path / "mine.csrbin"
    g.write_ripscsr1(str(mine))
    assert mine.read_bytes() == out_ref.read_bytes()
    # and on a larger random graph with duplicates
    rng = np.random.default_rng(1)
    lines = ["p sp 400 5000"] + [f"a {u} {v} {w:.6g}" for u, v, w in
                                 zip(rng.integers(1, 401, 5000), rng.integers(1, 401, 5000),
                                     rng.random(5000) * 100)]
    p.write_text("\n".join(lines) + "\n")
    subprocess.run([os.path.join(REF_BIN, "gr_to_csr"), str(p), str(out_ref)], cwd=tmp_path,
                   check=True, capture_output=True)
    M.read_dimacs(str(p)).write_ripscsr1(str(mine))
    assert mine.read_bytes() == out_ref.read_bytes()


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_BIN, "gr_to_adjacency")),
                    reason="reference converter not built (needs the reference tree)")
def test_ripsadj1_writer_matches_reference_converter(tmp_path):
    p = tmp_path / "t.gr"
    p.write_text(DIMACS)
    out_ref = tmp_path / "ref.adjbin"
    subprocess.run([os.path.join(REF_BIN, "gr_to_adjacency"), str(p), str(out_ref)],
                   cwd=tmp_path, check=True, capture_output=True)
    mine = tmp_path / "mine.adjbin"
    M.write_ripsadj1(M.read_dimacs(str(p)), str(mine))
    assert mine.read_bytes() == out_ref.read_bytes()
    A = M.read_ripsadj1(str(mine))
    assert A[1, 0] == np.float32(2.5) and np.isinf(A[0, 1])


def test_from_dense_outgoing_skips_diagonal_and_inf():
    inf = np.inf
    adj = np.array([[0.0, 1.0, inf], [inf, 5.0, 2.0], [3.0, inf, 0.0]], np.float32)
    g = M.from_dense(adj.ravel(), 3)
    assert g.rowptr.tolist() == [0, 1, 2, 3]
    assert g.col.tolist() == [1, 2, 0] and g.weights.tolist() == [1.0, 2.0, 3.0]
    with pytest.raises(M.Mi355Error):
        M.from_dense(np.array([0, -1, 0, 0], np.float32), 2)

"""The C++ compatibility layer (include/compat/*.hpp) at run time.

tests/compat/compat_check.cpp calls the drop-in functions as a caller of the
reference would (bf_csr_detail's sparse-vector Bellman-Ford of
BF/src/bf_hip_CSR.cpp:36-99, the CSR Bellman-Ford / delta-stepping entry
points with their progress callback, ONE's dense overloads and
dense_to_csr_outgoing of ONE/sssp_hip.cpp:320-348, and every documented error
path); this file compares what it returns with the oracle.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import mi355sssp as M
from oracle import orc

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "compat", "_bin", "compat_check")


def _run(*args, timeout=300):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (build() builds it)")
    print("run compat_check", args, flush=True)
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0 and "FAIL" not in r.stdout, r.stdout[-4000:] + r.stderr[-2000:]
    return r.stdout


def _write_incoming(path, g_out):
    """Raw incoming CSR (row v lists the edges u -> v) of an outgoing graph."""
    t = g_out.transpose()
    rp, col, w = t.rowptr.astype(np.int64), t.col.astype(np.int32), t.weights.astype(np.float32)
    with open(path, "wb") as f:
        f.write(struct.pack("<qq", t.n, len(col)))
        f.write(rp.tobytes())
        f.write(col.tobytes())
        f.write(w.tobytes())


def _field(out, name):
    for line in out.splitlines():
        parts = line.split()
        if parts and parts[0] not in ("ok", "FAIL") and name in parts[::2]:
            return int(parts[parts.index(name) + 1])
    raise AssertionError(f"{name} missing from output:\n{out}")


def test_dense_to_csr_outgoing_host():
    # host code only: runs without a GPU
    _run("host")


@pytest.mark.gpu
@pytest.mark.parametrize("scale,seed", [(9, 11), (11, 12)])
def test_bf_csr_detail_sparse_vector_bellman_ford(tmp_path, scale, seed):
    g = M.rmat(scale, 8, seed=seed)
    src = int(g.pick_sources(1, seed=seed)[0])
    gp, op = tmp_path / "g.bin", tmp_path / "d.bin"
    _write_incoming(gp, g)
    out = _run("detail", gp, src, op)
    got = np.fromfile(op, np.float32)
    want = orc.dijkstra(g.rowptr, g.col, g.weights, src)
    assert got.tobytes() == want.tobytes()
    t = g.transpose()
    _, it, conv = orc.bf_jacobi_in(t.rowptr, t.col, t.weights, src)
    assert conv and it > 3 and _field(out, "iterations") == it
    assert _field(out, "nnz") == int(np.isfinite(want).sum())


@pytest.mark.gpu
def test_bf_csr_entry_points_and_progress_bridge(tmp_path):
    g = M.rmat(11, 8, seed=21)
    gp, op = tmp_path / "g.bin", tmp_path / "d.bin"
    _write_incoming(gp, g)
    src = int(g.pick_sources(1, seed=5)[0])
    out = _run("bfcsr", gp, src, op)
    want = orc.dijkstra(g.rowptr, g.col, g.weights, src)
    assert np.fromfile(op, np.float32).tobytes() == want.tobytes()
    t = g.transpose()
    _, it, _ = orc.bf_jacobi_in(t.rowptr, t.col, t.weights, src)
    assert it > 3 and _field(out, "iterations") == it
    assert _field(out, "ds_buckets") == orc.ds_bucket_count(want, 0.25)


@pytest.mark.gpu
def test_one_dense_overloads(tmp_path):
    rng = np.random.default_rng(7)
    n = 48
    A = np.where(rng.random((n, n)) < 0.15, rng.integers(1, 20, (n, n)).astype(np.float32), np.inf)
    A = A.astype(np.float32)
    mp, op = tmp_path / "a.bin", tmp_path / "d.bin"
    A.tofile(mp)
    out = _run("dense", n, mp, 2, op)
    # the oracle on the same outgoing CSR (diagonal and +inf are not edges)
    mask = np.isfinite(A) & ~np.eye(n, dtype=bool)
    rowptr = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int64)
    col = np.nonzero(mask)[1].astype(np.int32)
    w = A[mask].astype(np.float32)
    want = orc.dijkstra(rowptr, col, w, 2)
    assert np.fromfile(op, np.float32).tobytes() == want.tobytes()
    assert _field(out, "edges") == int(mask.sum())


@pytest.mark.gpu
def test_compat_error_paths():
    _run("errors")

"""The reference's own test drivers, built unchanged against the compatibility
headers and libmi355sssp.so by tests/refdrivers/Makefile (from build() where
the reference sources are), run on the GPU: the reference's tests, passing on
the engine.

  BF/tests/test_correctness.cpp          14 known-answer cases (dense BF)
  DS/testing/verify_bf_delta_vs_dijkstra.cpp  BF and delta-stepping vs its CPU Dijkstra
  BF/tests/test_check_intervals.cpp      checked / unchecked / var-check BF and bf_csr_detail
  BF/tests/big_test.cpp                  RIPSCSR1 load, checked vs unchecked BF
  DS/testing/speed_compare_delta_vs_bf.cpp  BF vs delta-stepping distances (exit 2 on mismatch)
"""
import os
import re
import subprocess

import pytest

import mi355sssp as M

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "refdrivers", "_bin")


def _run(args, timeout=300):
    exe = os.path.join(BIN, args[0])
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (build() builds it where the reference sources are)")
    print("run", args, flush=True)
    return subprocess.run([exe] + [str(a) for a in args[1:]], capture_output=True, text=True, timeout=timeout)


def _csrbin(tmp_path, scale=10, seed=3):
    g = M.rmat(scale, 8, seed=seed).transpose()  # RIPSCSR1 holds incoming fp32 graphs
    p = tmp_path / "g.csrbin"
    g.write_ripscsr1(str(p))
    return g, p


def test_reference_bellman_ford_correctness_suite():
    r = _run(["test_correctness"])
    assert r.returncode == 0 and "failed=0" in r.stdout, r.stdout[-3000:] + r.stderr[-2000:]


def test_reference_delta_stepping_verifier():
    # n, average out-degree, delta, (warm-up), cases, seed, source
    r = _run(["verify_bf_delta_vs_dijkstra", 2048, 8, 4.0, 0, 4, 1, 0])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]


def test_reference_check_intervals(tmp_path):
    g, p = _csrbin(tmp_path)
    # graph, 1-based source, max iterations, check intervals
    r = _run(["test_check_intervals", p, 1, g.n, 1, 4, 16])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]


def test_reference_big_test(tmp_path):
    g, p = _csrbin(tmp_path, scale=12, seed=5)
    r = _run(["big_test", p, 1])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert re.search(r"^distance_mismatches: 0$", r.stdout, re.M), r.stdout[-3000:]
    assert re.search(r"^converged: true$", r.stdout, re.M), r.stdout[-3000:]


def test_reference_speed_compare():
    # n, average out-degree, delta, warm-up, repeats, seed, source
    r = _run(["speed_compare_delta_vs_bf", 4096, 8, 4.0, 1, 2, 7, 0])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "Distance mismatches > 1e-4: 0 " in r.stdout, r.stdout[-3000:]

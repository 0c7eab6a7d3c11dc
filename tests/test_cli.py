"""The sssp_hip command-line driver: argument handling (CPU) and a run on the
reference's tiny .gr fixture (GPU): output lines, INF, the fd 3 summary."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import mi355sssp as M
from conftest import PKG

CLI = os.path.join(PKG, "lib", "sssp_hip")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")


def _run(args, **kw):
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120, **kw)


@pytest.mark.parametrize("args,msg", [
    ([], "usage"),
    (["-s", "x", "g.gr"], "Invalid source vertex"),
    (["-s", "-3", "g.gr"], "Invalid source vertex"),
    (["-g", "1.5", "g.gr"], "Invalid GPU device"),
    (["-z", "g.gr"], "usage"),
    (["a.gr", "b.gr"], "Unexpected extra argument"),
    (["-q"], "usage"),
])
def test_argument_errors(args, msg):
    r = _run(args)
    assert r.returncode == 1
    assert msg in r.stderr


def _tiny_gr(tmp_path):
    ka = json.load(open(GOLDEN))["tiny_gr"]
    ends = np.array(ka["row_ends"], np.int64)
    rowptr = np.concatenate([[0], ends])
    g = M.from_csr(rowptr, np.array(ka["destinations"], np.int32), np.array(ka["weights"], np.float32))
    path = str(tmp_path / "tiny.gr")
    g.write_galois_gr(path)
    return path, ka


@pytest.mark.gpu
def test_tiny_graph_output_and_summary(tmp_path):
    path, ka = _tiny_gr(tmp_path)
    out = tmp_path / "dist.txt"
    summary = tmp_path / "fd3.txt"
    # a shell gives the driver its file descriptor 3
    r = subprocess.run(["bash", "-c", 'exec "$0" "$@" 3>"$FD3_FILE"', CLI, "-q", "-s", str(ka["source"]),
                        "-o", str(out), path], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, FD3_FILE=str(summary)))
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("processed vertices=") == 8  # ADDS_RUNS default
    lines = out.read_text().splitlines()
    want = [f"{i} INF" if d is None else f"{i} {d:.9g}" for i, d in enumerate(ka["expected"])]
    assert lines == want
    fields = summary.read_text().split()
    assert fields[0] == path and float(fields[1]) > 0 and int(fields[2]) >= 4


@pytest.mark.gpu
def test_runs_env_and_stdout_output(tmp_path):
    path, ka = _tiny_gr(tmp_path)
    # file descriptor 3 always points at a file of the test's own
    fd3 = tmp_path / "fd3.txt"
    r = subprocess.run(["bash", "-c", 'exec "$0" "$@" 3>"$FD3_FILE"', CLI, "-s", "0", "-o", "-", path],
                       capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, ADDS_RUNS="2", FD3_FILE=str(fd3)))
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("processed vertices=") == 2
    assert "0 0\n" in r.stdout and "4 INF" in r.stdout
    assert len(fd3.read_text().splitlines()) == 1
    r = _run(["-s", "7", path])
    assert r.returncode == 1 and "outside" in r.stderr


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus above 1 exits with a message before
    touching a device."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    assert "single MI355X" in r.stderr and r.stdout == ""


@pytest.mark.gpu
@pytest.mark.parametrize("name,value", [("ADDS_WORKER_BLOCKS", "0"), ("ADDS_WORKER_BLOCKS", "12x"),
                                        ("ADDS_WORKER_BLOCKS", "99999999"), ("ADDS_RESERVED_CUS", "0"),
                                        ("ADDS_RESERVED_CUS", "-1"), ("ADDS_WL_ENTRIES", "5")])
def test_invalid_knobs_end_the_program(tmp_path, name, value):
    # invalid ADDS_WORKER_BLOCKS / ADDS_RESERVED_CUS / ADDS_WL_ENTRIES exit
    # with the accepted range (ADDSF/kernel.hip:34-51, 468-510)
    path, _ = _tiny_gr(tmp_path)
    r = _run(["-s", "0", path], env=dict(os.environ, **{name: value}))
    assert r.returncode == 1, r.stdout + r.stderr
    assert f"Invalid {name}='{value}'; expected an integer in [" in r.stderr


@pytest.mark.gpu
def test_valid_knobs_and_invalid_runs(tmp_path):
    path, ka = _tiny_gr(tmp_path)
    out = tmp_path / "d.txt"
    env = dict(os.environ, ADDS_WORKER_BLOCKS="16", ADDS_RESERVED_CUS="8", ADDS_WL_ENTRIES="20000000",
               ADDS_RUNS="many")
    r = _run(["-s", str(ka["source"]), "-o", str(out), path], env=env)
    assert r.returncode == 0, r.stderr
    assert "warning: ignoring invalid ADDS_RUNS='many'; using 8" in r.stderr
    assert r.stdout.count("processed vertices=") == 8
    assert "workers: 16 workgroup(s)" in r.stdout and "reserved CUs=8" in r.stdout
    want = [f"{i} INF" if d is None else f"{i} {d:.9g}" for i, d in enumerate(ka["expected"])]
    assert out.read_text().splitlines() == want

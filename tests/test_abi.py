"""The C-ABI libraries load and export every function their headers declare
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

import mi355sssp as M

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^[A-Za-z_][\w\s\*]*?\b(mi355g?_\w+)\s*\(", re.M)


def _declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(DECL.findall(text))
    # typedef'd function-pointer types are not exported functions
    return {n for n in names if not n.endswith("_cb")}


@pytest.mark.parametrize("header,loader", [("mi355sssp.h", M.engine), ("mi355graph.h", M.graphlib)])
def test_every_declared_symbol_is_exported(header, loader):
    lib = loader()
    names = _declared(header)
    assert len(names) > 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported: {missing}"


def test_engine_reports_errors_without_a_device():
    lib = M.engine()
    # argument validation happens on the host before any HIP call
    assert lib.mi355_graph_create(None, 0, None, None) == -1
    assert b"null" in lib.mi355_last_error()


def test_phase_schedules_are_checked_on_the_host():
    """mi355_adds_schedule_set validates every row before it is kept."""
    from mi355sssp import learn as L

    for bad in ([(0.0, 1, 0.5)], [(1.0, 9, 0.5)], [(1.0, 1, 0.0)], [(1.0, 1, 1.5)], [(float("nan"), 1, 1)],
                [(1.0, 1, 1.0)] * 65):
        with pytest.raises(M.Mi355Error):
            L.set_schedule(bad)
    L.set_schedule([(0.25, 1, 0.5), (2.0, 4, 0.25)])
    L.set_schedule(None)


def test_compat_headers_compile_against_the_abi(tmp_path):
    """The C++ compatibility headers restating the reference signatures must
    compile (syntax only) with hipcc for gfx950."""
    compat = os.path.join(REPO, "include", "compat")
    headers = sorted(h for h in os.listdir(compat) if h.endswith(".hpp"))
    if not headers:
        pytest.skip("no compatibility headers")
    src = tmp_path / "all.cpp"
    src.write_text("".join(f'#include "{h}"\n' for h in headers) + "int main() { return 0; }\n")
    import subprocess

    r = subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-std=c++17", "-fsyntax-only",
                        "-I", os.path.join(REPO, "include"), "-I", compat, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


STRUCTS = [("CsrDesc", "mi355_csr_desc", "mi355sssp.h"), ("GraphInfo", "mi355_graph_info", "mi355sssp.h"),
           ("SsspOptions", "mi355_sssp_options", "mi355sssp.h"), ("SsspStats", "mi355_sssp_stats", "mi355sssp.h"),
           ("MlpParams", "mi355_mlp_params", "mi355sssp.h"),
           ("MlpTrainOptions", "mi355_mlp_train_options", "mi355sssp.h"),
           ("QueryBatch", "mi355_query_batch", "mi355sssp.h"), ("BatchOptions", "mi355_batch_options", "mi355sssp.h"),
           ("BatchStats", "mi355_batch_stats", "mi355sssp.h"), ("BfProgress", "mi355_bf_progress", "mi355sssp.h"),
           ("BfCsrStats", "mi355_bf_csr_stats", "mi355sssp.h"), ("OneOptions", "mi355_one_options", "mi355sssp.h"),
           ("OneResult", "mi355_one_result", "mi355sssp.h"), ("GView", "mi355g_view_t", "mi355graph.h"),
           ("IfmHeader", "mi355g_ifm_header_t", "mi355graph.h")]


def test_ctypes_structs_match_the_c_layout(tmp_path):
    """Every ctypes mirror has the C struct's size and, field by field (same
    names), the C offsets: compiled and printed by the host C compiler."""
    import ctypes as C
    from mi355sssp import _abi

    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "mi355sssp.h"', '#include "mi355graph.h"',
             "int main(void) {"]
    want = []
    for py, cname, _ in STRUCTS:
        cls = getattr(_abi, py)
        lines.append(f'  printf("%zu\\n", sizeof({cname}));')
        want.append(C.sizeof(cls))
        for f in cls._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({cname}, {f[0]}));')
            want.append(getattr(cls, f[0]).offset)
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == want


REF_KERNELS = "/root/reference/HIP_kernel"
REF_DRIVERS = ["bellman_ford/tests/test_correctness.cpp", "bellman_ford/tests/test_check_intervals.cpp",
               "bellman_ford/tests/big_test.cpp", "delta_stepping/testing/verify_bf_delta_vs_dijkstra.cpp",
               "delta_stepping/testing/speed_compare_delta_vs_bf.cpp"]


@pytest.mark.skipif(not os.path.isdir(REF_KERNELS), reason="reference sources are not on this machine")
@pytest.mark.parametrize("driver", REF_DRIVERS)
def test_reference_test_drivers_build_against_the_engine(tmp_path, driver):
    """The reference's own test drivers, unchanged, compile and link against
    the compatibility headers and libmi355sssp.so.  Each driver is linked (not
    copied) into a scratch tree whose `src/` holds the compatibility headers,
    so its `../src/*.hpp` includes resolve to them (the reference ships no
    `bf_hip_CSR_device_utils.hpp`; test_check_intervals.cpp needs it)."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    compat = os.path.join(root, "include", "compat")
    family = driver.split("/")[0]
    src = tmp_path / family / "src"
    src.mkdir(parents=True)
    for h in os.listdir(compat):
        (src / h).symlink_to(os.path.join(compat, h))
    link = tmp_path / driver
    link.parent.mkdir(parents=True, exist_ok=True)
    link.symlink_to(os.path.join(REF_KERNELS, driver))
    lib = os.path.join(root, "rips2026-amd_amd", "lib")
    r = subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-std=c++17", "-O0", "-I", compat,
                        "-I", os.path.join(root, "include"), str(link), "-L", lib, "-lmi355sssp",
                        "-o", str(tmp_path / "driver")], capture_output=True, text=True, timeout=600,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr[-2000:]

"""The C-ABI libraries load and export every function their headers declare
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

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

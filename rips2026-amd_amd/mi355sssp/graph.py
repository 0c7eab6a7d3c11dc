"""Host graphs: generators, file formats and CSR arrays (libmi355graph.so)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import GView, W_F32, W_U32, OUT, IN, gcheck, graphlib


class HostGraph:
    """Owning handle of a host CSR (int64 rowptr, int32 col, 32-bit weights)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        v = GView()
        gcheck(graphlib().mi355g_view(self._h, C.byref(v)))
        self.n = int(v.rows)
        self.cols = int(v.cols)
        self.m = int(v.nnz)
        self.weight_type = int(v.weight_type)
        self.orientation = int(v.orientation)
        self.declared_edges = int(v.declared_edges)
        self.loaded_edges = int(v.loaded_edges)
        # Views into the library-owned arrays (valid while self is alive).
        self.rowptr = np.ctypeslib.as_array(v.rowptr, shape=(self.n + 1,)) if self.n >= 0 else None
        if self.m > 0:
            self.col = np.ctypeslib.as_array(v.col, shape=(self.m,))
            self.wbits = np.ctypeslib.as_array(v.w, shape=(self.m,))
        else:
            self.col = np.zeros(0, np.int32)
            self.wbits = np.zeros(0, np.uint32)

    @property
    def weights(self) -> np.ndarray:
        return self.wbits.view(np.float32) if self.weight_type == W_F32 else self.wbits

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            graphlib().mi355g_free(h)
            self._h = None

    def transpose(self) -> "HostGraph":
        out = C.c_void_p()
        gcheck(graphlib().mi355g_transpose(self._h, C.byref(out)))
        return HostGraph(out)

    def sha256(self) -> str:
        buf = C.create_string_buffer(65)
        gcheck(graphlib().mi355g_sha256(self._h, buf))
        return buf.value.decode()

    def pick_sources(self, count: int, seed: int) -> np.ndarray:
        out = np.zeros(count, np.int32)
        gcheck(graphlib().mi355g_pick_sources(self._h, count, seed, out.ctypes.data))
        return out

    def write_ripscsr1(self, path: str) -> None:
        gcheck(graphlib().mi355g_write_ripscsr1(self._h, path.encode()))

    def write_galois_gr(self, path: str) -> None:
        gcheck(graphlib().mi355g_write_galois_gr(self._h, path.encode()))

    def out_degree(self) -> np.ndarray:
        return np.diff(self.rowptr)


def _make(fn, *args) -> HostGraph:
    out = C.c_void_p()
    gcheck(fn(*args, C.byref(out)))
    return HostGraph(out)


def grid4(rows: int, cols: int, wmin: int = 1, wmax: int = 100, seed: int = 1) -> HostGraph:
    return _make(graphlib().mi355g_gen_grid4, rows, cols, wmin, wmax, seed)


def rmat(scale: int, edgefactor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 1, permute: bool = True) -> HostGraph:
    return _make(graphlib().mi355g_gen_rmat, scale, edgefactor, a, b, c, seed, int(permute))


def routing_grid(side: int, p_delete: float = 0.10, p_diag: float = 0.02, mu: float = 3.0,
                 sigma: float = 1.0, seed: int = 1) -> HostGraph:
    return _make(graphlib().mi355g_gen_routing_grid, side, p_delete, p_diag, mu, sigma, seed)


def chain_random(n: int, avg_deg: int, wmin: float, wmax: float, weight_type: int = W_F32,
                 seed: int = 1) -> HostGraph:
    return _make(graphlib().mi355g_gen_chain_random, n, avg_deg, wmin, wmax, weight_type, seed)


def _wbits(w: np.ndarray, weight_type: int) -> np.ndarray:
    if weight_type == W_F32:
        return np.ascontiguousarray(w, dtype=np.float32).view(np.uint32)
    return np.ascontiguousarray(w, dtype=np.uint32)


def from_edges(n: int, src, dst, w, weight_type: int = W_F32, orientation: int = OUT,
               dedup_min: bool = False) -> HostGraph:
    s = np.ascontiguousarray(src, dtype=np.int32)
    d = np.ascontiguousarray(dst, dtype=np.int32)
    wb = _wbits(np.asarray(w), weight_type)
    return _make(graphlib().mi355g_from_edges, n, len(s), s.ctypes.data, d.ctypes.data,
                 wb.ctypes.data, weight_type, orientation, int(dedup_min))


def from_csr(rowptr, col, w, weight_type: int = W_F32, orientation: int = OUT,
             cols: int | None = None) -> HostGraph:
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    cl = np.ascontiguousarray(col, dtype=np.int32)
    wb = _wbits(np.asarray(w), weight_type)
    rows = len(rp) - 1
    return _make(graphlib().mi355g_from_csr, rows, rows if cols is None else cols, len(cl),
                 rp.ctypes.data, cl.ctypes.data, wb.ctypes.data, weight_type, orientation)


def from_dense(adj, n: int, orientation: int = OUT, skip_diagonal: bool = True,
               reject_negative: bool = True) -> HostGraph:
    a = np.ascontiguousarray(adj, dtype=np.float32)
    return _make(graphlib().mi355g_from_dense, a.ctypes.data, n, orientation, int(skip_diagonal),
                 int(reject_negative))


def write_ripsadj1(g: HostGraph, path: str) -> None:
    gcheck(graphlib().mi355g_write_ripsadj1(g._h, path.encode()))


def read_ripsadj1(path: str) -> np.ndarray:
    """Returns the dense incoming matrix as an (n, n) float32 array."""
    n = C.c_int64()
    gcheck(graphlib().mi355g_read_ripsadj1(path.encode(), None, 0, C.byref(n)))
    a = np.empty(n.value * n.value, np.float32)
    gcheck(graphlib().mi355g_read_ripsadj1(path.encode(), a.ctypes.data, a.size, C.byref(n)))
    return a.reshape(n.value, n.value)


def read_ripscsr1(path: str) -> HostGraph:
    return _make(graphlib().mi355g_read_ripscsr1, path.encode())


def read_dimacs(path: str) -> HostGraph:
    return _make(graphlib().mi355g_read_dimacs, path.encode())


def read_galois_gr(path: str) -> HostGraph:
    return _make(graphlib().mi355g_read_galois_gr, path.encode())


def rng_u64(seed: int, stream: int, index: int) -> int:
    return int(graphlib().mi355g_rng_u64(seed, stream, index))


def fill_unit_f32(seed: int, stream: int, first: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float32)
    graphlib().mi355g_fill_unit_f32(seed, stream, first, count, out.ctypes.data)
    return out

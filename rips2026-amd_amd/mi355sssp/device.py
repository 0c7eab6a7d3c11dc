"""Device-side API: graph handles, SSSP runs and the reference entry points."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import (ALG_DELTA, BfCsrStats, BfProgress, CsrDesc, DEVICE, GraphInfo, HOST, IN,
                   INF_IEEE, ORDER_AUTO, OUT, PROGRESS_CB, SsspOptions, SsspStats, W_F32, W_U32,
                   check, engine)
from .graph import HostGraph


class DeviceBuffer:
    """Raw device allocation made through the engine's allocator."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(engine().mi355_device_malloc(C.byref(self.ptr), max(self.nbytes, 4)))

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            engine().mi355_device_free(self.ptr)
            self.ptr = C.c_void_p()

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(engine().mi355_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes, None))
        return b

    def to_array(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        if out.nbytes:
            check(engine().mi355_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, None))
        return out


def desc_from_arrays(rowptr: np.ndarray, col: np.ndarray, wbits: np.ndarray, weight_type: int,
                     orientation: int, cols: int | None = None) -> CsrDesc:
    rows = len(rowptr) - 1
    d = CsrDesc()
    d.rows = rows
    d.cols = rows if cols is None else cols
    d.nnz = len(col)
    d.rowptr = rowptr.ctypes.data
    d.col = col.ctypes.data if len(col) else None
    d.w = wbits.ctypes.data if len(col) else None
    d.offset_bits = 64 if rowptr.dtype == np.int64 else 32
    d.weight_type = weight_type
    d.orientation = orientation
    d.memory = HOST
    return d


class DeviceGraph:
    """A device-resident graph handle (mi355_graph)."""

    def __init__(self, g: HostGraph, allow_negative: bool = False, order: int = ORDER_AUTO):
        """order: ORDER_AUTO (degree order for skewed graphs), ORDER_INPUT or
        ORDER_DEGREE (mi355_graph_create_ex)."""
        self.host = g
        d = desc_from_arrays(g.rowptr, g.col, g.wbits, g.weight_type, g.orientation)
        self._h = C.c_void_p()
        check(engine().mi355_graph_create_ex(C.byref(d), int(allow_negative), int(order), None,
                                             C.byref(self._h)))
        info = GraphInfo()
        check(engine().mi355_graph_info_get(self._h, C.byref(info)))
        self.info = info
        self.n = int(info.n)
        self.m = int(info.m)
        self.weight_type = int(info.weight_type)
        self.dist_buf = DeviceBuffer(4 * self.n)
        self.pred_buf = DeviceBuffer(4 * self.n)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            engine().mi355_graph_destroy(self._h)
            self._h = C.c_void_p()

    @property
    def handle(self):
        return self._h

    def options(self, algorithm: int = ALG_DELTA, delta: float = 0.0, max_iters: int = -1,
                **kw) -> SsspOptions:
        o = SsspOptions()
        engine().mi355_sssp_options_default(C.byref(o))
        o.algorithm = algorithm
        o.delta = delta
        o.max_iters = max_iters
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def sssp(self, source: int, algorithm: int = ALG_DELTA, delta: float = 0.0,
             max_iters: int = -1, want_pred: bool = True, **kw):
        """Returns (dist, pred, stats) as numpy arrays / SsspStats."""
        o = self.options(algorithm, delta, max_iters, **kw)
        st = SsspStats()
        check(engine().mi355_sssp(self._h, source, C.byref(o), self.dist_buf.ptr,
                                  self.pred_buf.ptr if want_pred else None, C.byref(st), None))
        dt = np.float32 if self.weight_type == W_F32 else np.uint32
        dist = self.dist_buf.to_array(dt, self.n)
        pred = self.pred_buf.to_array(np.int32, self.n) if want_pred else None
        return dist, pred, st

    def sssp_device(self, source: int, opts: SsspOptions, stats: bool = True):
        """One full SSSP leaving distances and parents on the device (in this
        graph's dist_buf / pred_buf); returns the stats (None without)."""
        st = SsspStats() if stats else None
        check(engine().mi355_sssp(self._h, source, C.byref(opts), self.dist_buf.ptr, self.pred_buf.ptr,
                                  C.byref(st) if st is not None else None, None))
        return st

    def round_log(self, cap: int) -> np.ndarray:
        """Array that the next BF / BF_JACOBI run on this graph fills with its
        per-round frontier sizes |F_0| .. |F_rounds| (at most cap entries; the
        rest stay -1).  Keep it alive until that run returns."""
        log = np.full(cap, -1, np.int32)
        check(engine().mi355_sssp_round_log(self._h, log.ctypes.data, int(cap)))
        return log

    def sssp_labels(self, source: int, opts: SsspOptions) -> SsspStats:
        st = SsspStats()
        check(engine().mi355_sssp_labels(self._h, source, C.byref(opts), C.byref(st), None))
        return st

    def download(self):
        rp = np.empty(self.n + 1, np.int64)
        col = np.empty(self.m, np.int32)
        w = np.empty(self.m, np.uint32)
        check(engine().mi355_graph_download(self._h, rp.ctypes.data, col.ctypes.data,
                                            w.ctypes.data))
        return rp, col, w


def _as_in_desc(g: HostGraph):
    """Incoming-orientation descriptor of a host graph (transposing if needed)."""
    if g.orientation != IN:
        g = g.transpose()
    return g, desc_from_arrays(g.rowptr, g.col, g.wbits, g.weight_type, IN)


def bf_csr(g: HostGraph, source: int, max_iters: int = -1, check_interval: int = 1,
           progress=None):
    """mi355_bf_csr_f32 on an incoming CSR: returns (dist, stats, progress list)."""
    keep, d = _as_in_desc(g)
    dist = np.empty(keep.n, np.float32)
    st = BfCsrStats()
    log = []

    def _cb(p, _user):
        log.append((p.contents.iteration, p.contents.max_iters, p.contents.convergence_checked,
                    p.contents.changed))

    cb = PROGRESS_CB(_cb)
    check(engine().mi355_bf_csr_f32(C.byref(d), source, max_iters, check_interval,
                                    dist.ctypes.data, C.byref(st), cb, None, None))
    return dist, st, log


def delta_stepping_csr(g: HostGraph, source: int, delta: float, max_iters: int = -1):
    """mi355_delta_stepping_csr_f32: returns (dist, iterations_used, converged, progress)."""
    keep, d = _as_in_desc(g)
    dist = np.empty(keep.n, np.float32)
    it = C.c_int()
    conv = C.c_int()
    log = []

    def _cb(p, _user):
        log.append((p.contents.iteration, p.contents.max_iters, p.contents.convergence_checked,
                    p.contents.changed))

    cb = PROGRESS_CB(_cb)
    check(engine().mi355_delta_stepping_csr_f32(C.byref(d), source, delta, max_iters,
                                                dist.ctypes.data, C.byref(it), C.byref(conv), cb,
                                                None, None))
    return dist, it.value, bool(conv.value), log

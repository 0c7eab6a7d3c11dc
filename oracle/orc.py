"""ctypes wrapper of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker / baseline, never as the
thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run `make -C oracle`")
        _lib = C.CDLL(path)
        P = C.c_void_p
        I64 = C.c_int64
        for name, res, args in [
            ("orc_dijkstra_f32", C.c_int, [I64, P, P, P, I64, P]),
            ("orc_dijkstra_u32", C.c_int, [I64, P, P, P, I64, P]),
            ("orc_canonical_pred_f32", None, [I64, P, P, P, I64, P, P]),
            ("orc_canonical_pred_u32", None, [I64, P, P, P, I64, P, P]),
            ("orc_check_pred_f32", I64, [I64, P, P, P, I64, P, P]),
            ("orc_check_pred_u32", I64, [I64, P, P, P, I64, P, P]),
            ("orc_bf_jacobi_in_f32", C.c_int, [I64, P, P, P, I64, C.c_int, C.c_int, P, P, P]),
            ("orc_bf_dense_f32", C.c_int, [P, C.c_int, C.c_int, C.c_int, P, P, P]),
            ("orc_bf_gauss_seidel_f32", C.c_int, [P, C.c_int, C.c_int, C.c_int, P, P]),
            ("orc_ds_bucket_count_f32", I64, [I64, P, C.c_float]),
            ("orc_minplus_gemm_f32", None, [P, P, P] + [C.c_int] * 6),
            ("orc_minplus_spgemm_f32", C.c_int, [I64, I64, P, P, P, I64, P, P, P, P, P, P, P]),
            ("orc_transpose", None, [I64, I64, P, P, P, P, P, P]),
            ("orc_dijkstra_box_f32", C.c_int, [I64, P, P, P] + [C.c_int32] * 5 + [P, C.c_int32, P]),
            ("orc_dijkstra_box_u32", C.c_int, [I64, P, P, P] + [C.c_int32] * 5 + [P, C.c_int32, P]),
            ("orc_canonical_pred_box", None, [I64, P, P, P, C.c_int32] + [C.c_int32] * 5 +
             [P, C.c_int32, P, P]),
        ]:
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def dijkstra(rowptr, col, w, src: int):
    """Outgoing CSR; float32 weights -> fp32 Dijkstra, uint32 -> saturating u32."""
    rp, cl = _c(rowptr, np.int64), _c(col, np.int32)
    n = len(rp) - 1
    if np.asarray(w).dtype == np.uint32:
        ww = _c(w, np.uint32)
        dist = np.empty(n, np.uint32)
        rc = lib().orc_dijkstra_u32(n, rp.ctypes.data, cl.ctypes.data, ww.ctypes.data, src,
                                    dist.ctypes.data)
    else:
        ww = _c(w, np.float32)
        dist = np.empty(n, np.float32)
        rc = lib().orc_dijkstra_f32(n, rp.ctypes.data, cl.ctypes.data, ww.ctypes.data, src,
                                    dist.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"orc_dijkstra failed: {rc}")
    return dist


def canonical_pred(rowptr, col, w, src: int, dist):
    rp, cl = _c(rowptr, np.int64), _c(col, np.int32)
    n = len(rp) - 1
    pred = np.empty(n, np.int32)
    if np.asarray(w).dtype == np.uint32:
        ww, dd = _c(w, np.uint32), _c(dist, np.uint32)
        lib().orc_canonical_pred_u32(n, rp.ctypes.data, cl.ctypes.data, ww.ctypes.data, src,
                                     dd.ctypes.data, pred.ctypes.data)
    else:
        ww, dd = _c(w, np.float32), _c(dist, np.float32)
        lib().orc_canonical_pred_f32(n, rp.ctypes.data, cl.ctypes.data, ww.ctypes.data, src,
                                     dd.ctypes.data, pred.ctypes.data)
    return pred


def check_pred(rowptr, col, w, src: int, dist, pred) -> int:
    rp, cl = _c(rowptr, np.int64), _c(col, np.int32)
    n = len(rp) - 1
    pr = _c(pred, np.int32)
    if np.asarray(w).dtype == np.uint32:
        ww, dd = _c(w, np.uint32), _c(dist, np.uint32)
        return int(lib().orc_check_pred_u32(n, rp.ctypes.data, cl.ctypes.data, ww.ctypes.data,
                                            src, dd.ctypes.data, pr.ctypes.data))
    ww, dd = _c(w, np.float32), _c(dist, np.float32)
    return int(lib().orc_check_pred_f32(n, rp.ctypes.data, cl.ctypes.data, ww.ctypes.data, src,
                                        dd.ctypes.data, pr.ctypes.data))


def bf_jacobi_in(rowptr, col, w, src: int, max_iters: int = -1, check_interval: int = 1):
    """Jacobi BF over an INCOMING CSR -> (dist, iterations_used, converged)."""
    rp, cl, ww = _c(rowptr, np.int64), _c(col, np.int32), _c(w, np.float32)
    n = len(rp) - 1
    dist = np.empty(n, np.float32)
    it, cv = C.c_int(), C.c_int()
    rc = lib().orc_bf_jacobi_in_f32(n, rp.ctypes.data, cl.ctypes.data, ww.ctypes.data, src,
                                    max_iters, check_interval, dist.ctypes.data, C.byref(it),
                                    C.byref(cv))
    if rc != 0:
        raise RuntimeError(f"orc_bf_jacobi_in_f32 failed: {rc}")
    return dist, it.value, bool(cv.value)


def bf_dense(A, n: int, src: int, max_iters: int = -1):
    a = _c(A, np.float32)
    dist = np.empty(n, np.float32)
    it, cv = C.c_int(), C.c_int()
    rc = lib().orc_bf_dense_f32(a.ctypes.data, n, src, max_iters, dist.ctypes.data, C.byref(it),
                                C.byref(cv))
    if rc != 0:
        raise RuntimeError(f"orc_bf_dense_f32 failed: {rc}")
    return dist, it.value, bool(cv.value)


def gauss_seidel(A, n: int, src: int, max_iters: int = -1):
    a = _c(A, np.float32)
    dist = np.empty(n, np.float32)
    it = C.c_int()
    rc = lib().orc_bf_gauss_seidel_f32(a.ctypes.data, n, src, max_iters, dist.ctypes.data,
                                       C.byref(it))
    if rc != 0:
        raise RuntimeError(f"orc_bf_gauss_seidel_f32 failed: {rc}")
    return dist, it.value


def ds_bucket_count(dist, delta: float) -> int:
    d = _c(dist, np.float32)
    return int(lib().orc_ds_bucket_count_f32(len(d), d.ctypes.data, C.c_float(delta)))


def minplus_gemm(A, B, M, N, K, lda=None, ldb=None, ldc=None):
    lda = K if lda is None else lda
    ldb = N if ldb is None else ldb
    ldc = N if ldc is None else ldc
    a, b = _c(A, np.float32), _c(B, np.float32)
    c = np.empty(M * ldc, np.float32)
    lib().orc_minplus_gemm_f32(a.ctypes.data, b.ctypes.data, c.ctypes.data, M, N, K, lda, ldb, ldc)
    return c


def minplus_spgemm(a_rowptr, a_col, a_val, a_cols, b_rowptr, b_col, b_val, b_cols):
    """Returns (rowptr, col, val) with every output row sorted by column."""
    arp, acl, av = _c(a_rowptr, np.int64), _c(a_col, np.int32), _c(a_val, np.float32)
    brp, bcl, bv = _c(b_rowptr, np.int64), _c(b_col, np.int32), _c(b_val, np.float32)
    rows = len(arp) - 1
    out_rp = np.empty(rows + 1, np.int64)
    nnz = C.c_int64()
    args = (rows, a_cols, arp.ctypes.data, acl.ctypes.data, av.ctypes.data, b_cols,
            brp.ctypes.data, bcl.ctypes.data, bv.ctypes.data, out_rp.ctypes.data)
    rc = lib().orc_minplus_spgemm_f32(*args, None, None, C.byref(nnz))
    if rc != 0:
        raise RuntimeError("orc_minplus_spgemm_f32 failed")
    col = np.empty(max(nnz.value, 1), np.int32)
    val = np.empty(max(nnz.value, 1), np.float32)
    rc = lib().orc_minplus_spgemm_f32(*args, col.ctypes.data, val.ctypes.data, C.byref(nnz))
    if rc != 0:
        raise RuntimeError("orc_minplus_spgemm_f32 failed")
    return out_rp, col[: nnz.value], val[: nnz.value]


def transpose(rowptr, col, wbits, cols: int):
    rp, cl, ww = _c(rowptr, np.int64), _c(col, np.int32), _c(wbits, np.uint32)
    rows = len(rp) - 1
    trp = np.empty(cols + 1, np.int64)
    tcl = np.empty(len(cl), np.int32)
    tw = np.empty(len(cl), np.uint32)
    lib().orc_transpose(rows, cols, rp.ctypes.data, cl.ctypes.data, ww.ctypes.data,
                        trp.ctypes.data, tcl.ctypes.data, tw.ctypes.data)
    return trp, tcl, tw


def dijkstra_box(rowptr, col, w, grid_cols: int, box, sources):
    """Multi-source Dijkstra on the subgraph induced by box = (r0, c0, r1, c1)
    of a grid with grid_cols columns; dist in the box's row-major order."""
    rp, cl = _c(rowptr, np.int64), _c(col, np.int32)
    n = len(rp) - 1
    r0, c0, r1, c1 = (int(b) for b in box)
    src = _c(sources, np.int32)
    area = (r1 - r0) * (c1 - c0)
    if np.asarray(w).dtype == np.uint32:
        ww, dist, fn = _c(w, np.uint32), np.empty(area, np.uint32), lib().orc_dijkstra_box_u32
    else:
        ww, dist, fn = _c(w, np.float32), np.empty(area, np.float32), lib().orc_dijkstra_box_f32
    rc = fn(n, rp.ctypes.data, cl.ctypes.data, ww.ctypes.data, grid_cols, r0, c0, r1, c1,
            src.ctypes.data, len(src), dist.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"orc_dijkstra_box failed: {rc}")
    return dist


def canonical_pred_box(rowptr, col, w, grid_cols: int, box, sources, dist):
    rp, cl = _c(rowptr, np.int64), _c(col, np.int32)
    n = len(rp) - 1
    r0, c0, r1, c1 = (int(b) for b in box)
    src = _c(sources, np.int32)
    is_f32 = np.asarray(w).dtype != np.uint32
    wb = _c(np.asarray(w, np.float32 if is_f32 else np.uint32), np.float32 if is_f32 else np.uint32)
    db = _c(dist, np.float32 if is_f32 else np.uint32)
    pred = np.empty(len(db), np.int32)
    lib().orc_canonical_pred_box(n, rp.ctypes.data, cl.ctypes.data, wb.ctypes.data, int(is_f32),
                                 grid_cols, r0, c0, r1, c1, src.ctypes.data, len(src),
                                 db.ctypes.data, pred.ctypes.data)
    return pred

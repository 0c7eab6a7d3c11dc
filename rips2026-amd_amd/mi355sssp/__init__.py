"""Python binding of the MI355X SSSP engine (ctypes over the C ABI).

Two shared libraries are loaded from ``../lib`` next to this package:

* ``libmi355sssp.so`` -- the HIP engine (include/mi355sssp.h)
* ``libmi355graph.so`` -- host graph ingest and workloads (include/mi355graph.h)

Loading fails loudly when a library is missing: there is no CPU fallback for
any engine entry point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import (ALG_ADDS, ALG_BF, ALG_BF_JACOBI, ALG_DELTA, INF_FLT_MAX, INF_IEEE,
                   W_F32, W_U32, OUT, IN, HOST, DEVICE, ORDER_AUTO, ORDER_INPUT, ORDER_DEGREE, SsspOptions, SsspStats, CsrDesc,
                   BfCsrStats, BfProgress, PROGRESS_CB, Mi355Error, engine, graphlib)
from .graph import HostGraph, grid4, rmat, routing_grid, chain_random, from_edges, \
    from_csr, from_dense, read_ripscsr1, write_ripsadj1, read_ripsadj1, read_dimacs, \
    read_galois_gr, rng_u64, fill_unit_f32
from .device import DeviceGraph, DeviceBuffer, bf_csr, delta_stepping_csr

__all__ = [
    "ALG_ADDS", "ALG_BF", "ALG_BF_JACOBI", "ALG_DELTA", "INF_FLT_MAX", "INF_IEEE", "W_F32",
    "W_U32", "OUT", "IN", "ORDER_AUTO", "ORDER_INPUT", "ORDER_DEGREE", "HOST", "DEVICE", "SsspOptions", "SsspStats", "CsrDesc", "BfCsrStats",
    "BfProgress", "PROGRESS_CB", "Mi355Error", "engine", "graphlib", "HostGraph", "grid4", "rmat",
    "routing_grid", "chain_random", "from_edges", "from_csr", "from_dense", "read_ripscsr1",
    "write_ripsadj1", "read_ripsadj1", "read_dimacs",
    "read_galois_gr", "rng_u64", "fill_unit_f32", "DeviceGraph", "DeviceBuffer", "bf_csr",
    "delta_stepping_csr",
]

from .kernels import (minplus_gemm, minplus_gemm_batched, minplus_autotune, minplus_tuning,  # noqa: E402
                      spgemm, bf_dense, gauss_seidel, one_sssp, hip_adds)
from ._abi import OneOptions, OneResult, AddsOptions, AddsStats  # noqa: E402
from . import batch, learn, routing  # noqa: E402,F401

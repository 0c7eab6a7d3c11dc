"""Route requests of an FPGA-interchange design on the engine.

The interchange converter (Routing/interchange_to_csr.cpp) writes the device
routing graph as RIPSCSR1 (incoming orientation: row v lists the edges u -> v,
columns ascending) and a RIPSIFM1 sidecar whose route requests name, for every
net, the graph nodes of its source and sink site pins
(Routing/interchange_to_csr.cpp:1200-1240, 1475-1496).  route() solves every
net as a multi-source shortest-path query: the nets go to the batched engine
in chunks, each query spanning the whole graph (a graph is the 1 x n grid whose
only rectangle is every node), then each sink's path is read back through the
canonical parents and every path edge is mapped to its PIP through the
sidecar's per-edge records (tile, pip) and pip table (wire0, wire1, forward),
which are aligned with the RIPSCSR1 edge order (:1470-1474).

This is one shortest-path pass per net (no congestion negotiation), the
consumer the converter's comments describe: "run shortest paths over the CSR
using these node IDs, then recover PIPs through edge_attrs and pip_data".
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._abi import IFM_HEADER_FIELDS, IfmHeader, gcheck, graphlib
from . import batch as B

NO_INDEX = (1 << 64) - 1


class RouteMetadata:
    """A RIPSIFM1 sidecar read by the host library (mi355g_read_ripsifm1)."""

    def __init__(self, path: str):
        h = C.c_void_p()
        gcheck(graphlib().mi355g_read_ripsifm1(str(path).encode(), C.byref(h)))
        self._h = h
        hdr = IfmHeader()
        gcheck(graphlib().mi355g_ifm_header(self._h, C.byref(hdr)))
        self.header = {f: int(getattr(hdr, f)) for f in IFM_HEADER_FIELDS}
        H = self.header

        def arr(n):
            return np.zeros(max(int(n), 0), np.uint64)

        self.node_device_ids = arr(H["node_count"])
        edge_attrs, pips, pins = arr(2 * H["edge_attr_count"]), arr(3 * H["pip_data_count"]), \
            arr(3 * H["site_pin_attr_count"])
        self.blocked_nodes, self.sink_stop_nodes = arr(H["blocked_node_count"]), arr(H["sink_stop_node_count"])
        gcheck(graphlib().mi355g_ifm_tables(self._h, self.node_device_ids.ctypes.data, edge_attrs.ctypes.data,
                                            pips.ctypes.data, pins.ctypes.data, self.blocked_nodes.ctypes.data,
                                            self.sink_stop_nodes.ctypes.data))
        self.edge_tile, self.edge_pip = edge_attrs.reshape(-1, 2).T.copy()
        self.pip_wire0, self.pip_wire1, self.pip_forward = pips.reshape(-1, 3).T.copy()
        self.site_pins = pins.reshape(-1, 3)
        nq = H["route_request_count"]
        nets, so, src, ko, snk = arr(2 * nq), arr(nq + 1), arr(3 * H["route_source_count"]), arr(nq + 1), \
            arr(3 * H["route_sink_count"])
        gcheck(graphlib().mi355g_ifm_requests(self._h, nets.ctypes.data, so.ctypes.data, src.ctypes.data,
                                              ko.ctypes.data, snk.ctypes.data))
        self.net_string, self.net_logical = nets.reshape(-1, 2).T.copy()
        self.src_offsets, self.sources = so.astype(np.int64), src.reshape(-1, 3)
        self.sink_offsets, self.sinks = ko.astype(np.int64), snk.reshape(-1, 3)
        cells, lnets, ports = arr(3 * H["logical_cell_count"]), arr(4 * H["logical_net_count"]), \
            arr(7 * H["logical_port_instance_count"])
        gcheck(graphlib().mi355g_ifm_logical(self._h, cells.ctypes.data, lnets.ctypes.data, ports.ctypes.data))
        self.logical_cells, self.logical_nets = cells.reshape(-1, 3), lnets.reshape(-1, 4)
        self.logical_ports = ports.reshape(-1, 7)

    def __del__(self):
        if getattr(self, "_h", None):
            graphlib().mi355g_ifm_free(self._h)
            self._h = None

    def string(self, index: int) -> str:
        if int(index) == NO_INDEX:
            return ""
        p, n = C.c_void_p(), C.c_uint64()
        gcheck(graphlib().mi355g_ifm_string(self._h, int(index), C.byref(p), C.byref(n)))
        return C.string_at(p, n.value).decode("utf-8", "replace") if n.value else ""

    def payloads(self) -> tuple[bytes, bytes]:
        """The raw physical and logical netlist bytes."""
        a, b = C.c_void_p(), C.c_void_p()
        gcheck(graphlib().mi355g_ifm_payloads(self._h, C.byref(a), C.byref(b)))
        pa, pb = self.header["physical_netlist_byte_count"], self.header["logical_netlist_byte_count"]
        return (C.string_at(a, pa) if pa else b""), (C.string_at(b, pb) if pb else b"")

    @property
    def request_count(self) -> int:
        return self.header["route_request_count"]

    def request_sources(self, q: int) -> np.ndarray:
        return self.sources[self.src_offsets[q]:self.src_offsets[q + 1], 0].astype(np.int64)

    def request_sinks(self, q: int) -> np.ndarray:
        return self.sinks[self.sink_offsets[q]:self.sink_offsets[q + 1]]


@dataclass
class RoutedSink:
    node: int
    site: str
    pin: str
    dist: float                 # +inf when no source reaches the sink
    nodes: np.ndarray           # path, source first (empty when unreached)
    edges: np.ndarray           # RIPSCSR1 edge index of each hop
    pips: list = field(default_factory=list)  # (tile, wire0, wire1, forward) per hop


@dataclass
class RoutedNet:
    request: int
    name: str
    logical_net: int            # NO_INDEX when the net has no logical match
    sources: np.ndarray
    sinks: list


def _edge_index(rowptr: np.ndarray, col: np.ndarray):
    """(v, u) -> RIPSCSR1 edge index lookup for edges u -> v (row v): sorted
    keys v * n + u and the edge index of each key."""
    n = len(rowptr) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    keys = rows * n + col.astype(np.int64)
    order = np.argsort(keys, kind="stable")
    return keys[order], order


def route(graph, meta: RouteMetadata, dg=None, max_label_slots: int = 1 << 30,
          resolve_names: bool = True, early_stop: bool = True, **batch_options) -> list[RoutedNet]:
    """Route every request of `meta` on `graph` (the RIPSCSR1 host graph the
    sidecar belongs to).  dg: an existing DeviceGraph of graph (built when
    None).  Nets go to the batched engine in chunks whose label storage
    (B.storage_slots: n per net) stays within max_label_slots.  The host side walks the parents of
    all sinks of a chunk together and looks every hop's edge up in one
    sorted search; names are resolved once per distinct string.  early_stop:
    each net's query stops once all of its sinks are settled (the sinks' paths
    only read labels below theirs, which are final by then)."""
    from .device import DeviceGraph

    n = int(graph.n)
    H = meta.header
    if H["node_count"] != n:
        raise ValueError(f"sidecar has {H['node_count']} nodes, graph {n}")
    if H["edge_attr_count"] != graph.m:
        raise ValueError(f"sidecar has {H['edge_attr_count']} edge records, graph {graph.m} edges")
    # the sidecar's edge records follow the file's (incoming) edge order
    if graph.orientation != 1:
        raise ValueError("route() needs the RIPSCSR1 graph in incoming orientation")
    if dg is None:
        dg = DeviceGraph(graph)
    keys, kidx = _edge_index(graph.rowptr, graph.col)
    names: dict[int, str] = {}

    def name(i) -> str:
        i = int(i)
        if i not in names:
            names[i] = meta.string(i)
        return names[i]

    # every net is a 1 x n rectangle, stored row-major (n slots, no padding)
    per_chunk = max(1, min(65535, max_label_slots // max(B.block_slots(1, n), 1)))
    routed: list[RoutedNet] = []
    todo = [q for q in range(meta.request_count) if len(meta.request_sources(q)) > 0]
    for c0 in range(0, len(todo), per_chunk):
        chunk = todo[c0:c0 + per_chunk]
        srcs = [np.unique(meta.request_sources(q)).astype(np.int32) for q in chunk]
        off = np.zeros(len(chunk) + 1, np.int32)
        off[1:] = np.cumsum([len(s) for s in srcs])
        boxes = np.tile(np.array([0, 0, 1, n], np.int32), (len(chunk), 1))
        qs = B.Queries(n, off, np.concatenate(srcs), boxes)
        # every sink of the chunk: its query's block and node
        sinks = [meta.request_sinks(q) for q in chunk]
        counts = np.array([len(s) for s in sinks], np.int64)
        rec = np.concatenate(sinks) if counts.sum() else np.zeros((0, 3), np.uint64)
        node = rec[:, 0].astype(np.int64)
        tgt = None
        if early_stop:
            t_off = np.zeros(len(chunk) + 1, np.int32)
            t_off[1:] = np.cumsum(counts)
            tgt = (t_off, node.astype(np.int32))
        d_dist, d_pred, offs, _ = B.run(dg, qs, want_pred=True, targets=tgt, device_out=True, **batch_options)
        # every sink's path, read on the device (source first; empty when unreached)
        qidx = np.repeat(np.arange(len(chunk), dtype=np.int32), counts)
        poff, tdist, wn = B.paths(dg, qs, offs, d_dist, d_pred, qidx, node.astype(np.int32))
        del d_dist, d_pred
        wn = wn.astype(np.int64)
        plen = np.diff(poff)
        d = tdist.astype(np.float64)
        d[plen == 0] = np.inf
        ws = np.repeat(np.arange(len(node), dtype=np.int64), plen)
        bounds = poff
        # every hop u -> v (row v, column u) in one sorted search
        same = np.zeros(len(ws), bool)
        if len(ws) > 1:
            same[1:] = ws[1:] == ws[:-1]
        hop_v, hop_u = wn[same], wn[np.flatnonzero(same) - 1]
        want = hop_v * n + hop_u
        pos = np.searchsorted(keys, want)
        if np.any(pos >= len(keys)) or np.any(keys[np.minimum(pos, len(keys) - 1)] != want):
            bad = int(ws[same][np.flatnonzero((pos >= len(keys)) | (keys[np.minimum(pos, len(keys) - 1)] != want))[0]])
            q = chunk[int(np.searchsorted(np.cumsum(counts), bad, side="right"))]
            raise RuntimeError(f"net {q}: path edge missing from the graph")
        hop_e = kidx[pos].astype(np.int64)
        hop_bounds = np.searchsorted(ws[same], np.arange(len(node) + 1))
        k = 0
        for i, q in enumerate(chunk):
            out_sinks = []
            for j in range(int(counts[i])):
                _, site, pin = (int(x) for x in rec[k])
                path = wn[bounds[k]:bounds[k + 1]].astype(np.int64)
                eidx = hop_e[hop_bounds[k]:hop_bounds[k + 1]]
                pips = []
                if resolve_names:
                    for e in eidx:
                        pd = int(meta.edge_pip[e])
                        pips.append((name(meta.edge_tile[e]), name(meta.pip_wire0[pd]), name(meta.pip_wire1[pd]),
                                     bool(meta.pip_forward[pd])))
                out_sinks.append(RoutedSink(int(node[k]), name(site) if resolve_names else "",
                                            name(pin) if resolve_names else "", float(d[k]), path, eidx, pips))
                k += 1
            routed.append(RoutedNet(q, name(meta.net_string[q]) if resolve_names else "",
                                    int(meta.net_logical[q]), srcs[i], out_sinks))
    return routed

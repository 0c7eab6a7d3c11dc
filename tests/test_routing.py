"""RIPSIFM1 route-request sidecar (reader) and the route consumer.

No sidecar file ships with the reference (its converter needs Cap'n Proto and
the FPGA-interchange schemas), so the fixtures are packed here byte by byte
from the writer's layout, Routing/interchange_to_csr.cpp:1393-1411 (header
and section order) and :1451-1587 (records): every field a little-endian
u64, strings length-prefixed, arrays raw after the header's counts."""
import struct

import numpy as np
import pytest

import mi355sssp as M
from mi355sssp import routing as R

NO = (1 << 64) - 1


def pack_sidecar(strings, node_ids, edge_attrs, pips, site_pins, requests, cells, nets, ports,
                 blocked, sink_stop, phys, logi, paths=(0, 1, 2, 3), version=2, orientation=1):
    """requests: (net string, logical net, [(node, site, pin)...] sources,
    [(node, site, pin)...] sinks)."""
    u = lambda *v: struct.pack(f"<{len(v)}Q", *v)  # noqa: E731
    out = [b"RIPSIFM1", u(version, orientation, len(strings), len(node_ids), len(edge_attrs), len(pips),
                         len(site_pins), len(requests), len(blocked), len(sink_stop), len(cells), len(nets),
                         len(ports), len(phys), len(logi), *paths)]
    for s in strings:
        b = s.encode()
        out.append(u(len(b)) + b)
    out += [u(*node_ids) if node_ids else b""]
    out += [u(*r) for r in edge_attrs] + [u(*r) for r in pips] + [u(*r) for r in site_pins]
    for net, lnet, srcs, sinks in requests:
        out.append(u(net, lnet, len(srcs)))
        out += [u(*s) for s in srcs]
        out.append(u(len(sinks)))
        out += [u(*s) for s in sinks]
    out += [u(*r) for r in cells] + [u(*r) for r in nets] + [u(*r) for r in ports]
    out += [u(*blocked) if blocked else b"", u(*sink_stop) if sink_stop else b"", bytes(phys), bytes(logi)]
    return b"".join(out)


def _small(tmp_path):
    strings = ["dev.device", "x.phys", "x.netlist", "top", "TILE_A", "TILE_B", "W0", "W1", "W2",
               "SITE0", "PIN_O", "PIN_I", "net_a", "net_b", "cellA", "port"]
    kw = dict(strings=strings, node_ids=[100, 101, 102, 103],
              edge_attrs=[(4, 0), (5, 1), (4, 1)], pips=[(6, 7, 1), (7, 8, 0)],
              site_pins=[(3, 9, 11)],
              requests=[(12, 0, [(0, 9, 10)], [(3, 9, 11), (2, 9, 11)]), (13, NO, [], [])],
              cells=[(14, 0, 1)], nets=[(12, 0, 0, 2)], ports=[(15, NO, 1, 2, 3, 1, 0), (15, 3, 0, 0, 0, 0, 1)],
              blocked=[1], sink_stop=[3], phys=b"\x01\x02\x03", logi=b"\x04\x05")
    path = tmp_path / "m.ifmeta.bin"
    path.write_bytes(pack_sidecar(**kw))
    return path, kw


def test_reader_returns_every_section(tmp_path):
    path, kw = _small(tmp_path)
    m = R.RouteMetadata(path)
    H = m.header
    assert (H["version"], H["orientation"], H["string_count"], H["node_count"]) == (2, 1, 16, 4)
    assert (H["route_request_count"], H["route_source_count"], H["route_sink_count"]) == (2, 1, 2)
    assert [m.string(i) for i in range(16)] == kw["strings"]
    assert m.string(NO) == ""
    assert (H["device_path_string"], H["logical_design_name_string"]) == (0, 3)
    assert m.node_device_ids.tolist() == kw["node_ids"]
    assert list(zip(m.edge_tile.tolist(), m.edge_pip.tolist())) == kw["edge_attrs"]
    assert list(zip(m.pip_wire0.tolist(), m.pip_wire1.tolist(), m.pip_forward.tolist())) == kw["pips"]
    assert [tuple(r) for r in m.site_pins.tolist()] == kw["site_pins"]
    assert m.net_string.tolist() == [12, 13] and m.net_logical.tolist() == [0, NO]
    assert m.request_sources(0).tolist() == [0] and m.request_sources(1).tolist() == []
    assert [tuple(r) for r in m.request_sinks(0).tolist()] == [(3, 9, 11), (2, 9, 11)]
    assert len(m.request_sinks(1)) == 0
    assert [tuple(r) for r in m.logical_cells.tolist()] == kw["cells"]
    assert [tuple(r) for r in m.logical_nets.tolist()] == kw["nets"]
    assert [tuple(r) for r in m.logical_ports.tolist()] == kw["ports"]
    assert m.blocked_nodes.tolist() == [1] and m.sink_stop_nodes.tolist() == [3]
    assert m.payloads() == (b"\x01\x02\x03", b"\x04\x05")


def test_reader_rejects_malformed_files(tmp_path):
    path, kw = _small(tmp_path)
    good = path.read_bytes()
    bad = {
        "magic": b"RIPSIFM2" + good[8:],
        "version": pack_sidecar(**kw, version=1),
        "orientation": pack_sidecar(**kw, orientation=0),
        "trailing": good + b"\x00",
        "pip index": pack_sidecar(**dict(kw, edge_attrs=[(4, 0), (5, 2), (4, 1)])),
        "string index": pack_sidecar(**dict(kw, pips=[(6, 99, 1), (7, 8, 0)])),
        "forward flag": pack_sidecar(**dict(kw, pips=[(6, 7, 2), (7, 8, 0)])),
        "sink node": pack_sidecar(**dict(kw, requests=[(12, 0, [(0, 9, 10)], [(4, 9, 11)])])),
        "logical net": pack_sidecar(**dict(kw, requests=[(12, 5, [(0, 9, 10)], [])])),
        "blocked node": pack_sidecar(**dict(kw, blocked=[7])),
        "header string": pack_sidecar(**kw, paths=(0, 1, 2, 16)),
    }
    # every proper prefix is truncated somewhere
    for cut in (4, 8, 60, 160, len(good) // 2, len(good) - 1):
        bad[f"cut {cut}"] = good[:cut]
    for name, data in bad.items():
        p = tmp_path / "bad.bin"
        p.write_bytes(data)
        with pytest.raises(M.Mi355Error) as e:
            R.RouteMetadata(p)
        assert e.value.code in (-4, -3), (name, str(e.value))  # MI355G_EFORMAT (or EIO)


def test_reader_refuses_huge_counts_before_allocating(tmp_path):
    path, kw = _small(tmp_path)
    data = bytearray(path.read_bytes())
    struct.pack_into("<Q", data, 8 + 8 * 3, 1 << 60)  # node_count
    p = tmp_path / "huge.bin"
    p.write_bytes(bytes(data))
    with pytest.raises(M.Mi355Error):
        R.RouteMetadata(p)


def device_fixture(tmp_path, n=1500, m=9000, nets=40, seed=3):
    """A synthetic device graph written as RIPSCSR1 (incoming, rows sorted by
    column, unit weights as the converter writes them) plus its sidecar: one
    (tile, pip) record per edge and nets of 1-3 sources and 1-5 sinks (some
    sinks isolated, so unreachable)."""
    rng = np.random.default_rng(seed)
    iso = n - 5  # the last five nodes have no incoming edges
    src = rng.integers(0, n, m)
    dst = rng.integers(0, iso, m)
    keep = src != dst
    g_out = M.from_edges(n, src[keep], dst[keep], np.ones(int(keep.sum()), np.float32), M.W_F32)
    g_in = g_out.transpose()
    csr_path = tmp_path / "dev.csrbin"
    g_in.write_ripscsr1(str(csr_path))
    g = M.read_ripscsr1(str(csr_path))
    strings = ["dev", "phys", "netlist", "top"] + [f"T{i}" for i in range(20)] + [f"W{i}" for i in range(30)] + \
        ["S0", "P0", "P1"] + [f"net{q}" for q in range(nets)]
    tiles, wires, s0 = range(4, 24), range(24, 54), 54
    npip = 200
    pips = [(int(rng.choice(wires)), int(rng.choice(wires)), int(rng.integers(0, 2))) for _ in range(npip)]
    edge_attrs = [(int(rng.choice(tiles)), int(rng.integers(0, npip))) for _ in range(g.m)]
    reqs = []
    for q in range(nets):
        ns, nk = int(rng.integers(1, 4)), int(rng.integers(1, 6))
        srcs = [(int(v), s0, s0 + 1) for v in rng.integers(0, iso, ns)]
        sinks = [(int(v), s0, s0 + 2) for v in rng.integers(0, n, nk)]
        if q % 7 == 0:
            sinks.append((n - 1, s0, s0 + 2))  # no incoming edges: unreachable
        reqs.append((57 + q, NO, srcs, sinks))
    side = pack_sidecar(strings=strings, node_ids=list(range(1000, 1000 + n)), edge_attrs=edge_attrs, pips=pips,
                        site_pins=[], requests=reqs, cells=[], nets=[], ports=[], blocked=[], sink_stop=[],
                        phys=b"", logi=b"")
    meta_path = tmp_path / "dev.csrbin.ifmeta.bin"
    meta_path.write_bytes(side)
    return g, R.RouteMetadata(meta_path)


@pytest.mark.gpu
def test_route_requests_against_the_oracle(tmp_path):
    from oracle import orc

    g, meta = device_fixture(tmp_path)
    out = g.transpose()  # outgoing copy for the oracle
    nets = R.route(g, meta)
    assert len(nets) == meta.request_count
    again = R.route(g, meta, max_label_slots=3 * g.n)  # 3 nets per chunk
    # nets run to the end (no per-net stop at the last sink) route the same
    full = R.route(g, meta, early_stop=False)
    for net, net3 in zip(nets, full):
        for sk, sk3 in zip(net.sinks, net3.sinks):
            assert np.float32(sk.dist).tobytes() == np.float32(sk3.dist).tobytes()
            assert np.array_equal(sk.nodes, sk3.nodes) and np.array_equal(sk.edges, sk3.edges)
    rowp, col = g.rowptr, g.col
    reached = unreached = 0
    for net, net2 in zip(nets, again):
        srcs = np.unique(meta.request_sources(net.request))
        want = orc.dijkstra_box(out.rowptr, out.col, out.weights, g.n, (0, 0, 1, g.n), srcs)
        wpred = orc.canonical_pred_box(out.rowptr, out.col, out.weights, g.n, (0, 0, 1, g.n), srcs, want)
        assert net.name == meta.string(meta.net_string[net.request])
        for sk, sk2 in zip(net.sinks, net2.sinks):
            assert np.float32(sk.dist).tobytes() == want[sk.node].tobytes()
            assert np.array_equal(sk.nodes, sk2.nodes) and sk.pips == sk2.pips
            if not np.isfinite(sk.dist):
                unreached += 1
                assert len(sk.nodes) == 0 and len(sk.edges) == 0
                continue
            reached += 1
            assert sk.nodes[0] in srcs and sk.nodes[-1] == sk.node
            # every hop follows the canonical parent, and its edge / PIP records
            for i, e in enumerate(sk.edges):
                u, v = int(sk.nodes[i]), int(sk.nodes[i + 1])
                assert wpred[v] == u
                assert rowp[v] <= e < rowp[v + 1] and col[e] == u
                pd = int(meta.edge_pip[e])
                assert sk.pips[i] == (meta.string(meta.edge_tile[e]), meta.string(meta.pip_wire0[pd]),
                                      meta.string(meta.pip_wire1[pd]), bool(meta.pip_forward[pd]))
            assert len(sk.edges) == sk.dist  # unit weights: hops = distance
    assert reached > 20 and unreached > 0, (reached, unreached)


def test_route_rejects_a_sidecar_of_another_graph(tmp_path):
    g, meta = device_fixture(tmp_path, n=300, m=1200, nets=3)
    small = M.from_edges(10, np.array([0]), np.array([1]), np.ones(1, np.float32), M.W_F32)
    with pytest.raises(ValueError):  # other node count
        R.route(small, meta, dg=object())
    with pytest.raises(ValueError):  # outgoing orientation: edge records would not line up
        R.route(g.transpose(), meta, dg=object())


def _host_paths(dg, qs, off, dist, pred, query, target):
    """A stand-in for B.paths over host arrays (1 x n rectangles): each
    target's parent chain, source first, empty when unreached."""
    poff, tdist, nodes = [0], [], []
    for q, t in zip(query, target):
        b = int(off[q])
        tdist.append(dist[b + t])
        if not np.isfinite(dist[b + t]):
            poff.append(poff[-1])
            continue
        chain = [int(t)]
        while pred[b + chain[-1]] >= 0:
            chain.append(int(pred[b + chain[-1]]))
        nodes.extend(chain[::-1])
        poff.append(len(nodes))
    return np.array(poff, np.int64), np.array(tdist, np.float32), np.array(nodes, np.int32)


def _oracle_run(out):
    """A stand-in for the batched engine in CPU tests: the oracle's
    multi-source Dijkstra and canonical parents for every query."""
    from oracle import orc

    def run(dg, qs, want_pred=True, download=True, **kw):
        dist, pred, off = [], [], [0]
        for q in range(qs.count):
            box = tuple(int(x) for x in qs.boxes[q])
            srcs = qs.query_sources(q)
            d = orc.dijkstra_box(out.rowptr, out.col, out.weights, qs.grid_cols, box, srcs)
            dist.append(d)
            pred.append(orc.canonical_pred_box(out.rowptr, out.col, out.weights, qs.grid_cols, box, srcs, d))
            off.append(off[-1] + len(d))
        return np.concatenate(dist), np.concatenate(pred), np.array(off), None
    return run


def test_route_paths_and_pips_on_the_cpu(tmp_path, monkeypatch):
    """route()'s host side (chunking, path walk through the parents, edge
    lookup, PIP records, names) with the oracle standing in for the device."""
    g, meta = device_fixture(tmp_path, n=600, m=3000, nets=25, seed=11)
    out = g.transpose()
    run = _oracle_run(out)
    chunk_slots, chunk_targets = [], []

    def recording_run(dg, qs, *a, **kw):
        chunk_slots.append(R.B.storage_slots(qs))
        chunk_targets.append(kw.get("targets"))
        return run(dg, qs, *a, **kw)

    monkeypatch.setattr(R.B, "run", recording_run)
    monkeypatch.setattr(R.B, "paths", _host_paths)
    nets = R.route(g, meta, dg=object())
    # the targets handed to the engine: every net's sinks, per query
    todo = [q for q in range(meta.request_count) if len(meta.request_sources(q))]
    t_counts = np.concatenate([np.diff(t[0]) for t in chunk_targets])
    assert np.array_equal(t_counts, [len(meta.request_sinks(q)) for q in todo])
    assert np.array_equal(np.concatenate([t[1] for t in chunk_targets]),
                          np.concatenate([meta.request_sinks(q)[:, 0] for q in todo]).astype(np.int32))
    chunk_targets.clear()
    R.route(g, meta, dg=object(), early_stop=False)
    assert all(t is None for t in chunk_targets)
    chunk_slots.clear()
    chunked = R.route(g, meta, dg=object(), max_label_slots=4 * g.n)
    # each chunk's label storage (tile padding included) stays within the budget
    assert len(chunk_slots) > 1 and max(chunk_slots) <= 4 * g.n
    bare = R.route(g, meta, dg=object(), resolve_names=False)
    assert [n.request for n in nets] == [q for q in range(meta.request_count) if len(meta.request_sources(q))]
    rowp, col = g.rowptr, g.col
    reached = unreached = 0
    for net, net2, net3 in zip(nets, chunked, bare):
        assert net.name == meta.string(meta.net_string[net.request]) and net3.name == ""
        assert np.array_equal(net.sources, np.unique(meta.request_sources(net.request)))
        assert [s.node for s in net.sinks] == [int(x) for x in meta.request_sinks(net.request)[:, 0]]
        for sk, sk2, sk3 in zip(net.sinks, net2.sinks, net3.sinks):
            assert np.array_equal(sk.nodes, sk2.nodes) and np.array_equal(sk.edges, sk3.edges)
            assert sk.pips == sk2.pips and sk3.pips == [] and sk3.site == ""
            if not np.isfinite(sk.dist):
                unreached += 1
                assert len(sk.nodes) == 0 and len(sk.edges) == 0
                continue
            reached += 1
            assert sk.nodes[0] in net.sources and sk.nodes[-1] == sk.node
            assert len(sk.edges) == len(sk.nodes) - 1 == sk.dist  # unit weights
            for i, e in enumerate(sk.edges):
                u, v = int(sk.nodes[i]), int(sk.nodes[i + 1])
                assert rowp[v] <= e < rowp[v + 1] and col[e] == u
                pd = int(meta.edge_pip[e])
                assert sk.pips[i] == (meta.string(meta.edge_tile[e]), meta.string(meta.pip_wire0[pd]),
                                      meta.string(meta.pip_wire1[pd]), bool(meta.pip_forward[pd]))
    assert reached > 20 and unreached > 0, (reached, unreached)

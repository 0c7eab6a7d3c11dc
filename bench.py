#!/usr/bin/env python3
"""Benchmark: single-source shortest paths on R-MAT scale 22 (one MI355X).

Workload (the reference's headline configuration that fits one GPU): R-MAT
scale 22, edge factor 16, (a,b,c,d) = (.57,.19,.19,.05), self loops and
duplicate pairs removed, symmetrised, fp32 weights in (0,1], 64 seeded
sources with out-degree > 0.  The graph is generated on the host, uploaded
once, and stays resident in HBM; a step is one full SSSP (label init
included) from the next source with plain ADDS (persistent worklist kernel,
fixed heuristic delta).  Delta-stepping and Bellman-Ford are timed beside it
on the same sources.  Distances of the first --check sources are checked bit
for bit against the CPU oracle outside the timed region.

Prints one JSON line (the bench contract) with roofline and cpu_baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "rips2026-amd_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithm_id(name: str) -> int:
    import mi355sssp as M

    return {"bf": M.ALG_BF, "bf_jacobi": M.ALG_BF_JACOBI, "delta": M.ALG_DELTA,
            "adds": M.ALG_ADDS}[name]


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def profile_traffic(path: str):
    """HBM bytes per launch of the persistent ADDS kernel on this workload,
    from the committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE per
    dispatch, scaled by the copy-kernel calibration it records); None if the
    summary is absent."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


TRAFFIC_FILE = os.path.join(REPO, "profiles", "k_adds_rmat_s22_traffic.md")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--sources", type=int, default=64)
    ap.add_argument("--algorithm", default="adds", choices=["bf", "bf_jacobi", "delta", "adds"])
    ap.add_argument("--delta", type=float, default=0.0, help="bucket width (0: heuristic)")
    ap.add_argument("--dynamic-delta", type=int, default=0,
                    help="ADDS: 0 = plain fixed-heuristic delta, 1 = feedback policy")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--check", type=int, default=4, help="sources checked against the oracle")
    ap.add_argument("--compare-steps", type=int, default=16,
                    help="SSSPs timed for each other algorithm (reported, not the value)")
    ap.add_argument("--cpu-sources", type=int, default=8, help="oracle Dijkstra runs timed")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="one device-to-device copy of this size (counter calibration runs)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: the SSSP engine runs on a single MI355X "
              "(multi-GPU sharding is described in DESIGN.md, not run)", file=sys.stderr)
        return 2

    import mi355sssp as M

    t0 = time.perf_counter()
    g = M.rmat(args.scale, args.edgefactor, seed=args.seed)
    log(f"R-MAT scale {args.scale}: n={g.n} m={g.m} generated in {time.perf_counter() - t0:.1f}s")
    sources = g.pick_sources(args.sources, seed=args.seed)
    t0 = time.perf_counter()
    dg = M.DeviceGraph(g)
    log(f"uploaded in {time.perf_counter() - t0:.1f}s; max degree {dg.info.max_degree}, "
        f"hub chunks {dg.info.large_chunks}")

    if args.calib_copy_mib > 0:
        nbytes = args.calib_copy_mib << 20
        a = M.DeviceBuffer(nbytes)
        b = M.DeviceBuffer(nbytes)
        M.engine().mi355_memset_d8(a.ptr, 1, nbytes, None)
        M._abi.check(M.engine().mi355_memcpy_d2d(b.ptr, a.ptr, nbytes, None))
        M.engine().mi355_stream_sync(None)
        log(f"calibration copy of {nbytes} bytes done")
        del a, b

    def opts_for(name):
        kw = {"dynamic_delta": args.dynamic_delta} if name == "adds" else {}
        return dg.options(algorithm_id(name), delta=args.delta, **kw), kw

    alg = algorithm_id(args.algorithm)
    opts, kw = opts_for(args.algorithm)

    # reached edges (m_r) and vertices (n_r) per source, for the metric
    mr = np.zeros(len(sources), np.int64)
    nr = np.zeros(len(sources), np.int64)
    for i, s in enumerate(sources):
        _, _, st = dg.sssp(int(s), alg, delta=args.delta, want_pred=False, **kw)
        mr[i], nr[i] = st.reached_edges, st.reached
        if i == 0:
            log(f"source {s}: iterations {st.iterations_used} rounds {st.rounds} "
                f"delta {st.final_delta:.4g} reached {st.reached} m_r {st.reached_edges}")

    # parity spot check (outside the timed region)
    if args.check > 0:
        from oracle import orc

        for s in sources[: args.check]:
            dist, pred, _ = dg.sssp(int(s), alg, delta=args.delta, **kw)
            want = orc.dijkstra(g.rowptr, g.col, g.weights, int(s))
            ok = dist.tobytes() == want.tobytes()
            bad = orc.check_pred(g.rowptr, g.col, g.weights, int(s), dist, pred)
            log(f"check source {s}: dist bit-exact={ok} pred violations={bad}")
            if not ok or bad:
                print("parity check failed", file=sys.stderr)
                return 1

    def timed(o, steps, warmup):
        for i in range(warmup):
            dg.sssp_labels(int(sources[i % len(sources)]), o)
        M.engine().mi355_stream_sync(None)
        dev_ms = kern_ms = 0.0
        t0 = time.perf_counter()
        for i in range(steps):
            st = dg.sssp_labels(int(sources[i % len(sources)]), o)
            dev_ms += st.device_ms
            kern_ms += st.kernel_ms
        M.engine().mi355_stream_sync(None)
        return time.perf_counter() - t0, dev_ms, kern_ms

    wall, dev_ms, kern_ms = timed(opts, args.steps, args.warmup)
    idx = np.arange(args.steps) % len(sources)
    edges = float(mr[idx].sum())
    bytes_alg = float((12 * mr[idx] + 20 * nr[idx] + 8 * g.n).sum())
    ms_per = wall * 1e3 / args.steps
    gteps = edges / wall / 1e9
    # dominant kernel: the persistent worklist kernel for ADDS (one launch per
    # SSSP); the whole HIP-event span of the run for the multi-launch ones
    kernel_ms = kern_ms if kern_ms > 0 else dev_ms
    achieved = bytes_alg / (kernel_ms * 1e-3) / 1e9
    log(f"{args.algorithm}: {ms_per:.3f} ms/SSSP wall, {dev_ms / args.steps:.3f} ms device, "
        f"{kernel_ms / args.steps:.3f} ms kernel, {gteps:.2f} GTEPS")

    others = {}
    if args.compare_steps > 0:
        cidx = np.arange(args.compare_steps) % len(sources)
        for name in ("adds", "delta", "bf"):
            if name == args.algorithm:
                continue
            o, _ = opts_for(name)
            w2, d2, _ = timed(o, args.compare_steps, 2)
            others[name] = {"ms_per_sssp": w2 * 1e3 / args.compare_steps,
                            "device_ms_per_sssp": d2 / args.compare_steps,
                            "gteps": float(mr[cidx].sum()) / w2 / 1e9}
            log(f"{name}: {others[name]['ms_per_sssp']:.3f} ms/SSSP, {others[name]['gteps']:.2f} GTEPS")

    # CPU baseline: the oracle's fp32 Dijkstra, single thread, same graph
    cpu = None
    if args.cpu_sources > 0:
        from oracle import orc

        cs = sources[: args.cpu_sources]
        t0 = time.perf_counter()
        for s in cs:
            orc.dijkstra(g.rowptr, g.col, g.weights, int(s))
        cwall = time.perf_counter() - t0
        cedges = float(mr[: len(cs)].sum())
        cpu = {"value": cedges / cwall / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": f"{len(cs)} sources of the same graph, {cwall / len(cs) * 1e3:.0f} ms/SSSP "
                         f"(oracle fp32 binary-heap Dijkstra, 1 thread, {cpu_model()}, "
                         f"{os.cpu_count()} host CPUs)",
               "ms_per_sssp": cwall / len(cs) * 1e3}

    traffic = profile_traffic(TRAFFIC_FILE) if args.algorithm == "adds" and args.scale == 22 else None
    out = {
        "metric": "SSSP GTEPS + ms/query (learned-delta & batched vs plain ADDS), bit-exact dist",
        "value": gteps,
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded R-MAT generator)",
        "config": {"workload": f"rmat_s{args.scale}_ef{args.edgefactor}_sym_f32_sssp",
                   "algorithm": args.algorithm,
                   "policy": ("feedback" if args.dynamic_delta else "plain fixed-heuristic delta")
                   if args.algorithm == "adds" else None,
                   "n": g.n, "m": g.m,
                   "sources": len(sources), "graph_sha256": g.sha256(),
                   "avg_reached_edges": float(mr.mean()),
                   "graph500_gteps": gteps / 2.0,
                   "device_ms_per_sssp": dev_ms / args.steps,
                   "kernel_ms_per_sssp": kernel_ms / args.steps,
                   "others": others},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

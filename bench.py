#!/usr/bin/env python3
"""Benchmark of the MI355X treedepth solver; prints ONE JSON line.

Default workload (the n = 30 subset-DP configuration): 8 graphs, 4 x G(30,0.2)
seeds 0-3 and 4 x 5x6 grids with 5 of their 49 edges dropped, seeds 0-3.  One
step = the full subset DP (all 2^30 subsets) plus the elimination-order walk
and tree replay for each of the 8 graphs.  value = subsets resolved per second
over the timed steps (wall clock; every call is synchronous).

roofline: the level kernel (DPLevelKernel) against HBM peak, with
achieved = algorithmic bytes / GPU time of the level launches (HIP events on
the library's stream), algorithmic bytes = 2^(n-1) (n+2) per graph (k one-byte
reads of level k-1 and one write per k-subset).

cpu_baseline: the oracle's flat-table DP (OpenMP over each level) timed on the
host on a bounded sample: the subgraph induced by the first m vertices of the
first workload graph, m chosen so the run takes about 10-20 s.

Other workloads: --workload dp20 | dpbatch.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="dp30", choices=["dp30", "dp20", "dpbatch"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--batch-mode", type=int, default=0)
    ap.add_argument("--low-bits", type=int, default=None)
    ap.add_argument("--tiles-per-lane", type=int, default=None)
    ap.add_argument("--block", type=int, default=None)
    return ap.parse_args()


def cpu_baseline_dp(g, target_s=12.0):
    """Oracle flat-table DP on the subgraph induced by the first m vertices."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib
    from gisbdw_amd import graphs as G

    o = oracle_lib.Oracle()
    threads = min(os.cpu_count() or 1, 16)

    def induced(m):
        mask = (1 << m) - 1
        return G.Graph(m, [g.adj[v] & mask for v in range(m)], f"{g.name}[:{m}]")

    m = 22
    t0 = time.perf_counter()
    o.dp_fast(induced(m), threads=threads)
    dt = time.perf_counter() - t0
    rate = (1 << m) / dt
    # pick the largest m <= n whose estimated time fits the target
    while m < g.n and (1 << (m + 1)) / rate * 1.15 < target_s:
        m += 1
    t0 = time.perf_counter()
    o.dp_fast(induced(m), threads=threads)
    dt = time.perf_counter() - t0
    return {"value": (1 << m) / dt, "unit": "subsets/s", "cores": threads, "kind": "port",
            "sample": f"oracle flat-table DP (faithful recurrence of dynamic_gpu.cu, OpenMP "
                      f"within each level) on the subgraph induced by vertices 0..{m - 1} of "
                      f"{g.name}: 2^{m} subsets in {dt:.2f} s"}


def profile_traffic():
    """HBM bytes per level launch from the committed rocprofv3 --pmc summary,
    if present (see profiles/)."""
    p = ROOT / "profiles" / "dp_level_kernel_n30_traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("traffic_bytes_per_launch")
        except Exception:
            return None
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); "
              "the multi-GPU design is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    from gisbdw_amd import api
    from gisbdw_amd import graphs as G

    ctx = api.Context(0)
    if args.low_bits:
        ctx.set_param("dp.low_bits", args.low_bits)
    if args.tiles_per_lane:
        ctx.set_param("dp.tiles_per_lane", args.tiles_per_lane)
    if args.block:
        ctx.set_param("dp.block", args.block)

    if args.workload in ("dp30", "dp20"):
        graphs = G.workload_dp_n30() if args.workload == "dp30" else [G.workload_dp_single(0)]
        dp = api.DynamicGPU(ctx)

        def step():
            lvl_ms = 0.0
            alg = 0
            sub = 0
            for g in graphs:
                dp(g)
                lvl_ms += dp.stats.level_ms
                alg += dp.stats.alg_bytes
                sub += dp.stats.subsets
            return lvl_ms, alg, sub

        for _ in range(args.warmup):
            step()
        print(f"[bench] warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        lvl_ms = 0.0
        alg = sub = 0
        launches = 0
        for i in range(args.steps):
            a, b, c = step()
            lvl_ms += a
            alg += b
            sub += c
            print(f"[bench] step {i} level_ms={a:.3f}", file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        nlev = sum(max(g.n - ctx.get_param("dp.low_bits"), 0) + 1 for g in graphs)
        achieved = alg / (lvl_ms * 1e-3) / 1e9
        n = graphs[0].n
        out = {
            "metric": f"DP subsets/sec (exact subset DP, n={n})",
            "value": sub / wall,
            "unit": "subsets/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": wall / args.steps * 1e3,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "u8",
            "data": "synthetic",
            "config": {
                "workload": ("dp-n30: 4 x G(30,0.2) seeds 0-3 + 4 x 5x6 grid minus 5 of 49 "
                             "edges seeds 0-3" if args.workload == "dp30" else
                             "dp-n20: one G(20,0.3) seed 0 (reseeded until connected)"),
                "graphs": len(graphs),
                "n": n,
                "low_bits": ctx.get_param("dp.low_bits"),
                "tiles_per_lane": ctx.get_param("dp.tiles_per_lane"),
                "block": ctx.get_param("dp.block"),
            },
            "roofline": {
                "bound": "hbm",
                "achieved": achieved,
                "peak": HBM_PEAK_GBPS,
                "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBPS,
                "traffic": profile_traffic(),
                "kernel": "DPLevelKernel",
                "alg_bytes_per_launch": alg / (args.steps * nlev),
                "avg_launch_ms": lvl_ms / (args.steps * nlev),
            },
            "extra": {
                "graphs_per_s": len(graphs) * args.steps / wall,
                "level_kernel_ms_per_step": lvl_ms / args.steps,
                "subsets_per_s_kernel": sub / (lvl_ms * 1e-3),
                "compulsory_GBps": 2 * sub / (lvl_ms * 1e-3) / 1e9,
            },
        }
    else:
        graphs = G.workload_dp_batch(4096, 0)

        def step():
            tds, pars, st = api.dp_solve_batch(ctx, graphs, args.batch_mode)
            return st

        for _ in range(args.warmup):
            step()
        t0 = time.perf_counter()
        lvl_ms = 0.0
        alg = sub = 0
        for i in range(args.steps):
            st = step()
            lvl_ms += st.level_ms
            alg += st.alg_bytes
            sub += st.subsets
            print(f"[bench] step {i} level_ms={st.level_ms:.3f}", file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        achieved = alg / (lvl_ms * 1e-3) / 1e9
        out = {
            "metric": "DP graphs/sec (batched exact subset DP)",
            "value": len(graphs) * args.steps / wall,
            "unit": "graphs/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": wall / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u8", "data": "synthetic",
            "config": {"workload": "dp-batch: 4096 G(n,p), n~U{12..24}, p~U[0.15,0.5], seed 0",
                       "graphs": len(graphs), "mode": args.batch_mode},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": None},
            "extra": {"subsets_per_s": sub / wall},
        }
    if not args.no_cpu:
        g0 = graphs[0]
        out["cpu_baseline"] = cpu_baseline_dp(g0)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

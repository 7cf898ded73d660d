"""Benchmark: BFS GTEPS on R-MAT scale 26 / edge factor 32 with the device
edge budget capped at 25% of the edge bytes (the rest read zero-copy from
pinned host memory), partitions of 2^16 edges; plus the CC pass of the same
configuration and the multithreaded CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 26] ...

One step = one BFS from vertex 1 over the whole graph (first launch to
empty frontier; graph generation and store set-up are outside the timed
region).  Each engine call returns only after its HIP stream has drained, so
the host clock around the K steps brackets all device work.  Prints ONE
JSON line.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12       # MI355X HBM3E, bytes/s
LINK_PEAK = 64.0e9      # PCIe 5.0 x16, one direction, bytes/s
METRIC = "GTEPS at capped HBM edge budget; batched BFS queries/s + p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def mixed_roofline_ms(hbm_bytes, host_bytes):
    return max(hbm_bytes / HBM_PEAK, host_bytes / LINK_PEAK) * 1e3


def cpu_baseline(g, src, seconds, label):
    """BFS on the host cores with the oracle's multithreaded restatement, on
    the same graph, bounded to roughly `seconds` of work."""
    from oracle import oracle as O
    off, dst, _, _ = g.export()
    og = O.CSR(off, dst)
    runs, total, trav = 0, 0.0, 0
    while total < seconds or runs == 0:
        t0 = time.perf_counter()
        lv, tr = O.bfs_mt(og, src)
        total += time.perf_counter() - t0
        trav += tr
        runs += 1
        if runs >= 20:
            break
    return {"value": trav / total / 1e9, "unit": "GTEPS", "cores": O.threads(),
            "kind": "port",
            "sample": f"{runs} full BFS run(s) from vertex {src} on the same {label} graph "
                      f"(oracle/oracle.c orc_bfs_mt, OpenMP top-down), {total:.1f} s"}, lv


def pct(xs, p):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(round(p / 100.0 * (len(xs) - 1))))] if xs else None


def serve_stream(T, g, queries, rate, seed):
    """Submits (algo, src) queries to a fresh server, at Poisson arrivals of
    `rate` per second (None: all at once); returns throughput and latency."""
    srv = T.Server(g, max_lanes=64)
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(queries)) if rate else np.zeros(len(queries))
    t0 = time.perf_counter()
    due = 0.0
    tickets = []
    for (a, s), gap in zip(queries, gaps):
        due += gap
        while rate:
            left = due - (time.perf_counter() - t0)
            if left <= 0:
                break
            time.sleep(min(0.0005, left))
        tickets.append((a, srv.submit(a, s)))
    srv.drain()
    wall = time.perf_counter() - t0
    res = [(a, srv.poll(t)) for a, t in tickets]
    srv.stop()
    assert all(r["status"] == 0 for _, r in res)
    first = min(r["submit_ms"] for _, r in res)
    last = max(r["done_ms"] for _, r in res)
    out = {"queries": len(res), "qps": round(len(res) / ((last - first) / 1e3), 2),
           "wall_s": round(wall, 3)}
    for name in ("all", "bfs", "sssp"):
        lat = [r["done_ms"] - r["submit_ms"] for a, r in res if name in ("all", a)]
        if lat:
            out[f"{name}_p50_ms"] = round(pct(lat, 50), 3)
            out[f"{name}_p99_ms"] = round(pct(lat, 99), 3)
    return out, res


def serving_bench(T, args):
    """Resident R-MAT scale 25 / ef 16 with weights in [1, 64]: one query per
    launch vs. 64-lane batches vs. the continuous-batching server."""
    t0 = time.perf_counter()
    g = T.Graph.rmat(args.serve_scale, 16, seed=2, symmetric=True, weights=True)
    g.store_config(None, 1 << 16)
    log(f"serving graph: n={g.n} m={g.m} ({time.perf_counter() - t0:.1f} s)")
    rng = np.random.default_rng(3)
    out = {"graph": {"scale": args.serve_scale, "edge_factor": 16, "vertices": g.n,
                     "edges": g.m, "edge_bytes": 8, "resident": True}}
    # the same 64 uniform random sources: one query per launch sequence
    # (plain single-source kernels) vs. one 64-lane batch
    b64 = rng.integers(0, g.n, 64).astype(np.uint32)
    deg = np.diff(g.offsets()).astype(np.int64)
    out["sources_isolated"] = int((deg[b64] == 0).sum())
    def plain_qps(fn):
        fn(int(b64[0]), out=False)
        t = time.perf_counter()
        for s in b64:
            fn(int(s), out=False)
        return round(64 / (time.perf_counter() - t), 2)

    # plain BFS two ways: the reference's one-query-per-launch top-down
    # kernel (the serving target's baseline), and this engine's own
    # direction-optimising single query
    g.set_option("direction", 0)
    out["plain_bfs_topdown_qps"] = plain_qps(g.bfs)
    g.set_option("direction", 1)
    out["plain_bfs_qps"] = plain_qps(g.bfs)
    out["plain_sssp_qps"] = plain_qps(g.sssp)
    for algo in ("bfs", "sssp"):
        r = g.batch(algo, b64)
        out[f"batch64_{algo}_qps"] = round(64 / (r["stats"].ms / 1e3), 2)
        out[f"batch64_{algo}_levels"] = r["stats"].iterations
        out[f"batch64_{algo}_vs_plain"] = round(out[f"batch64_{algo}_qps"] /
                                                out[f"plain_{algo}_qps"], 2)
    out["batch64_bfs_vs_plain_topdown"] = round(out["batch64_bfs_qps"] /
                                                out["plain_bfs_topdown_qps"], 2)
    log(f"plain/batch: {out}")
    # BFS-only stream: saturation, then Poisson at 70% of it
    nb = args.serve_queries // 2
    qb = [("bfs", int(s)) for s in rng.integers(0, g.n, nb)]
    sat, _ = serve_stream(T, g, qb, None, 4)
    pois, _ = serve_stream(T, g, qb, 0.7 * sat["qps"], 5)
    out["serve_bfs_saturated"] = sat
    out["serve_bfs_poisson70"] = pois
    out["serve_bfs_speedup_vs_plain"] = round(sat["qps"] / out["plain_bfs_qps"], 2)
    out["serve_bfs_speedup_vs_plain_topdown"] = round(sat["qps"] /
                                                      out["plain_bfs_topdown_qps"], 2)
    log(f"bfs stream: {sat} / {pois}")
    # the mixed 50/50 BFS+SSSP stream of the configuration
    qm = [("bfs" if rng.random() < 0.5 else "sssp", int(s))
          for s in rng.integers(0, g.n, args.serve_queries)]
    msat, _ = serve_stream(T, g, qm, None, 6)
    mpois, _ = serve_stream(T, g, qm, 0.7 * msat["qps"], 7)
    out["serve_mix_saturated"] = msat
    out["serve_mix_poisson70"] = mpois
    log(f"mix stream: {msat} / {mpois}")
    g.close()
    return out


def measured_traffic(workload, kernel):
    """Memory traffic per launch of `kernel` in `workload` from the committed
    counter summary profiles/traffic.json: rocprofv3 FETCH_SIZE and
    WRITE_SIZE per dispatch (KiB), each divided by what the counter reads
    for a plain copy of known size (thievory_amd/calibrate.py).  None when
    no such measurement is on file."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            e = json.load(f).get(workload, {}).get(kernel)
    except (OSError, ValueError):
        return None
    return int(e["bytes_per_launch"]) if e else None


def bfs_roofline(stats, workload=None):
    """Roofline of the BFS steps.  Kernel view: the traversal kernels'
    algorithmic bytes over their HIP-event time -- the bottom-up probe
    (the first level's sweep or a later level's candidate-list probe, then
    the short-list scan: one event window) against HBM (its reads are the
    visited / residency / hint / degree words, candidates, short lists and
    frontier words, all in HBM), the top-down expansion and the early-exit
    bottom-up window against the host link.
    The dominant kernel (most time) is the headline.  Mixed views: T_roof =
    max(HBM bytes / 8 TB/s, host bytes / 64 GB/s) against the step time,
    once with the bytes this run read and once with the bytes a
    top-down-only traversal of the same levels reads."""
    med = statistics.median(s.ms for s in stats)
    kern = {}
    for name, ms_key, launches, nbytes, bound in (
            ("k_bu_probe|k_bu_probe_list+k_bu_short", "probe_ms", lambda s: s.probe_launches,
             lambda s: s.probe_hbm_bytes, "hbm"),
            ("k_expand<BFS>", "expand_ms", lambda s: s.expand_launches,
             lambda s: s.host_bytes - s.bottom_up_host_bytes, "host-link"),
            ("bottom-up window (k_bottom_up)", "bottom_up_ms",
             lambda s: s.bottom_up_levels, lambda s: s.bottom_up_host_bytes, "host-link")):
        t = sum(getattr(s, ms_key) for s in stats)
        n = sum(launches(s) for s in stats)
        b = sum(nbytes(s) for s in stats)
        if n:
            peak = HBM_PEAK if bound == "hbm" else LINK_PEAK
            gbps = b / (t * 1e-3) / 1e9 if t else 0.0
            kern[name] = {"bound": bound, "ms_total": round(t, 3), "launches": n,
                          "avg_launch_ms": round(t / n, 4), "bytes_per_launch": b // n,
                          "achieved_GBps": round(gbps, 2),
                          "frac": round(gbps * 1e9 / peak, 4)}
    dom = max(kern, key=lambda k: kern[k]["ms_total"])
    d = kern[dom]
    t_act = statistics.median(mixed_roofline_ms(s.hbm_bytes, s.host_bytes) for s in stats)
    t_td = statistics.median(mixed_roofline_ms(s.topdown_hbm_bytes, s.topdown_host_bytes)
                             for s in stats)
    return {
        "bound": d["bound"], "kernel": dom, "achieved": d["achieved_GBps"],
        "peak": (HBM_PEAK if d["bound"] == "hbm" else LINK_PEAK) / 1e9, "unit": "GB/s",
        "frac": d["frac"], "traffic": measured_traffic(workload, dom),
        "avg_launch_ms": d["avg_launch_ms"], "launches": d["launches"],
        "bytes_per_launch": d["bytes_per_launch"],
        "kernels": kern,
        "mixed": {"t_roof_ms": round(t_act, 3), "t_ms": round(med, 3),
                  "frac": round(t_act / med, 4), "bytes": "read by this run",
                  "host_bytes": stats[0].host_bytes, "hbm_bytes": stats[0].hbm_bytes},
        "mixed_topdown": {"t_roof_ms": round(t_td, 3), "t_ms": round(med, 3),
                          "frac": round(t_td / med, 4),
                          "bytes": "top-down algorithmic (every reached vertex's edges)",
                          "host_bytes": stats[0].topdown_host_bytes,
                          "hbm_bytes": stats[0].topdown_hbm_bytes},
        "bottom_up_levels": stats[0].bottom_up_levels,
        "edges_examined": stats[0].edges_examined,
        "hint_hits": stats[0].hint_hits,
    }


def step_summary(stats, wall, kernel):
    """Throughput, mixed roofline and expansion-kernel link figures of K runs."""
    exp_ms = sum(s.expand_ms for s in stats)
    launches = sum(s.expand_launches for s in stats)
    host_all = sum(s.host_bytes for s in stats)
    hbm_all = sum(s.hbm_bytes for s in stats)
    med = statistics.median(s.ms for s in stats)
    t_roof = statistics.median(mixed_roofline_ms(s.hbm_bytes, s.host_bytes) for s in stats)
    link = host_all / (exp_ms * 1e-3) / 1e9 if exp_ms else 0.0
    return {
        "edges_per_run": stats[0].edges_traversed,
        "iterations": stats[0].iterations,
        "ms_median": round(med, 3),
        "gteps": round(sum(s.edges_traversed for s in stats) / wall / 1e9, 4),
        "roofline": {"bound": "host-link", "kernel": kernel, "achieved": round(link, 2),
                     "peak": LINK_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(link * 1e9 / LINK_PEAK, 4),
                     "avg_launch_ms": round(exp_ms / launches, 4) if launches else None,
                     "launches": launches,
                     "mixed": {"t_roof_ms": round(t_roof, 3), "t_ms": round(med, 3),
                               "frac": round(t_roof / med, 4)}},
        "host_bytes_per_run": stats[0].host_bytes,
        "hbm_bytes_per_run": stats[0].hbm_bytes,
    }


def sssp_bench(T, args):
    """SSSP from a seeded random source with degree > 0 on R-MAT scale 26 /
    ef 16, 8-byte {dst, weight} records, weights in [1, 64], 30% HBM budget."""
    g = T.Graph.rmat(args.scale, 16, seed=1, symmetric=True, weights=True)
    g.store_config(int(0.30 * g.m * 8), args.partition_edges)
    deg = np.diff(g.offsets())
    rng = np.random.default_rng(1)
    src = int(rng.choice(np.nonzero(deg)[0]))
    g.sssp(src, out=False)
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(g.sssp(src, out=False)[1])
    out = {"workload": f"sssp-rmat{args.scale}-ef16-w8-budget30pct", "source": src,
           "vertices": g.n, "edges": g.m, "store": g.store_info()}
    out.update(step_summary(stats, time.perf_counter() - t0, "k_expand<SSSP>"))
    dist, _ = g.sssp(src)
    # algorithm-independent rate: edges out of the reached vertices per
    # second (delta-stepping relaxes fewer edges than Bellman-Ford, so the
    # processed-edge rate alone would understate a faster run)
    reached_edges = int(deg[dist != g.n + 1].sum())
    out["reached_edges"] = reached_edges
    out["reached_edge_gteps"] = round(reached_edges / out["ms_median"] / 1e6, 4)
    out["sssp_delta"] = g.get_option("sssp_delta")
    if not args.no_cpu:
        from oracle import oracle as O
        off, dst, w, _ = g.export()
        t = time.perf_counter()
        cd = O.sssp_mt(O.CSR(off, dst, w), src)
        ct = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(stats[0].edges_traversed / ct / 1e9, 4),
                               "unit": "G processed edges/s", "cores": O.threads(),
                               "kind": "port", "seconds": round(ct, 2),
                               "sample": "1 frontier Bellman-Ford run (orc_sssp_mt)"}
        out["cpu_matches_gpu"] = bool(np.array_equal(cd, dist))
    g.close()
    return out


def pr_bench(T, args):
    """PageRank push (CSR) and pull (CSC) on directed R-MAT scale 26 / ef 16,
    25% HBM budget, fp32, stop at max |delta| < 1e-7 or 100 iterations;
    normalised L1 against the FP64 fixpoint."""
    out = {"workload": f"pr-rmat{args.scale}-ef16-directed-budget25pct"}
    res = {}
    for mode, transpose in (("push", False), ("pull", True)):
        g = T.Graph.rmat(args.scale, 16, seed=1, symmetric=False, transpose=transpose)
        g.store_config(int(0.25 * g.m * 4), args.partition_edges)
        t0 = time.perf_counter()
        r, st = g.pr(mode)
        wall = time.perf_counter() - t0
        s = step_summary([st], wall, f"k_expand<PR_{mode.upper()}>")
        s["vertices"], s["edges"] = g.n, g.m
        out[mode] = s
        res[mode] = r
        if transpose and not args.no_cpu:
            from oracle import oracle as O
            off, src, _, od = g.export()
            t = time.perf_counter()
            ex, it = O.pr_exact_mt(O.CSR(off, src, None, od), tol=1e-10, max_iter=300)
            out["fp64_oracle"] = {"iterations": it, "seconds": round(time.perf_counter() - t, 1),
                                  "cores": O.threads()}
            for k in res:
                out[k]["l1_vs_fp64"] = O.normalized_l1(res[k], ex)
        g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--edge-factor", type=int, default=32)
    ap.add_argument("--budget", type=float, default=0.25)
    ap.add_argument("--partition-edges", type=int, default=1 << 16)
    ap.add_argument("--source", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cc", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--serve-scale", type=int, default=25)
    ap.add_argument("--serve-queries", type=int, default=2000)
    ap.add_argument("--workload", choices=["bfs", "sssp", "pr"], default="bfs",
                    help="bfs: the headline line (BFS + CC + serving); sssp / pr: the "
                         "SSSP and PageRank configurations, printed as their own line")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): this build runs "
              "on a single MI355X; multi-GPU sharding is a design note, not a code path.",
              file=sys.stderr)
        sys.exit(2)

    import thievory_amd as T
    dev = T.device_info()
    log(f"device {dev}")
    if args.serve_only:
        print(json.dumps({"serving": serving_bench(T, args)}), flush=True)
        return
    if args.workload == "sssp":
        print(json.dumps(sssp_bench(T, args)), flush=True)
        return
    if args.workload == "pr":
        print(json.dumps(pr_bench(T, args)), flush=True)
        return
    t0 = time.perf_counter()
    g = T.Graph.rmat(args.scale, args.edge_factor, seed=1, symmetric=True)
    gen_s = time.perf_counter() - t0
    log(f"generated R-MAT scale {args.scale} ef {args.edge_factor}: n={g.n} m={g.m} "
        f"in {gen_s:.1f} s")
    budget = int(args.budget * g.m * g.edge_bytes)
    t0 = time.perf_counter()
    g.store_config(budget, args.partition_edges)
    store = g.store_info()
    log(f"store: {store} (budget {budget} B) in {time.perf_counter() - t0:.1f} s")

    for i in range(args.warmup):
        _, st = g.bfs(args.source, out=False)
        log(f"warmup {i}: {st.ms:.1f} ms, {st.iterations} levels")

    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, st = g.bfs(args.source, out=False)
        stats.append(st)
        log(f"step {i}: {st.ms:.2f} ms, {st.edges_traversed / st.ms / 1e6:.2f} GTEPS, "
            f"probe {st.probe_ms:.2f} ms, expand {st.expand_ms:.2f} ms, "
            f"bottom-up {st.bottom_up_ms:.2f} ms, "
            f"levels {st.iterations} ({st.bottom_up_levels} bottom-up)")
    wall = time.perf_counter() - t0

    trav = sum(s.edges_traversed for s in stats)
    value = trav / wall / 1e9
    ms_per_step = wall * 1e3 / args.steps
    workload = f"bfs-rmat{args.scale}-ef{args.edge_factor}-budget{int(args.budget * 100)}pct"
    roof = bfs_roofline(stats, workload)
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded R-MAT, generated on the GPU)",
        "config": {"workload": workload,
                   "scale": args.scale, "edge_factor": args.edge_factor, "vertices": g.n,
                   "edges": g.m, "edge_bytes": g.edge_bytes, "budget_frac": args.budget,
                   "partition_edges": args.partition_edges, "source": args.source,
                   "store": store, "graph500_gteps": round(value / 2, 4)},
        "roofline": roof,
    }
    if not args.no_cc:
        cc, cst = g.cc()
        # union-find CC reads a small part of the edges: the graph's edge
        # count over the time is the algorithm-independent rate
        out["cc"] = {"ms": round(cst.ms, 2), "edges_processed": cst.edges_traversed,
                     "iterations": cst.iterations,
                     "gteps": round(cst.edges_traversed / cst.ms / 1e6, 3),
                     "graph_edges_per_s_g": round(g.m / cst.ms / 1e6, 3),
                     "method": "union-find" if g.get_option("cc_union") else "label propagation",
                     "host_bytes": cst.host_bytes,
                     "mixed_roofline_frac": round(
                         mixed_roofline_ms(cst.hbm_bytes, cst.host_bytes) / cst.ms, 4)}
        log(f"cc: {out['cc']}")
    lv, _ = g.bfs(args.source)  # levels for the CPU cross-check (untimed)
    if not args.no_cpu:
        cb, cpu_lv = cpu_baseline(g, args.source, args.cpu_seconds,
                                  f"R-MAT {args.scale}/{args.edge_factor}")
        out["cpu_baseline"] = cb
        out["cpu_matches_gpu"] = bool(np.array_equal(cpu_lv, lv))
        log(f"cpu baseline: {cb}")
    g.close()
    if not args.no_serve:
        out["serving"] = serving_bench(T, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

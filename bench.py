#!/usr/bin/env python3
"""Benchmark: ms per 7-vertex pattern on one MI355X, exact count.

Workload (one "step" = one full match of the pattern over the graph):
  RMAT scale 16 (65,536 vertices), edge factor 16, (a,b,c,d) =
  (0.57,0.19,0.19,0.05), seed 1; symmetrized, deduplicated, self-loops
  removed, adjacency sorted (u32 ids, u64 offsets); 4 uniform labels
  (seed 2); labeled pattern K4{0,1,2,3} + path 3-4-5-6, every vertex label 1.
The graph is resident in HBM before the timed region.  Each step is
synchronous (the library waits on its stream before returning), so the
wall-clock bracket around K steps covers exactly K complete matches.

The JSON line carries the roofline of the dominant kernel (algorithmic
bytes per launch / its mean duration from hipEvents on the launch stream)
and a CPU baseline: the oracle (oracle/oracle.c, OpenMP) timed on a bounded
sample of the same jobs and scaled to the whole job set by the ratio of
matches (the oracle's work grows with the number of matches it enumerates).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

C1_EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (3, 4), (4, 5), (5, 6)]
HBM_PEAK_GBS = 8000.0


def pattern_text(k, edges, label=1):
    lines = ["# t 0"] + [f"v {i} {label}" for i in range(k)]
    for a, b in edges:
        lines += [f"e {a} {b} 1", f"e {b} {a} 1"]
    return "\n".join(lines) + "\n"


def cpu_baseline(g, text, per_edge, total_count, seconds, threads):
    """Time the oracle on evenly spaced edge ranges until ~`seconds` of CPU
    work, then scale by total matches / sampled matches."""
    import numpy as np

    import oracle

    rp, ci, lb = g.csr()
    m = ci.size
    p = oracle.parse_g(text, True)
    ranges = 256
    width = max(64, m // (ranges * 32))
    starts = [int(i * m // ranges) for i in range(ranges)]
    order = np.random.default_rng(0).permutation(ranges)
    t_used, c_sampled, edges_sampled, jobs_sampled = 0.0, 0, 0, 0
    for i in order:
        lo = starts[i]
        hi = min(m, lo + width)
        est = int(per_edge[lo:hi].sum())
        t0 = time.perf_counter()
        c, jobs, _ = oracle.count(rp, ci, lb, p, True, edge_lo=lo, edge_hi=hi, threads=threads)
        t_used += time.perf_counter() - t0
        assert c == est, f"oracle/GPU mismatch on edges [{lo},{hi}): {c} vs {est}"
        c_sampled += c
        edges_sampled += hi - lo
        jobs_sampled += jobs
        if t_used >= seconds:
            break
    if c_sampled == 0:
        return None
    ms_full = t_used * 1e3 * total_count / c_sampled
    return {
        "value": ms_full,
        "unit": "ms",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle (oracle/oracle.c, OpenMP) on {edges_sampled} edges in "
                   f"{width}-edge ranges spread over the CSR ({jobs_sampled} jobs, "
                   f"{c_sampled} matches, {t_used:.1f} s), scaled by total/sampled matches; "
                   f"sampled counts equal the GPU per-edge counts"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the matcher shards by job ranges across GPUs "
              "(see DESIGN.md), but only --gpus 1 is run here", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import pumatch_amd as pm

    t0 = time.perf_counter()
    g = pm.Graph.rmat(16, 16, a=0.57, b=0.19, c=0.19, seed=1)
    g.set_uniform_labels(4, 2)
    text = pattern_text(7, C1_EDGES)
    p = pm.Pattern.parse(text, True)
    print(f"[bench] graph n={g.n} m={g.m} built in {time.perf_counter() - t0:.1f}s", flush=True)

    first = pm.match(g, p, per_edge=True)  # warm-up with per-edge counts for the baseline
    for _ in range(max(0, args.warmup - 1)):
        pm.match(g, p)
    print(f"[bench] warm-up done count={first.count}", flush=True)

    results = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        results.append(pm.match(g, p))
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    counts = {r.count for r in results} | {first.count}
    assert len(counts) == 1, f"count changed between steps: {counts}"
    count = first.count

    # roofline of the dominant kernel phase
    names = ["k_fold", "k_match_edges", "k_match_tasks"]
    ph_ms = np.mean([r.phase_ms for r in results], axis=0)
    ph_bytes = results[-1].phase_alg_bytes
    dom = int(np.argmax(ph_ms))
    achieved = ph_bytes[dom] / (ph_ms[dom] * 1e-3) / 1e9 if ph_ms[dom] > 0 else 0.0
    kernel_ms = float(np.mean([r.kernel_ms for r in results]))
    roofline = {
        "bound": "hbm",
        "kernel": names[dom],
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": None,
        "kernel_ms": round(float(ph_ms[dom]), 4),
        "alg_bytes": int(ph_bytes[dom]),
    }

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        print(f"[bench] cpu baseline ({threads} threads, ~{args.cpu_seconds}s)", flush=True)
        cpu = cpu_baseline(g, text, first.per_edge, count, args.cpu_seconds, threads)

    out = {
        "metric": "ms per 7-vertex pattern (plain GPU), exact count, HBM GB/s",
        "value": round(ms_per_step, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": ("rmat16-ef16-labeled4: RMAT scale 16, edge factor 16, (.57,.19,.19,.05) "
                         "seed 1, 4 uniform labels seed 2, labeled K4 + path 3-4-5-6 (all label 1)"),
            "n": g.n,
            "m_directed": g.m,
            "pattern_vertices": 7,
            "labeled": True,
        },
        "count": count,
        "jobs": first.jobs,
        "kernel_ms": round(kernel_ms, 4),
        "alg_GBps": round(results[-1].alg_bytes / (kernel_ms * 1e-3) / 1e9, 2),
        "phase_ms": [round(float(x), 4) for x in ph_ms],
        "tasks": results[-1].tasks,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

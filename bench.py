#!/usr/bin/env python3
"""Benchmark: ms per 7-vertex pattern on one MI355X, exact count.

Main line (one "step" = one full match of the pattern over the graph):
  RMAT scale 16 (65,536 vertices), edge factor 16, (a,b,c,d) =
  (0.57,0.19,0.19,0.05), seed 1; symmetrized, deduplicated, self-loops
  removed, adjacency sorted (u32 ids, u64 offsets); 4 uniform labels
  (seed 2); labeled pattern K4{0,1,2,3} + path 3-4-5-6, every vertex label 1.
The graph is resident in HBM before the timed region.  Launch granularities
come from the autotuner (pm_autotune, sampled launches before timing); the
untuned time is reported beside.  Each step is synchronous (the library
waits on its stream before returning), so the wall-clock bracket around K
steps covers exactly K complete matches.

The JSON line carries the roofline of the dominant kernel (algorithmic
bytes per launch / its mean duration from hipEvents on the launch stream;
HBM traffic from the rocprofv3 FETCH_SIZE/WRITE_SIZE runs summarised in
profiles/) and a CPU baseline: the oracle (oracle/oracle.c, OpenMP) timed on
a bounded sample of the same jobs and scaled to the whole job set by the
ratio of matches.

Two more objects compare plain and learned+batched schedules:
  "batched": 64 random connected labeled patterns (3-7 vertices, density
    0.3-0.8, seed 6) on RMAT scale 14 with 8 labels (seed 5): plain =
    one pm_match per pattern in file order with default options; batched =
    one pm_match_batch launch with autotuned orders and granularities.
  "learned": the MLP trained on traces of seeds 7-9 and evaluated on seed
    10: learned job order on permuted RMAT scale 15 (resident), and learned
    prefetch + eviction, alone and in one batched launch, out of core on
    permuted Chung-Lu graphs (200k vertices, max degree 5000) with 40% of
    the rows in HBM, each against the plain schedule (LRU, rows of <= 1024
    entries read from host memory when not cached); the deferral-only
    schedule (no direct reads) is timed beside it.
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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_match_edges_rmat16_c1.txt")


def pattern_text(k, edges, labels=None):
    labels = labels or [1] * k
    lines = ["# t 0"] + [f"v {i} {labels[i]}" for i in range(k)]
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


def batched_section(pm):
    from pumatch_amd.workloads import batch_patterns

    g = pm.Graph.rmat(14, 16, seed=5)
    g.set_uniform_labels(8, 5)
    texts = batch_patterns(64, seed=6, nlabels=8)
    pats = [pm.Pattern.parse(t, True) for t in texts]
    n7 = sum(1 for t in texts if t.count("\nv ") == 7)
    plain_ms, plain = 0.0, []
    for p in pats:
        r = pm.match(g, p, order="file")
        plain.append(r.count)
        plain_ms += r.kernel_ms
    t0 = time.perf_counter()
    opts, rep = pm.autotune(g, pats, sample=4)
    tune_s = time.perf_counter() - t0
    rs = pm.match_batch(g, pats, repeats=3, **opts)
    assert [r.count for r in rs] == plain, "batched counts differ from single-pattern runs"
    batch_ms = rs[0].min_ms
    # the same launch as a captured hipGraph, replayed
    rg = pm.match_batch(g, pats, repeats=3, graph=True, **opts)
    assert [r.count for r in rg] == plain, "graph replays differ from single-pattern runs"
    return {
        "workload": ("rmat14-ef16-labeled8: RMAT scale 14, edge factor 16, seed 5, 8 labels; "
                     "64 random connected labeled patterns, 3-7 vertices, density 0.3-0.8, seed 6"),
        "patterns": len(pats),
        "patterns_7_vertices": n7,
        "plain_sequential_ms": round(plain_ms, 3),
        "plain_ms_per_pattern": round(plain_ms / len(pats), 4),
        "batched_tuned_ms": round(batch_ms, 3),
        "batched_ms_per_pattern": round(batch_ms / len(pats), 4),
        "batched_graph_replay_ms": round(rg[0].min_ms, 3),
        "speedup": round(plain_ms / batch_ms, 3),
        "counts_equal": True,
        "autotune_s": round(tune_s, 2),
        "autotune": {"chunk": opts["chunk"], "split": opts["split"],
                     "hub_factor": opts["hub_factor"], "reordered_patterns": rep["reordered"],
                     "sample": rep["sample"]},
    }


def learned_section(pm):
    """Train the MLP on traces of plain runs over graphs of seeds 7-9 and
    evaluate on seed 10: (1) learned job order on permuted RMAT scale 15,
    whole chip with default granularities and with 1024 waves and no
    splitting; (2) learned package ranking (prefetch + eviction) out of core
    on permuted Chung-Lu graphs against LRU; (3) one batched out-of-core
    launch of the same patterns with the learned policy against the plain
    runs one by one."""
    import numpy as np

    from pumatch_amd.workloads import permuted

    def text(k, edges, labels=None):
        return pattern_text(k, edges, labels)

    tri = pm.Pattern.parse(text(3, [(0, 1), (1, 2), (0, 2)]), False)
    k4 = pm.Pattern.parse(text(4, [(a, b) for a in range(4) for b in range(a)]), False)
    c1 = pm.Pattern.parse(text(7, C1_EDGES), True)
    out = {}

    # (1) learned job order
    def rmat(seed):
        g0 = pm.Graph.rmat(15, 16, seed=seed, device=False)
        g = permuted(g0, seed + 100)
        g.set_uniform_labels(4, seed)
        return g

    pats = [("k4", k4), ("triangle", tri), ("c1", c1)]
    Xs, Ys, Ws = [], [], []
    for seed in (7, 8, 9):
        g = rmat(seed)
        rw = np.zeros(g.n)
        for _, p in pats:
            rw += pm.match(g, p, traces=True).root_work
        Xs.append(pm.features(g))
        Ys.append(np.stack([np.log2(1 + rw), np.zeros(g.n)], 1))
        Ws.append(np.stack([np.ones(g.n), np.zeros(g.n)], 1))
    m = pm.Model(1)
    t0 = time.perf_counter()
    loss = m.train(np.concatenate(Xs), np.concatenate(Ys).astype(np.float32),
                   np.concatenate(Ws).astype(np.float32), epochs=30, lr=3e-3)
    train_s = time.perf_counter() - t0
    g = rmat(10)
    order = {"workload": ("rmat15-perm: RMAT scale 15, edge factor 16, ids permuted, 4 labels; "
                          "unlabeled K4 and triangle, labeled K4 + path"),
             "train": f"root-work traces of plain runs on seeds 7-9, {train_s:.2f} s on the GPU, "
                      f"final loss {loss:.3f}", "eval_seed": 10}
    for tag, kw in (("default", {}), ("1024_waves_no_split", dict(waves=1024, split=-1))):
        tp = tl = 0.0
        per = {}
        for name, p in pats:
            a = pm.match(g, p, repeats=5, **kw)
            b = pm.match(g, p, repeats=5, model=m, learned=pm.PM_LEARN_ORDER, **kw)
            assert a.count == b.count
            per[name] = [round(a.min_ms, 3), round(b.min_ms, 3)]
            tp += a.min_ms
            tl += b.min_ms
        order[tag] = {"plain_ms": round(tp, 3), "learned_ms": round(tl, 3),
                      "speedup": round(tp / tl, 3), "per_pattern_plain_learned_ms": per}
    out["order"] = order

    # (2) learned package ranking out of core, (3) batched.  Max degree
    # 5000 as in the 65.6M-vertex configuration, so the hub rows (> 1024
    # entries, the direct-read limit) page through the cache
    N, DMAX, MEAN, PKG = 200000, 5000, 30, 16384

    def chunglu(seed):
        g0 = pm.Graph.chung_lu(N, 2.2, DMAX, MEAN, seed=seed, device=False)
        g = permuted(g0, seed + 100)
        g.set_uniform_labels(8, seed)
        return g

    ppats = [("triangle", tri), ("k4", k4)]
    Xs, Ys = [], []
    for seed in (7, 8, 9):
        g = chunglu(seed)
        cache = int(g.m * 4 * 0.4) // PKG * PKG
        rw, pu = np.zeros(g.n), np.zeros(g.n)
        for _, p in ppats:
            r = pm.match(g, p, traces=True, cache_bytes=cache, package_bytes=PKG)
            rw += r.root_work
            pu += r.package_use
        Xs.append(pm.features(g))
        Ys.append(np.stack([np.log2(1 + rw), np.log2(1 + pu)], 1))
    mp = pm.Model(1)
    mp.train(np.concatenate(Xs), np.concatenate(Ys).astype(np.float32), epochs=30, lr=3e-3)
    g = chunglu(10)
    cache = int(g.m * 4 * 0.4) // PKG * PKG
    paging = {"workload": (f"chunglu200k-perm: Chung-Lu n={N}, exponent 2.2, max degree {DMAX}, "
                           f"mean {MEAN}, ids permuted; {cache} B HBM cache of {PKG} B packages "
                           "(40% of rows); rows of <= 1024 entries read from pinned host "
                           "memory when not cached; unlabeled triangle and K4"), "eval_seed": 10,
              "patterns": {}}
    tp = tl = 0.0
    for name, p in ppats:
        res = pm.match(g, p, repeats=3)
        a = pm.match(g, p, cache_bytes=cache, package_bytes=PKG)
        b = pm.match(g, p, cache_bytes=cache, package_bytes=PKG, policy="learned", prefetch=True,
                     model=mp)
        c = pm.match(g, p, cache_bytes=cache, package_bytes=PKG, direct_degree=-1)
        assert a.count == b.count == c.count == res.count
        paging["patterns"][name] = {
            "resident_ms": round(res.min_ms, 2),
            "plain_ms": round(a.kernel_ms, 2), "learned_ms": round(b.kernel_ms, 2),
            "no_direct_reads_ms": round(c.kernel_ms, 2),
            "plain_rounds": a.rounds, "learned_rounds": b.rounds, "no_direct_rounds": c.rounds,
            "plain_link_MB": round(a.link_bytes / 1e6, 2), "learned_link_MB": round(b.link_bytes / 1e6, 2),
            "link_GBps_during_copies": round(a.link_bytes / max(a.link_ms, 1e-9) / 1e6, 2),
            "count": a.count}
        tp += a.kernel_ms
        tl += b.kernel_ms
    paging["plain_ms"] = round(tp, 2)
    paging["learned_ms"] = round(tl, 2)
    paging["speedup"] = round(tp / tl, 3)
    ps = [p for _, p in ppats]
    rb = pm.match_batch(g, ps, cache_bytes=cache, package_bytes=PKG, policy="learned", prefetch=True,
                        model=mp, learned=pm.PM_LEARN_ORDER | pm.PM_LEARN_PAGING)
    assert [r.count for r in rb] == [paging["patterns"][n]["count"] for n, _ in ppats]
    paging["learned_batched_ms"] = round(rb[0].kernel_ms, 2)
    paging["learned_batched_speedup_vs_plain_sequential"] = round(tp / rb[0].kernel_ms, 3)
    out["paging"] = paging
    return out


def copy_calibration():
    """A plain device-to-device copy of known size (the runtime's blit
    kernel), for scaling the FETCH_SIZE and WRITE_SIZE counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = 1 << 30  # bytes
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched and learned objects")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched object")
    ap.add_argument("--no-tune", action="store_true", help="time default launch options")
    ap.add_argument("--copy-calibration", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X (--gpus 1): jobs are independent CSR-edge ranges, so "
              "the matcher would shard by edge ranges with one u64 sum per pattern at the end, "
              "but several GPUs are not run here", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    import numpy as np

    import pumatch_amd as pm

    t0 = time.perf_counter()
    g = pm.Graph.rmat(16, 16, a=0.57, b=0.19, c=0.19, seed=1)
    g.set_uniform_labels(4, 2)
    text = pattern_text(7, C1_EDGES)
    p = pm.Pattern.parse(text, True)
    print(f"[bench] graph n={g.n} m={g.m} built in {time.perf_counter() - t0:.1f}s", flush=True)

    first = pm.match(g, p, per_edge=True)  # warm-up with per-edge counts for the baseline
    opts, tune = {}, None
    if not args.no_tune:
        t0 = time.perf_counter()
        opts, tune = pm.autotune(g, [p], sample=1)
        tune["seconds"] = round(time.perf_counter() - t0, 2)
    for _ in range(max(0, args.warmup - 1)):
        pm.match(g, p, **opts)
    print(f"[bench] warm-up done count={first.count} opts={opts}", flush=True)

    results = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        results.append(pm.match(g, p, **opts))
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    counts = {r.count for r in results} | {first.count}
    assert len(counts) == 1, f"count changed between steps: {counts}"
    count = first.count
    # the same steps as captured-hipGraph replays (first call captures)
    pm.match(g, p, graph=True, **opts)
    t0 = time.perf_counter()
    graph_runs = [pm.match(g, p, graph=True, **opts) for _ in range(args.steps)]
    graph_ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    assert all(r.count == first.count for r in graph_runs)
    untuned = [pm.match(g, p) for _ in range(3)]
    assert all(r.count == count for r in untuned)
    untuned_ms = float(np.mean([r.kernel_ms for r in untuned]))

    # roofline of the dominant kernel phase
    names = ["k_fold", "k_match_edges", "k_match_tasks"]
    ph_ms = np.mean([r.phase_ms for r in results], axis=0)
    ph_bytes = results[-1].phase_alg_bytes
    dom = int(np.argmax(ph_ms))
    achieved = ph_bytes[dom] / (ph_ms[dom] * 1e-3) / 1e9 if ph_ms[dom] > 0 else 0.0
    kernel_ms = float(np.mean([r.kernel_ms for r in results]))
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and dom == 1:
        # measured offline: rocprofv3 FETCH_SIZE / WRITE_SIZE runs, scaled by
        # the same counters on a copy of known size (the file gives the commands)
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch:"):
                    traffic = int(line.split(":")[1])
    roofline = {
        "bound": "hbm",
        "kernel": names[dom],
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": traffic,
        "kernel_ms": round(float(ph_ms[dom]), 4),
        "alg_bytes": int(ph_bytes[dom]),
    }

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        print(f"[bench] cpu baseline ({threads} threads, ~{args.cpu_seconds}s)", flush=True)
        cpu = cpu_baseline(g, text, first.per_edge, count, args.cpu_seconds, threads)

    extras = {}
    if not args.no_extras and not args.no_batched:
        print("[bench] batched: 64 patterns, plain sequential vs tuned batch", flush=True)
        extras["batched"] = batched_section(pm)
    if not args.no_extras:
        print("[bench] learned paging: train on seeds 7-9, evaluate on seed 10", flush=True)
        extras["learned"] = learned_section(pm)

    out = {
        "metric": "ms per 7-vertex pattern (learned+batched vs plain GPU), exact count, HBM GB/s",
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
        "graph_replay_ms_per_step": round(graph_ms_per_step, 4),
        "untuned_kernel_ms": round(untuned_ms, 4),
        "tuned_options": opts,
        "autotune": tune,
        "alg_GBps": round(results[-1].alg_bytes / (kernel_ms * 1e-3) / 1e9, 2),
        "phase_ms": [round(float(x), 4) for x in ph_ms],
        "tasks": results[-1].tasks,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

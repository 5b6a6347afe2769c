"""GPU counts through libpumatch.so against the CPU oracle (bit-exact).

Every test here runs the HIP kernels on the MI355X."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import pumatch_amd as pm
from tests.helpers import (C1_EDGES, C7_EDGES, HT_EDGES, K5P_EDGES, complete_graph, cycle_graph,
                           edges_to_adj, falling, g_text, random_pattern, star_graph)

pytestmark = pytest.mark.gpu

CONFIG_PATTERNS = {"c1": C1_EDGES, "k5p": K5P_EDGES, "ht": HT_EDGES, "c7": C7_EDGES}


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


@pytest.fixture(scope="module")
def rmat9():
    g = pm.Graph.rmat(9, 8, seed=11)
    g.set_uniform_labels(2, 5)
    return g


@pytest.fixture(scope="module")
def rmat8():
    g = pm.Graph.rmat(8, 8, seed=11)
    g.set_uniform_labels(2, 5)
    return g


def check(g, text, labeled, **kw):
    print(f"[gpu parity] n={g.n} m={g.m} labeled={labeled} {kw}", flush=True)
    r = pm.match(g, pm.Pattern.parse(text, labeled), per_edge=True, **kw)
    rp, ci, lb = g.csr()
    want, jobs, pe = oracle.count(rp, ci, lb, oracle.parse_g(text, labeled), labeled, per_edge=True)
    assert r.jobs == jobs
    assert np.array_equal(r.per_edge, pe), "per-edge counts differ"
    assert r.count == want
    return r


@pytest.mark.parametrize("labeled", [True, False])
@pytest.mark.parametrize("i", range(1, 21))
def test_reference_patterns(rmat9, patterns_dir, i, labeled):
    check(rmat9, open(os.path.join(patterns_dir, f"{i}.g")).read(), labeled)


@pytest.mark.parametrize("labeled", [True, False])
@pytest.mark.parametrize("name", sorted(CONFIG_PATTERNS))
def test_config_patterns(rmat8, name, labeled):
    check(rmat8, g_text(7, edges_to_adj(7, CONFIG_PATTERNS[name])), labeled)


@pytest.mark.parametrize("seed", range(16))
def test_random_patterns(rmat9, seed):
    k = 3 + seed % 5
    adj = random_pattern(k, 0.3 + 0.03 * seed, seed)
    labs = [int(x) for x in np.random.default_rng(seed).integers(0, 2, size=k)]
    check(rmat9, g_text(k, adj, labs), True)
    if k <= 6:
        check(rmat9, g_text(k, adj), False)


def test_closed_forms():
    for n in (6, 9):
        rp, ci = complete_graph(n)
        g = pm.Graph.from_csr(rp, ci, labels=np.ones(n, np.uint32))
        for edges, k in ((C1_EDGES, 7), (K5P_EDGES, 7), ([(0, 1), (1, 2)], 3)):
            r = pm.match(g, pm.Pattern.parse(g_text(k, edges_to_adj(k, edges)), False))
            assert r.count == falling(n, k)
        r = pm.match(g, pm.Pattern.parse(g_text(7, edges_to_adj(7, C1_EDGES)), True))
        assert r.count == n * (n - 1) * (n - 2) * (n - 3) * (n - 1) ** 3
    n = 7
    rp, ci = cycle_graph(n)
    g = pm.Graph.from_csr(rp, ci, labels=np.ones(n, np.uint32))
    for k in range(3, 8):
        text = g_text(k, edges_to_adj(k, [(i, (i + 1) % k) for i in range(k)]))
        assert pm.match(g, pm.Pattern.parse(text, False)).count == (2 * n if k == n else 0)
        walks = round(sum((2 * np.cos(2 * np.pi * j / n)) ** k for j in range(n)))
        assert pm.match(g, pm.Pattern.parse(text, True)).count == walks
    rp, ci = star_graph(6)
    g = pm.Graph.from_csr(rp, ci, labels=np.ones(7, np.uint32))
    text = g_text(3, edges_to_adj(3, [(0, 1), (1, 2)]))
    assert pm.match(g, pm.Pattern.parse(text, False)).count == 30
    r = pm.match(g, pm.Pattern.parse(text, True))
    assert (r.count, r.jobs) == (36, 6)


def test_schedule_invariance():
    g = pm.Graph.rmat(10, 8, seed=2)
    g.set_uniform_labels(2, 3)
    texts = [(g_text(7, edges_to_adj(7, C1_EDGES)), True), (g_text(7, edges_to_adj(7, K5P_EDGES)), False),
             (g_text(7, edges_to_adj(7, HT_EDGES)), True)]
    for text, labeled in texts:
        p = pm.Pattern.parse(text, labeled)
        base = pm.match(g, p, per_edge=True)
        for kw in (dict(waves=64), dict(waves=3000, chunk=1), dict(chunk=7, split=-1),
                   dict(split=1), dict(split=5, hub_factor=1), dict(hub_factor=1000),
                   dict(order="file")):
            print(f"[invariance] labeled={labeled} {kw}", flush=True)
            r = pm.match(g, p, per_edge=True, **kw)
            assert r.count == base.count
            assert np.array_equal(r.per_edge, base.per_edge)


def test_c1_configuration_sampled():
    """The benchmark workload (RMAT scale 16, 4 labels, labeled K4 + path):
    per-edge counts on sampled edge ranges against the oracle, and the total
    equal to the sum of per-edge counts."""
    g = pm.Graph.rmat(16, 16, seed=1)
    g.set_uniform_labels(4, 2)
    text = g_text(7, edges_to_adj(7, C1_EDGES))
    p = pm.Pattern.parse(text, True)
    full = pm.match(g, p, per_edge=True)
    assert int(full.per_edge.sum()) == full.count
    rp, ci, lb = g.csr()
    m = g.m
    for lo in (m // 3, m // 2, (3 * m) // 4, m - 3000):
        hi = lo + 1500
        print(f"[c1 sample] edges [{lo}, {hi})", flush=True)
        want, _, pe = oracle.count(rp, ci, lb, oracle.parse_g(text, True), True, edge_lo=lo,
                                   edge_hi=hi, per_edge=True)
        assert np.array_equal(pe[lo:hi], full.per_edge[lo:hi])
        part = pm.match(g, p, edge_lo=lo, edge_hi=hi)
        assert part.count == want


def test_edge_cases():
    # no edges at all
    g = pm.Graph.from_csr(np.zeros(5, np.uint64), np.zeros(0, np.uint32), labels=np.zeros(4, np.uint32))
    p = pm.Pattern.parse(g_text(3, edges_to_adj(3, [(0, 1), (1, 2)])), True)
    r = pm.match(g, p)
    assert (r.count, r.jobs) == (0, 0)
    # labels that no job can satisfy
    rp, ci = complete_graph(6)
    g = pm.Graph.from_csr(rp, ci, labels=np.zeros(6, np.uint32))
    r = pm.match(g, pm.Pattern.parse(g_text(7, edges_to_adj(7, C1_EDGES)), True))
    assert (r.count, r.jobs) == (0, 0)
    # a labeled path folds to the job edge (core of two vertices)
    g = pm.Graph.rmat(9, 8, seed=4)
    g.set_uniform_labels(2, 1)
    for edges, k in (([(0, 1), (1, 2), (2, 3)], 4), ([(0, 1), (0, 2), (0, 3), (1, 4)], 5),
                     ([(0, 1), (1, 2)], 3), ([(0, 1), (1, 2), (0, 2)], 3)):
        check(g, g_text(k, edges_to_adj(k, edges), [1, 0, 1, 0, 1][:k]), True)
        check(g, g_text(k, edges_to_adj(k, edges)), False)


def test_cli_matches_oracle_cli(tmp_path, patterns_dir):
    g = pm.Graph.rmat(9, 8, seed=6, device=False)
    g.set_uniform_labels(2, 9)
    prefix = str(tmp_path / "rmat9")
    g.write(prefix)
    pat = os.path.join(patterns_dir, "7.g")
    for extra in ([], ["--unlabeled"], ["--vertex-induced"], ["--unlabeled", "--vertex-induced"]):
        a = subprocess.run([pm.CLI_PATH, prefix, pat, "--no-result-file"] + extra,
                           capture_output=True, text=True, timeout=120)
        b = subprocess.run([oracle.CLI_PATH, prefix, pat] + extra, capture_output=True, text=True,
                           timeout=120)
        assert a.returncode == 0, a.stderr
        assert b.returncode == 0, b.stderr
        ga = a.stdout.strip().splitlines()[-1].split()
        gb = b.stdout.strip().splitlines()[-1].split()
        assert ga[1] == gb[1]
        float(ga[0])
    r = subprocess.run([pm.CLI_PATH, prefix, pat], capture_output=True, text=True, timeout=120,
                       cwd=tmp_path)
    line = open(tmp_path / "result.txt").read().strip().split("\t")
    assert line[0] == prefix and line[1] == pat and line[3] == r.stdout.split()[-1]


def test_vertex_relabelling_invariance():
    g = pm.Graph.rmat(10, 8, seed=8, device=False)
    rp, ci, _ = g.csr()
    n = rp.size - 1
    perm = np.random.default_rng(1).permutation(n).astype(np.int64)
    src = np.repeat(np.arange(n), np.diff(rp).astype(np.int64))
    a, b = perm[src], perm[ci.astype(np.int64)]
    order = np.lexsort((b, a))
    rp2 = np.zeros(n + 1, np.uint64)
    np.add.at(rp2, a + 1, 1)
    rp2 = np.cumsum(rp2).astype(np.uint64)
    h1 = pm.Graph.from_csr(rp, ci)
    h2 = pm.Graph.from_csr(rp2, b[order].astype(np.uint32))
    for edges in (HT_EDGES, K5P_EDGES):
        p = pm.Pattern.parse(g_text(7, edges_to_adj(7, edges)), False)
        assert pm.match(h1, p).count == pm.match(h2, p).count


@pytest.mark.parametrize("seed", range(6))
def test_core_order_keeps_per_edge_counts(rmat9, seed):
    """Labeled plans may match vertices 2.. in any order; per-job counts are
    the same in file order and in the planner's order, and equal the
    oracle's (file order)."""
    from pumatch_amd.workloads import batch_patterns
    text = batch_patterns(6, seed=40 + seed, nlabels=2, kmin=6, kmax=7)[seed]
    p = pm.Pattern.parse(text, True)
    a = pm.match(rmat9, p, per_edge=True, order="file")
    b = check(rmat9, text, True)
    assert a.count == b.count and np.array_equal(a.per_edge, b.per_edge)

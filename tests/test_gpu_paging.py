"""Out-of-core matching: rows in pinned host memory, paged into a small HBM
package cache between rounds.  Counts must equal the resident run and the
oracle exactly, whatever the cache size, package size and policy."""
import os

import numpy as np
import pytest

import oracle
import pumatch_amd as pm
from tests.helpers import C1_EDGES, HT_EDGES, K5P_EDGES, edges_to_adj, g_text

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graph():
    # 0.8 MB of rows, max degree 168: every row fits one 4 KB package, so a
    # 120-slot (480 KB) cache holds 60% of the graph and must page
    g = pm.Graph.chung_lu(20000, 2.2, 150, 10, seed=4)
    g.set_uniform_labels(2, 4)
    return g


PKG = 4096
SLOTS = 120

CASES = [("c1", C1_EDGES, True), ("k5p", K5P_EDGES, False), ("ht", HT_EDGES, True),
         ("c5", [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)], False),
         ("tri", [(0, 1), (1, 2), (0, 2)], False),
         ("k4", [(a, b) for a in range(4) for b in range(a)], True)]


@pytest.mark.parametrize("policy", ["lru", "fifo", "learned"])
@pytest.mark.parametrize("name,edges,labeled", CASES)
def test_paged_equals_resident(graph, name, edges, labeled, policy):
    k = max(max(e) for e in edges) + 1
    p = pm.Pattern.parse(g_text(k, edges_to_adj(k, edges)), labeled)
    base = pm.match(graph, p, per_edge=True)
    print(f"[paging] {name} labeled={labeled} policy={policy}", flush=True)
    # direct reads off: every missing row defers its search
    r = pm.match(graph, p, per_edge=True, cache_bytes=SLOTS * PKG, package_bytes=PKG, policy=policy,
                 direct_degree=-1)
    print(f"[paging]   rounds={r.rounds} misses={r.misses} deferred={r.deferred} "
          f"link={r.link_bytes} ms={r.kernel_ms:.1f}", flush=True)
    assert r.count == base.count
    assert np.array_equal(r.per_edge, base.per_edge)
    assert r.packages > SLOTS and r.link_bytes > SLOTS * PKG // 2


def test_paging_forces_rounds_and_matches_oracle(graph):
    rp, ci, lb = graph.csr()
    for edges, labeled in ((HT_EDGES, True), ([(0, 1), (1, 2), (0, 2)], False)):
        text = g_text(7 if len(edges) > 3 else 3, edges_to_adj(7 if len(edges) > 3 else 3, edges))
        p = pm.Pattern.parse(text, labeled)
        r = pm.match(graph, p, per_edge=True, cache_bytes=SLOTS * PKG, package_bytes=PKG,
                     direct_degree=-1)
        want, jobs, pe = oracle.count(rp, ci, lb, oracle.parse_g(text, labeled), labeled,
                                      per_edge=True)
        assert r.count == want and r.jobs == jobs and np.array_equal(r.per_edge, pe)
        assert r.rounds > 1 and r.misses > 0 and r.deferred > 0
        # the whole graph fits: one round, no misses
        big = pm.match(graph, p, cache_bytes=64 << 20, package_bytes=PKG)
        assert big.count == want and big.rounds == 1 and big.misses == 0


def test_paged_host_only_graph(tmp_path, patterns_dir):
    """A graph opened without an HBM copy of its rows runs out of core."""
    g = pm.Graph.rmat(10, 8, seed=5, device=False)
    g.set_uniform_labels(2, 1)
    prefix = str(tmp_path / "g")
    g.write(prefix)
    h = pm.Graph.open(prefix, labels=True, device=False)
    p = pm.Pattern.load(os.path.join(patterns_dir, "14.g"), True)
    r = pm.match(h, p, cache_bytes=(72 + 8) * 16384, package_bytes=16384)
    rp, ci, lb = h.csr()
    text = open(os.path.join(patterns_dir, "14.g")).read()
    want, _, _ = oracle.count(rp, ci, lb, oracle.parse_g(text, True), True)
    assert r.count == want


def test_cache_too_small_is_an_error(graph):
    p = pm.Pattern.parse(g_text(3, edges_to_adj(3, [(0, 1), (1, 2)])), True)
    with pytest.raises(pm.PumatchError):
        pm.match(graph, p, cache_bytes=8 * 4096, package_bytes=4096)


@pytest.mark.parametrize("direct", [16, 64, 1024])
@pytest.mark.parametrize("name,edges,labeled", CASES)
def test_direct_short_rows_equal_resident(graph, name, edges, labeled, direct):
    """Missing rows of <= direct entries are read from pinned host memory
    (no deferral for them); counts and per-edge counts stay exact."""
    k = max(max(e) for e in edges) + 1
    p = pm.Pattern.parse(g_text(k, edges_to_adj(k, edges)), labeled)
    base = pm.match(graph, p, per_edge=True)
    print(f"[paging direct] {name} labeled={labeled} direct={direct}", flush=True)
    r = pm.match(graph, p, per_edge=True, cache_bytes=SLOTS * PKG, package_bytes=PKG,
                 direct_degree=direct)
    print(f"[paging direct]   rounds={r.rounds} misses={r.misses} deferred={r.deferred}",
          flush=True)
    assert r.count == base.count
    assert np.array_equal(r.per_edge, base.per_edge)
    if direct >= 168:  # every row of the graph is short enough: nothing misses
        assert r.misses == 0 and r.deferred == 0

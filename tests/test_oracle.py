"""The CPU oracle pinned against direct enumeration and closed forms.

Nothing here needs a GPU; the oracle is oracle/oracle.c (liboracle.so)."""
import os

import numpy as np
import pytest

import oracle
from tests import bruteforce as bf
from tests.helpers import (C1_EDGES, C7_EDGES, HT_EDGES, K5P_EDGES, complete_graph, cycle_graph,
                           edges_to_adj, falling, g_text, random_graph, random_pattern,
                           star_graph)


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


def run(rowptr, colidx, labels, k, adj, labeled, plab=None, **kw):
    p = oracle.parse_g(g_text(k, adj, plab), labeled)
    return oracle.count(rowptr, colidx, labels, p, labeled, **kw)


@pytest.mark.parametrize("seed", range(12))
def test_unlabeled_matches_embedding_enumeration(seed):
    k = 3 + seed % 3
    rowptr, colidx = random_graph(8, 0.55, 100 + seed)
    adj = random_pattern(k, 0.4, seed)
    got, _, _ = run(rowptr, colidx, None, k, adj, False)
    assert got == bf.count_embeddings(rowptr, colidx, k, adj)


@pytest.mark.parametrize("seed", range(12))
def test_labeled_matches_homomorphism_enumeration(seed):
    k = 3 + seed % 3
    rowptr, colidx = random_graph(7, 0.5, 200 + seed)
    rng = np.random.default_rng(seed)
    labels = rng.integers(0, 2, size=7).astype(np.uint32)
    adj = random_pattern(k, 0.35, 50 + seed, adjacent_21=bool(seed % 2))
    plab = [int(x) for x in rng.integers(0, 2, size=k)]
    got, _, _ = run(rowptr, colidx, labels, k, adj, True, plab)
    assert got == bf.count_labeled(rowptr, colidx, labels, k, adj, plab)


def test_seven_vertex_patterns_on_a_small_graph():
    rowptr, colidx = random_graph(8, 0.6, 7)
    for edges in (C1_EDGES, K5P_EDGES, HT_EDGES, C7_EDGES):
        adj = edges_to_adj(7, edges)
        got, _, _ = run(rowptr, colidx, None, 7, adj, False)
        assert got == bf.count_embeddings(rowptr, colidx, 7, adj)


@pytest.mark.parametrize("n", [4, 6, 9])
def test_complete_graph_closed_forms(n):
    rowptr, colidx = complete_graph(n)
    labels = np.ones(n, dtype=np.uint32)
    # unlabeled: every injective map is an embedding -> n!/(n-k)!
    for edges, k in ((C1_EDGES, 7), (K5P_EDGES, 7), ([(0, 1), (1, 2)], 3), ([(0, 1), (1, 2), (0, 2), (2, 3)], 4)):
        got, _, _ = run(rowptr, colidx, None, k, edges_to_adj(k, edges), False)
        assert got == falling(n, k)
    # labeled with one label: homomorphisms into K_n are proper n-colourings,
    # the chromatic polynomial (here every pattern degree fits in n-1)
    if n >= 6:
        c1 = n * (n - 1) * (n - 2) * (n - 3) * (n - 1) ** 3
        got, _, _ = run(rowptr, colidx, labels, 7, edges_to_adj(7, C1_EDGES), True)
        assert got == c1
        cyc = (n - 1) ** 7 - (n - 1)
        got, _, _ = run(rowptr, colidx, labels, 7, edges_to_adj(7, C7_EDGES), True)
        assert got == cyc


@pytest.mark.parametrize("n", [5, 7, 10])
def test_cycle_graph_closed_forms(n):
    rowptr, colidx = cycle_graph(n)
    labels = np.ones(n, dtype=np.uint32)
    for k in range(3, 8):
        edges = [(i, (i + 1) % k) for i in range(k)]
        adj = edges_to_adj(k, edges)
        got, _, _ = run(rowptr, colidx, None, k, adj, False)
        assert got == (2 * n if k == n else 0)
        # closed walks of length k: trace(A^k) = sum_j (2 cos(2 pi j / n))^k
        walks = round(sum((2 * np.cos(2 * np.pi * j / n)) ** k for j in range(n)))
        got, _, _ = run(rowptr, colidx, labels, k, adj, True)
        assert got == walks


def test_star_graph_paths():
    rowptr, colidx = star_graph(6)
    adj = edges_to_adj(3, [(0, 1), (1, 2)])
    got, _, _ = run(rowptr, colidx, None, 3, adj, False)
    assert got == 6 * 5 * 1 * 1 * 2 // 2  # ordered pairs of distinct leaves
    labels = np.ones(7, dtype=np.uint32)
    # job (leaf, centre) needs deg(leaf) >= 1, deg(centre) >= 2, some
    # neighbour of the centre with degree >= 1: 6 jobs x 6 leaves
    got, jobs, _ = run(rowptr, colidx, labels, 3, adj, True)
    assert (got, jobs) == (36, 6)


def test_per_edge_counts_sum_to_total():
    rowptr, colidx = random_graph(9, 0.5, 3)
    adj = edges_to_adj(5, [(0, 1), (1, 2), (0, 2), (2, 3), (3, 4), (1, 4)])
    total, _, pe = run(rowptr, colidx, None, 5, adj, False, per_edge=True)
    p = oracle.parse_g(g_text(5, adj), False)
    aut = oracle.plan_info(p, False)["aut"]
    assert int(pe.sum()) * aut == total


def test_vertex_relabelling_invariance():
    rng = np.random.default_rng(5)
    n = 9
    rowptr, colidx = random_graph(n, 0.55, 11)
    perm = rng.permutation(n)
    edges = [(perm[v], perm[int(x)]) for v in range(n) for x in colidx[rowptr[v]:rowptr[v + 1]]]
    from tests.helpers import csr_from_edges
    rp2, ci2 = csr_from_edges(n, edges)
    for kk, edges_p in ((7, HT_EDGES), (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2)])):
        adj = edges_to_adj(kk, edges_p)
        a, _, _ = run(rowptr, colidx, None, kk, adj, False)
        b, _, _ = run(rp2, ci2, None, kk, adj, False)
        assert a == b


def test_automorphisms_of_reference_patterns(patterns_dir):
    for i in range(1, 21):
        text = open(os.path.join(patterns_dir, f"{i}.g")).read()
        p = oracle.parse_g(text, False)
        adj = [[p.adj[a][b] for b in range(p.k)] for a in range(p.k)]
        assert oracle.plan_info(p, False)["aut"] == bf.automorphisms(p.k, adj)


def test_known_automorphism_counts():
    cases = {(7, tuple(C7_EDGES)): 14, (7, tuple(K5P_EDGES)): 12, (7, tuple(C1_EDGES)): 6,
             (7, tuple(HT_EDGES)): 4}
    for (k, edges), want in cases.items():
        p = oracle.parse_g(g_text(k, edges_to_adj(k, list(edges))), False)
        assert oracle.plan_info(p, False)["aut"] == want


def test_symmetry_pairs_cover_each_orbit_once():
    # K_k: vertex u must be below every earlier vertex
    for k in range(3, 8):
        adj = edges_to_adj(k, [(a, b) for a in range(k) for b in range(a)])
        info = oracle.plan_info(oracle.parse_g(g_text(k, adj), False), False)
        assert info["B"][1:] == [(1 << u) - 1 for u in range(1, k)]
        assert info["rc_order"] == 1


def test_rejects_patterns_the_matcher_cannot_order():
    bad = edges_to_adj(4, [(0, 1), (2, 3), (1, 3)])  # vertex 2 has no earlier neighbour
    p = oracle.parse_g(g_text(4, bad), False)
    with pytest.raises(ValueError):
        oracle.plan_info(p, False)

"""Vertex-induced counts of the CPU oracle (EDGE_INDUCED=false,
src/gpu_match.cu:466-470,570-587,734-738), pinned against direct
enumeration, closed forms on complete graphs and the inclusion-exclusion
identity that links them to edge-induced counts.  No GPU needed."""
from itertools import combinations

import numpy as np
import pytest

import oracle
from tests import bruteforce as bf
from tests.helpers import (C1_EDGES, C7_EDGES, HT_EDGES, K5P_EDGES, complete_graph, cycle_graph,
                           edges_to_adj, falling, g_text, random_graph, random_pattern)


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


def run(rowptr, colidx, labels, k, adj, labeled, plab=None, induced=True, **kw):
    p = oracle.parse_g(g_text(k, adj, plab), labeled)
    return oracle.count(rowptr, colidx, labels, p, labeled, induced=induced, **kw)


@pytest.mark.parametrize("seed", range(12))
def test_unlabeled_induced_matches_enumeration(seed):
    k = 3 + seed % 3
    rowptr, colidx = random_graph(8, 0.5, 300 + seed)
    adj = random_pattern(k, 0.3, 70 + seed)
    got, _, _ = run(rowptr, colidx, None, k, adj, False)
    assert got == bf.count_embeddings(rowptr, colidx, k, adj, induced=True)


@pytest.mark.parametrize("seed", range(12))
def test_labeled_induced_matches_enumeration(seed):
    k = 3 + seed % 3
    rowptr, colidx = random_graph(7, 0.45, 400 + seed)
    rng = np.random.default_rng(seed + 40)
    labels = rng.integers(0, 2, size=7).astype(np.uint32)
    adj = random_pattern(k, 0.3, 90 + seed, adjacent_21=bool(seed % 2))
    plab = [int(x) for x in rng.integers(0, 2, size=k)]
    got, _, _ = run(rowptr, colidx, labels, k, adj, True, plab)
    assert got == bf.count_labeled(rowptr, colidx, labels, k, adj, plab, induced=True)


def test_seven_vertex_induced_on_a_small_graph():
    rowptr, colidx = random_graph(8, 0.55, 17)
    for edges in (C1_EDGES, K5P_EDGES, HT_EDGES, C7_EDGES):
        adj = edges_to_adj(7, edges)
        got, _, _ = run(rowptr, colidx, None, 7, adj, False)
        assert got == bf.count_embeddings(rowptr, colidx, 7, adj, induced=True)


def complete_multipartite_parts(k, adj):
    """Number of parts when the non-edges of the pattern form disjoint
    cliques (the pattern is complete multipartite), else None."""
    parts = []
    for u in range(k):
        for part in parts:
            if all(not adj[u][w] for w in part):
                part.append(u)
                break
        else:
            parts.append([u])
    for part in parts:
        if any(adj[a][b] for a in part for b in part if a != b):
            return None
    for a, b in combinations(range(k), 2):
        same = any(a in part and b in part for part in parts)
        if same == bool(adj[a][b]):
            return None
    return len(parts)


@pytest.mark.parametrize("n", [5, 7])
def test_complete_graph_induced_closed_forms(n):
    rowptr, colidx = complete_graph(n)
    labels = np.ones(n, dtype=np.uint32)
    cases = [(4, [(0, 1), (1, 2), (2, 3), (3, 0)]),             # C4 = K_{2,2}
             (4, [(0, 1), (1, 2), (0, 2), (2, 3)]),             # paw: not multipartite
             (5, [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (0, 4), (1, 4), (2, 4)]),
             (4, [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]),  # K4
             (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)])]     # C5
    seen_nonzero = seen_zero = False
    for k, edges in cases:
        adj = edges_to_adj(k, edges)
        # unlabeled: only the complete pattern embeds induced in K_n
        complete = len(edges) == k * (k - 1) // 2
        got, _, _ = run(rowptr, colidx, None, k, adj, False)
        assert got == (falling(n, k) if complete else 0)
        # labeled homomorphisms into K_n: adjacent vertices take distinct
        # images, non-adjacent ones the same image, so the count is n!/(n-c)!
        # with c parts when the pattern is complete multipartite, else 0
        c = complete_multipartite_parts(k, adj)
        got, _, _ = run(rowptr, colidx, labels, k, adj, True)
        assert got == (falling(n, c) if c else 0)
        seen_nonzero |= bool(c) and c < k
        seen_zero |= c is None
    assert seen_nonzero and seen_zero


@pytest.mark.parametrize("n", [6, 9])
def test_cycle_graph_induced_paths(n):
    # induced paths with k vertices in C_n (n > k): 2n ordered copies, as
    # non-induced; the k-cycle pattern is induced only when k == n
    rowptr, colidx = cycle_graph(n)
    for k in range(3, 6):
        path = edges_to_adj(k, [(i, i + 1) for i in range(k - 1)])
        got, _, _ = run(rowptr, colidx, None, k, path, False)
        assert got == 2 * n
    # a triangle-with-tail pattern never occurs in a cycle
    got, _, _ = run(rowptr, colidx, None, 4, edges_to_adj(4, [(0, 1), (1, 2), (0, 2), (2, 3)]), False)
    assert got == 0


@pytest.mark.parametrize("seed", range(6))
def test_inclusion_exclusion_over_added_edges(seed):
    # #induced embeddings of P = sum over sets S of non-edges of
    # (-1)^|S| * #embeddings of P + S  (each non-edge is a condition
    # "phi(i) not adjacent to phi(j)")
    k = 4 + seed % 2
    rowptr, colidx = random_graph(12, 0.45, 500 + seed)
    adj = random_pattern(k, 0.25, 110 + seed)
    non = [(i, j) for i in range(k) for j in range(i) if not adj[i][j]]
    want = 0
    for r in range(len(non) + 1):
        for S in combinations(non, r):
            a2 = [row[:] for row in adj]
            for i, j in S:
                a2[i][j] = a2[j][i] = 1
            emb, _, _ = run(rowptr, colidx, None, k, a2, False, induced=False)
            want += (-1) ** r * emb
    got, _, _ = run(rowptr, colidx, None, k, adj, False)
    assert got == want
    assert len(non) > 0


def test_induced_never_exceeds_edge_induced_and_equals_it_for_cliques():
    rowptr, colidx = random_graph(14, 0.5, 9)
    rng = np.random.default_rng(3)
    labels = rng.integers(0, 2, size=14).astype(np.uint32)
    for k, edges in ((4, [(0, 1), (1, 2), (2, 3), (3, 0)]), (5, [(0, 1), (1, 2), (0, 2), (2, 3), (3, 4)])):
        adj = edges_to_adj(k, edges)
        for labeled in (False, True):
            lab = labels if labeled else None
            a, _, _ = run(rowptr, colidx, lab, k, adj, labeled, induced=True)
            b, _, _ = run(rowptr, colidx, lab, k, adj, labeled, induced=False)
            assert a <= b
    k4 = edges_to_adj(4, [(a, b) for a in range(4) for b in range(a)])
    for labeled in (False, True):
        lab = labels if labeled else None
        a, _, _ = run(rowptr, colidx, lab, 4, k4, labeled, induced=True)
        b, _, _ = run(rowptr, colidx, lab, 4, k4, labeled, induced=False)
        assert a == b


def test_per_edge_induced_counts_sum_to_total():
    rowptr, colidx = random_graph(10, 0.5, 13)
    adj = edges_to_adj(5, [(0, 1), (1, 2), (2, 3), (3, 4), (1, 4)])
    total, _, pe = run(rowptr, colidx, None, 5, adj, False, per_edge=True)
    aut = oracle.plan_info(oracle.parse_g(g_text(5, adj), False), False)["aut"]
    assert int(pe.sum()) * aut == total


def test_oracle_cli_modes_match_the_library(tmp_path):
    """pm_oracle <prefix> <pattern.g> [--unlabeled] [--vertex-induced] prints
    the same totals as the oracle library for the four modes."""
    import subprocess

    import pumatch_amd as pm
    g = pm.Graph.rmat(8, 8, seed=12, device=False)
    g.set_uniform_labels(2, 3)
    prefix = str(tmp_path / "g")
    g.write(prefix)
    text = g_text(5, edges_to_adj(5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2)]),
                  [0, 1, 0, 1, 1])
    pat = tmp_path / "house.g"
    pat.write_text(text)
    rp, ci, lb = g.csr()
    for labeled in (True, False):
        for induced in (False, True):
            args = [oracle.CLI_PATH, prefix, str(pat)]
            if not labeled:
                args.append("--unlabeled")
            if induced:
                args.append("--vertex-induced")
            out = subprocess.run(args, capture_output=True, text=True, timeout=120)
            assert out.returncode == 0, out.stderr
            got = int(out.stdout.strip().splitlines()[-1].split()[1])
            want, _, _ = oracle.count(rp, ci, lb, oracle.parse_g(text, labeled), labeled,
                                      induced=induced)
            assert got == want, (labeled, induced)

"""Small graph and pattern builders shared by the tests (numpy only)."""
import math
import random

import numpy as np


def csr_from_edges(n, edges):
    """Symmetric, deduplicated, loop-free CSR (rowptr u64, colidx u32)."""
    nb = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            nb[a].add(b)
            nb[b].add(a)
    rowptr = np.zeros(n + 1, dtype=np.uint64)
    for v in range(n):
        rowptr[v + 1] = rowptr[v] + len(nb[v])
    colidx = np.array([x for v in range(n) for x in sorted(nb[v])], dtype=np.uint32)
    return rowptr, colidx


def random_graph(n, p, seed):
    rng = random.Random(seed)
    return csr_from_edges(n, [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < p])


def complete_graph(n):
    return csr_from_edges(n, [(a, b) for a in range(n) for b in range(a + 1, n)])


def cycle_graph(n):
    return csr_from_edges(n, [(i, (i + 1) % n) for i in range(n)])


def star_graph(leaves):
    return csr_from_edges(leaves + 1, [(0, i) for i in range(1, leaves + 1)])


def random_pattern(k, density, seed, adjacent_21=True):
    """Connected pattern in a valid matching order: 1 ~ 0, every u >= 2 has
    an earlier neighbour (and 2 ~ 1 when asked)."""
    rng = random.Random(seed)
    adj = [[0] * k for _ in range(k)]

    def link(a, b):
        adj[a][b] = adj[b][a] = 1

    link(0, 1)
    for u in range(2, k):
        link(u, 1 if (u == 2 and adjacent_21) else rng.randrange(u))
    for a in range(k):
        for b in range(a):
            if rng.random() < density:
                link(a, b)
    return adj


def g_text(k, adj, labels=None):
    labels = labels or [1] * k
    lines = ["# t 0"] + [f"v {i} {labels[i]}" for i in range(k)]
    for a in range(k):
        for b in range(k):
            if a != b and adj[a][b]:
                lines.append(f"e {a} {b} {adj[a][b]}")
    return "\n".join(lines) + "\n"


def edges_to_adj(k, edges):
    adj = [[0] * k for _ in range(k)]
    for a, b in edges:
        adj[a][b] = adj[b][a] = 1
    return adj


def falling(n, k):
    return math.factorial(n) // math.factorial(n - k) if n >= k else 0


# Patterns of the benchmark configurations, in valid matching orders.
C1_EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (3, 4), (4, 5), (5, 6)]
C7_EDGES = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 6), (6, 0)]
K5P_EDGES = [(a, b) for a in range(5) for b in range(a + 1, 5)] + [(3, 5), (4, 6)]
HT_EDGES = [(0, 1), (0, 2), (1, 2), (1, 3), (3, 4), (2, 4), (0, 5), (0, 6), (5, 6)]

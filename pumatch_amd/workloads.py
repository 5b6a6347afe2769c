"""Synthetic pattern sets for the batched workload.

Patterns are written in the reference's .g layout (pattern/*.g: '# t 0',
'v <id> <label>', every edge in both directions) and are valid matching
orders for the reference's planner (src/pattern.h:67-100 reads them in file
order): vertex 1 is adjacent to vertex 0 and every later vertex has an
earlier neighbour.
"""
import random


def random_connected_pattern(k: int, density: float, rng: random.Random):
    """Adjacency matrix of a connected pattern in a valid matching order with
    max(k-1, round(density * k(k-1)/2)) edges: a random spanning tree grown
    in vertex order (u >= 1 picks an earlier parent), then uniformly chosen
    extra pairs."""
    adj = [[0] * k for _ in range(k)]
    for u in range(1, k):
        v = 0 if u == 1 else rng.randrange(u)
        adj[u][v] = adj[v][u] = 1
    want = max(k - 1, round(density * k * (k - 1) / 2))
    rest = [(a, b) for a in range(k) for b in range(a) if not adj[a][b]]
    rng.shuffle(rest)
    for a, b in rest[:want - (k - 1)]:
        adj[a][b] = adj[b][a] = 1
    return adj


def pattern_g_text(adj, labels) -> str:
    k = len(adj)
    lines = ["# t 0"] + [f"v {i} {labels[i]}" for i in range(k)]
    for a in range(k):
        for b in range(k):
            if a != b and adj[a][b]:
                lines.append(f"e {a} {b} 1")
    return "\n".join(lines) + "\n"


def batch_patterns(count: int = 64, seed: int = 6, nlabels: int = 8, kmin: int = 3,
                   kmax: int = 7, dmin: float = 0.3, dmax: float = 0.8):
    """`count` random connected labeled patterns of kmin..kmax vertices with
    edge density (edges / vertex pairs) drawn uniformly from [dmin, dmax] and
    labels uniform over nlabels, all from one seeded stream.  Returns the .g
    texts."""
    rng = random.Random(seed)
    out = []
    for _ in range(count):
        k = rng.randint(kmin, kmax)
        dens = rng.uniform(dmin, dmax)
        adj = random_connected_pattern(k, dens, rng)
        labels = [rng.randrange(nlabels) for _ in range(k)]
        out.append(pattern_g_text(adj, labels))
    return out


def permuted(g, seed: int, device: bool = True):
    """Copy of graph g with its vertex ids relabelled by a seeded random
    permutation (labels move with their vertices).  Generated graphs number
    vertices by expected degree; real inputs do not."""
    import numpy as np

    from . import Graph

    rp, ci, lb = g.csr()
    n = rp.size - 1
    perm = np.random.default_rng(seed).permutation(n).astype(np.int64)
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp.astype(np.int64)))
    a, b = perm[src], perm[ci.astype(np.int64)]
    order = np.lexsort((b, a))
    rp2 = np.zeros(n + 1, np.uint64)
    np.add.at(rp2, a + 1, 1)
    rp2 = np.cumsum(rp2).astype(np.uint64)
    labels = None
    if lb is not None and lb.size:
        labels = np.zeros(n, np.uint32)
        labels[perm] = lb
    return Graph.from_csr(rp2, b[order].astype(np.uint32), labels=labels, device=device)

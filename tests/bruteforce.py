"""Direct enumeration of the count definition on tiny graphs (no plan, no
symmetry breaking, no slots): the independent check the oracle is pinned
against.

labeled:   sum over directed edges (r,c) passing the job filter
           (src/job_queue.h:44-61) of the number of maps phi with
           phi(0)=r, phi(1)=c, lab(phi(u)) == l_P(u) for u >= 2 and every
           pattern edge mapped onto a graph edge (a homomorphism count).
unlabeled: the number of injective maps carrying every pattern edge onto a
           graph edge (embeddings), which the reference's
           |Aut| x symmetry-broken count equals when vertex 2 is adjacent to
           vertex 1 (test.cu:198).
induced:   (either mode) every pattern non-edge must also land on a pair of
           non-adjacent graph vertices (EDGE_INDUCED=false).
"""
from itertools import permutations, product


def adjacency(rowptr, colidx):
    n = len(rowptr) - 1
    return [set(int(x) for x in colidx[rowptr[v]:rowptr[v + 1]]) for v in range(n)]


def pattern_edges(adj, k):
    return [(i, j) for i in range(k) for j in range(i) if adj[i][j]]


def pattern_non_edges(adj, k):
    return [(i, j) for i in range(k) for j in range(i) if not adj[i][j]]


def count_labeled(rowptr, colidx, labels, k, padj, plab, induced=False):
    nb = adjacency(rowptr, colidx)
    n = len(nb)
    deg = [len(s) for s in nb]
    pdeg = [sum(1 for j in range(k) if padj[i][j] > 0) for i in range(k)]
    edges = pattern_edges(padj, k)
    non = pattern_non_edges(padj, k) if induced else []
    total = 0
    for r in range(n):
        for c in sorted(nb[r]):
            if labels[r] != plab[0] or labels[c] != plab[1]:
                continue
            if deg[r] < pdeg[0] or deg[c] < pdeg[1]:
                continue
            if not any(deg[w] >= pdeg[2] for w in nb[c]):
                continue
            cand = [v for v in range(n)]
            for rest in product(cand, repeat=k - 2):
                phi = (r, c) + rest
                if any(labels[phi[u]] != plab[u] for u in range(2, k)):
                    continue
                if all(phi[j] in nb[phi[i]] for i, j in edges) and \
                        not any(phi[j] in nb[phi[i]] for i, j in non):
                    total += 1
    return total


def count_embeddings(rowptr, colidx, k, padj, induced=False):
    nb = adjacency(rowptr, colidx)
    n = len(nb)
    edges = pattern_edges(padj, k)
    non = pattern_non_edges(padj, k) if induced else []
    total = 0
    for phi in permutations(range(n), k):
        if all(phi[j] in nb[phi[i]] for i, j in edges) and \
                not any(phi[j] in nb[phi[i]] for i, j in non):
            total += 1
    return total


def automorphisms(k, padj):
    """Permutations preserving adjacency (every edge label 1)."""
    e = set(pattern_edges(padj, k))
    e |= {(j, i) for i, j in e}
    return sum(1 for p in permutations(range(k)) if {(p[i], p[j]) for i, j in e} == e)

"""Batched launches: many patterns of uneven size in one persistent launch.
Each pattern's count must equal its separate single-pattern run (and the
oracle), whatever the batch composition, order or schedule."""
import numpy as np
import pytest

import oracle
import pumatch_amd as pm
from pumatch_amd.workloads import batch_patterns
from tests.helpers import C1_EDGES, HT_EDGES, K5P_EDGES, edges_to_adj, g_text

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


@pytest.fixture(scope="module")
def rmat10():
    g = pm.Graph.rmat(10, 8, seed=5)
    g.set_uniform_labels(4, 7)
    return g


def mixed_set():
    """Random labeled patterns (3-7 vertices) plus unlabeled ones."""
    texts = [(t, True) for t in batch_patterns(24, seed=6, nlabels=4)]
    # unlabeled: K5 + pendants (2.9e11 matches, cheap through the clique) and
    # a house (5-cycle + chord); the unlabeled house + triangle would mean
    # 2.7e11 matches by enumeration here, minutes for the oracle
    house = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2)]
    texts += [(g_text(7, edges_to_adj(7, K5P_EDGES)), False),
              (g_text(5, edges_to_adj(5, house)), False),
              (g_text(7, edges_to_adj(7, HT_EDGES)), True)]
    texts += [(g_text(3, edges_to_adj(3, [(0, 1), (1, 2), (0, 2)])), False),
              (g_text(4, edges_to_adj(4, [(0, 1), (1, 2), (2, 3), (3, 0)])), False),
              (g_text(7, edges_to_adj(7, C1_EDGES)), True)]
    return texts


def test_batch_equals_single_and_oracle(rmat10):
    texts = mixed_set()
    pats = [pm.Pattern.parse(t, lab) for t, lab in texts]
    print(f"[batch] {len(pats)} patterns n={rmat10.n} m={rmat10.m}", flush=True)
    rs = pm.match_batch(rmat10, pats)
    rp, ci, lb = rmat10.csr()
    for i, ((t, lab), p, r) in enumerate(zip(texts, pats, rs)):
        single = pm.match(rmat10, p)
        want, jobs, _ = oracle.count(rp, ci, lb, oracle.parse_g(t, lab), lab)
        assert (r.count, r.jobs) == (single.count, single.jobs) == (want, jobs), f"pattern {i}"


def test_batch_order_and_schedule_invariance(rmat10):
    texts = mixed_set()
    pats = [pm.Pattern.parse(t, lab) for t, lab in texts]
    base = [r.count for r in pm.match_batch(rmat10, pats)]
    perm = np.random.default_rng(3).permutation(len(pats))
    for kw in (dict(), dict(waves=97, chunk=1), dict(split=3), dict(split=-1, chunk=33),
               dict(hub_factor=1)):
        print(f"[batch invariance] {kw}", flush=True)
        rs = pm.match_batch(rmat10, [pats[i] for i in perm], **kw)
        assert [r.count for r in rs] == [base[i] for i in perm]
    # duplicates, a batch of one
    rs = pm.match_batch(rmat10, [pats[0], pats[0], pats[5]])
    assert [r.count for r in rs] == [base[0], base[0], base[5]]
    assert pm.match_batch(rmat10, [pats[7]])[0].count == base[7]
    assert pm.match_batch(rmat10, []) == []


def test_batch_paged():
    # 0.8 MB of rows in 4 KB packages, 120-slot cache (60% of the graph)
    g = pm.Graph.chung_lu(20000, 2.2, 150, 10, seed=4)
    g.set_uniform_labels(2, 4)
    texts = [(t, True) for t in batch_patterns(8, seed=9, nlabels=2, kmax=5)]
    texts += [(g_text(3, edges_to_adj(3, [(0, 1), (1, 2), (0, 2)])), False),
              (g_text(4, edges_to_adj(4, [(0, 1), (1, 2), (2, 3), (3, 0)])), False)]
    pats = [pm.Pattern.parse(t, lab) for t, lab in texts]
    base = [r.count for r in pm.match_batch(g, pats)]
    print(f"[batch paged] {len(pats)} patterns", flush=True)
    rs = pm.match_batch(g, pats, cache_bytes=120 * 4096, package_bytes=4096, direct_degree=-1)
    assert [r.count for r in rs] == base
    assert rs[0].rounds > 1 and rs[0].link_bytes > 0
    singles = [pm.match(g, p, cache_bytes=120 * 4096, package_bytes=4096).count for p in pats]
    # and with short rows read directly (the default)
    assert [r.count for r in pm.match_batch(g, pats, cache_bytes=120 * 4096,
                                            package_bytes=4096)] == singles
    assert singles == base


def test_batch_per_edge_rejected(rmat10):
    p = pm.Pattern.parse(g_text(3, edges_to_adj(3, [(0, 1), (1, 2)])), True)
    o = pm.Opts()
    pm.lib().pm_opts_default(pm.C.byref(o))
    pe = np.zeros(rmat10.m, np.uint64)
    o.per_edge = pe.ctypes.data_as(pm.C.POINTER(pm.C.c_uint64))
    hs = (pm.C.c_void_p * 2)(p._h.value, p._h.value)
    rs = (pm.Result * 2)()
    assert pm.lib().pm_match_batch(rmat10._h, hs, 2, pm.C.byref(o), rs) == pm.PM_ERR_ARG


def test_sampled_run_counts_sampled_jobs(rmat10):
    """A sampled launch (every s-th block of 4096 edges) counts exactly the
    jobs of those blocks: the sum of the full run's per-edge counts there."""
    g = pm.Graph.rmat(13, 8, seed=5)
    g.set_uniform_labels(4, 7)
    text = batch_patterns(3, seed=6, nlabels=4, kmin=5, kmax=5)[2]
    p = pm.Pattern.parse(text, True)
    full = pm.match(g, p, per_edge=True)
    for s in (2, 3, 7):
        mask = (np.arange(g.m) // 4096) % s == 0
        want = int(full.per_edge[mask].sum())
        print(f"[sample] s={s}", flush=True)
        assert pm.match(g, p, sample=s).count == want
        assert pm.match_batch(g, [p, p], sample=s)[1].count == want


def test_autotune_keeps_counts(rmat10):
    texts = mixed_set()[:12]
    pats = [pm.Pattern.parse(t, lab) for t, lab in texts]
    base = [r.count for r in pm.match_batch(rmat10, pats, order="file")]
    opts, rep = pm.autotune(rmat10, pats, max_orders=3)
    print(f"[autotune] {opts} {rep}", flush=True)
    assert rep["trials"] > 0 and rep["best_ms"] > 0
    assert [r.count for r in pm.match_batch(rmat10, pats, **opts)] == base
    assert [pm.match(rmat10, p, **opts).count for p in pats] == base

"""Resident re-split (pm_opts.resplit_us): a search past its time budget
hands its untried candidates back as pieces that further rounds resume over
all waves.  Counts and per-edge counts must stay exact for every budget,
including budgets so small that nearly every search splits."""
import numpy as np
import pytest

import oracle
import pumatch_amd as pm
from pumatch_amd.workloads import batch_patterns
from tests.helpers import C1_EDGES, C7_EDGES, HT_EDGES, K5P_EDGES, edges_to_adj, g_text

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


@pytest.fixture(scope="module")
def rmat():
    g = pm.Graph.rmat(8, 8, seed=31)
    g.set_uniform_labels(2, 6)
    return g


_WANT = {}


def oracle_count(g, text, labeled):
    # one oracle run per pattern, shared by the budgets
    key = (text, labeled)
    if key not in _WANT:
        rp, ci, lb = g.csr()
        _WANT[key] = oracle.count(rp, ci, lb, oracle.parse_g(text, labeled), labeled,
                                  per_edge=True)
    return _WANT[key]


CASES = [("c7", C7_EDGES, False), ("k5p", K5P_EDGES, False), ("ht", HT_EDGES, True),
         ("ht", HT_EDGES, False), ("c5", [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)], False),
         ("c1", C1_EDGES, True),  # K4 core: splits inside the fused step only
         ("c4", [(0, 1), (1, 2), (2, 3), (3, 0)], False)]


@pytest.mark.parametrize("budget", [1, 20, 0, -1])
@pytest.mark.parametrize("name,edges,labeled", CASES)
def test_resplit_counts_match_oracle(rmat, name, edges, labeled, budget):
    k = max(max(e) for e in edges) + 1
    text = g_text(k, edges_to_adj(k, edges))
    print(f"[resplit] {name} labeled={labeled} budget={budget}", flush=True)
    r = pm.match(rmat, pm.Pattern.parse(text, labeled), per_edge=True, resplit_us=budget)
    want, jobs, pe = oracle_count(rmat, text, labeled)
    print(f"[resplit]   rounds={r.rounds}", flush=True)
    assert r.count == want and r.jobs == jobs
    assert np.array_equal(r.per_edge, pe)
    if budget == 1 and name in ("c7", "k5p"):
        assert r.rounds > 1  # the tiny budget did split


def test_resplit_in_batches_and_graph_mode(rmat):
    texts = batch_patterns(12, seed=41, nlabels=2, kmin=5, kmax=7)
    pats = [pm.Pattern.parse(t, True) for t in texts]
    base = [r.count for r in pm.match_batch(rmat, pats, resplit_us=-1)]
    for budget in (1, 30):
        print(f"[resplit] batch budget={budget}", flush=True)
        assert [r.count for r in pm.match_batch(rmat, pats, resplit_us=budget)] == base
        assert [r.count for r in pm.match_batch(rmat, pats, resplit_us=budget, chunk=1,
                                                split=64)] == base
        for _ in range(2):
            got = [r.count for r in pm.match_batch(rmat, pats, resplit_us=budget, graph=True)]
            assert got == base
            assert pm.match(rmat, pats[0], resplit_us=-1).count == base[0]

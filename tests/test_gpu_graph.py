"""Graph mode (pm_opts.graph): a run's launch sequence captured once as a
hipGraph and replayed.  Every replay must equal the eager run, with other
eager work and synchronisations between the replays."""
import numpy as np
import pytest

import oracle
import pumatch_amd as pm
from tests.helpers import C1_EDGES, edges_to_adj, g_text

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat():
    g = pm.Graph.rmat(11, 8, seed=21)
    g.set_uniform_labels(3, 4)
    return g


TEXTS = [(g_text(7, edges_to_adj(7, C1_EDGES)), True),
         (g_text(4, edges_to_adj(4, [(0, 1), (1, 2), (2, 3), (3, 0)])), False),
         (g_text(5, edges_to_adj(5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2)]),
                 [0, 1, 0, 2, 1]), True)]


def test_replays_equal_eager_runs(rmat):
    rp, ci, lb = rmat.csr()
    pats = [pm.Pattern.parse(t, lab) for t, lab in TEXTS]
    want = [oracle.count(rp, ci, lb, oracle.parse_g(t, lab), lab)[0] for t, lab in TEXTS]
    assert [pm.match(rmat, p).count for p in pats] == want
    for rnd in range(3):
        for i, p in enumerate(pats):
            print(f"[graph] round {rnd} pattern {i}", flush=True)
            r = pm.match(rmat, p, graph=True, repeats=2)
            assert r.count == want[i]
            # eager work of another pattern in between (other tables, arena
            # contents and counters)
            other = pats[(i + 1) % len(pats)]
            assert pm.match(rmat, other).count == want[(i + 1) % len(pats)]
            r = pm.match(rmat, p, graph=True)
            assert r.count == want[i]


def test_batched_replays(rmat):
    pats = [pm.Pattern.parse(t, lab) for t, lab in TEXTS]
    eager = [r.count for r in pm.match_batch(rmat, pats, chunk=4)]
    for rnd in range(4):
        print(f"[graph] batch round {rnd}", flush=True)
        got = [r.count for r in pm.match_batch(rmat, pats, chunk=4, graph=True, repeats=2)]
        assert got == eager
        assert pm.match(rmat, pats[rnd % 3]).count == eager[rnd % 3]
    # other granularities: a new capture, the same counts
    got = [r.count for r in pm.match_batch(rmat, pats, chunk=1, split=-1, graph=True)]
    assert got == eager


def test_graph_mode_falls_back_where_it_does_not_apply(rmat):
    p = pm.Pattern.parse(*TEXTS[0])
    base = pm.match(rmat, p, per_edge=True)
    r = pm.match(rmat, p, per_edge=True, graph=True)  # per-edge counts: eager
    assert r.count == base.count and np.array_equal(r.per_edge, base.per_edge)
    assert r.phase_ms[1] > 0
    g = pm.Graph.chung_lu(20000, 2.2, 150, 10, seed=4)
    g.set_uniform_labels(3, 4)
    q = pm.Pattern.parse(*TEXTS[1])
    want = pm.match(g, q).count
    r = pm.match(g, q, graph=True, cache_bytes=120 * 4096, package_bytes=4096)  # out of core: eager
    assert r.count == want

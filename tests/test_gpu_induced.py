"""Vertex-induced matching (PM_VERTEX_INDUCED, the reference's
EDGE_INDUCED=false mode, src/gpu_match.cu:466-470,570-587,734-738) on the
GPU against the CPU oracle, bit-exact, through libpumatch.so."""
import os

import numpy as np
import pytest

import oracle
import pumatch_amd as pm
from tests.helpers import (C1_EDGES, C7_EDGES, HT_EDGES, K5P_EDGES, complete_graph, edges_to_adj,
                           falling, g_text, random_pattern)

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
    print(f"[gpu induced] n={g.n} m={g.m} labeled={labeled} {kw}", flush=True)
    r = pm.match(g, pm.Pattern.parse(text, labeled, induced=True), per_edge=True, **kw)
    rp, ci, lb = g.csr()
    want, jobs, pe = oracle.count(rp, ci, lb, oracle.parse_g(text, labeled), labeled,
                                  per_edge=True, induced=True)
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
    adj = random_pattern(k, 0.2 + 0.03 * seed, 200 + seed)
    labs = [int(x) for x in np.random.default_rng(seed).integers(0, 2, size=k)]
    check(rmat9, g_text(k, adj, labs), True)
    if k <= 6:
        check(rmat9, g_text(k, adj), False)


@pytest.mark.parametrize("kw", [dict(split=-1), dict(split=64), dict(waves=1), dict(waves=97, chunk=1),
                                dict(chunk=16, hub_factor=2), dict(order="file")])
def test_launch_options(rmat9, kw):
    house = g_text(5, edges_to_adj(5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2)]))
    check(rmat9, house, True, **kw)
    check(rmat9, house, False, **kw)
    c4 = g_text(4, edges_to_adj(4, [(0, 1), (1, 2), (2, 3), (3, 0)]))
    check(rmat9, c4, False, **kw)


def test_cliques_equal_edge_induced(rmat9):
    for k in (3, 4, 5):
        text = g_text(k, edges_to_adj(k, [(a, b) for a in range(k) for b in range(a)]))
        for labeled in (True, False):
            a = pm.match(rmat9, pm.Pattern.parse(text, labeled, induced=True)).count
            b = pm.match(rmat9, pm.Pattern.parse(text, labeled)).count
            assert a == b and a > 0
    path = g_text(4, edges_to_adj(4, [(0, 1), (1, 2), (2, 3)]))
    a = pm.match(rmat9, pm.Pattern.parse(path, False, induced=True)).count
    b = pm.match(rmat9, pm.Pattern.parse(path, False)).count
    assert 0 < a < b


def test_complete_graph_closed_forms():
    n = 7
    rp, ci = complete_graph(n)
    g = pm.Graph.from_csr(rp, ci, labels=np.ones(n, np.uint32))
    k4 = g_text(4, edges_to_adj(4, [(a, b) for a in range(4) for b in range(a)]))
    c4 = g_text(4, edges_to_adj(4, [(0, 1), (1, 2), (2, 3), (3, 0)]))
    paw = g_text(4, edges_to_adj(4, [(0, 1), (1, 2), (0, 2), (2, 3)]))
    assert pm.match(g, pm.Pattern.parse(k4, False, induced=True)).count == falling(n, 4)
    assert pm.match(g, pm.Pattern.parse(c4, False, induced=True)).count == 0
    # labeled homomorphisms of C4 = K_{2,2} into K_n: each side maps to one
    # vertex, the two sides to distinct ones
    assert pm.match(g, pm.Pattern.parse(c4, True, induced=True)).count == falling(n, 2)
    assert pm.match(g, pm.Pattern.parse(paw, True, induced=True)).count == 0


def test_batch_mixes_induced_and_edge_induced(rmat9):
    texts = [g_text(4, edges_to_adj(4, [(0, 1), (1, 2), (2, 3), (3, 0)])),
             g_text(5, edges_to_adj(5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2)]), [0, 1, 0, 1, 1]),
             g_text(3, edges_to_adj(3, [(0, 1), (1, 2)]))]
    pats, want = [], []
    for t in texts:
        for labeled in (True, False):
            for induced in (True, False):
                p = pm.Pattern.parse(t, labeled, induced=induced)
                pats.append(p)
                want.append(pm.match(rmat9, p).count)
    got = [r.count for r in pm.match_batch(rmat9, pats, chunk=4)]
    assert got == want


def test_out_of_core_is_rejected(rmat9):
    p = pm.Pattern.parse(g_text(4, edges_to_adj(4, [(0, 1), (1, 2), (2, 3)])), False, induced=True)
    with pytest.raises(pm.PumatchError, match="resident"):
        pm.match(rmat9, p, cache_bytes=1 << 20, package_bytes=4096)

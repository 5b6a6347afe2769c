"""Learned parts on the GPU: features, BF16-MFMA inference and FP32
training kernels against the float64 oracle, and learned scheduling /
paging keeping every count exact."""
import numpy as np
import pytest

import pumatch_amd as pm
from oracle import mlp
from pumatch_amd.workloads import batch_patterns
from tests.helpers import C1_EDGES, HT_EDGES, edges_to_adj, g_text

pytestmark = pytest.mark.gpu


def rand_model(seed, scale=0.25):
    m = pm.Model(seed)
    m.params = np.random.default_rng(seed).normal(0, scale, pm.N_PARAMS).astype(np.float32)
    return m


def test_features_match_oracle():
    g = pm.Graph.rmat(9, 8, seed=3)
    g.set_uniform_labels(8, 1)
    rp, ci, lb = g.csr()
    got = pm.features(g)
    want = mlp.features(rp, ci, lb)
    assert got.shape == (g.n, 16)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-5)


def test_inference_matches_oracle():
    rng = np.random.default_rng(4)
    m = rand_model(5)
    X = rng.normal(0, 1, (1000, 16)).astype(np.float32)
    Y = m.predict(X)
    want = mlp.forward_bf16(m.params, X)
    exact, _ = mlp.forward(m.params, X)
    # tolerance: BF16 operands, float32 accumulation in another order
    assert np.allclose(Y, want, rtol=2e-3, atol=2e-3)
    assert np.abs(Y - exact).max() < 0.1 * np.abs(exact).max()


def test_gradient_matches_oracle():
    rng = np.random.default_rng(6)
    m = rand_model(7)
    for n in (1, 64, 300):
        X = rng.normal(0, 1, (n, 16)).astype(np.float32)
        Y = rng.normal(0, 1, (n, 2)).astype(np.float32)
        W = rng.uniform(0, 1, (n, 2)).astype(np.float32)
        loss, g = m.grad(X, Y, W)
        wl, wg = mlp.loss_grad(m.params, X, Y, W)
        # float32 kernels vs float64: 1e-4 of the largest component
        assert abs(loss - wl) <= 1e-4 * max(1.0, abs(wl))
        assert np.abs(g - wg).max() <= 1e-4 * np.abs(wg).max(), n


def test_training_fits_a_known_function():
    rng = np.random.default_rng(8)
    X = rng.normal(0, 1, (8192, 16)).astype(np.float32)
    X[:, 15] = 1.0
    Y = np.stack([np.maximum(X[:, 0] + 0.5 * X[:, 1], 0) + 0.3 * X[:, 2],
                  np.abs(X[:, 3]) - 0.2 * X[:, 4]], 1).astype(np.float32)
    m = pm.Model(9)
    before, _ = m.grad(X, Y)
    after = m.train(X, Y, epochs=40, batch=1024, lr=3e-3, seed=2)
    check, _ = m.grad(X, Y)
    print(f"[train] loss {before:.4f} -> {after:.4f} ({check:.4f})", flush=True)
    assert check < 0.1 * before
    r = np.corrcoef(m.predict(X)[:, 0], Y[:, 0])[0, 1]
    assert r > 0.95


def test_learned_order_and_paging_keep_counts():
    g = pm.Graph.chung_lu(20000, 2.2, 150, 10, seed=4)
    g.set_uniform_labels(2, 4)
    m = rand_model(11, 0.1)
    for edges, labeled in ((HT_EDGES, True), ([(0, 1), (1, 2), (0, 2)], False),
                           (C1_EDGES, True)):
        k = max(max(e) for e in edges) + 1
        p = pm.Pattern.parse(g_text(k, edges_to_adj(k, edges)), labeled)
        base = pm.match(g, p, per_edge=True)
        print(f"[learned] k={k} labeled={labeled}", flush=True)
        r = pm.match(g, p, per_edge=True, model=m, learned=pm.PM_LEARN_ORDER)
        assert r.count == base.count and np.array_equal(r.per_edge, base.per_edge)
        for prefetch in (False, True):
            r = pm.match(g, p, cache_bytes=120 * 4096, package_bytes=4096, policy="learned",
                         prefetch=prefetch, model=m, direct_degree=-1)
            assert r.count == base.count and r.rounds > 1


def test_traces():
    g = pm.Graph.chung_lu(20000, 2.2, 150, 10, seed=4)
    g.set_uniform_labels(2, 4)
    p = pm.Pattern.parse(g_text(7, edges_to_adj(7, HT_EDGES)), True)
    r = pm.match(g, p, traces=True, per_edge=True)
    rp, ci, _ = g.csr()
    src = np.repeat(np.arange(g.n), np.diff(rp.astype(np.int64)))
    roots = np.unique(src[r.per_edge > 0])
    assert r.root_work.sum() > 0 and r.root_work.sum() <= r.alg_bytes
    assert np.all(r.root_work[roots] > 0)
    assert not np.any(r.package_use)  # resident: no packages
    q = pm.match(g, p, traces=True, cache_bytes=120 * 4096, package_bytes=4096, direct_degree=-1)
    assert q.count == r.count
    assert q.package_use.max() >= 1 and q.package_use.max() <= q.rounds


def test_trained_model_gain_on_a_fresh_graph():
    """Train on root-work traces of plain runs over permuted RMAT graphs of
    seeds 7-9 and evaluate on seed 10: the network predicts the held-out
    per-root work far better than its untrained self, and the learned job
    order (heaviest predicted roots first) finishes the edge phase sooner
    than CSR order when jobs are not split, with identical counts."""
    from pumatch_amd.workloads import permuted

    def graph(seed):
        g0 = pm.Graph.rmat(14, 16, seed=seed, device=False)
        g = permuted(g0, seed + 100)
        g.set_uniform_labels(4, seed)
        return g

    pats = [pm.Pattern.parse(g_text(4, edges_to_adj(4, [(a, b) for a in range(4)
                                                        for b in range(a)])), False),
            pm.Pattern.parse(g_text(7, edges_to_adj(7, C1_EDGES)), True)]

    def traces(g):
        rw = np.zeros(g.n)
        for p in pats:
            rw += pm.match(g, p, traces=True).root_work
        y = np.stack([np.log2(1 + rw), np.zeros(g.n)], 1).astype(np.float32)
        w = np.stack([np.ones(g.n), np.zeros(g.n)], 1).astype(np.float32)
        return pm.features(g), y, w

    data = [traces(graph(s)) for s in (7, 8, 9)]
    X = np.concatenate([d[0] for d in data])
    Y = np.concatenate([d[1] for d in data])
    W = np.concatenate([d[2] for d in data])
    m = pm.Model(1)
    g = graph(10)
    Xe, Ye, We = traces(g)
    untrained, _ = m.grad(Xe, Ye, We)
    m.train(X, Y, W, epochs=30, lr=3e-3)
    trained, _ = m.grad(Xe, Ye, We)
    r = np.corrcoef(m.predict(Xe)[:, 0], Ye[:, 0])[0, 1]
    print(f"[learned] held-out loss {untrained:.3f} -> {trained:.3f}, r={r:.3f}", flush=True)
    assert trained < 0.2 * untrained and r > 0.8
    tp = tl = 0.0
    for p in pats:
        a = pm.match(g, p, repeats=5, waves=1024, split=-1)
        b = pm.match(g, p, repeats=5, waves=1024, split=-1, model=m, learned=pm.PM_LEARN_ORDER)
        print(f"[learned order] plain {a.min_ms:.3f} ms learned {b.min_ms:.3f} ms", flush=True)
        assert a.count == b.count
        tp += a.min_ms
        tl += b.min_ms
    assert tl < 0.95 * tp

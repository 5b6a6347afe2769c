"""Randomised cross-checks: random graphs, random labeled and unlabeled
patterns of 3-7 vertices, random launch options (waves, chunk, split, hub
factor, order, sampling, batching, out-of-core cache, learned policy) -- every
count against the oracle or against the default run."""
import numpy as np
import pytest

import oracle
import pumatch_amd as pm
from pumatch_amd.workloads import batch_patterns, permuted
from tests.helpers import edges_to_adj, g_text, random_pattern

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


@pytest.mark.parametrize("seed", range(6))
def test_random_options_match_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    g0 = pm.Graph.chung_lu(int(rng.integers(2000, 6000)), 2.2, int(rng.integers(40, 120)), 8,
                           seed=seed, device=False)
    g = permuted(g0, seed)
    g.set_uniform_labels(int(rng.integers(2, 5)), seed)
    rp, ci, lb = g.csr()
    texts = [(t, True, False) for t in batch_patterns(3, seed=200 + seed, nlabels=4, kmax=6)]
    k = int(rng.integers(3, 6))
    texts.append((g_text(k, random_pattern(k, 0.5, seed)), False, False))
    # vertex-induced variants of one labeled and one unlabeled pattern
    texts.append((texts[0][0], True, True))
    texts.append((texts[-2][0], False, True))
    pats = [pm.Pattern.parse(t, lab, induced=ind) for t, lab, ind in texts]
    want = [oracle.count(rp, ci, lb, oracle.parse_g(t, lab), lab, induced=ind)[0]
            for t, lab, ind in texts]
    for trial in range(3):
        kw = {}
        if rng.random() < 0.5:
            kw["waves"] = int(rng.choice([64, 300, 2048]))
        if rng.random() < 0.5:
            kw["chunk"] = int(rng.choice([1, 3, 17, 64]))
        if rng.random() < 0.5:
            kw["split"] = int(rng.choice([-1, 1, 64, 4096]))
        if rng.random() < 0.5:
            kw["hub_factor"] = int(rng.choice([1, 2, 16]))
        if rng.random() < 0.3:
            kw["order"] = "file"
        if rng.random() < 0.4:
            kw["graph"] = True
        if rng.random() < 0.5:
            kw["resplit_us"] = int(rng.choice([-1, 1, 5, 100]))
        print(f"[stress] seed={seed} trial={trial} {kw}", flush=True)
        got = [r.count for r in pm.match_batch(g, pats, **kw)]
        assert got == want, kw
        one = pm.match(g, pats[trial % len(pats)], **kw)
        assert one.count == want[trial % len(pats)]


@pytest.mark.parametrize("seed", range(3))
def test_random_out_of_core_and_learned(seed):
    rng = np.random.default_rng(300 + seed)
    g0 = pm.Graph.chung_lu(20000, 2.2, 150, 10, seed=seed, device=False)
    g = permuted(g0, seed)
    g.set_uniform_labels(3, seed)
    texts = [(t, True) for t in batch_patterns(3, seed=400 + seed, nlabels=3, kmax=5)]
    texts.append((g_text(3, edges_to_adj(3, [(0, 1), (1, 2), (0, 2)])), False))
    pats = [pm.Pattern.parse(t, lab) for t, lab in texts]
    base = [r.count for r in pm.match_batch(g, pats)]
    m = pm.Model(seed)
    slots = int(rng.integers(100, 180))
    for kw in (dict(policy="lru"), dict(policy="fifo"),
               dict(policy="learned", prefetch=True, model=m),
               dict(model=m, learned=pm.PM_LEARN_ORDER)):
        kw["direct_degree"] = int(rng.choice([-1, -1, 0, 8, 40]))
        print(f"[stress ooc] seed={seed} slots={slots} {kw}", flush=True)
        got = [r.count for r in pm.match_batch(g, pats, cache_bytes=slots * 4096,
                                                package_bytes=4096, **kw)]
        assert got == base, kw
    got = [r.count for r in pm.match_batch(g, pats, model=m, learned=pm.PM_LEARN_ORDER)]
    assert got == base

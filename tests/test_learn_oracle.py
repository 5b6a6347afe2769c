"""Pins the CPU oracle of the learned parts (oracle/mlp.py): its gradient
against central finite differences of its own loss, a hand-computed case,
BF16 rounding, and closed forms of the features."""
import numpy as np

from oracle import mlp
from tests.helpers import complete_graph, star_graph


def rand_model(seed, scale=0.3):
    return np.random.default_rng(seed).normal(0, scale, mlp.NPARAMS)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    w = rand_model(2)
    X = rng.normal(0, 1, (23, 16))
    Y = rng.normal(0, 1, (23, 2))
    Wt = rng.uniform(0, 2, (23, 2))
    _, g = mlp.loss_grad(w, X, Y, Wt)
    idx = rng.choice(mlp.NPARAMS, 300, replace=False)
    eps = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        num = (mlp.loss_grad(wp, X, Y, Wt)[0] - mlp.loss_grad(wm, X, Y, Wt)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)), i
    # every parameter block is exercised
    for name, (off, size) in mlp.OFF.items():
        assert np.abs(g[off:off + size]).max() > 0, name


def test_hand_computed_case():
    w = np.zeros(mlp.NPARAMS)
    o = mlp.OFF["b3"][0]
    w[o:o + 2] = [1.0, 2.0]
    loss, g = mlp.loss_grad(w, np.ones((1, 16)), np.zeros((1, 2)))
    assert loss == 2.5  # ((1-0)^2 + (2-0)^2) / 2
    assert np.array_equal(g[o:o + 2], [1.0, 2.0])
    assert np.count_nonzero(g) == 2  # dead ReLUs pass nothing back


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.14159], np.float32)
    r = mlp.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # halfway: ties to even
    assert r[2] == 1.0 + 2 ** -6  # 1 + 3/256 -> 1 + 4/256
    assert r[3] == -2.5 and abs(r[4] - 3.140625) < 1e-12


def test_features_closed_forms():
    n = 9
    rp, ci = complete_graph(n)
    f = mlp.features(rp, ci, np.zeros(n, np.uint32))
    assert np.allclose(f[:, 0], np.log2(n))
    assert np.allclose(f[:, 1], np.log2(1 + (n - 1) ** 2))
    assert np.allclose(f[:, 3], np.log2(1 + (n - 1) * (n - 2) / 2))  # triangles at v
    assert np.allclose(f[:, 6], 1.0) and np.allclose(f[:, 14], np.log2(n))
    rp, ci = star_graph(5)
    f = mlp.features(rp, ci)
    assert f[0, 0] == np.log2(6) and f[0, 3] == 0.0 and np.allclose(f[1:, 2], np.log2(6))
    assert np.allclose(f[:, 15], 1.0) and np.allclose(f[:, 6:15], 0.0)

"""Pins the CPU reference of the learned proposal by central finite
differences in exact FP64 (no rounding), on small graphs."""
import numpy as np

import mlp_ref
from gisbdw_amd import graphs as G


def test_finite_differences_exact_fp64():
    rng = np.random.default_rng(0)
    gs = []
    for s in range(3):
        g = G.gnp_connected(7 + s, 0.4, s)
        adj = np.zeros(64, np.uint64)
        adj[: g.n] = g.adj_array()
        tgt = (1 << 1) | (1 << 3)
        gs.append((g.n, adj, tgt))
    w = mlp_ref.init_weights(3)
    L, grad = mlp_ref.batch_loss_grad(w, gs, emulate=False)
    idx = []
    for k, sh in mlp_ref.SHAPES.items():
        size = int(np.prod(sh))
        idx += list(mlp_ref.OFF[k] + rng.choice(size, size=min(size, 12), replace=False))
    eps = 1e-6
    for p in idx:
        wp, wm = w.copy(), w.copy()
        wp[p] += eps
        wm[p] -= eps
        lp, _ = mlp_ref.batch_loss_grad(wp, gs, emulate=False)
        lm, _ = mlp_ref.batch_loss_grad(wm, gs, emulate=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad[p]) <= 1e-6 + 1e-4 * abs(fd), (p, fd, grad[p])


def test_features_small_cases():
    x = mlp_ref.features(4, G.path(4).adj_array())
    # P4: degrees 1,2,2,1; cores 1; distance-2 sets 2,3,3,2; ecc 3,2,2,3
    assert np.allclose(x[:4, 0] * 4, [1, 2, 2, 1])
    assert np.allclose(x[:4, 1] * 4, [1, 1, 1, 1])
    assert np.allclose(x[:4, 2] * 16, [2, 3, 3, 2])
    assert np.allclose(x[:4, 3] * 8, [3, 2, 2, 3])
    x = mlp_ref.features(5, G.complete(5).adj_array())
    assert np.allclose(x[:5, 1] * 4, 4)

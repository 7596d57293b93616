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


def _curve_fixture():
    import json
    from pathlib import Path

    return json.loads((Path(__file__).resolve().parent / "golden" / "mlp_loss_curve.json").read_text())


def test_loss_curve_fixture(oracle):
    """The committed full-batch Adam trajectory equals the FP64 reference run
    now, on the same training set and the oracle's optimal-root labels; the
    loss falls.  Also pins Adam itself: after the first step every weight with
    a gradient above 1e-4 moves by lr against its gradient's sign."""
    fx = _curve_fixture()
    gs = mlp_ref.curve_dataset()
    assert [g.name for g in gs] == [d["name"] for d in fx["graphs"]]
    assert [[str(a) for a in g.adj] for g in gs] == [d["adj"] for d in fx["graphs"]]
    targets = [mlp_ref.optimal_root_set(oracle, g) for g in gs]
    assert [str(t) for t in targets] == fx["targets"]
    data = [(g.n, np.array(g.adj + [0] * (64 - g.n), np.uint64), t) for g, t in zip(gs, targets)]
    w0 = mlp_ref.init_weights(fx["init_seed"]).astype(np.float32).astype(np.float64)
    curve, w = mlp_ref.adam_train(w0, data, fx["steps"], fx["lr"])
    assert np.allclose(curve, fx["loss_curve"], rtol=1e-12, atol=0)
    assert curve[-1] < curve[0]
    _, g = mlp_ref.batch_loss_grad(w0, data)
    _, w1 = mlp_ref.adam_train(w0, data, 1, fx["lr"])
    big = np.abs(g) > 1e-4
    assert np.allclose((w1 - w0)[big], -fx["lr"] * np.sign(g[big]), rtol=1e-3)

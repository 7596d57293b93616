"""CPU FP64 reference of the learned proposal (test infrastructure).

Mirrors csrc/mlp.hip: the same features, the same network, the same loss,
and -- with ``emulate=True`` -- bf16 rounding at the same points as the GPU
(weights at MFMA inputs, every stored activation, every pre-activation
gradient fed to an MFMA).  With ``emulate=False`` everything is exact FP64,
which is what the finite-difference check differentiates.
"""
import numpy as np

H = 64
OFF = dict(W0=0, B0=256, WS1=320, WN1=4416, B1=8512, WS2=8576, WN2=12672, B2=16768,
           W3=16832, B3=20928, W4=20992, B4=25088, W5=25152, B5=25216)
NPARAMS = 25217
SHAPES = dict(W0=(64, 4), B0=(64,), WS1=(64, 64), WN1=(64, 64), B1=(64,), WS2=(64, 64),
              WN2=(64, 64), B2=(64,), W3=(64, 64), B3=(64,), W4=(64, 64), B4=(64,),
              W5=(64,), B5=(1,))


def bf16(x):
    """Round to bfloat16 (nearest even) through float32, return float64."""
    a = np.asarray(x, np.float64).astype(np.float32).view(np.uint32).astype(np.uint64)
    a = ((a + 0x7FFF + ((a >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return a.view(np.float32).astype(np.float64)


def unpack(w):
    w = np.asarray(w, np.float64)
    return {k: w[OFF[k]: OFF[k] + int(np.prod(SHAPES[k]))].reshape(SHAPES[k]) for k in OFF}


def pack(d):
    out = np.zeros(NPARAMS)
    for k, v in d.items():
        out[OFF[k]: OFF[k] + v.size] = np.ravel(v)
    return out


def features(n, adj):
    """deg/4, core number/4, |vertices within distance 2|/16, ecc/8."""
    rows = [int(adj[v]) for v in range(n)]
    deg = [bin(r).count("1") for r in rows]
    # core numbers by repeated peeling
    core = [0] * n
    k = 1
    while True:
        alive = set(range(n))
        changed = True
        while changed:
            changed = False
            for v in list(alive):
                d = sum(1 for u in alive if rows[v] >> u & 1)
                if d < k:
                    alive.discard(v)
                    changed = True
        if not alive:
            break
        for v in alive:
            core[v] = k
        k += 1
    x = np.zeros((64, 4))
    for v in range(n):
        n2 = rows[v]
        m = rows[v]
        while m:
            u = (m & -m).bit_length() - 1
            m &= m - 1
            n2 |= rows[u]
        n2 &= ~(1 << v)
        seen, front, ecc = 1 << v, 1 << v, 0
        while True:
            nb = 0
            f = front
            while f:
                u = (f & -f).bit_length() - 1
                f &= f - 1
                nb |= rows[u]
            nb &= ~seen
            if not nb:
                break
            seen |= nb
            front = nb
            ecc += 1
        x[v] = [deg[v] * 0.25, core[v] * 0.25, bin(n2).count("1") * 0.0625, ecc * 0.125]
    return x


def forward_backward(w, x, n, adj, target, emulate=True, scale=1.0):
    """Loss of one graph and the gradient of scale * loss (flat vector)."""
    P = unpack(w)
    R = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    A = np.zeros((64, 64))
    for v in range(n):
        for u in range(n):
            if int(adj[v]) >> u & 1:
                A[v, u] = 1.0
    deg = A.sum(1)
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    Amean = A * inv[:, None]

    W0, w5 = R(P["W0"]), R(P["W5"])
    Ws1, Wn1, Ws2, Wn2, W3, W4 = (R(P[k]) for k in ("WS1", "WN1", "WS2", "WN2", "W3", "W4"))
    relu = lambda a: np.maximum(a, 0.0)  # noqa: E731
    h0 = R(relu(x @ W0.T + P["B0"]))
    m1 = R(Amean @ h0)
    h1 = R(relu(h0 @ Ws1.T + m1 @ Wn1.T + P["B1"]))
    m2 = R(Amean @ h1)
    h2 = R(relu(h1 @ Ws2.T + m2 @ Wn2.T + P["B2"]))
    z1 = R(relu(h2 @ W3.T + P["B3"]))
    z2 = R(relu(z1 @ W4.T + P["B4"]))
    s = z2 @ w5 + P["B5"][0]

    mask = np.arange(64) < n
    sm = np.where(mask, s, -np.inf)
    mx = sm.max()
    e = np.where(mask, np.exp(sm - mx), 0.0)
    z = e.sum()
    p = e / z
    t = np.array([1.0 if (int(target) >> v & 1) else 0.0 for v in range(64)])
    t /= t.sum()
    logp = np.where(mask, sm - mx - np.log(z), 0.0)
    loss = -np.sum(t * logp)

    ds = np.where(mask, (p - t) * scale, 0.0)
    G = {}
    G["W5"] = ds @ z2
    G["B5"] = np.array([ds.sum()])
    gA = R(np.outer(ds, w5) * (z2 > 0))

    def dense(gA, X, W, kw, kb):
        G[kw] = gA.T @ X
        G[kb] = gA.sum(0)
        return gA @ W

    F0 = dense(gA, z1, W4, "W4", "B4")
    gA = R(F0 * (z1 > 0))
    F0 = dense(gA, h2, W3, "W3", "B3")
    gA = R(F0 * (h2 > 0))

    def mp(gA, hin, mag, Ws, Wn, ks, kn, kb):
        F = dense(gA, hin, Ws, ks, kb)
        G[kn] = gA.T @ mag
        F1 = gA @ Wn
        return F + Amean.T @ F1

    F0 = mp(gA, h1, m2, Ws2, Wn2, "WS2", "WN2", "B2")
    gA = R(F0 * (h1 > 0))
    F0 = mp(gA, h0, m1, Ws1, Wn1, "WS1", "WN1", "B1")
    da0 = F0 * (h0 > 0)
    G["W0"] = da0.T @ x
    G["B0"] = da0.sum(0)
    return loss, pack(G), s[:n]


def batch_loss_grad(w, graphs, emulate=True):
    """graphs: list of (n, adj rows, target).  Mean loss, gradient of it."""
    L = 0.0
    g = np.zeros(NPARAMS)
    for n, adj, tgt in graphs:
        x = features(n, adj)
        l, gg, _ = forward_backward(w, x, n, adj, tgt, emulate, 1.0 / len(graphs))
        L += l
        g += gg
    return L / len(graphs), g


def init_weights(seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(NPARAMS)
    for k, sh in SHAPES.items():
        if k.startswith("B"):
            w[OFF[k]: OFF[k] + int(np.prod(sh))] = rng.uniform(-0.1, 0.1, int(np.prod(sh)))
        else:
            fan = sh[-1]
            w[OFF[k]: OFF[k] + int(np.prod(sh))] = rng.uniform(-1, 1, int(np.prod(sh))) * np.sqrt(6 / fan)
    return w


def adam_train(w, graphs, steps, lr, b1=0.9, b2=0.999, eps=1e-8, emulate=True):
    """Full-batch Adam with bias correction (Kingma and Ba, 2015, Algorithm 1)
    in FP64: one step per pass over `graphs`.  Returns the loss before each
    step and the final weights."""
    w = np.asarray(w, np.float64).copy()
    m = np.zeros(NPARAMS)
    v = np.zeros(NPARAMS)
    curve = []
    for t in range(1, steps + 1):
        loss, g = batch_loss_grad(w, graphs, emulate)
        curve.append(loss)
        m = b1 * m + (1.0 - b1) * g
        v = b2 * v + (1.0 - b2) * g * g
        w = w - lr * (m / (1.0 - b1 ** t)) / (np.sqrt(v / (1.0 - b2 ** t)) + eps)
    return curve, w


def optimal_root_set(oracle, g):
    """{v : td(G \\ v) = td(G) - 1} of a connected graph, from the oracle DP."""
    from gisbdw_amd import graphs as G

    td = oracle.dp_fast(g)[0]
    pos = 0
    for v in range(g.n):
        keep = [u for u in range(g.n) if u != v]
        rest = G.Graph(g.n - 1, [sum(1 << i for i, u in enumerate(keep) if g.adj[x] >> u & 1)
                                 for x in keep])
        if oracle.dp_fast(rest)[0] == td - 1:
            pos |= 1 << v
    return pos


def curve_dataset():
    """The fixed training set of the loss-curve fixture: 24 connected
    G(n, p), n = 6..13 cycling, p = 0.25 / 0.35 / 0.45 cycling, seeds 0-23."""
    from gisbdw_amd import graphs as G

    return [G.gnp_connected(6 + i % 8, (0.25, 0.35, 0.45)[i % 3], i) for i in range(24)]

"""Pins of the FP64 numpy network reference (CPU): central finite differences
of the loss against the analytic gradient on small random graphs, and
invariances the architecture fixes."""
import numpy as np

from oracle import mpnn_ref as R


def small_graph(n=9, seed=0):
    rng = np.random.default_rng(seed)
    rows = [sorted(set(rng.choice(n, size=rng.integers(1, 4), replace=False)) - {i}) for i in range(n)]
    rows[3] = []  # an isolated node (deg 0)
    # symmetrise: j in N(i) <=> i in N(j)
    adj = [set(r) for r in rows]
    for i, r in enumerate(rows):
        for j in r:
            adj[j].add(i)
    adj[3] = set()
    for i in range(n):
        adj[i].discard(3)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    row_ptr[1:] = np.cumsum([len(a) for a in adj])
    col = np.concatenate([sorted(a) for a in adj]).astype(np.int64)
    efeat = rng.standard_normal((len(col), R.E_IN))
    x = rng.standard_normal((n, R.F_IN))
    t = rng.standard_normal((n, R.OUT))
    w = (rng.random(n) > 0.2).astype(np.float64)
    return row_ptr, col, efeat, x, t, w


def test_gradient_matches_central_differences():
    row_ptr, col, efeat, x, t, w = small_graph()
    rng = np.random.default_rng(1)
    theta = rng.standard_normal(R.n_params()) * 0.3
    loss, g = R.loss_and_grad(theta, row_ptr, col, efeat, x, t, w)
    idx = rng.choice(R.n_params(), size=60, replace=False)
    # every parameter block is probed at least once
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in R.shapes()])
    idx = np.unique(np.concatenate([idx, offs[:-1]]))
    h = 1e-6
    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        lp = R.loss_and_grad(tp, row_ptr, col, efeat, x, t, w)[0]
        lm = R.loss_and_grad(tm, row_ptr, col, efeat, x, t, w)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 2e-5 * max(1.0, abs(fd)), (k, fd, g[k])


def test_isolated_node_and_permutation_equivariance():
    row_ptr, col, efeat, x, t, w = small_graph(seed=3)
    theta = np.random.default_rng(2).standard_normal(R.n_params()) * 0.2
    y = R.forward(theta, row_ptr, col, efeat, x)
    # node 3 has no edges: its message is zero, so only its own features matter
    p = R.unpack(theta)
    h = R.silu(x[3] @ p["W0"] + p["b0"])
    for l in range(R.N_LAYERS):
        h = h + R.silu(np.concatenate([h, np.zeros(R.HID)]) @ p[f"W2_{l}"] + p[f"b2_{l}"])
    assert np.allclose(y[3], h @ p["Wd"] + p["bd"], rtol=1e-12)
    # relabelling nodes permutes the output
    n = len(row_ptr) - 1
    perm = np.random.default_rng(5).permutation(n)
    inv = np.argsort(perm)
    new_rows = [sorted(inv[col[row_ptr[perm[i]]:row_ptr[perm[i] + 1]]]) for i in range(n)]
    new_eidx = []
    for i in range(n):
        old = perm[i]
        seg = list(range(row_ptr[old], row_ptr[old + 1]))
        seg.sort(key=lambda k: inv[col[k]])
        new_eidx += seg
    rp2 = np.zeros(n + 1, dtype=np.int64)
    rp2[1:] = np.cumsum([len(r) for r in new_rows])
    col2 = np.concatenate([np.array(r, dtype=np.int64) for r in new_rows])
    y2 = R.forward(theta, rp2, col2, efeat[new_eidx], x[perm])
    assert np.allclose(y2, y[perm], rtol=1e-12, atol=1e-12)

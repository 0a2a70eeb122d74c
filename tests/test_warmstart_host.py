"""Host-side feature construction of the learned ADR warm start."""
import numpy as np

from pd_gpu1_amd import warmstart as W


def test_damage_is_broken_fraction_per_row():
    rp = np.array([0, 3, 3, 5])
    mask = np.array([1, 0, 1, 0, 0], np.uint8)
    np.testing.assert_allclose(W.damage(rp, mask), [1.0 / 3.0, 0.0, 1.0])


def test_node_and_edge_features_layout():
    pos = np.array([[-1.0, 0.0, 1.0], [0.5, 1.0, -1.0]])
    dp = np.array([[1e-6, 0.0, 0.0], [0.0, -1e-6, 0.0]])
    top = np.array([False, True, False])
    bot = np.array([False, False, True])
    f = W.node_features(pos, dp, np.array([0.0, 0.5, 1.0]), top, bot)
    assert f.shape == (3, 8) and f.dtype == np.float32
    np.testing.assert_allclose(f[:, 0:2], pos.T)
    np.testing.assert_allclose(f[:, 2], [1.0, 0.0, 0.0])
    np.testing.assert_allclose(f[:, 3], [0.0, -1.0, 0.0])
    np.testing.assert_allclose(f[:, 4], [0.0, 0.5, 1.0])
    np.testing.assert_array_equal(f[:, 5:8], [[0, 0, 1], [1, 0, 1], [0, 1, 1]])
    x = np.array([0.0, 1e-3, 0.0])
    y = np.array([0.0, 0.0, 2e-3])
    rp = np.array([0, 2, 3, 4])
    col = np.array([1, 2, 0, 0])
    ef = W.edge_features(x, y, rp, col, np.array([1, 0, 1, 1], np.uint8), delta=2e-3)
    np.testing.assert_allclose(ef, [[0.5, 0, 1], [0, 1, 0], [-0.5, 0, 1], [0, -1, 1]])


def test_positions_are_centred_and_scaled():
    x, y, n_int, _, _ = W.plate_layout(10)
    pos = W.normalized_positions(x, y, n_int)
    np.testing.assert_allclose(pos[:, :n_int].mean(axis=1), 0.0, atol=1e-12)
    np.testing.assert_allclose(np.abs(pos[:, :n_int]).max(axis=1), 1.0)
    assert pos[1, n_int:n_int + 30].min() > 1.0 and pos[1, n_int + 30:].max() < -1.0


def test_samples_are_first_increments_and_post_break_increments():
    n = 4
    x, y, n_int, top, bot = W.plate_layout(n)
    N = len(x)
    rp = np.arange(N + 1, dtype=np.int32)  # one edge per node
    col = np.roll(np.arange(N, dtype=np.int32), 1)
    rng = np.random.default_rng(0)
    us = [np.zeros((2, N))]
    for k in range(4):
        us.append(us[-1] + W.DU * (1.0 + 0.1 * k) * rng.uniform(size=(2, N)))
    intact = np.ones(N, np.uint8)
    broken = intact.copy()
    broken[3] = 0
    # a bond breaks during the second increment (between states 1 and 2)
    rec = dict(n=n, n_int=n_int, rp=rp, col=col, x=x, y=y, top=top, bot=bot, u=us,
               mask=[intact, intact, broken, broken, broken])
    out = list(W.samples(rec, delta=1.5e-3))
    assert len(out) == 2  # increment 1 (no history) and increment 3 (after the break)
    base = W.base_increment(x, y, n_int, top)
    x0, ef0, t0, w = out[0]
    np.testing.assert_allclose(t0, ((us[1] - us[0] - base) / W.DU).T, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(x0[:, 2:4], 0.0)
    x1, ef1, t1, _ = out[1]
    np.testing.assert_allclose(t1, ((us[3] - us[2] - base) / W.DU).T, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(x1[:, 2:4], ((us[2] - us[1]) / W.DU).T, rtol=1e-5)
    assert ef1[3, 2] == 0 and ef0[3, 2] == 1
    assert w[:n_int].all() and not w[n_int:].any()
    assert W.uses_network(0, intact, None) and not W.uses_network(2, broken, broken)


def test_base_increment_moves_strips_by_du():
    x, y, n_int, top, bot = W.plate_layout(12)
    b = W.base_increment(x, y, n_int, top)
    assert np.all(b[0] == 0.0)
    np.testing.assert_allclose(b[1, top].max(), W.DU)
    np.testing.assert_allclose(b[1, bot].min(), -W.DU)
    assert np.all(np.abs(b[1]) <= W.DU) and np.all(np.abs(b[1, :n_int]) < W.DU)
    order = np.argsort(y, kind="stable")
    assert np.all(np.diff(b[1, order]) >= 0.0)
    np.testing.assert_allclose(b[1, :n_int].mean(), 0.0, atol=1e-20)

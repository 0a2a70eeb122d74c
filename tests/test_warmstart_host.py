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
    crack = np.arange(24, dtype=np.float64).reshape(8, 3)
    f = W.node_features(pos, dp, np.array([0.0, 0.5, 1.0]), top, bot, crack)
    assert f.shape == (3, 16) and f.dtype == np.float32
    np.testing.assert_allclose(f[:, 8:16], crack.T)
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
    rec = dict(n=n, n_int=n_int, rp=rp, col=col, x=x, y=y, top=top, bot=bot, u=us, clen=1e-3, angle=0.3,
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


def test_crack_features_follow_the_crack_frame():
    half = 2.0e-3
    clen, ang = 2.0e-3, 0.4  # crack half-length a = 0.5 half sides
    c, s = np.cos(ang), np.sin(ang)
    # along the crack: its centre, a tip, beyond the tip; across it: just above and below the centre
    along = np.array([0.0, 0.5, 1.0])
    pos = np.concatenate([np.stack([along * c, along * s]), np.array([[-s * 0.05, s * 0.05], [c * 0.05, -c * 0.05]])],
                         axis=1)
    f = W.crack_features(pos, clen, ang, half)
    np.testing.assert_allclose(f[0, :3], [0.0, 1.0 / 3.0, 2.0 / 3.0], atol=1e-12)   # s / 3
    np.testing.assert_allclose(f[1, 3:], [0.1 / 3.0, -0.1 / 3.0], atol=1e-12)      # t / 3
    np.testing.assert_allclose(f[2, 3:], [np.sqrt(1.0), -np.sqrt(1.0)], atol=1e-12)  # opening: +1 above, -1 below
    assert f[2, 2] == 0.0                                                          # beyond the tip
    np.testing.assert_allclose(f[3], 0.5)
    np.testing.assert_allclose(f[4:6, 0], [c, s])
    np.testing.assert_allclose(f[6, 1], 1.0)                                       # at the tip
    np.testing.assert_allclose(f[7, 0], 1.0)                                       # at the centre
    assert not W.crack_features(pos, 0.0, ang, half).any()


def test_mirror_maps_are_point_reflections():
    from types import SimpleNamespace
    n = 6
    x, y, n_int, top, bot = W.plate_layout(n)
    P = np.stack([x, y], 1)
    d2 = ((P[:, None] - P[None]) ** 2).sum(-1)
    N = len(x)
    rows = [np.nonzero((d2[i] < (3.015 * W.DX) ** 2) & (np.arange(N) != i))[0] for i in range(N)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    plate = SimpleNamespace(x=x, y=y, n_int=n_int, rp=rp, col=col, _mir=None)
    mir, bmir = W.Plate.mirror_maps(plate)
    cx, cy = x[:n_int].mean(), y[:n_int].mean()
    np.testing.assert_allclose(x[mir] - cx, -(x - cx), atol=1e-15)
    np.testing.assert_allclose(y[mir] - cy, -(y - cy), atol=1e-15)
    assert np.array_equal(mir[mir], np.arange(N)) and np.all(mir[:n_int] < n_int)
    assert np.array_equal(mir[top], np.sort(np.nonzero(bot)[0])[::-1]) or set(mir[top]) == set(np.nonzero(bot)[0])
    r = np.repeat(np.arange(N), np.diff(rp))
    assert np.array_equal(r[bmir], mir[r]) and np.array_equal(col[bmir], mir[col])
    assert np.array_equal(bmir[bmir], np.arange(len(col)))


class _FakeProb:
    """Stands in for a handle: converges after `need` iterations."""

    def __init__(self, need, delay=0.0):
        self.need, self.delay, self.calls, self.u = need, delay, [], None

    def set(self, name, u):
        self.u = u

    def adr_solve(self, ct0, n, tol):
        import time
        time.sleep(self.delay)
        self.calls.append((ct0, n))
        left = self.need - ct0
        if left <= n:
            return left, tol * 0.5, True
        return n, 1.0, False


class _FakePlate:
    def __init__(self, need, delay=0.0):
        self.prob = _FakeProb(need, delay)


def test_race_returns_the_first_run_to_converge():
    a, b = _FakePlate(700, 0.002), _FakePlate(300, 0.002)
    its, r, b_won = W.race(a, "ga", b, "gb", 1e-6, 10000, chunk=100)
    assert b_won and its == 300 and r < 1e-6
    assert a.prob.u == "ga" and b.prob.u == "gb"
    # chunks continue the iteration count; the loser stopped early
    assert [c[0] for c in b.prob.calls] == [0, 100, 200]
    assert len(a.prob.calls) < 7


def test_race_keeps_the_guess_when_it_wins_or_both_fail():
    a, b = _FakePlate(250), _FakePlate(900, 0.01)
    its, r, b_won = W.race(a, "ga", b, "gb", 1e-6, 10000, chunk=100)
    assert not b_won and its == 250
    a, b = _FakePlate(10**6), _FakePlate(10**6)
    its, r, b_won = W.race(a, "ga", b, "gb", 1e-6, 500, chunk=128)
    assert not b_won and its == 500
    assert [c[1] for c in a.prob.calls] == [128, 128, 128, 116]


def test_compact_first_increment_sample_equals_the_full_record():
    n = 4
    x, y, n_int, top, bot = W.plate_layout(n)
    N = len(x)
    P = np.stack([x, y], 1)
    d2 = ((P[:, None] - P[None]) ** 2).sum(-1)
    rows = [np.nonzero((d2[i] < (3.015 * W.DX) ** 2) & (np.arange(N) != i))[0] for i in range(N)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    rng = np.random.default_rng(3)
    mask = (rng.uniform(size=len(col)) > 0.2).astype(np.uint8)
    u1 = W.DU * rng.uniform(size=(2, N))
    delta = 3.015 * W.DX
    rec = dict(n=n, n_int=n_int, rp=rp, col=col, x=x, y=y, top=top, bot=bot, u=[np.zeros((2, N)), u1],
               clen=1e-3, angle=0.2, mask=[mask, mask])
    x_full, ef_full, t_full, w_full = next(W.samples(rec, delta))
    side = W.SideGraph(n, rp, col, delta)
    comp = dict(n=n, clen=1e-3, angle=0.2, mask=np.packbits(mask, bitorder="little"), nb=len(col),
                t=((u1 - side.base) / W.DU).T.astype(np.float32))
    xc, efc, tc, wc = W.first_sample(comp, side)
    np.testing.assert_array_equal(xc, x_full)
    np.testing.assert_array_equal(efc, ef_full)
    np.testing.assert_array_equal(tc, t_full)
    np.testing.assert_array_equal(wc, w_full)


def test_epoch_order_is_a_permutation_sorted_by_side_within_windows():
    recs = [dict(n=int(v)) for v in np.random.default_rng(5).integers(40, 200, size=1000)]
    order = W.epoch_order(recs, np.random.default_rng(0), window=128)
    assert sorted(order) == list(range(1000))
    for k in range(0, 1000, 128):
        sides = [recs[i]["n"] for i in order[k:k + 128]]
        assert sides == sorted(sides)


class _ProbeProb:
    """Stands in for a handle: the residual after k updates from start `name`
    is r0 * f**k; ADR stops once it is <= tol."""

    def __init__(self, rates):
        self.rates, self.cur, self.k, self.masks = rates, None, 0, 0

    def set(self, field, u):
        self.cur, self.k = u, 0

    def set_fail_mask(self, mask):
        self.masks += 1

    def adr_solve(self, ct0, n, tol, nonfinite_ok=False):
        assert ct0 == self.k  # a continuation carries on the iteration count
        r0, f = self.rates[self.cur]
        for i in range(n):
            if r0 * f ** self.k <= tol:
                return i, r0 * f ** self.k, True
            self.k += 1
        return n, r0 * f ** self.k, False


def test_probe_rule_keeps_a_better_guess_and_counts_every_iteration():
    prob = _ProbeProb({"plain": (1e-2, 0.99), "guess": (1e-3, 0.98)})
    its, r, fell_back = W.probe_then_solve(prob, "guess", "plain", None, 50, 1e-6, 100000)
    need = int(np.ceil(np.log(1e-3) / np.log(0.98)))  # updates from the guess to reach 1e-6
    assert not fell_back and r <= 1e-6 and its == 50 + need and prob.cur == "guess"


def test_probe_rule_falls_back_when_the_guess_residual_is_larger():
    prob = _ProbeProb({"plain": (1e-2, 0.99), "guess": (1e-1, 0.995)})
    its, r, fell_back = W.probe_then_solve(prob, "guess", "plain", None, 50, 1e-6, 100000)
    need = int(np.ceil(np.log(1e-4) / np.log(0.99)))
    assert fell_back and r <= 1e-6 and its == 50 + 50 + need and prob.cur == "plain"
    assert prob.masks == 2  # the fail mask is reset before the guess's run and before the plain rerun
    # a plain probe that converges ends the increment
    prob = _ProbeProb({"plain": (1e-5, 0.5), "guess": (1.0, 0.5)})
    its, r, fell_back = W.probe_then_solve(prob, "guess", "plain", None, 50, 1e-6, 100000)
    assert fell_back and its == 4 and prob.cur == "plain"

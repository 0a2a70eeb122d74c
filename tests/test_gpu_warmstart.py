"""Learned ADR warm start end to end on small plates: batched-ADR training
data, training on the GPU, and warm-started ADR that reaches the same
tolerance as plain ADR in fewer iterations."""
import numpy as np
import pytest

from pd_gpu1_amd import batch as PB
from pd_gpu1_amd import warmstart as W
from pd_gpu1_amd.mpnn import Mpnn

pytestmark = pytest.mark.gpu


def test_warm_start_converges_in_fewer_iterations():
    specs = PB.plate_family(12, seed_base=0x5EED0004, n_min=24, n_max=40)
    recs, iters, _ = W.generate(specs, 3)
    assert iters.shape == (3, 12) and (iters > 0).all()
    with Mpnn(seed=11) as net:
        p = PB.plate_params(W.DX)
        hist = W.train(net, recs, p.delta, epochs=30, lr=3e-3, lr_end=3e-4)
        print("loss per epoch", [round(h, 4) for h in hist])
        assert hist[-1] < 0.5 * hist[0]
        rows = W.evaluate(net, [(36, 36, 0.3 * 36 * W.DX, 0.2)], n_inc=4, log=print)
    rows = np.array(rows, dtype=float)
    first, later = rows[rows[:, 1] == 0], rows[rows[:, 1] >= 1]
    assert first[:, 10].all() and not later[:, 10].any()  # network only without history
    # from the second increment on, the previous increment predicts the next
    assert later[:, 3].sum() < 0.7 * later[:, 2].sum()
    assert later[:, 4].sum() < 0.7 * later[:, 2].sum()


def test_warm_and_plain_reach_the_same_state():
    n = 30
    plain = W.Plate(n, 0.3 * n * W.DX, -0.3)
    warm = W.Plate(n, 0.3 * n * W.DX, -0.3)
    try:
        with Mpnn(seed=3) as net:
            net.set_graph(warm.rp, warm.col)
            for _ in range(3):
                _, r_p, r_probe, _ = plain.step_plain()
                _, r_w, _, _, _, _ = warm.step_warm(net, r_probe, use_net=True)
                assert r_p <= 1e-6 and r_w <= 1e-6
        up, uw = plain.prob.get("U"), warm.prob.get("U")
        assert np.max(np.abs(up - uw)) < 1e-3 * np.max(np.abs(up))
        assert np.array_equal(plain.prob.fail_mask(), warm.prob.fail_mask())
    finally:
        plain.close()
        warm.close()

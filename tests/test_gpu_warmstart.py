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
                _, r_w, _, _, _, _ = warm.step_warm(net, use_net=True)
                assert r_p <= 1e-6 and r_w <= 1e-6
        up, uw = plain.prob.get("U"), warm.prob.get("U")
        assert np.max(np.abs(up - uw)) < 1e-3 * np.max(np.abs(up))
        assert np.array_equal(plain.prob.fail_mask(), warm.prob.fail_mask())
    finally:
        plain.close()
        warm.close()


def test_two_handles_run_from_two_threads():
    # each handle keeps its copies on its own stream, so one thread may
    # capture a handle's step graph while another copies into its handle;
    # the results equal those of the same calls made one after the other
    import threading
    n = 40
    plates = [W.Plate(n, 0.3 * n * W.DX, 0.1 * k) for k in (1, 2, 1, 2)]
    tols = (1e-6, 1e-7)

    def load(P, tol, out):
        for _ in range(3):
            u = P.prob.get("U")
            u[1, P.top] += W.DU
            u[1, P.bot] -= W.DU
            P.prob.set("U", u)
            its, r, ok = P.prob.adr_solve(0, 200000, tol)
            assert ok, (its, r)
            out.append(its)

    try:
        seq = [[], []]
        for k in range(2):
            load(plates[k], tols[k], seq[k])
        par = [[], []]
        th = [threading.Thread(target=load, args=(plates[2 + k], tols[k], par[k])) for k in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert par == seq
        for k in range(2):
            assert np.array_equal(plates[k].prob.get("U"), plates[2 + k].prob.get("U"))
            assert np.array_equal(plates[k].prob.fail_mask(), plates[2 + k].prob.fail_mask())
    finally:
        for P in plates:
            P.close()


def test_shipped_network_on_a_large_unseen_plate():
    # the shipped weights on the first plate of the unseen evaluation family
    # (315 x 315 free particles, 101,115 with the strips): the network's start
    # reaches tol in fewer iterations than the uniform-stretch start, and both
    # end in the state plain ADR reaches, with the same broken bonds
    import os
    wpath = os.path.join(os.path.dirname(W.__file__), "data", "mpnn_warmstart.npy")
    n, _, clen, ang = W.eval_specs(1)[0]
    assert n == 315
    plates = [W.Plate(n, clen, ang) for _ in range(3)]
    plain, extrap, warm = plates
    try:
        with Mpnn() as net:
            net.set_params(np.load(wpath))
            net.set_graph(warm.rp, warm.col)
            ip, rp_, _, _ = plain.step_plain()
            ie, re_, _, fe, _, used_e = extrap.step_warm(None, use_net=False)
            iw, rw, _, fw, _, used_w = warm.step_warm(net, use_net=True)
        print("iterations plain", ip, "uniform stretch", ie, "network", iw)
        assert used_w and not used_e and not fw and not fe
        assert max(rp_, re_, rw) <= 1e-6
        assert iw < ie and iw < ip
        up = plain.prob.get("U")
        for P in (extrap, warm):
            d = np.max(np.abs(P.prob.get("U") - up)) / np.max(np.abs(up))
            print("max |u - u_plain| / max |u_plain|", d)
            assert d < 1e-3
            assert np.array_equal(P.prob.fail_mask(), plain.prob.fail_mask())
    finally:
        for P in plates:
            P.close()

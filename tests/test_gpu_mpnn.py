"""The BF16/MFMA message-passing network against the FP64 numpy network on
the same parameters, graph and features.

BF16 operands carry 8 significant bits, so forward outputs are compared at
3e-2 of the output scale and the parameter gradient by direction (cosine
> 0.99) and norm (within 10 %).  Training must reduce the loss."""
import numpy as np
import pytest

from pd_gpu1_amd import _lib
from pd_gpu1_amd.mpnn import N_FEAT, Mpnn
from oracle import mpnn_ref as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _graph(n=14, dx=1e-3):
    x, y = O.coord_plate(n, n, dx)
    rp, col, _ = O.neighbors_cell(2, x, y, None, 3.015 * dx)
    rows = np.repeat(np.arange(len(x)), np.diff(rp))
    dvec = np.stack([x[col] - x[rows], y[col] - y[rows]], axis=1) / (3.015 * dx)
    rng = np.random.default_rng(3)
    intact = (rng.uniform(size=len(col)) > 0.1).astype(np.float64)
    ef = np.concatenate([dvec, intact[:, None]], axis=1)
    xf = np.concatenate([rng.uniform(-1, 1, (len(x), N_FEAT - 1)), np.ones((len(x), 1))], axis=1)
    return rp, col, ef, xf


def test_forward_matches_fp64_network():
    rp, col, ef, xf = _graph()
    with Mpnn(seed=7) as net:
        theta = net.get_params().astype(np.float64)
        assert len(theta) == R.n_params()
        net.set_graph(rp, col)
        net.set_features(xf, ef)
        y = net.forward()
    y_ref = R.forward(theta, rp, col, ef, xf)
    scale = np.max(np.abs(y_ref))
    assert np.max(np.abs(y - y_ref)) / scale < 3e-2, np.max(np.abs(y - y_ref)) / scale
    assert np.corrcoef(y.ravel(), y_ref.ravel())[0, 1] > 0.999


def test_gradient_matches_fp64_network():
    rp, col, ef, xf = _graph()
    rng = np.random.default_rng(4)
    tgt = rng.normal(size=(len(xf), 2))
    w = rng.uniform(0.5, 1.5, len(xf))
    with Mpnn(seed=8) as net:
        theta = net.get_params().astype(np.float64)
        net.set_graph(rp, col)
        net.set_features(xf, ef, tgt, w)
        loss, g = net.gradient()
    loss_ref, g_ref = R.loss_and_grad(theta, rp, col, ef, xf, tgt, w)
    assert abs(loss - loss_ref) / loss_ref < 3e-2
    cos = float(g @ g_ref / (np.linalg.norm(g) * np.linalg.norm(g_ref)))
    assert cos > 0.99, cos
    assert abs(np.linalg.norm(g) / np.linalg.norm(g_ref) - 1.0) < 0.1
    # every parameter block receives gradient
    o = 0
    for name, sh in R.shapes():
        n = int(np.prod(sh))
        blk, blk_ref = g[o:o + n], g_ref[o:o + n]
        o += n
        c = float(blk @ blk_ref / (np.linalg.norm(blk) * np.linalg.norm(blk_ref) + 1e-300))
        assert c > 0.97, (name, c)


def test_training_reduces_loss_and_params_roundtrip():
    rp, col, ef, xf = _graph()
    rows = np.repeat(np.arange(len(xf)), np.diff(rp))
    # a smooth target the network can fit: neighbourhood means of two features
    tgt = np.zeros((len(xf), 2))
    np.add.at(tgt, rows, xf[col, :2])
    tgt /= np.maximum(np.diff(rp), 1)[:, None]
    with Mpnn(seed=9) as net:
        net.set_graph(rp, col)
        net.set_features(xf, ef, tgt)
        l0 = net.gradient()[0]
        for _ in range(200):
            last = net.train_step(2e-3)
        assert last < 0.2 * l0, (l0, last)
        th = net.get_params()
        with Mpnn(seed=1) as other:
            other.set_params(th)
            other.set_graph(rp, col)
            other.set_features(xf, ef)
            net.set_features(xf, ef)
            assert np.array_equal(other.forward(), net.forward())


def test_rejects_asymmetric_graph():
    with Mpnn() as net:
        with pytest.raises(_lib.PdError):
            net.set_graph(np.array([0, 1, 1], np.int32), np.array([1], np.int32))

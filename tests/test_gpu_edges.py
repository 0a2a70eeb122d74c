"""GPU edge cases through the C-ABI: ragged rows (isolated particles next to
full rows), problems with no bonds at all (down to a single particle), and
the argument errors the entry points report instead of running.

The reference has no tests of its own; its kernels run over whatever rows
the neighbour search produced (Neighbor.cu:42-80 appends nothing for an
isolated particle, Force_Bond.cu:169-196 then sums no bonds, and
Base_Function.cu:220-238 divides 0 by 0 for its damage).
"""
import ctypes as C

import numpy as np
import pytest

import pd_gpu1_amd as pd
from pd_gpu1_amd import _lib, scenarios
from pd_gpu1_amd._lib import dptr
from pd_gpu1_amd.engine import params
from oracle.runner import OracleModel

pytestmark = pytest.mark.gpu

DX = 1.0e-3


def _pmb2d(precision=_lib.PD_PREC_FP64):
    p, _, _ = params(2, _lib.PD_MODEL_PMB, 1.0e-8, DX, thick=DX, length=0.02, height=0.02, emod=192e9,
                     pratio=1.0 / 3.0, dens=8000.0, s0=0.004)
    p.precision = precision
    return p


def _ragged_cloud(seed=5):
    """A 14 x 12 lattice with a fifth of its points removed at random (rows of
    many lengths) plus four points farther than the horizon from everything
    (empty rows)."""
    rng = np.random.default_rng(seed)
    gx, gy = np.meshgrid(np.arange(14) * DX, np.arange(12) * DX, indexing="ij")
    keep = rng.random(gx.size) > 0.2
    x = gx.ravel()[keep] + rng.uniform(-0.05, 0.05, int(keep.sum())) * DX
    y = gy.ravel()[keep] + rng.uniform(-0.05, 0.05, int(keep.sum())) * DX
    far = np.array([[0.05, 0.05], [0.06, 0.05], [-0.05, 0.02], [0.03, -0.04]])
    return np.concatenate([x, far[:, 0]]), np.concatenate([y, far[:, 1]])


def _problem(p, x, y):
    prob = pd.Problem(p, 0)
    prob.set_positions(x, y, None, n_int=len(x))
    prob.build_neighbors()
    prob.init_state()
    return prob


def test_ragged_rows_and_isolated_particles_fp64():
    p = _pmb2d()
    x, y = _ragged_cloud()
    n = len(x)
    m = OracleModel(scenarios.params_dict(p), x, y, None, n)
    m.build_neighbors()
    m.init_state()
    u = np.random.default_rng(9).uniform(-3e-5, 3e-5, size=(2, n))
    with _problem(p, x, y) as prob:
        rp, col = prob.csr()
        assert np.array_equal(rp, m.row_ptr) and np.array_equal(col, m.col)
        lens = np.diff(rp)
        assert (lens == 0).sum() >= 4 and lens.max() >= 20 and len(np.unique(lens)) > 8
        prob.set("U", u)
        m.U = [u[0].copy(), u[1].copy()]
        prob.force()
        m.force()
        f = prob.get("F")
        assert np.array_equal(f, m.stack("F"))
        assert np.array_equal(prob.fail_mask(), m.fail)
        assert (m.fail == 0).sum() > 0  # the displacements break bonds
        assert np.all(f[:, lens == 0] == 0.0)
        prob.damage()
        m.damage()
        dmg = prob.get("DMG")
        assert np.array_equal(dmg, m.dmg, equal_nan=True)
        assert np.isnan(dmg[lens == 0]).all() and not np.isnan(dmg[lens > 0]).any()


def test_ragged_rows_mixed_lds_and_gathering_agree():
    x, y = _ragged_cloud()
    u = np.random.default_rng(9).uniform(-3e-5, 3e-5, size=(2, len(x)))
    outs = []
    for lds in (0, 1):
        with _problem(_pmb2d(_lib.PD_PREC_MIXED), x, y) as prob:
            prob.set_lds(lds)
            prob.set_tile(4)  # several tiles, the far points alone in theirs
            prob.set("U", u)
            prob.force()
            lens = np.diff(prob.csr()[0])
            outs.append((prob.get("F"), prob.fail_mask(), prob.lds_info()[0]))
    (f0, m0, l0), (f1, m1, l1) = outs
    assert not l0 and l1
    assert np.array_equal(m0, m1) and np.array_equal(f0, f1)
    assert np.all(f1[:, lens == 0] == 0.0)
    with _problem(_pmb2d(), x, y) as ref:  # FP64 break decisions are the same
        ref.set("U", u)
        ref.force()
        assert np.array_equal(ref.fail_mask(), m1)
        f64 = ref.get("F")
    assert np.max(np.abs(f1 - f64)) <= 1e-5 * np.max(np.abs(f64))


@pytest.mark.parametrize("precision", [_lib.PD_PREC_FP64, _lib.PD_PREC_MIXED])
@pytest.mark.parametrize("n", [1, 3])
def test_problem_without_bonds(precision, n):
    p = _pmb2d(precision)
    x, y = 0.01 * np.arange(n), np.zeros(n)  # 10 mm apart: no pair within the horizon
    v = np.stack([np.linspace(1.0, 2.0, n), np.linspace(-3.0, 0.5, n)])
    with _problem(p, x, y) as prob:
        assert prob.num_bonds == 0
        assert prob.csr()[0].tolist() == [0] * (n + 1)
        assert prob.fail_mask().size == 0
        prob.set("V", v)
        prob.run_verlet(4)
        assert np.all(prob.get("F") == 0.0) and np.all(prob.get("A") == 0.0)
        uref = np.zeros_like(v)
        for _ in range(4):  # Integrate.cu:447-481 with a = 0
            uref = uref + v * p.dt
        assert np.array_equal(prob.get("U"), uref) and np.array_equal(prob.get("V"), v)
        prob.damage()
        assert np.isnan(prob.get("DMG")).all()


def test_entry_points_reject_invalid_arguments():
    L = _lib.lib()
    p = _pmb2d()
    x, y = np.zeros(4), np.zeros(4)
    out = np.zeros(8)
    with pd.Problem(p, 0) as prob:
        h = prob._h
        assert L.pd_set_positions(h, dptr(x), dptr(y), None, 5, 4) == _lib.PD_E_INVALID  # n_int > n
        assert L.pd_set_positions(h, dptr(x), dptr(y), None, 0, 4) == _lib.PD_E_INVALID
        assert L.pd_set_positions(h, None, dptr(y), None, 4, 4) == _lib.PD_E_INVALID
        assert L.pd_force(h) == _lib.PD_E_INVALID  # no bonds yet
        assert L.pd_run_verlet(h, 3) == _lib.PD_E_INVALID
        assert L.pd_last_error(h)
        prob.set_positions(0.01 * np.arange(4), y, None, n_int=4)
        prob.build_neighbors()
        assert L.pd_run_verlet(h, -1) == _lib.PD_E_INVALID
        assert L.pd_get_field(h, 99, dptr(out)) == _lib.PD_E_INVALID
        assert L.pd_get_field(h, _lib.FIELDS["U"], None) == _lib.PD_E_INVALID
        prob.run_verlet(1)  # the handle still works after the rejected calls
    hb = C.c_void_p()
    assert L.pd_batch_create(C.byref(p), None, 0, 0, C.byref(hb)) == _lib.PD_E_INVALID
    p3, _, _ = params(3, _lib.PD_MODEL_PMB, 1.0e-8, DX, length=0.01, height=0.01, width=0.01)
    spec = (_lib.PdSpecimen * 1)(_lib.PdSpecimen(20, 20, 0.005, 0.0))
    assert L.pd_batch_create(C.byref(p3), spec, 1, 0, C.byref(hb)) == _lib.PD_E_INVALID


def test_rows_longer_than_256_neighbours_are_a_capacity_error():
    # delta = 4.5 dx on a cubic lattice gives interior rows of 380 bonds, more
    # than the 256 a row may hold (max_neighbors is capped there: the CSR row
    # sort holds 256 entries)
    p, _, _ = params(3, _lib.PD_MODEL_PMB, 1.0e-8, DX, length=0.012, height=0.012, width=0.012, emod=192e9,
                     pratio=0.25, dens=8000.0, s0=0.004, delta_factor=4.5)
    p.max_neighbors = 257
    with pytest.raises(pd.PdError) as ei:
        pd.Problem(p, 0)
    assert ei.value.code == _lib.PD_E_INVALID
    p.max_neighbors = 256
    g = (np.arange(12) - 5.5) * DX
    x, y, z = (a.ravel().copy() for a in np.meshgrid(g, g, g, indexing="ij"))
    prob = pd.Problem(p, 0)
    prob.set_positions(x, y, z)
    with pytest.raises(pd.PdError) as ei:
        prob.build_neighbors()
    assert ei.value.code == _lib.PD_E_CAPACITY and "exceeds max_neighbors 256" in str(ei.value)
    prob.close()

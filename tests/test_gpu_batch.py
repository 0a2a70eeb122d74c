"""Batched quasi-static ADR (persistent launch, device work queue) against the
FP64 oracle run specimen by specimen with the same fixed reduction tree."""
import numpy as np
import pytest

from pd_gpu1_amd import batch as PB
from pd_gpu1_amd import scenarios
from pd_gpu1_amd._lib import PdError
from oracle.runner import adr_increment, plate_specimen

pytestmark = pytest.mark.gpu


def oracle_specimen(p, n, clen, ang):
    return plate_specimen(scenarios.params_dict(p), n, clen, ang)


def oracle_increment(m, du, tol, max_iters, T):
    return adr_increment(m, du, tol, max_iters, T)


@pytest.mark.parametrize("block,group,parts", [(256, 1, 1), (1024, 1, 1), (256, 8, 1), (1024, 32, 1),
                                              (256, 1, 4), (1024, 4, 2), (1024, 8, 16)])
def test_batch_adr_bit_exact_against_oracle(block, group, parts):
    p = PB.plate_params()
    specs = [(12, 12, 0.0025, 0.3), (17, 17, 0.003, -0.2), (9, 9, 0.0, 0.0), (21, 21, 0.004, 0.5),
             (14, 14, 0.002, 0.0)]
    ms = [oracle_specimen(p, n, cl, a) for n, _, cl, a in specs]
    du, tol, max_it = 3.0e-6, 1e-6, 600
    with PB.Batch(p, specs) as b:
        b.set_group(group)
        b.set_parts(parts)
        for s, m in enumerate(ms):
            rp, col, mask = b.csr(s)
            assert np.array_equal(rp, m.row_ptr) and np.array_equal(col, m.col) and np.array_equal(mask, m.fail)
        for inc in range(3):
            its, res, _ = b.increment(du, tol, max_it, block=block, grid=3 * parts)
            assert b.parts_used() == parts
            for s, m in enumerate(ms):
                it_o, r_o = oracle_increment(m, du, tol, max_it, block)
                assert its[s] == it_o and res[s] == r_o, (inc, s, its[s], it_o, res[s], r_o)
                assert np.array_equal(b.get(s, "U"), m.stack("U")), (inc, s)
                assert np.array_equal(b.csr(s)[2], m.fail), (inc, s)
        assert np.all(np.isfinite(res)) and np.all(its > 0)


def test_batch_adr_converges_to_tolerance():
    p = PB.plate_params()
    specs = PB.plate_family(16, n_min=20, n_max=60)
    with PB.Batch(p, specs) as b:
        for inc in range(4):
            its, res, ms = b.increment(1.0e-6, 1e-6, 200000, block=1024)
            print("increment", inc, "iterations", its.tolist(), "ms", round(ms, 2), flush=True)
            assert np.all(res <= 1e-6), res


def test_batch_warm_start_reaches_the_same_states_in_fewer_iterations():
    p = PB.plate_params()
    specs = PB.plate_family(12, n_min=24, n_max=60)
    its_p, its_w = [], []
    with PB.Batch(p, specs) as plain, PB.Batch(p, specs) as warm:
        warm.set_warm(1)
        for inc in range(4):
            a, ra, _ = plain.increment(1.0e-6, 1e-6, 200000, block=1024)
            b, rb, _ = warm.increment(1.0e-6, 1e-6, 200000, block=1024)
            assert np.all(ra <= 1e-6) and np.all(rb <= 1e-6)
            its_p.append(a)
            its_w.append(b)
        for s in range(len(specs)):
            up, uw = plain.get(s, "U"), warm.get(s, "U")
            assert np.max(np.abs(up - uw)) < 1e-3 * np.max(np.abs(up)), s
            assert np.array_equal(plain.csr(s)[2], warm.csr(s)[2])
    assert np.array_equal(its_p[0], its_w[0])  # no history: the first increment is plain
    assert sum(x.sum() for x in its_w[1:]) < 0.5 * sum(x.sum() for x in its_p[1:])


def test_batch_parts_match_one_workgroup_per_specimen():
    # several workgroups per specimen (chosen to fill the device) against one,
    # with the warm start: the same iterations, residuals, states and bonds
    p = PB.plate_params()
    specs = PB.plate_family(12, n_min=24, n_max=60)
    with PB.Batch(p, specs) as one, PB.Batch(p, specs) as many:
        one.set_parts(1)
        many.set_parts(0)
        for b in (one, many):
            b.set_warm(1)
        for inc in range(3):
            a, ra, _ = one.increment(1.0e-6, 1e-6, 200000, block=1024)
            b, rb, _ = many.increment(1.0e-6, 1e-6, 200000, block=1024)
            assert many.parts_used() > 1 and one.parts_used() == 1
            assert np.array_equal(a, b) and np.array_equal(ra, rb), inc
        for s in range(len(specs)):
            assert np.array_equal(one.get(s, "U"), many.get(s, "U")), s
            assert np.array_equal(one.csr(s)[2], many.csr(s)[2]), s


def test_batch_retire_skips_only_unconverged_specimens():
    p = PB.plate_params()
    specs = PB.plate_family(6, n_min=24, n_max=40)
    with PB.Batch(p, specs) as cut, PB.Batch(p, specs) as ok:
        cut.set_retire(1)
        cut.set_parts(0)  # the skip runs in every workgroup of a specimen's group
        ok.set_retire(1)
        a, ra, _ = cut.increment(1.0e-6, 1e-6, 20, block=1024)  # 20 iterations reach no specimen's tol
        assert np.all(a == 20) and np.all(ra > 1e-6)
        before = [cut.get(s, "U").copy() for s in range(len(specs))]
        a2, _, _ = cut.increment(1.0e-6, 1e-6, 20, block=1024)
        assert np.all(a2 == 0)
        for s in range(len(specs)):
            assert np.array_equal(cut.get(s, "U"), before[s]), s
        for inc in range(2):  # converged increments never retire a specimen
            b, rb, _ = ok.increment(1.0e-6, 1e-6, 200000, block=1024)
            assert np.all(rb <= 1e-6) and np.all(b > 0)


def test_batch_group_barrier_timeout_is_reported():
    # a one-tick bound on the group barriers' wait: every workgroup leaves the
    # kernel at its next wait and the increment reports the timeout; the
    # specimens are restored to their state before the increment, so the
    # same handle then runs the increment as a fresh batch does
    p = PB.plate_params()
    specs = PB.plate_family(8, n_min=40, n_max=80)
    with PB.Batch(p, specs) as ref:
        ref.set_parts(1)
        ref.increment(1.0e-6, 1e-6, 200000, block=1024)
        its_ref, res_ref, _ = ref.increment(1.0e-6, 1e-6, 200000, block=1024)
        u_ref = [ref.get(s) for s in range(len(specs))]
        fail_ref = [ref.csr(s)[2] for s in range(len(specs))]
    with PB.Batch(p, specs) as b:
        b.set_parts(1)
        b.increment(1.0e-6, 1e-6, 200000, block=1024)
        before = [(b.get(s), b.get(s, "VHO"), b.get(s, "FO"), b.csr(s)[2]) for s in range(len(specs))]
        b.set_parts(4)
        b.set_barrier_timeout(1e-9)
        with pytest.raises(PdError, match="restored"):
            b.increment(1.0e-6, 1e-6, 2000, block=1024)
        for s in range(len(specs)):
            after = (b.get(s), b.get(s, "VHO"), b.get(s, "FO"), b.csr(s)[2])
            assert all(np.array_equal(x, y) for x, y in zip(before[s], after)), s
        b.set_barrier_timeout(20.0)  # the default bound: the same increment converges
        its, res, _ = b.increment(1.0e-6, 1e-6, 200000, block=1024)
        assert b.parts_used() == 4 and np.all(res <= 1e-6)
        assert np.array_equal(its, its_ref) and np.array_equal(res, res_ref)
        for s in range(len(specs)):
            assert np.array_equal(b.get(s), u_ref[s]) and np.array_equal(b.csr(s)[2], fail_ref[s]), s


def test_batch_warm_start_through_fracture():
    # 2 um increments: bonds start to break at the sixth increment and the
    # cracks run over the last three; the previous-increment start still
    # saves iterations in total and every increment reaches tol.  Once bonds
    # break the broken-bond sets of the two runs differ (the bonds that break
    # during an ADR run depend on the path from its start), so only their
    # sizes are compared.
    p = PB.plate_params()
    specs = PB.plate_family(12, n_min=24, n_max=48)
    tot, broke = {}, {}
    for mode in (0, 1):
        with PB.Batch(p, specs) as b:
            b.set_warm(mode)
            its_all, nb = [], []
            for inc in range(10):
                its, res, _ = b.increment(2.0e-6, 1e-6, 40000, block=1024)
                assert np.all(res <= 1e-6), (mode, inc, res)
                its_all.append(int(its.sum()))
                nb.append(sum(int((b.csr(s_)[2] == 0).sum()) for s_ in range(len(specs))))
            tot[mode], broke[mode] = np.array(its_all), np.array(nb)
            print("mode", mode, "iterations per increment", its_all, "broken", nb, flush=True)
    assert broke[0][-1] > 2 * broke[0][0] and broke[1][-1] > 2 * broke[1][0]
    assert tot[1].sum() < 0.8 * tot[0].sum()
    assert abs(broke[1][-1] - broke[0][-1]) < 0.15 * broke[0][-1]


@pytest.mark.parametrize("warm,du,parts", [(0, 1.0e-6, 1), (1, 2.0e-6, 2), (1, 2.0e-6, 4)])
def test_batch_tail_launch_matches_one_launch(warm, du, parts):
    # with few workgroups per specimen the device is mostly idle once the
    # queue is empty: the running specimens pause and continue in launches
    # with more workgroups each, resuming at the iteration they paused at,
    # so every result equals the one-launch run (2 um increments break bonds
    # from the sixth increment on)
    p = PB.plate_params()
    specs = PB.plate_family(12, n_min=24, n_max=100)
    out = {}
    for tail in (0, 64):
        with PB.Batch(p, specs) as b:
            b.set_warm(warm)
            b.set_parts(parts)
            b.set_tail(tail)
            rows = []
            for inc in range(8):
                its, res, _ = b.increment(du, 1e-6, 40000, block=1024)
                rows.append((its.copy(), res.copy(), b.tail_parts()))
            out[tail] = (rows, [b.get(s) for s in range(len(specs))], [b.csr(s)[2] for s in range(len(specs))])
    (ra, ua, fa), (rb, ub, fb) = out[0], out[64]
    print("tail launch workgroups per specimen", [r[2] for r in rb])
    assert all(r[2] == 0 for r in ra) and all(r[2] >= 16 for r in rb)
    for (ia, sa, _), (ib, sb, _) in zip(ra, rb):
        assert np.array_equal(ia, ib) and np.array_equal(sa, sb)
    for s in range(len(specs)):
        assert np.array_equal(ua[s], ub[s]) and np.array_equal(fa[s], fb[s]), s


def test_batch_separation_retires_fractured_plates():
    # 5 um increments fracture most plates through within a few increments;
    # with separation detection those end their increment once the strip
    # reaction collapses (flag 2) instead of running to the iteration cap,
    # and the specimens that never separate are unchanged
    p = PB.plate_params()
    specs = PB.plate_family(12, n_min=24, n_max=48)
    runs = {}
    for ratio in (0.0, 0.05):
        with PB.Batch(p, specs) as b:
            b.set_warm(1)
            b.set_retire(1)
            b.set_separation(ratio)
            its_all = []
            for inc in range(8):
                its, res, _ = b.increment(5.0e-6, 1e-6, 20000, block=1024)
                its_all.append(its.copy())
            runs[ratio] = (np.array(its_all), b.retired(), [b.get(s) for s in range(len(specs))])
    its0, ret0, u0 = runs[0.0]
    its1, ret1, u1 = runs[0.05]
    print("retired without separation", ret0.tolist(), "with", ret1.tolist())
    print("iterations", its0.sum(), its1.sum())
    assert (ret0 == 2).sum() == 0 and (ret1 == 2).sum() >= 3
    assert its1.sum() < its0.sum()
    for s in np.nonzero(ret1 == 0)[0]:
        assert np.array_equal(its0[:, s], its1[:, s]) and np.array_equal(u0[s], u1[s])


def test_batch_recomputed_coefficients_match_streamed():
    """The force phase recomputes each bond's {kappa, |xi|} from the reference
    positions (default, after the set-up check) or streams them: the same
    iterations, residuals, states and broken bonds, bit for bit, through
    increments that break bonds; an unknown mode is refused."""
    p = PB.plate_params()
    specs = PB.plate_family(10, n_min=24, n_max=70)
    with PB.Batch(p, specs) as rc, PB.Batch(p, specs) as st:
        assert rc.coef_recomputed()
        st.set_coef_mode(0)
        assert not st.coef_recomputed()
        for inc, du in enumerate([2.0e-6, 1.0e-5, 1.0e-5]):
            a, ra, _ = rc.increment(du, 1e-6, 20000)
            b, rb, _ = st.increment(du, 1e-6, 20000)
            assert np.array_equal(a, b) and np.array_equal(ra, rb), inc
            for s in range(len(specs)):
                assert np.array_equal(rc.get(s, "U"), st.get(s, "U")), (inc, s)
                assert np.array_equal(rc.csr(s)[2], st.csr(s)[2]), (inc, s)
        assert sum(int((rc.csr(s)[2] == 0).sum()) for s in range(len(specs))) > 0
        with pytest.raises(PdError):
            st.set_coef_mode(2)

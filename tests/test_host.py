"""Host-side helpers that need no GPU."""
import numpy as np

from pd_gpu1_amd import batch as PB


def test_plate_family_is_seeded_and_in_range():
    a = PB.plate_family(8)
    assert a == PB.plate_family(8)
    fam = PB.plate_family(512)
    assert all(40 <= s[0] <= 200 for s in fam)
    assert min(s[0] for s in fam) < 50 and max(s[0] for s in fam) > 190
    assert all(abs(s[3]) <= np.radians(30) + 1e-12 for s in fam)
    assert all(0.1 * s[0] * 0.5e-3 - 1e-15 <= s[2] <= 0.5 * s[0] * 0.5e-3 + 1e-15 for s in fam)


def test_splitmix64_reference_values():
    # first outputs of SplitMix64 seeded with 0 (published test vector)
    v, st = PB.splitmix64(0)
    assert v == 0xE220A8397B1DCDAF
    v, st = PB.splitmix64(st)
    assert v == 0x6E789E6AA1B965F4


def test_crack_half_length_and_speed():
    from pd_gpu1_amd import observables as ob
    nx, ny, dx = 20, 9, 1e-3
    dmg = np.zeros(nx * ny)
    c = (ny // 2) * nx
    dmg[c + 6: c + 14] = 0.5          # damaged band centred on the plate
    dmg[c + 5] = 0.30                 # below the threshold
    assert ob.crack_half_length(dmg, nx, ny, dx) == (nx // 2 - 6) * dx
    assert ob.crack_half_length(np.zeros(nx * ny), nx, ny, dx) == 0.0
    t = np.array([0.0, 1e-6, 2e-6, 3e-6])
    assert abs(ob.crack_speed(t, 0.004 + 500.0 * t) - 500.0) < 1e-6


def test_kw_crack_angles_recover_synthetic_cracks():
    from pd_gpu1_amd import observables as ob
    rng = np.random.default_rng(3)
    rad, dx = 0.025, 1e-3
    xs, ys, dm = [], [], []
    for side, ang in ((-1.0, 68.0), (1.0, 72.0)):
        t = np.linspace(3 * dx, 0.035, 60)
        a = np.radians(ang)
        xs.append(side * rad + side * t * np.sin(a) + rng.normal(0, 0.2 * dx, t.size))
        ys.append(-t * np.cos(a) + rng.normal(0, 0.2 * dx, t.size))
        dm.append(np.full(t.size, 0.5))
    # undamaged lattice and damage above the tips (the notch itself) are ignored
    gx, gy = np.meshgrid(np.linspace(-0.1, 0.1, 60), np.linspace(-0.05, 0.05, 30))
    xs.append(gx.ravel()); ys.append(gy.ravel()); dm.append(np.zeros(gx.size))
    xs.append(np.full(20, rad)); ys.append(np.linspace(0.001, 0.04, 20)); dm.append(np.full(20, 0.45))
    left, right = ob.kw_crack_angles(np.concatenate(xs), np.concatenate(ys), np.concatenate(dm), rad, dx)
    assert abs(left - 68.0) < 1.5 and abs(right - 72.0) < 1.5
    # a crack turning towards the centre line has a negative angle
    t = np.linspace(3 * dx, 0.03, 40)
    inward = ob.crack_angle_from_tip(rad - t * np.sin(0.5), -t * np.cos(0.5), rad, 0.0, 1.0)
    assert abs(inward + np.degrees(0.5)) < 1e-6
    assert np.isnan(ob.crack_angle_from_tip([], [], 0.0, 0.0, 1.0))


def test_batch_choose_parts_rule():
    # pd_batch_choose_parts: the smallest power of two W (<= wmax) with
    # largest work * capacity <= total work * W, work = bonds * sqrt(particles);
    # host arithmetic only, so it runs without a device
    from pd_gpu1_amd.batch import choose_parts
    n = np.full(256, 10000)
    nb = np.full(256, 280000)
    assert choose_parts(nb, n, 256) == 1                 # one specimen per workgroup fills the device
    assert choose_parts(nb[:64], n[:64], 256) == 4       # 64 equal specimens: 4 workgroups each
    assert choose_parts(nb[:1], n[:1], 256) == 16        # one specimen: capped at wmax
    assert choose_parts(nb[:1], n[:1], 256, wmax=4) == 4
    # one specimen with 8x the work of each of 255 others (2x the bonds, 16x
    # the particles): 8 * 256 <= 263 * 8 first at W = 8 (bonds alone would give 2)
    nb2, n2 = nb.copy(), n.copy()
    nb2[0] *= 2
    n2[0] *= 16
    assert choose_parts(nb2, n2, 256) == 8
    # retired specimens do not count: all but one retired -> wmax
    ret = np.ones(64, dtype=np.int32)
    ret[5] = 0
    assert choose_parts(nb[:64], n[:64], 256, retired=ret) == 16
    assert choose_parts(nb[:64], n[:64], 256, retired=np.zeros(64, np.int32)) == 4
    assert choose_parts(nb[:0], n[:0], 256) == 1         # nothing to run

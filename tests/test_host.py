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

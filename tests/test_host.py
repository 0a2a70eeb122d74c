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

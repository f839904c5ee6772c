"""Polar-sweep case set (host only): the fixed-seed chord draw and the
uneven grid shapes (2c, 0.75c, 0.9c)."""
import numpy as np

from foil_cfd_amd import group


def test_lowbias32_known_values():
    # lowbias32 (Wellons): x ^= x >> 16; x *= 0x7feb352d; x ^= x >> 15;
    # x *= 0x846ca68b; x ^= x >> 16 -- hash of 0 is 0, and the hash is a
    # bijection on 32-bit words (checked on a sample)
    assert group.lowbias32(0) == 0
    xs = np.random.default_rng(0).integers(0, 2**32, 20000, dtype=np.uint64)
    hs = {group.lowbias32(int(x)) for x in xs}
    assert len(hs) == len(set(int(x) for x in xs))


def test_polar_chords_fixed_and_in_range():
    a = group.polar_chords()
    assert a == group.polar_chords()
    assert len(a) == 16 and all(96 <= c <= 192 for c in a)
    assert len(set(a)) > 8  # uneven grids


def test_polar_sweep_grid_shapes():
    cs = group.polar_sweep(4, -4.0, 16.0, chords=[40, 48, 56, 64], ramp=False)
    for c, ch in zip(cs, [40, 48, 56, 64]):
        assert c.dims == (2 * ch, int(round(0.75 * ch)), int(round(0.9 * ch)))
        assert c.flags.size == c.dims[0] * c.dims[1] * c.dims[2]

"""CPU: the training-data oracle (compress.py, the data loader's shift) on
the reference fixtures, and its pins."""
import os

import numpy as np

from oracle import dataset as OD

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_oracle_compress_on_reference_fixtures():
    """compress.py over the reference's own m3d outputs (one clip): the pupil
    order arrays have the fixture's 1901 / 1749 entries, the amplitudes are
    the float32 of the text values, the mask the block means."""
    bits = open(os.path.join(GOLD, "ref_mask.bin"), "rb").read()
    maskall, arrs, lo, mo, los, mos = OD.compress(bits, os.path.join(GOLD, "ref_inputxx.csv"),
                                                 os.path.join(GOLD, "ref_inputlm.csv"), 1)
    assert maskall.shape == (1, 512, 512) and maskall.dtype == np.float16
    assert [a.shape[1] for a in arrs] == [1901, 1901, 1749, 1749, 1749, 1749]
    assert len(lo) == 1901 and len(los) == 1749
    first = open(os.path.join(GOLD, "ref_inputxx.csv")).readline().split(",")
    assert arrs[0][0, 0] == np.float32(float(first[0]))
    # every mask value is a multiple of 1/16 and their mean is the bit density
    assert np.all((maskall.astype(np.float64) * 16) % 1 == 0)
    assert abs(maskall.astype(np.float64).mean() - 0.63946533203125) < 1e-12


def test_oracle_shift_is_a_phase_rotation():
    """rshift + i ishift = (re + i im) exp(2 pi i (l ls + m ms) / 512) to
    float32 rounding; the zero shift is the identity, bit for bit."""
    rng = np.random.default_rng(2)
    re = rng.standard_normal((2, 50)).astype(np.float32)
    im = rng.standard_normal((2, 50)).astype(np.float32)
    lo, mo = rng.integers(-20, 21, 50), rng.integers(-20, 21, 50)
    for ls, ms in ((0, 0), (3, -7), (511, 200)):
        r = OD.rshift(50, lo, mo, re, im, 1, ls, ms)
        i = OD.ishift(50, lo, mo, re, im, 1, ls, ms)
        want = (re[1] + 1j * im[1]) * np.exp(2j * np.pi * ((lo * ls + mo * ms) % 512) / 512)
        assert np.abs((r + 1j * i) - want).max() < 1e-6 * np.abs(want).max()
        if ls == ms == 0:
            assert np.array_equal(r, re[1]) and np.array_equal(i, im[1])


def test_oracle_roll_direction():
    """np.roll(x, (ms, -ls)): out[i][j] = x[i - ms][j + ls] (periodic)."""
    x = np.arange(16 * 16, dtype=np.float16).reshape(1, 16, 16)
    y = OD.getitem_mask(x, 0, 3, 5)
    assert y[7, 2] == x[0, 2, 5] and y[0, 0] == x[0, (0 - 5) % 16, 3]

"""CPU: the mask producers' oracle pinned to the reference fixture, and the
host side of the producers (the band layouts drawn by the C ABI's planner and
by the Python generators) against it.

Layouts are rasterised here by a small NumPy painter; the product rasterises
them on the GPU (tests/test_gpu_maskgen.py).  Bit-exact throughout.
"""
import os
import random
import subprocess

import numpy as np
import pytest

from oracle import maskgen as OM

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "ref_mask.bin")
REF_MASK_OUT = os.path.join(ROOT, "oracle", "_ref", "mask.out")


def golden():
    with open(GOLDEN, "rb") as f:
        return f.read()


def paint(lay):
    """NumPy raster of a MaskLayout -> mask.bin bytes."""
    from euvlitho_amd.masks import CONST_ONE, CONST_ZERO
    nx, ny = lay.ndivx, lay.ndivy
    img = np.ones((ny, nx), dtype=np.uint8)
    for l, ax in enumerate(lay.axis):
        lay_img = np.zeros((ny, nx), dtype=np.uint8)
        for pos, s in enumerate(lay.sel[l][: nx if ax == 0 else ny]):
            v = 1 if s == CONST_ONE else (0 if s == CONST_ZERO else lay.patterns[l][s])
            if ax == 0:
                lay_img[:, pos] = v
            else:
                lay_img[pos, :] = v
        img &= lay_img
    if lay.invert:
        img ^= 1
    return np.packbits(img.reshape(-1)).tobytes()


# ------------------------------------------------------------------ oracle pins
def test_oracle_cpp_program_reproduces_fixture():
    """mask.cpp + mVBF.h with the C library's rand() from srand(1) (the
    program never seeds) == the reference's mask.bin, bit for bit."""
    assert OM.mask_program(2048, 2048, 1, seed=1) == golden()


def test_oracle_line_pattern_reproduces_fixture():
    """LinePattern("H", "BF") fed the C program's draws (randint(0, 1) =
    rand() % 2, randint(0, 4) = rand() % 5) is the C mask: every 0/1 and
    tail convention of the Python generator inverts the C one, and BF inverts
    back.  Its array M[x][y'] is the fixture without the reader's y flip."""
    m = OM.line_pattern(2048, 2048, OM.CRand(1), direction="H", field_type="BF")
    assert OM.pattern_to_bits(m[:, ::-1]) == golden()


class _Const:
    def __init__(self, i, u):
        self.i, self.u = i, u

    def randint(self, a, b):
        return self.i

    def random(self):
        return self.u


@pytest.mark.parametrize("n,cd", [(64, 8), (70, 8), (96, 12), (100, 12)])
def test_oracle_random_line_space_closed_form(n, cd):
    """Constant draws (randint -> 0, random -> 0): every band and segment is
    cd wide, bands alternate line/space from a line, segments 1/0 from a 1;
    T = ceil((n - cd) / cd) of each are drawn and the tail follows the parity
    of T (a segment tail is 1 after an odd count, a band tail repeats the last
    line after an odd count).  The second pass's tail loop turns space to 1 at
    its first position (mask_pattern.py:183-185), so after an even count only
    that position is left alone.  M[a][b] = f1(b, a) & f2(a, b)."""
    T = -(-(n - cd) // cd)
    q = np.arange(n)
    seg = np.where(q < T * cd, (q // cd) % 2 == 0, T % 2 == 1).astype(int)       # randmask
    line1 = np.where(q < T * cd, (q // cd) % 2 == 0, T % 2 == 1)                 # band is a line band
    line2 = np.where(q < T * cd, (q // cd) % 2 == 0, (T % 2 == 1) | (q > T * cd))
    A, B = np.meshgrid(q, q, indexing="ij")
    want = np.where(line1[B], seg[A], 1) & np.where(line2[A], seg[B], 1)
    got = OM.random_line_space(n, n, _Const(0, 0.0), cd=cd, fac=1.0, fac1d=3.0, field_type="BF")
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(OM.random_line_space(n, n, _Const(0, 0.0), cd=cd, field_type="DF"), 1 - want)


def test_oracle_random_line_space_block_structure():
    """Real draws with fac = fac1d = 0: every band and segment is exactly cd
    wide (int(cd e^0) = cd), so away from the last band the clip is constant
    on cd x cd blocks, and it takes both values."""
    n, cd = 192, 16
    for seed in (5, 6):
        m = OM.random_line_space(n, n, random.Random(seed), cd=cd, fac=0.0, fac1d=0.0)
        assert set(np.unique(m)) == {0, 1}
        core = m[: n - cd, : n - cd].reshape((n - cd) // cd, cd, (n - cd) // cd, cd)
        assert (core == core[:, :1, :, :1]).all()


@pytest.mark.skipif(not os.path.exists(REF_MASK_OUT), reason="oracle/_ref/mask.out not built (no reference sources)")
def test_reference_program_matches_fixture_and_oracle(tmp_path):
    """The reference's own mask program, compiled from its sources into
    oracle/_ref, writes the fixture mask.bin; its maskimage.csv is the 0/1
    text of the same clip."""
    subprocess.run([REF_MASK_OUT], cwd=tmp_path, check=True, timeout=120)
    assert (tmp_path / "mask.bin").read_bytes() == golden()
    m = OM.maskgen(2048, 2048, OM.CRand(1)).reshape(2048, 2048)
    want = "".join("".join(f"{v}," for v in row) + "\n" for row in m[:64])
    assert (tmp_path / "maskimage.csv").read_text()[: len(want)] == want


# ------------------------------------------------------------------ host side of the producers
@pytest.mark.parametrize("seed,nx,ny,n,cd,gap", [(1, 2048, 2048, 1, 56, 80), (7, 1024, 512, 3, 56, 80),
                                                 (0, 264, 128, 2, 24, 16), (123456789, 512, 520, 2, 40, 33)])
def test_plan_linespace_matches_oracle(seed, nx, ny, n, cd, gap):
    """euvl_mask_linespace_plan (the host draws of the C ABI, its own rand()
    implementation) painted == the oracle's program with the C library's
    rand() (seed 0 is seed 1 in both)."""
    from euvlitho_amd.masks import line_space_plan
    lays = line_space_plan(nx, ny, n, seed, cd, gap)
    got = b"".join(paint(l) for l in lays)
    rng = OM.CRand(seed)
    want = b"".join(np.packbits(OM.maskgen(nx, ny, rng, cd, gap).astype(np.uint8)).tobytes() for _ in range(n))
    assert got == want
    if (seed, nx, ny, n) == (1, 2048, 2048, 1):
        assert got == golden()


def test_plan_linespace_rejects_bad_sizes():
    from euvlitho_amd import _lib
    from euvlitho_amd.masks import line_space_plan
    with pytest.raises(_lib.EuvlError):
        line_space_plan(100, 64, 1, 1, 8, 8)      # ndivx not a multiple of 8
    with pytest.raises(_lib.EuvlError):
        line_space_plan(48, 64, 1, 1, 56, 8)      # ndivx <= cd


@pytest.mark.parametrize("direction,field,w,h,seed", [("H", "BF", 256, 256, 3), ("V", "DF", 256, 128, 4),
                                                      ("H", "DF", 136, 200, 5), ("V", "BF", 64, 320, 6)])
def test_line_pattern_layout_matches_oracle(direction, field, w, h, seed):
    from euvlitho_amd.masks import LinePattern
    lay = LinePattern(cd=24, gap=16, direction=direction, field_type=field).layout(w, h, random.Random(seed))
    want = OM.line_pattern(w, h, random.Random(seed), cd=24, gap=16, direction=direction, field_type=field)
    assert paint(lay) == OM.pattern_to_bits(want)


def test_line_pattern_layout_with_c_draws_is_fixture():
    from euvlitho_amd.masks import LinePattern
    lay = LinePattern(direction="H", field_type="BF").layout(2048, 2048, OM.CRand(1))
    bits = np.unpackbits(np.frombuffer(paint(lay), dtype=np.uint8)).reshape(2048, 2048)
    assert np.packbits(bits[::-1].reshape(-1)).tobytes() == golden()


@pytest.mark.parametrize("field,n,cd,seed", [("BF", 128, 12, 1), ("DF", 160, 16, 2), ("BF", 96, 8, 3)])
def test_random_line_space_layout_matches_oracle(field, n, cd, seed):
    from euvlitho_amd.masks import RandomLineSpacePattern
    lay = RandomLineSpacePattern(cd=cd, fac=1.0, fac1d=3.0, field_type=field).layout(n, n, random.Random(seed))
    want = OM.random_line_space(n, n, random.Random(seed), cd=cd, fac=1.0, fac1d=3.0, field_type=field)
    assert paint(lay) == OM.pattern_to_bits(want)


def test_random_line_space_needs_square():
    from euvlitho_amd.masks import RandomLineSpacePattern
    with pytest.raises(ValueError):
        RandomLineSpacePattern().layout(128, 64, random.Random(0))

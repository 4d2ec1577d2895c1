"""GPU: the mask producers through the C ABI (euvl_mask_linespace,
euvl_mask_raster) against the reference fixture and the oracle, bit-exact,
and their clips fed to the EM path."""
import os
import random
import sys

import numpy as np
import pytest

from oracle import maskgen as OM

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU test without a GPU")


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_mask.bin")


def golden():
    with open(GOLDEN, "rb") as f:
        return f.read()


def test_linespace_program_is_fixture():
    """The reference mask program's clip (seed 1, 2048 x 2048, cd 56, gap 80)
    rasterised on the GPU == the fixture mask.bin."""
    from euvlitho_amd.masks import line_space_bits
    say("line_space_bits 2048x2048 seed 1")
    bits = line_space_bits(2048, 2048, 1, 1).cpu().numpy().tobytes()
    assert bits == golden()


@pytest.mark.parametrize("seed,nx,ny,n,cd,gap", [(7, 1024, 512, 3, 56, 80), (99, 264, 128, 5, 24, 16),
                                                 (2, 512, 1000, 2, 40, 33)])
def test_linespace_matches_oracle(seed, nx, ny, n, cd, gap):
    """Several clips from one stream == the oracle's ndata loop with the C
    library's rand()."""
    from euvlitho_amd.masks import line_space_bits
    say(f"line_space_bits {nx}x{ny} n={n} seed {seed} cd {cd} gap {gap}")
    got = line_space_bits(nx, ny, n, seed, cd, gap).cpu().numpy().tobytes()
    rng = OM.CRand(seed)
    want = b"".join(np.packbits(OM.maskgen(nx, ny, rng, cd, gap).astype(np.uint8)).tobytes() for _ in range(n))
    assert got == want


@pytest.mark.parametrize("direction,field,w,h,seed", [("H", "BF", 256, 256, 3), ("V", "DF", 256, 128, 4),
                                                      ("V", "BF", 64, 320, 6)])
def test_line_pattern_matches_oracle(direction, field, w, h, seed):
    """LinePattern on the GPU returns the reference's array (the oracle's
    pixel loops with the same random draws), and its mask.bin bytes are
    that array written in mask.bin orientation."""
    from euvlitho_amd.masks import LinePattern
    say(f"LinePattern {direction} {field} {w}x{h} seed {seed}")
    pat = LinePattern(cd=24, gap=16, direction=direction, field_type=field)
    got = pat(w, h, rng=random.Random(seed))
    want = OM.line_pattern(w, h, random.Random(seed), cd=24, gap=16, direction=direction, field_type=field)
    np.testing.assert_array_equal(got, want)
    bits = pat.bits(w, h, 2, rng=random.Random(seed)).cpu().numpy().tobytes()
    rng = random.Random(seed)
    want2 = b"".join(OM.pattern_to_bits(OM.line_pattern(w, h, rng, cd=24, gap=16, direction=direction,
                                                        field_type=field)) for _ in range(2))
    assert bits == want2


@pytest.mark.parametrize("field,n,cd,seed", [("BF", 256, 24, 1), ("DF", 200, 16, 2)])
def test_random_line_space_matches_oracle(field, n, cd, seed):
    from euvlitho_amd.masks import RandomLineSpacePattern
    say(f"RandomLineSpacePattern {field} {n} cd {cd} seed {seed}")
    got = RandomLineSpacePattern(cd=cd, field_type=field)(n, n, rng=random.Random(seed))
    want = OM.random_line_space(n, n, random.Random(seed), cd=cd, field_type=field)
    np.testing.assert_array_equal(got, want)


def test_raster_rejects_out_of_pool_index():
    """A layout that selects a pattern beyond its pool is refused on the host
    (no launch)."""
    from euvlitho_amd import _lib
    from euvlitho_amd.masks import MaskLayout, raster_layouts
    lay = MaskLayout(64, 32, [0], [np.full(64, 1, dtype=np.int32)], [[np.ones(32, dtype=np.uint8)]])
    with pytest.raises(_lib.EuvlError):
        raster_layouts([lay])
    ok = MaskLayout(64, 32, [0], [np.zeros(64, dtype=np.int32)], [[np.ones(32, dtype=np.uint8)]])
    assert bytes(raster_layouts([ok]).cpu().numpy()) == b"\xff" * (64 * 32 // 8)


def test_generated_clip_feeds_the_em_path():
    """A clip rasterised on the device goes to the mask spectrum as it is:
    the same spectrum as the oracle's clip uploaded from the host."""
    from euvlitho_amd.config import preset
    from euvlitho_amd.engine import Engine
    from euvlitho_amd.masks import line_space_bits
    cfg = preset("c1")
    eng = Engine(cfg)
    dev = line_space_bits(cfg.ndivx, cfg.ndivy, 1, 11, 56, 80)
    host = eng.upload_mask(OM.mask_program(cfg.ndivx, cfg.ndivy, 1, seed=11))
    s1 = eng.mask_spectrum(dev, 20, 20).cpu().numpy()
    s2 = eng.mask_spectrum(host, 20, 20).cpu().numpy()
    eng.close()
    assert np.array_equal(s1, s2)


def test_cli_mask_writes_reference_files(tmp_path):
    """python -m euvlitho_amd mask: mask.bin == the fixture, maskimage.csv ==
    the text mask.cpp:49-59 writes for the first clip."""
    from euvlitho_amd.__main__ import main
    assert main(["mask", "--out", str(tmp_path)]) == 0
    assert (tmp_path / "mask.bin").read_bytes() == golden()
    m = OM.maskgen(2048, 2048, OM.CRand(1)).reshape(2048, 2048)
    want = "".join("".join(f"{v}," for v in row) + "\n" for row in m)
    assert (tmp_path / "maskimage.csv").read_text() == want

"""GPU: the training-data steps through the library against the oracle,
bit-exact, and the generator's chain against the file route."""
import os
import sys

import numpy as np
import pytest

from oracle import dataset as OD

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU test without a GPU")


def test_compress_matches_oracle_on_fixtures(tmp_path):
    from euvlitho_amd.dataset import compress
    say("compress on the reference fixtures")
    compress(os.path.join(GOLD, "ref_mask.bin"), os.path.join(GOLD, "ref_inputxx.csv"),
             os.path.join(GOLD, "ref_inputlm.csv"), 1, str(tmp_path))
    bits = open(os.path.join(GOLD, "ref_mask.bin"), "rb").read()
    maskall, arrs, lo, mo, los, mos = OD.compress(bits, os.path.join(GOLD, "ref_inputxx.csv"),
                                                 os.path.join(GOLD, "ref_inputlm.csv"), 1)
    ld = lambda *p: np.load(os.path.join(tmp_path, *p))
    got = ld("maskdata", "mask.npy")
    assert got.dtype == np.float16 and np.array_equal(got, maskall)
    for name, a in zip(("re0", "im0", "rex", "imx", "rey", "imy"), arrs):
        g = ld("ampdata", f"{name}.npy")
        assert g.dtype == np.float32 and np.array_equal(g, a)
    for name, a in (("lorder", lo), ("morder", mo), ("lorders", los), ("morders", mos)):
        g = ld("inputlm", f"{name}.npy")
        assert g.dtype == a.dtype and np.array_equal(g, a)


def test_augment_matches_oracle():
    from euvlitho_amd.dataset import augment
    rng = np.random.default_rng(3)
    n, ncut = 5, 300
    masks = (rng.integers(0, 17, (n, 512, 512)) / 16).astype(np.float16)
    re = rng.standard_normal((n, ncut)).astype(np.float32)
    im = rng.standard_normal((n, ncut)).astype(np.float32)
    lo, mo = rng.integers(-25, 26, ncut), rng.integers(-25, 26, ncut)
    ls, ms = rng.integers(0, 512, n), rng.integers(0, 512, n)
    om, orr, oi = augment(masks, re, im, lo, mo, ls, ms)
    om, orr, oi = om.cpu().numpy(), orr.cpu().numpy(), oi.cpu().numpy()
    for k in range(n):
        assert np.array_equal(om[k], OD.getitem_mask(masks, k, ls[k], ms[k]))
        assert np.array_equal(orr[k], OD.rshift(ncut, lo, mo, re, im, k, ls[k], ms[k]))
        assert np.array_equal(oi[k], OD.ishift(ncut, lo, mo, re, im, k, ls[k], ms[k]))


def test_generate_equals_the_file_route(tmp_path):
    """generate (masks, EM + M3D and compmask on the GPU, no text files)
    writes the npy files that compress makes from the intermediate mask.bin,
    inputxx.csv and inputlm.csv it also writes."""
    from euvlitho_amd.config import high_na, ANNULAR
    from euvlitho_amd.dataset import compress, generate
    cfg = high_na(256, 6, ANNULAR, ndivs=1, image=(64, 64))
    say("generate 2 clips at 256 nm")
    direct = tmp_path / "direct"
    generate(2, str(direct), cfg, seed=5, cd=24, gap=16, nout=64, write_intermediates=True)
    files = tmp_path / "files"
    compress(str(direct / "mask.bin"), str(direct / "inputxx.csv"), str(direct / "inputlm.csv"), 2, str(files),
             ndiv=256, nout=64)
    for sub, names in (("maskdata", ["mask"]), ("ampdata", ["re0", "im0", "rex", "imx", "rey", "imy"]),
                       ("inputlm", ["lorder", "morder", "lorders", "morders"])):
        for nm in names:
            a = np.load(direct / sub / f"{nm}.npy")
            b = np.load(files / sub / f"{nm}.npy")
            assert a.dtype == b.dtype and np.array_equal(a, b), (sub, nm)

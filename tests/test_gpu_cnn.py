"""GPU: the CNN predictor's kernels through the C ABI against the oracle
(FP32 on the GPU, float64 in the oracle; bars stated per test)."""
import sys

import numpy as np
import pytest

from oracle import cnn as OC

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU test without a GPU")


def test_compmask_of_reference_mask_is_exact():
    """The reference mask.bin through euvl_cnn_compmask == the block means
    (multiples of 1/16, exact in FP32)."""
    import os
    from euvlitho_amd.cnn import compmask
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_mask.bin")
    bits = open(path, "rb").read()
    got = compmask(path).cpu().numpy()
    assert got.shape == (1, 1, 512, 512)
    np.testing.assert_array_equal(got[0, 0], OC.compmask(bits).astype(np.float32))
    two = compmask(bits + bits[::-1]).cpu().numpy()        # two clips
    np.testing.assert_array_equal(two[0], got[0])


@pytest.mark.parametrize("i,hw,batch", [(1, 64, 2), (2, 32, 3), (3, 48, 1), (4, 32, 2), (5, 32, 2)])
def test_conv_block_matches_oracle(i, hw, batch):
    """One conv block (conv i of the reference's channel ladder) at an
    hw x hw input: FP32 within 2e-5 of the largest output."""
    from euvlitho_amd.cnn import CNNmodel
    say(f"conv block {i} at {hw}x{hw}, batch {batch}")
    p = OC.random_params(1901, 10 + i, nin=64)
    net = CNNmodel(p)
    rng = np.random.default_rng(i)
    cin = OC_CH[i - 1]
    x = rng.standard_normal((batch, cin, hw, hw)).astype(np.float32)
    w, b, sc, sh = net.conv[i - 1]
    dx = torch.from_numpy(x).cuda()
    y = torch.empty((batch, OC_CH[i], hw // 2, hw // 2), dtype=torch.float32, device="cuda")
    from euvlitho_amd import _lib
    from euvlitho_amd.cnn import _handle
    lib, h = _handle(dx.device)
    _lib.check(lib.euvl_cnn_conv_block(h, batch, cin, OC_CH[i], hw, hw, _lib.ptr(dx), _lib.ptr(w), _lib.ptr(b),
                                       _lib.ptr(sc), _lib.ptr(sh), _lib.ptr(y)), "conv")
    got = y.cpu().numpy()
    want = np.array([OC.conv_block(x[k].astype(np.float64), p, i) for k in range(batch)])
    e = np.abs(got - want).max() / np.abs(want).max()
    say(f"  max err / max = {e:.2e}")
    assert e < 2e-5


OC_CH = (1, 16, 32, 64, 128, 256)


@pytest.mark.parametrize("nin,n_out,batch", [(64, 1901, 3), (32, 1749, 2)])
def test_network_matches_oracle(nin, n_out, batch):
    from euvlitho_amd.cnn import CNNmodel
    say(f"network nin {nin} outputs {n_out} batch {batch}")
    p = OC.random_params(n_out, 20 + nin, nin=nin)
    x = np.random.default_rng(nin).random((batch, 1, nin, nin)).astype(np.float32)
    got = CNNmodel(p)(torch.from_numpy(x).cuda()).cpu().numpy()
    want = OC.forward(p, x.astype(np.float64))
    e = np.abs(got - want).max() / np.abs(want).max()
    say(f"  max err / max = {e:.2e}")
    assert got.shape == (batch, n_out) and e < 1e-4


def test_full_size_network_on_reference_mask():
    """The reference's input size (512 x 512 from mask.bin, fc1 65536 -> 4096)
    with seeded FP32 weights: within 1e-4 of the largest output."""
    import os
    from euvlitho_amd.cnn import CNNmodels, compmask
    say("full-size network, reference mask")
    p = OC.random_params(1749, 99, nin=512, dtype=np.float32)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_mask.bin")
    x = compmask(path)
    got = CNNmodels(p)(x).cpu().numpy()
    want = OC.forward({k: v.astype(np.float64) for k, v in p.items()}, x.cpu().numpy().astype(np.float64))
    e = np.abs(got - want).max() / np.abs(want).max()
    say(f"  max err / max = {e:.2e}")
    assert e < 1e-4


def test_predict_m3d_and_checkpoints(tmp_path):
    """Fullmodel / predict_m3d with six networks loaded from checkpoint files
    (a Lightning-style dict with "state_dict", read with weights_only=True):
    the outputs divided by the factor rows == the oracle's."""
    from euvlitho_amd.cnn import CNNmodel, CNNmodels, predict_m3d
    nin = 64
    ps = [OC.random_params(1901 if k < 2 else 1749, 40 + k, nin=nin) for k in range(6)]
    models = []
    for k, p in enumerate(ps):
        path = tmp_path / f"m{k}.ckpt"
        torch.save({"state_dict": {n: torch.from_numpy(v.astype(np.float32)) for n, v in p.items()},
                    "epoch": 3}, path)
        models.append((CNNmodel if k < 2 else CNNmodels).load_from_checkpoint(str(path)))
    rng = np.random.default_rng(5)
    f0, fx, fy = rng.uniform(0.5, 2.0, 1901), rng.uniform(0.5, 2.0, 1749), rng.uniform(0.5, 2.0, 1749)
    x = rng.random((2, 1, nin, nin)).astype(np.float32)
    got = predict_m3d(torch.from_numpy(x).cuda(), *models, f0, fx, fy)
    p32 = [{n: v.astype(np.float32).astype(np.float64) for n, v in p.items()} for p in ps]
    want = OC.predict_m3d(x.astype(np.float64), p32, f0, fx, fy)
    for g, w in zip(got, want):
        assert g.dtype == np.float64 and g.shape == w.shape
        assert np.abs(g - w).max() / np.abs(w).max() < 1e-4


def test_cli_predict_writes_inputxx(tmp_path):
    """python -m euvlitho_amd predict: checkpoints and factor files laid out
    as predict.py reads them (here at a 64 x 64 network input, so that the
    six checkpoints stay small) -> inputxx.csv rows (np.savetxt) == the
    oracle's predict_m3d on the reference mask."""
    import os
    from euvlitho_amd.__main__ import main
    names = ("re0", "im0", "rex", "imx", "rey", "imy")
    ps = [OC.random_params(1901 if k < 2 else 1749, 60 + k, nin=64) for k in range(6)]
    rng = np.random.default_rng(9)
    facs = {"re0/factor0.csv": rng.uniform(0.5, 2, 1901), "rex/factorx.csv": rng.uniform(0.5, 2, 1749),
            "rey/factory.csv": rng.uniform(0.5, 2, 1749)}
    for name, p in zip(names, ps):
        (tmp_path / "model" / name).mkdir(parents=True)
        torch.save({"state_dict": {n: torch.from_numpy(v.astype(np.float32)) for n, v in p.items()}},
                   tmp_path / "model" / name / "model.ckpt")
    for f, v in facs.items():
        np.savetxt(tmp_path / "model" / f, v[None], delimiter=",")
    mask = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_mask.bin")
    assert main(["predict", "--mask", mask, "--model-dir", str(tmp_path / "model"), "--nout", "64",
                 "--out", str(tmp_path)]) == 0
    rows = [np.array(l.split(","), dtype=np.float64) for l in open(tmp_path / "inputxx.csv")]
    assert len(rows) == 6
    x = OC.compmask(open(mask, "rb").read(), 2048, 64)[None, None]
    f = [facs["re0/factor0.csv"], facs["rex/factorx.csv"], facs["rey/factory.csv"]]
    p32 = [{n: v.astype(np.float32).astype(np.float64) for n, v in p.items()} for p in ps]
    want = OC.predict_m3d(x, p32, *f)
    for r, w in zip(rows, want):
        assert np.abs(r - w[0]).max() / np.abs(w).max() < 1e-4

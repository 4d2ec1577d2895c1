"""GPU tests at the workload sizes: the 1024 nm / +-20-order / 200-point clip
(N = 1681, four source divisions) in FP64 against the oracle, and the
precision forms against FP64 on that clip.  Tolerances: FP64 amplitudes and
M3D parameters 1e-9 of the largest value, Abbe image 1e-10 of the peak; the
precision forms 1e-3 of the peak for the images (the metric's bar), with the
amplitude and M3D errors printed per quantity; FP32 imaging alone on FP64
amplitudes 1e-5 of the peak."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c2():
    from euvlitho_amd.config import preset
    from euvlitho_amd.masks import manhattan
    cfg = preset("c2")
    return cfg, manhattan(cfg.ndivx, cfg.ndivy, 0.5, 0)


def gpu_clip(cfg, m, prec, socs=True):
    from euvlitho_amd.engine import Engine
    eng = Engine(cfg, precision=prec)
    bits = eng.upload_mask(m)
    Ax = eng.amp_s(bits, "X")
    a0, ax, ay = eng.m3d(bits, Ax0=Ax[0])
    img = eng.abbe(bits, Ax=Ax)
    out = [Ax, a0, ax, ay, img]
    if socs:
        out.append(eng.socs(bits, a0, ax, ay, nsocs=24, nsocsxy=20))
    out = [t.cpu().numpy() for t in out]
    info = eng.last_info
    eng.close()
    return out, info


def test_config2_clip_vs_oracle(c2):
    """One whole clip of the 1024 nm workload: the EM solves of all four
    divisions, the M3D fit and the 256 x 256 Abbe image over the 200 source
    points, against the oracle."""
    from oracle import abbe as OA, em as OE, geometry as G, problems as OP
    cfg, m = c2
    say("config 2 clip: N=1681, 4 divisions, 200 source points")
    (Ax, a0, ax, ay, img), info = gpu_clip(cfg, m, 0, socs=False)
    say(f"  GPU info {info.tolist()}")
    p = OP.hna(cfg.ndivx, cfg.orders[1], None)
    d = G.derived(p)
    li, mi, L, M = cfg.order_set()
    li, mi = li.astype(np.int64), mi.astype(np.int64)
    offs = cfg.offsets()
    amps = [OE.amp_s("X", m, p, d, li, mi, L, M, o[0], o[1]) for o in offs]
    for q in range(len(offs)):
        e = np.abs(Ax[q] - amps[q]).max() / np.abs(amps[q]).max()
        say(f"  amplitudes, division {q}: {e:.2e}")
        assert e < 1e-9
    vc, ab = OA.blank_amplitudes(p, d)
    r0, rx, ry = OA.m3d_params(p, d, m, li, mi, amps[0], vc, ab)
    s = np.abs(r0).max()
    for g, r, name in ((a0, r0, "a0"), (ax, rx, "ax"), (ay, ry, "ay")):
        e = np.abs(g - r).max() / s
        say(f"  {name}: {e:.2e} of max|a0|")
        assert e < 1e-9
    divs = [(tuple(o), l0, m0) for o, (_, _, l0, m0) in zip(offs, cfg.source_points())]
    ref = OA.abbe_image(p, d, m, li, mi, L, M, divs, 256, 256, amps=amps)
    e = np.abs(img - ref).max() / ref.max()
    say(f"  Abbe image: {e:.2e} of peak")
    assert e < 1e-10


@pytest.mark.parametrize("form", ["mixed", "fast"])
def test_config4_precision_forms_vs_fp64(c2, form):
    """The precision forms on the 1024 nm clip against the FP64 drop-in:
    'mixed' (coupling and top-solve products and the imaging in FP32; matrix
    functions and multilayer in FP64) and 'fast' (the matrix functions in FP32
    too)."""
    from euvlitho_amd import _lib
    cfg, m = c2
    prec = _lib.PREC_MIXED if form == "mixed" else _lib.PREC_FAST
    (A64, a064, ax64, ay64, I64, S64), _ = gpu_clip(cfg, m, 0)
    (A32, a032, ax32, ay32, I32, S32), info = gpu_clip(cfg, m, prec)
    errs = {"amplitudes": np.abs(A32 - A64).max() / np.abs(A64).max(),
            "a0": np.abs(a032 - a064).max() / np.abs(a064).max(),
            "ax": np.abs(ax32 - ax64).max() / np.abs(ax64).max(),
            "ay": np.abs(ay32 - ay64).max() / np.abs(ay64).max(),
            "abbe image": np.abs(I32 - I64).max() / I64.max(),
            "socs image": np.abs(S32 - S64).max() / np.abs(S64).max()}
    say(f"  {form}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()) + f"; info {info[0].tolist()}")
    assert errs["abbe image"] < 1e-3 and errs["socs image"] < 1e-3
    assert errs["amplitudes"] < 1e-2 and errs["a0"] < 1e-2
    assert errs["abbe image"] > 0.0


def test_fp32_imaging_alone(c2):
    """FP32 imaging on the same FP64 amplitudes and M3D parameters: the
    Abbe image (x pass on FP32 MFMAs, compensated source sums) and the SOCS
    image within 1e-5 of their FP64 peaks."""
    from euvlitho_amd import _lib
    from euvlitho_amd.engine import Engine
    cfg, m = c2
    out = {}
    for name, prec in (("fp64", 0), ("fp32", _lib.PREC_FP32_IMAGING)):
        eng = Engine(cfg, precision=prec)
        bits = eng.upload_mask(m)
        if name == "fp64":
            Ax = eng.amp_s(bits, "X")
            a0, ax, ay = eng.m3d(bits, Ax0=Ax[0])
        out[name] = (eng.abbe(bits, Ax=Ax).cpu().numpy(),
                     eng.socs(bits, a0, ax, ay, nsocs=24, nsocsxy=20).cpu().numpy())
    e1 = np.abs(out["fp32"][0] - out["fp64"][0]).max() / out["fp64"][0].max()
    e2 = np.abs(out["fp32"][1] - out["fp64"][1]).max() / np.abs(out["fp64"][1]).max()
    say(f"  FP32 imaging: Abbe {e1:.2e}, SOCS {e2:.2e} of peak")
    assert 0.0 < e1 < 1e-5 and 0.0 < e2 < 1e-5

"""GPU end to end at the reference-default optics against the reference's own
outputs: m3d/inputxx.csv (M3D parameters, N = 8705 modes) and abbe/emint.csv
(rigorous Abbe image, 512 x 512).  The fixtures hold 6 significant digits."""
import gzip
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_rows(path):
    rows = []
    with open(path) as f:
        for line in f:
            v = [x for x in line.strip().split(",") if x]
            if v:
                rows.append(np.array([float(x) for x in v]))
    return rows


def test_reference_default_matches_fixtures():
    from euvlitho_amd.config import reference_default
    from euvlitho_amd.engine import Engine
    print("reference default: N=8705 EM solve, M3D fit, Abbe 512x512", flush=True)
    cfg = reference_default()
    eng = Engine(cfg)
    with open(os.path.join(GOLDEN, "ref_mask.bin"), "rb") as f:
        bits = eng.upload_mask(f.read())
    Ax = eng.amp_s(bits, "X")
    print("  amp_s info", eng.last_info.tolist(), flush=True)
    a0, ax, ay = [t.cpu().numpy() for t in eng.m3d(bits, Ax0=Ax[0])]
    _, _, snum = cfg.pupil
    x = snum >= 8
    got = [a0.real, a0.imag, ax[x].real, ax[x].imag, ay[x].real, ay[x].imag]
    ref = read_rows(os.path.join(GOLDEN, "ref_inputxx.csv"))
    for i, (g, r) in enumerate(zip(got, ref)):
        e = np.abs(g - r).max() / np.abs(r).max()
        print(f"  inputxx row {i}: max err / row max = {e:.3e}", flush=True)
        assert e < 2e-5
    img = eng.abbe(bits, Ax=Ax).cpu().numpy()
    with gzip.open(os.path.join(GOLDEN, "ref_emint.csv.gz"), "rt") as f:
        rows = [np.array([float(v) for v in line.strip().split(",") if v]) for line in f if line.strip()]
    R = np.array(rows)  # row r = y index 511 - r
    ref_img = R[::-1, :].T
    e = np.abs(img - ref_img).max() / ref_img.max()
    print(f"  emint max err / peak = {e:.3e}", flush=True)
    assert e < 2e-5

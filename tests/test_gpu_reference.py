"""GPU end to end at the reference-default optics against the reference's own
outputs: m3d/inputxx.csv (M3D parameters, N = 8705 modes) and abbe/emint.csv
(rigorous Abbe image, 512 x 512).

The fixtures carry 6 significant digits, so the bar is 2e-5 of each row's
maximum (and of the image peak).  This build does not meet it: the FP64 oracle
(eigendecomposition route, LAPACK) and this GPU path (matrix-function route)
agree with each other to ~1e-12 on this problem, and both miss the fixtures by
0.83 % (Re a0), 0.46 % (Im a0), 0.35 % / 0.26 % (ax), 3.5 % / 3.6 % (ay) of
the row maxima and 2.5 % of the image peak.  The test is marked as an
expected failure (strict), so the gap shows in every report, and an
agreement reached by a later fix turns it into an unexpected pass.

What the residual looks like: divided by the thin-mask shadow phase, the
a0 difference is the spectrum of a feature at the mask's y wrap (pattern rows
2047/0), full width in x (the l = 0 column carries most of it) and, off that
column, 75 % correlated with the mask's own boundary row; a single-row model
at the wrap explains 65 % of its energy.  Variants ruled out on this GPU path:
the commented-out 0.5 nm grid of the reference drivers, a truncated or
zero-padded mask file, flips, transposes, shifts and incidence-plane changes.
"""
import gzip
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# 6 significant digits in the fixtures
ROW_BAR = 2e-5
IMAGE_BAR = 2e-5
MEASURED = ("inputxx rows miss by 8.3e-3 / 4.6e-3 / 3.5e-3 / 2.6e-3 / 3.5e-2 / 3.6e-2 of the row maxima and "
            "emint.csv by 2.5e-2 of peak (oracle and GPU agree to 1e-12); the cause is open")


def read_rows(path):
    rows = []
    with open(path) as f:
        for line in f:
            v = [x for x in line.strip().split(",") if x]
            if v:
                rows.append(np.array([float(x) for x in v]))
    return rows


@pytest.mark.xfail(strict=True, reason=MEASURED)
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
    assert [len(r) for r in ref] == [len(g) for g in got]
    errs = []
    for i, (g, r) in enumerate(zip(got, ref)):
        e = np.abs(g - r).max() / np.abs(r).max()
        c = np.corrcoef(g, r)[0, 1]
        print(f"  inputxx row {i}: max err / row max = {e:.3e}, correlation {c:.6f}", flush=True)
        errs.append(e)
    img = eng.abbe(bits, Ax=Ax).cpu().numpy()
    with gzip.open(os.path.join(GOLDEN, "ref_emint.csv.gz"), "rt") as f:
        rows = [np.array([float(v) for v in line.strip().split(",") if v]) for line in f if line.strip()]
    R = np.array(rows)  # row r holds y index 511 - r
    ref_img = R[::-1, :].T
    e = np.abs(img - ref_img).max() / ref_img.max()
    c = np.corrcoef(img.ravel(), ref_img.ravel())[0, 1]
    print(f"  emint max err / peak = {e:.3e}, correlation {c:.6f}", flush=True)
    assert max(errs) < ROW_BAR and e < IMAGE_BAR

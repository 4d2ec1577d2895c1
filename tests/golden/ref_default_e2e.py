"""Oracle end to end at the reference-default optics on the reference mask.

Runs the FP64 oracle (one N = 8705 eigenproblem plus the two N = 1009 blank
solves) on the masks as the fixtures' program revision read them
(fixture_reads.py: every mask array read one row off in y) and compares with
the reference fixtures ref_inputxx.csv (m3d.cpp output) and
ref_emint.csv.gz (intensity.cpp output), both written by the reference with
6 significant digits.  With reading="committed" it runs the committed
sources' reading instead (uniform blanks, unshifted clip).  Writes
ref_default_e2e.json beside this file and, when asked, the oracle's arrays
(a0, ax, ay, image) to an npz file.  Usage: python tests/golden/ref_default_e2e.py
[committed] (about 12 minutes on 8 cores).
"""
import gzip
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import abbe as OA  # noqa: E402
from oracle import em, geometry as G, maskio, problems as P  # noqa: E402
from tests.golden.fixture_reads import fixture_reads  # noqa: E402


def compare_inputxx(rows_ref, a0, ax, ay, xflag):
    got = [a0.real, a0.imag, ax[xflag].real, ax[xflag].imag, ay[xflag].real, ay[xflag].imag]
    rel = [float(np.abs(g - r).max() / np.abs(r).max()) for g, r in zip(got, rows_ref)]
    return rel


def run(golden=HERE, out=os.path.join(HERE, "ref_default_e2e.json"), arrays=None, reading="fixture"):
    t0 = time.time()
    p = P.REF_DEFAULT
    d = G.derived(p)
    with open(os.path.join(golden, "ref_mask.bin"), "rb") as f:
        mask = maskio.unpack(f.read(), 2048, 2048)
    if reading == "fixture":
        em_mask, thin, b_vc, b_ab = [m.reshape(-1) for m in fixture_reads(mask.reshape(2048, 2048))]
        blank_masks = (b_vc, b_ab)
    else:
        em_mask, thin, blank_masks = mask, mask, None
    li, mi, L, M = G.order_set(p, "diamond", cutoff=6.0)
    Ax = em.amp_s("X", em_mask, p, d, li, mi, L, M, d["kx0"], d["ky0"])
    t1 = time.time()
    vc, ab = OA.blank_amplitudes(p, d, masks=blank_masks)
    a0, ax, ay = OA.m3d_params(p, d, thin, li, mi, Ax, vc, ab)
    lin, min_, snum = G.pupil_points(p, d)
    rows = maskio.read_inputxx(os.path.join(golden, "ref_inputxx.csv"))
    rel_rows = compare_inputxx(rows, a0, ax, ay, snum >= 8)
    nd = G.ndivs_rule(p, d)
    src = G.source_lattice(p, d, nd, 1, 0.9, 0.55, 90.0)
    divs = [((d["kx0"], d["ky0"], 0.0, 0.0),) + src[(0, 0)]]
    img = OA.abbe_image(p, d, em_mask, li, mi, L, M, divs, 512, 512, amps=[Ax])
    with gzip.open(os.path.join(golden, "ref_emint.csv.gz"), "rt") as f:
        tmp = os.path.join(HERE, "_emint_tmp.csv")
        with open(tmp, "w") as g:
            g.write(f.read())
    ref_img = maskio.read_image_csv(tmp)
    os.remove(tmp)
    rel_img = float(np.abs(img - ref_img).max() / np.abs(ref_img).max())
    rep = dict(
        reading=reading, N=int(len(li)), ncut=int(len(lin)), n_src=int(len(src[(0, 0)][0])),
        inputxx_rel_per_row=rel_rows, inputxx_max_rel_row=max(rel_rows), emint_max_rel=rel_img,
        seconds_em=t1 - t0, seconds_total=time.time() - t0,
    )
    with open(out, "w") as f:
        json.dump(rep, f, indent=1)
    if arrays:
        np.savez_compressed(arrays, a0=a0, ax=ax, ay=ay, img=img)
    return rep


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "committed":
        print(json.dumps(run(out=os.path.join(HERE, "ref_default_e2e_committed.json"), reading="committed"),
                         indent=1))
    else:
        print(json.dumps(run(), indent=1))

"""Training-data steps restated (TEST INFRASTRUCTURE ONLY).

* ``compress``: cnn/intensity/m3d/compress.py:16-110 with its loops, the
  float16 mask arrays from ``oracle.cnn.compmask`` (pinned to the script's
  own float16 dots), csv.reader for the text files;
* ``rshift`` / ``ishift`` / ``getitem``: cnn/cnn/model/re0/re0.py:18-62
  (float32 tables and arithmetic, np.roll).
"""
import csv
import math

import numpy as np

from . import cnn as OC


def compress(mask_bytes, inputxx_path, inputlm_path, ndata, ndiv=2048, nout=512):
    """Returns (maskall, [re0, im0, rex, imx, rey, imy], lorder, morder,
    lorders, morders) as compress.py builds them."""
    per = ndiv * ndiv // 8
    maskall = np.empty((ndata, nout, nout), dtype="float16")
    for idx in range(ndata):
        maskall[idx] = OC.compmask(mask_bytes[idx * per:(idx + 1) * per], ndiv, nout)
    with open(inputlm_path, newline="") as finputlm:
        content = [[int(row[0]), int(row[1]), int(row[2])] for row in csv.reader(finputlm)]
    ncut = len(content)
    lorder = np.empty(ncut, dtype=int)
    morder = np.empty(ncut, dtype=int)
    xorder = np.empty(ncut, dtype=int)
    for i in range(ncut):
        lorder[i], morder[i], xorder[i] = content[i]
    ncuts = int((xorder == 1).sum())
    arrays = [np.empty((ndata, n), dtype="float32") for n in (ncut, ncut, ncuts, ncuts, ncuts, ncuts)]
    with open(inputxx_path, newline="") as finputxx:
        freader = csv.reader(finputxx)
        for idx in range(ndata):
            for a in arrays:
                row = next(freader)
                for i in range(a.shape[1]):
                    a[idx][i] = float(row[i])
    lorders = np.empty(ncuts, dtype=int)
    morders = np.empty(ncuts, dtype=int)
    nc = 0
    for i in range(ncut):
        if xorder[i] == 1:
            lorders[nc] = lorder[i]
            morders[nc] = morder[i]
            nc = nc + 1
    return maskall, arrays, lorder, morder, lorders, morders


def tables(nout=512):
    sinlm = np.empty(nout, dtype="float32")
    coslm = np.empty(nout, dtype="float32")
    for i in range(nout):
        sinlm[i] = math.sin(2. * math.pi * i / nout)
        coslm[i] = math.cos(2. * math.pi * i / nout)
    return sinlm, coslm


def rshift(ncut, lorder, morder, rea, ima, idx, lshift, mshift, nout=512):
    sinlm, coslm = tables(nout)
    ramp = np.empty(ncut, dtype="float32")
    for i in range(ncut):
        lm = (lorder[i] * lshift + morder[i] * mshift) % nout
        ramp[i] = rea[idx][i] * coslm[lm] - ima[idx][i] * sinlm[lm]
    return ramp


def ishift(ncut, lorder, morder, rea, ima, idx, lshift, mshift, nout=512):
    sinlm, coslm = tables(nout)
    iamp = np.empty(ncut, dtype="float32")
    for i in range(ncut):
        lm = (lorder[i] * lshift + morder[i] * mshift) % nout
        iamp[i] = rea[idx][i] * sinlm[lm] + ima[idx][i] * coslm[lm]
    return iamp


def getitem_mask(maskall, idx, lshift, mshift):
    """The mask half of MaskampDataset.__getitem__ (re0.py:50-62)."""
    return np.roll(maskall[idx], (mshift, -lshift), axis=(0, 1))

"""Training data for the CNN M3D predictor.

* ``compress``: cnn/intensity/m3d/compress.py:16-110 -- a mask.bin of ndata
  clips, the inputxx.csv that m3d.cpp wrote for them and inputlm.csv become
  maskdata/mask.npy (float16 [ndata, 512, 512], compmask on the GPU),
  ampdata/{re0,im0,rex,imx,rey,imy}.npy (float32) and
  inputlm/{lorder,morder,lorders,morders}.npy, the same files.
* ``generate``: the whole chain mask.cpp -> m3d.cpp -> compress.py on the
  GPU without the text files in between: line/space clips rasterised on the
  device, the rigorous M3D parameters of each clip (EM solves + fit), the
  down-sampled masks; optionally the intermediate files as well.
* ``augment``: the data loader's random shift (MaskampDataset.__getitem__,
  cnn/cnn/model/re0/re0.py:24-62) for a batch of samples at once.
"""
import ctypes
import math
import os

import numpy as np
import torch

from . import _lib, csvio

NDIV, NDIVM = 2048, 512


def _handle(device):
    from .cnn import _handle as h
    return h(device)


def _save(path, arr):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, arr)


def down_sampled_masks(bits, ndiv=NDIV, nout=NDIVM, device="cuda:0", chunk=64):
    """mask.bin bytes (host or device) of n clips -> float16 [n, nout, nout]
    host array (compmask on the GPU; every value is a multiple of
    (nout/ndiv)^2, exact in float16)."""
    from .cnn import compmask
    if isinstance(bits, (bytes, bytearray)):
        bits = torch.from_numpy(np.frombuffer(bits, dtype=np.uint8).copy()).to(device)
    per = ndiv * ndiv // 8
    n = bits.numel() // per
    out = np.empty((n, nout, nout), dtype=np.float16)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        out[s:e] = compmask(bits[s * per:e * per], ndiv, nout)[:, 0].half().cpu().numpy()
    return out


def _write_npy(out, masks, amps, lorder, morder, xorder):
    _save(os.path.join(out, "maskdata", "mask.npy"), masks)
    for name, a in zip(("re0", "im0", "rex", "imx", "rey", "imy"), amps):
        _save(os.path.join(out, "ampdata", f"{name}.npy"), a)
    sel = np.asarray(xorder) == 1
    lorder, morder = np.asarray(lorder, dtype=int), np.asarray(morder, dtype=int)
    _save(os.path.join(out, "inputlm", "lorder.npy"), lorder)
    _save(os.path.join(out, "inputlm", "morder.npy"), morder)
    _save(os.path.join(out, "inputlm", "lorders.npy"), lorder[sel])
    _save(os.path.join(out, "inputlm", "morders.npy"), morder[sel])


def compress(mask="maska9398.bin", inputxx="inputxx.csv", inputlm="inputlm.csv", ndata=2000, out=".",
             ndiv=NDIV, nout=NDIVM, device="cuda:0"):
    """compress.py:16-110.  ncut and ncuts come from inputlm.csv (1901 and
    1749 in the reference); each clip's six inputxx rows are read over their
    first ncut / ncuts values and stored as float32."""
    lorder, morder, xorder = csvio.read_inputlm(inputlm)
    ncut, ncuts = len(lorder), int((xorder == 1).sum())
    with open(mask, "rb") as f:
        bits = f.read((ndiv * ndiv // 8) * ndata)
    if len(bits) < (ndiv * ndiv // 8) * ndata:
        raise ValueError(f"{mask}: fewer than {ndata} clips of {ndiv}x{ndiv}")
    masks = down_sampled_masks(bits, ndiv, nout, device)
    lines = [l for l in open(inputxx, newline="")]
    if len(lines) < 6 * ndata:
        raise ValueError(f"{inputxx}: {len(lines)} rows for {ndata} clips")
    amps = [np.empty((ndata, n), dtype=np.float32) for n in (ncut, ncut, ncuts, ncuts, ncuts, ncuts)]
    for idx in range(ndata):
        for r in range(6):
            vals = lines[6 * idx + r].rstrip("\r\n").split(",")
            a = amps[r]
            a[idx] = np.array([float(v) for v in vals[:a.shape[1]]], dtype=np.float64)
    _write_npy(out, masks, amps, lorder, morder, xorder)
    return masks, amps


def generate(ndata, out=".", cfg=None, seed=1, cd=56, gap=80, nout=NDIVM, device="cuda:0",
             write_intermediates=False, progress=None):
    """mask.cpp -> m3d.cpp -> compress.py for ndata line/space clips of the
    problem cfg (default: the reference drivers' optics), all on the GPU.
    Writes the npy files of ``compress``; with write_intermediates also
    mask.bin, inputlm.csv and inputxx.csv (exact number form), which
    ``compress`` turns into the same npy files."""
    from .config import reference_default
    from .engine import Engine
    from .masks import line_space_bits
    cfg = cfg or reference_default()
    eng = Engine(cfg, device)
    bits = line_space_bits(cfg.ndivx, cfg.ndivy, ndata, seed, cd, gap, device)
    per = cfg.ndivx * cfg.ndivy // 8
    lin, min_, snum = cfg.pupil
    xinput = (np.asarray(snum) >= 8).astype(np.int64)  # inputlm.csv x: at least 8 lattice sources
    sel = xinput.astype(bool)
    ncut, ncuts = len(lin), int(sel.sum())
    amps = [np.empty((ndata, n), dtype=np.float32) for n in (ncut, ncut, ncuts, ncuts, ncuts, ncuts)]
    if write_intermediates:
        os.makedirs(out, exist_ok=True)
        csvio.write_inputlm(os.path.join(out, "inputlm.csv"), lin, min_, xinput)
    for idx in range(ndata):
        clip = bits[idx * per:(idx + 1) * per]
        a0, ax, ay = (t.cpu().numpy() for t in eng.m3d(clip))
        ax, ay = ax[sel], ay[sel]
        for r, v in enumerate((a0.real, a0.imag, ax.real, ax.imag, ay.real, ay.imag)):
            amps[r][idx] = v
        if write_intermediates:
            csvio.write_inputxx(os.path.join(out, "inputxx.csv"), a0, ax, ay, xinput, "repr", append=idx > 0)
        if progress:
            progress(idx + 1, ndata)
    if write_intermediates:
        with open(os.path.join(out, "mask.bin"), "wb") as f:
            f.write(bits.cpu().numpy().tobytes())
    masks = down_sampled_masks(bits, cfg.ndivx, nout, device) if cfg.ndivx == cfg.ndivy else None
    _write_npy(out, masks, amps, lin, min_, xinput)
    eng.close()
    return masks, amps


def trig_tables(nout=NDIVM):
    """The float32 sinlm / coslm tables of re0.py:18-22 (math.sin/cos in
    double, rounded to float32)."""
    s = np.empty(nout, dtype=np.float32)
    c = np.empty(nout, dtype=np.float32)
    for i in range(nout):
        s[i] = math.sin(2.0 * math.pi * i / nout)
        c[i] = math.cos(2.0 * math.pi * i / nout)
    return c, s


def augment(masks, re, im, lorder, morder, lshift, mshift, device="cuda:0"):
    """MaskampDataset.__getitem__ for a batch: masks float16 [n, nout, nout]
    rolled by (mshift, -lshift); targets rshift(re, im) and ishift(re, im)
    (float32 [n, ncut]).  Inputs may be host arrays or device tensors;
    returns device tensors (masks, ramp, iamp).  (re0.py trains on ramp;
    im0.py, imx.py and imy.py on iamp.)"""
    dev = torch.device(device)
    to = lambda a, dt: (a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a))).to(
        device=dev, dtype=dt).contiguous()
    m = to(masks, torch.float16) if masks is not None else None
    r, i = to(re, torch.float32), to(im, torch.float32)
    n, ncut = r.shape
    nout = m.shape[-1] if m is not None else NDIVM
    shifts = np.stack([np.asarray(lshift, dtype=np.int32).reshape(n), np.asarray(mshift, dtype=np.int32).reshape(n)],
                      axis=1).copy()
    lo, mo = np.ascontiguousarray(lorder, dtype=np.int32), np.ascontiguousarray(morder, dtype=np.int32)
    cs, sn = trig_tables(nout)
    om = torch.empty_like(m) if m is not None else None
    orr, oi = torch.empty_like(r), torch.empty_like(i)
    lib, h = _handle(dev)
    p = lambda t: _lib.ptr(t) if t is not None else None
    _lib.check(lib.euvl_augment_batch(h, n, nout, p(m), shifts.ctypes.data, ncut, lo.ctypes.data, mo.ctypes.data,
                                      p(r), p(i), cs.ctypes.data, sn.ctypes.data, p(om), p(orr), p(oi)),
               "euvl_augment_batch")
    return om, orr, oi

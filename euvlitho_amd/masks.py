"""Mask producers and the mask.bin container.

mask.bin: ``NDIVX*NDIVY/8`` bytes per clip, clips concatenated, MSB-first
bits, bit index ``NDIVX*j + i`` (i = x fastest, j = row from the top), bit 1 =
vacuum (reflecting), 0 = absorber.  Format of cnn/intensity/mask/mask.cpp:9-47,
read back by cnn/intensity/abbe/intensity.cpp:20-29,197-207.
"""
import numpy as np


def manhattan(ndivx, ndivy, density, seed, side=(16, 64)):
    """Seeded random Manhattan clip: vacuum rectangles with integer sides drawn
    uniformly from ``side`` (nm = px), uniform positions with periodic wrap, added
    until the vacuum fraction reaches ``density``.  Returns mask2d (int8,
    length ndivx*ndivy, index ndivx*j + i)."""
    rng = np.random.default_rng(seed)
    m = np.zeros((ndivy, ndivx), dtype=np.int8)
    target = int(round(density * ndivx * ndivy))
    filled = 0
    while filled < target:
        w = int(rng.integers(side[0], side[1] + 1))
        h = int(rng.integers(side[0], side[1] + 1))
        x0 = int(rng.integers(0, ndivx))
        y0 = int(rng.integers(0, ndivy))
        ys = (y0 + np.arange(h)) % ndivy
        xs = (x0 + np.arange(w)) % ndivx
        blk = m[np.ix_(ys, xs)]
        filled += int(blk.size - blk.sum())
        m[np.ix_(ys, xs)] = 1
    return m.reshape(-1)


def pack_bits(mask2d):
    """MSB-first packing of one clip (mask.cpp:9-26)."""
    return np.packbits(np.asarray(mask2d, dtype=np.uint8).reshape(-1), bitorder="big").tobytes()


def unpack_bits(buf, ndivx, ndivy, sample=0):
    """Inverse of ``pack_bits`` for clip ``sample`` of a concatenated file."""
    nbytes = ndivx * ndivy // 8
    raw = np.frombuffer(buf, dtype=np.uint8, count=nbytes, offset=sample * nbytes)
    return np.unpackbits(raw, bitorder="big").astype(np.int8)


def read_mask_bin(path, ndivx, ndivy, sample=0):
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < (sample + 1) * ndivx * ndivy // 8:
        raise ValueError(f"{path}: {len(buf)} bytes, too short for clip {sample} of {ndivx}x{ndivy}")
    return unpack_bits(buf, ndivx, ndivy, sample)


def write_mask_bin(path, masks):
    with open(path, "wb") as f:
        for m in masks:
            f.write(pack_bits(m))


# ---------------------------------------------------------------- line/space generators
#
# The reference's line/space generators draw a few hundred random numbers per
# clip and fill the clip pixel by pixel.  Here the draws become a band layout
# on the host (which bands, which 1-D pattern each band repeats) and the GPU
# rasterises the layout straight into packed mask.bin bits
# (euvl_mask_raster / euvl_mask_linespace); no pixel loop runs on the host.
#
# Orientation: the Python generators return arrays M[x][y'] whose second index
# runs bottom-up (the reader linear.py:203-215 turns mask.bin into that form
# with reshape(NDIVY, NDIVX), flipud and a transpose), so M[x][y'] is mask.bin
# pixel (i = x, j = ndivy - 1 - y').

import random as _random
from dataclasses import dataclass, field

CONST_ONE, CONST_ZERO = -1, -2


@dataclass
class MaskLayout:
    """One clip as the AND of layers (then inverted if ``invert``): layer l
    cuts x (axis 0: column i selects a pattern over y) or y (axis 1: row j
    selects a pattern over x); ``sel[l][pos]`` is a pattern index into
    ``patterns[l]`` or CONST_ONE / CONST_ZERO."""
    ndivx: int
    ndivy: int
    axis: list
    sel: list                      # per layer: int32 array over its axis
    patterns: list                 # per layer: list of uint8 arrays over the other axis
    invert: bool = False


def _layer_pool(lay):
    pools, offs, off = [], [], 0
    for l, pats in enumerate(lay.patterns):
        offs.append(off)
        plen = lay.ndivy if lay.axis[l] == 0 else lay.ndivx
        for p in pats:
            p = np.asarray(p, dtype=np.uint8)
            if p.shape != (plen,):
                raise ValueError(f"layer {l}: pattern of length {p.shape} where {plen} is needed")
            pools.append(p)
            off += plen
    return pools, offs, off


def raster_layouts(layouts, device="cuda:0", handle=None):
    """Rasterise clips of one size and layer count on the GPU; returns the
    packed mask.bin bytes of all clips (uint8 tensor on ``device``)."""
    import ctypes
    import torch
    from . import _lib
    if not layouts:
        raise ValueError("no layouts")
    nx, ny, nl = layouts[0].ndivx, layouts[0].ndivy, len(layouts[0].axis)
    if any((l.ndivx, l.ndivy, len(l.axis), l.invert) != (nx, ny, nl, layouts[0].invert) for l in layouts):
        raise ValueError("layouts of one call must share size, layer count and inversion")
    stride = max(nx, ny)
    axis = np.zeros(len(layouts) * nl, dtype=np.int32)
    sel = np.full((len(layouts) * nl, stride), CONST_ZERO, dtype=np.int32)
    offs = np.zeros(len(layouts) * nl, dtype=np.int64)
    chunks, base = [], 0
    for m, lay in enumerate(layouts):
        pools, loffs, size = _layer_pool(lay)
        for l in range(nl):
            L = m * nl + l
            axis[L] = lay.axis[l]
            s = np.asarray(lay.sel[l], dtype=np.int32)
            sel[L, :len(s)] = s
            offs[L] = base + loffs[l]
        chunks.extend(pools)
        base += size
    pool = np.concatenate(chunks) if chunks else np.zeros(1, dtype=np.uint8)
    lib = _lib.load()
    out = torch.empty(len(layouts) * nx * ny // 8, dtype=torch.uint8, device=device)
    own = handle is None
    h = ctypes.c_void_p() if own else handle
    if own:
        _lib.check(lib.euvl_create(ctypes.byref(h)), "euvl_create")
    try:
        _lib.check(lib.euvl_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)),
                   "euvl_set_stream")
        _lib.check(lib.euvl_mask_raster(h, nx, ny, len(layouts), nl, axis.ctypes.data, sel.ctypes.data,
                                        offs.ctypes.data, pool.ctypes.data, int(base), int(layouts[0].invert),
                                        _lib.ptr(out)), "euvl_mask_raster")
    finally:
        if own:
            lib.euvl_destroy(h)
    return out


def line_space_bits(ndivx=2048, ndivy=None, n_masks=1, seed=1, cd=56, gap=80, device="cuda:0", handle=None):
    """The reference's mask program (cnn/intensity/mask/mask.cpp:28-62 with
    mVBF.h:1-70) on the GPU: ``n_masks`` clips from one C-library rand()
    stream seeded with ``seed`` (the reference's own masks: seed 1).  Returns
    the mask.bin bytes as a uint8 tensor on ``device``."""
    import ctypes
    import torch
    from . import _lib
    ndivy = ndivx if ndivy is None else ndivy
    lib = _lib.load()
    out = torch.empty(n_masks * ndivx * ndivy // 8, dtype=torch.uint8, device=device)
    own = handle is None
    h = ctypes.c_void_p() if own else handle
    if own:
        _lib.check(lib.euvl_create(ctypes.byref(h)), "euvl_create")
    try:
        _lib.check(lib.euvl_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)),
                   "euvl_set_stream")
        _lib.check(lib.euvl_mask_linespace(h, seed, ndivx, ndivy, cd, gap, n_masks, _lib.ptr(out)),
                   "euvl_mask_linespace")
    finally:
        if own:
            lib.euvl_destroy(h)
    return out


def line_space_plan(ndivx=2048, ndivy=None, n_masks=1, seed=1, cd=56, gap=80):
    """Host-only layouts of ``line_space_bits`` (euvl_mask_linespace_plan):
    one MaskLayout per clip."""
    import ctypes
    from . import _lib
    ndivy = ndivx if ndivy is None else ndivy
    stride = max(ndivx, ndivy)
    sel = np.zeros((n_masks, stride), dtype=np.int32)
    offs = np.zeros(n_masks, dtype=np.int64)
    cap = n_masks * (ndivx // max(cd, 1) + 2) * ndivy
    pool = np.zeros(max(cap, 1), dtype=np.uint8)
    need = ctypes.c_int64()
    lib = _lib.load()
    _lib.check(lib.euvl_mask_linespace_plan(seed, ndivx, ndivy, cd, gap, n_masks, sel.ctypes.data, offs.ctypes.data,
                                            pool.ctypes.data, cap, ctypes.byref(need)), "euvl_mask_linespace_plan")
    ends = list(offs[1:]) + [need.value]
    out = []
    for m in range(n_masks):
        npat = (ends[m] - offs[m]) // ndivy
        pats = [pool[offs[m] + q * ndivy: offs[m] + (q + 1) * ndivy].copy() for q in range(npat)]
        out.append(MaskLayout(ndivx, ndivy, [0], [sel[m, :ndivx].copy()], [pats]))
    return out


def bits_to_pattern(bits, ndivx, ndivy, sample=0):
    """mask.bin bytes (numpy or tensor) -> the Python generators' array
    M[x][y'] (int), as linear.py:203-215 reads a mask.bin."""
    raw = bits.cpu().numpy() if hasattr(bits, "cpu") else np.asarray(bits)
    m = unpack_bits(raw.tobytes(), ndivx, ndivy, sample).reshape(ndivy, ndivx)
    return np.flipud(m).T.astype(int)


class LinePattern:
    """elitho/mask_pattern.py:18-94: column (direction "H") or row ("V")
    bands of width cd that alternate between empty and a line/space pattern
    whose segments are multiples of gap; "BF" inverts.  Same constructor,
    same draws from ``rng`` (default: the ``random`` module, as the
    reference), same result; the raster runs on the GPU."""

    def __init__(self, cd: int = 56, gap: int = 80, direction: str = "H", field_type: str = "BF"):
        self.cd, self.gap, self.direction, self.field_type = cd, gap, direction, field_type

    def _line(self, n, rng):
        # rand_mask_line (mask_pattern.py:32-54): 1 on "line" segments
        m = np.zeros(n, dtype=np.uint8)
        line = rng.randint(0, 1)
        s = 0
        while s < n - self.gap:
            a = self.gap * rng.randint(0, 4)
            if line == 1:
                a = a + 3 * self.gap
            m[s:min(s + a, n)] = 1 if line == 1 else 0
            line = 1 - line
            s += a
        if line == 0:
            m[s:] = 1
        return m

    def _bands(self, n_outer, n_inner, rng):
        # _generate (mask_pattern.py:56-78) over outer positions: per position a
        # pattern index or a constant, and the patterns over the inner axis
        sel = np.full(n_outer, CONST_ZERO, dtype=np.int32)
        pats, last = [], None
        space = rng.randint(0, 1)
        s = 0
        while s < n_outer - self.cd:
            if space == 1:
                space = 0
            else:
                pats.append(self._line(n_inner, rng))
                last = len(pats) - 1
                sel[s:min(s + self.cd, n_outer)] = last
                space = 1
            s += self.cd
        if space == 0:
            sel[s:] = CONST_ONE
        else:
            sel[s:] = last
        return sel, pats

    def layout(self, width, height, rng=None):
        rng = _random if rng is None else rng
        if self.direction == "H":
            sel, pats = self._bands(width, height, rng)          # M[x][y'] = pattern over y'
            lay = MaskLayout(width, height, [0], [sel], [[p[::-1].copy() for p in pats]])
        elif self.direction == "V":
            sel, pats = self._bands(height, width, rng)          # M[x][y'] = G[y'][x]
            lay = MaskLayout(width, height, [1], [sel[::-1].copy()], [pats])
        else:
            raise ValueError("Invalid line pattern direction")
        lay.invert = self.field_type == "BF"
        return lay

    def bits(self, width, height, n=1, rng=None, device="cuda:0", handle=None):
        """``n`` clips drawn in sequence -> packed mask.bin bytes on ``device``."""
        return raster_layouts([self.layout(width, height, rng) for _ in range(n)], device, handle)

    def __call__(self, width, height, rng=None, device="cuda:0"):
        return bits_to_pattern(self.bits(width, height, 1, rng, device), width, height)


class RandomLineSpacePattern:
    """elitho/mask_pattern.py:97-196: the AND of bands along y' and bands
    along x of widths int(cd exp(fac u)), each empty or a line/space pattern
    of segments int(cd exp(fac1d u)); "DF" inverts.  Square clips only (the
    reference indexes its (height, width) array as [x][y'])."""

    def __init__(self, cd: int = 60, fac: float = 1.0, fac1d: float = 3.0, field_type: str = "BF"):
        self.cd, self.fac, self.fac1d, self.field_type = cd, fac, fac1d, field_type

    def _randmask(self, n, rng):
        # randmask (mask_pattern.py:111-133): 0 on "line" segments
        m = np.zeros(n, dtype=np.uint8)
        line = rng.randint(0, 1)
        s = 0
        while s < n - self.cd:
            a = int(self.cd * np.exp(self.fac1d * rng.random()))
            m[s:min(s + a, n)] = 0 if line == 1 else 1
            line = 1 - line
            s += a
        m[s:] = 0 if line == 0 else 1
        return m

    def _pass(self, n_outer, n_inner, rng, first):
        # one direction of _generate (mask_pattern.py:140-188): outer positions
        # select a pattern over the inner axis (or a constant 1).  The second
        # pass's tail loop sets space = 1 at its first position
        # (mask_pattern.py:183-185), so after a space band only that first
        # tail position keeps the first pass's value.
        sel = np.full(n_outer, CONST_ONE, dtype=np.int32)
        pats = [np.zeros(n_inner, dtype=np.uint8)]  # the array a tail reads before any draw
        last = 0
        space = rng.randint(0, 1)
        s = 0
        while s < n_outer - self.cd:
            a = int(self.cd * np.exp(self.fac * rng.random()))
            if space == 1:
                space = 0
            else:
                pats.append(self._randmask(n_inner, rng))
                last = len(pats) - 1
                sel[s:min(s + a, n_outer)] = last
                space = 1
            s += a
        if space == 1:
            sel[s:] = last
        elif not first:
            sel[s + 1:] = last
        return sel, pats

    def layout(self, width, height, rng=None):
        if width != height:
            raise ValueError("RandomLineSpacePattern needs a square clip (width == height)")
        rng = _random if rng is None else rng
        n = width
        # first pass: bands over y' (M's second index), patterns over x
        sel1, pats1 = self._pass(n, n, rng, True)
        # second pass: bands over x, patterns over y' (min with the first)
        sel2, pats2 = self._pass(n, n, rng, False)
        lay = MaskLayout(n, n, [1, 0], [sel1[::-1].copy(), sel2], [pats1, [p[::-1].copy() for p in pats2]])
        lay.invert = self.field_type == "DF"
        return lay

    def bits(self, width, height, n=1, rng=None, device="cuda:0", handle=None):
        return raster_layouts([self.layout(width, height, rng) for _ in range(n)], device, handle)

    def __call__(self, width, height, rng=None, device="cuda:0"):
        return bits_to_pattern(self.bits(width, height, 1, rng, device), width, height)

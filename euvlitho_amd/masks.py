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

"""mask.bin and CSV formats.

Restates decompressBits (cnn/intensity/abbe/intensity.cpp:20-29,197-207,
cnn/intensity/m3d/m3d.cpp:17-26,197-209), compressBits
(cnn/intensity/mask/mask.cpp:9-26), the inputxx.csv reader
(linear.py:229-242) and the image CSV layout (intensity.cpp:390-407).
"""
import numpy as np


def unpack(buf, ndivx, ndivy, sample=0):
    """mask2d[NDIVX*j + i] (i = x fastest); MSB-first bits, bit 1 = vacuum."""
    nbytes = ndivx * ndivy // 8
    raw = np.frombuffer(buf, dtype=np.uint8)[sample * nbytes:(sample + 1) * nbytes]
    out = np.zeros(ndivx * ndivy, dtype=np.int64)
    for b in range(8):  # intensity.cpp:23-24: bit i = 7..0 of each byte
        out[b::8] = (raw >> (7 - b)) & 1
    return out


def pack(mask2d):
    """compressBits: MSB-first, zero-padded last byte (mask.cpp:9-26)."""
    bits = np.asarray(mask2d, dtype=np.uint8).ravel()
    pad = (-len(bits)) % 8
    if pad:
        bits = np.concatenate([bits, np.zeros(pad, np.uint8)])
    out = np.zeros(len(bits) // 8, dtype=np.uint8)
    for b in range(8):
        out |= bits[b::8] << (7 - b)
    return out.tobytes()


def read_inputxx(path):
    """Six rows, trailing empty field dropped (linear.py:229-242)."""
    rows = []
    with open(path) as f:
        for line in f:
            vals = [v for v in line.strip().split(",") if v != ""]
            if vals:
                rows.append(np.array([float(v) for v in vals]))
    return rows


def read_image_csv(path):
    """Image CSV: row r holds y index n-1-r, columns are x (intensity.cpp:390-407)."""
    rows = read_inputxx(path)
    a = np.array(rows)  # [r, i]
    ny = a.shape[0]
    img = np.empty((a.shape[1], ny))
    for r in range(ny):
        img[:, ny - 1 - r] = a[r]
    return img

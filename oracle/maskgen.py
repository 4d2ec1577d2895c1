"""Mask generators, restated loop by loop (TEST INFRASTRUCTURE ONLY).

* ``crand``: the C library's own rand()/srand() through ctypes (the function
  mVBF.h calls; mask.cpp:31 leaves srand commented out, so the reference
  stream is seed 1).
* ``maskgen``: maskgen / randmask of cnn/intensity/mask/mVBF.h:2-39,41-70,
  one clip as mask2d[NDIVX*j + i]; ``mask_program`` is the ndata loop of
  cnn/intensity/mask/mask.cpp:41-60.
* ``line_pattern`` / ``random_line_space``: LinePattern and
  RandomLineSpacePattern of elitho/mask_pattern.py:18-94 and :97-196, with
  their pixel loops, returning the reference's arrays M[x][y'].  ``rng`` is
  anything with randint(a, b) and random() (the reference uses the
  ``random`` module).
"""
import ctypes

import numpy as np


class CRand:
    """randint/random facade over the C library's rand() (seeded by srand)."""

    def __init__(self, seed=1):
        self.libc = ctypes.CDLL("libc.so.6")
        self.libc.srand(ctypes.c_uint(seed))

    def rand(self):
        return self.libc.rand()

    def randint(self, a, b):  # rand() % (b - a + 1) + a: the C program's rand() % 2, rand() % 5
        return a + self.libc.rand() % (b - a + 1)


def randmask(n, gap, rng):
    """mVBF.h:41-70 (rand() % 2 and rand() % 5 as randint(0, 1), randint(0, 4))."""
    m = np.full(n, -1, dtype=np.int64)
    line = rng.randint(0, 1)
    s = 0
    while s < n - gap:
        a = gap * rng.randint(0, 4)
        if line == 1:
            a = a + 3 * gap
        for i in range(s, min(s + a, n)):
            m[i] = 0 if line == 1 else 1
        line = 0 if line == 1 else 1
        s = s + a
    for i in range(s, n):
        m[i] = 0 if line == 0 else 1
    return m


def maskgen(ndivx, ndivy, rng, cd=56, gap=80):
    """mVBF.h:2-39; returns mask2d[NDIVX*j + i]."""
    m2 = np.full((ndivy, ndivx), -1, dtype=np.int64)
    m1y = None
    space = rng.randint(0, 1)
    s = 0
    while s < ndivx - cd:
        if space == 1:
            for i in range(s, min(s + cd, ndivx)):
                m2[:, i] = 1
            space = 0
        else:
            m1y = randmask(ndivy, gap, rng)
            for i in range(s, min(s + cd, ndivx)):
                m2[:, i] = m1y
            space = 1
        s = s + cd
    for i in range(s, ndivx):
        m2[:, i] = 0 if space == 0 else m1y
    assert (m2 >= 0).all()
    return m2.reshape(-1)


def mask_program(ndivx, ndivy, ndata, seed=1, cd=56, gap=80):
    """mask.cpp:41-60: ndata clips from one rand() stream, concatenated
    mask.bin bytes (MSB-first, mask.cpp:9-26)."""
    rng = CRand(seed)
    out = [np.packbits(maskgen(ndivx, ndivy, rng, cd, gap).astype(np.uint8)) for _ in range(ndata)]
    return np.concatenate(out).tobytes()


def rand_mask_line(n, gap, rng):
    """mask_pattern.py:32-54."""
    m = np.zeros(n, dtype=int)
    line = rng.randint(0, 1)
    s = 0
    while s < n - gap:
        a = gap * rng.randint(0, 4)
        if line == 1:
            a = a + 3 * gap
        for i in range(s, min(s + a, n)):
            m[i] = 1 if line == 1 else 0
        line = 1 - line
        s += a
    for i in range(s, n):
        if line == 0:
            m[i] = 1
    return m


def _line_generate(width, height, cd, gap, rng):
    """LinePattern._generate (mask_pattern.py:56-78)."""
    m2 = np.zeros((width, height), dtype=int)
    space = rng.randint(0, 1)
    s = 0
    m1y = None
    while s < width - cd:
        if space == 1:
            space = 0
        else:
            m1y = rand_mask_line(height, gap, rng)
            for i in range(s, min(s + cd, width)):
                m2[i, :] = m1y
            space = 1
        s += cd
    for i in range(s, width):
        if space == 0:
            m2[i, :] = np.ones(height, dtype=int)
        else:
            m2[i, :] = m1y
    return m2


def line_pattern(width, height, rng, cd=56, gap=80, direction="H", field_type="BF"):
    """LinePattern.__call__ (mask_pattern.py:80-94)."""
    if direction == "H":
        m = _line_generate(width, height, cd, gap, rng)
    elif direction == "V":
        m = _line_generate(height, width, cd, gap, rng).T
    else:
        raise ValueError(direction)
    return 1 - m if field_type == "BF" else m


def _ls_randmask(n, cd, fac1d, rng):
    """RandomLineSpacePattern.randmask (mask_pattern.py:111-133)."""
    m = np.zeros(n, dtype=int)
    line = rng.randint(0, 1)
    s = 0
    while s < n - cd:
        a = int(cd * np.exp(fac1d * rng.random()))
        for i in range(s, min(s + a, n)):
            m[i] = 0 if line == 1 else 1
        line = 1 - line
        s += a
    for i in range(s, n):
        m[i] = 0 if line == 0 else 1
    return m


def random_line_space(width, height, rng, cd=60, fac=1.0, fac1d=3.0, field_type="BF"):
    """RandomLineSpacePattern._generate + __call__ (mask_pattern.py:135-196)."""
    m2 = np.zeros((height, width), dtype=int)
    m1x = np.zeros(width, dtype=int)
    m1y = np.zeros(height, dtype=int)
    space = rng.randint(0, 1)
    s = 0
    while s < height - cd:
        a = int(cd * np.exp(fac * rng.random()))
        if space == 1:
            for i in range(s, min(s + a, height)):
                for j in range(width):
                    m2[j, i] = 1
            space = 0
        else:
            m1x = _ls_randmask(width, cd, fac1d, rng)
            for i in range(s, min(s + a, height)):
                for j in range(width):
                    m2[j, i] = m1x[j]
            space = 1
        s += a
    for i in range(s, height):
        if space == 0:
            for j in range(width):
                m2[j, i] = 1
        else:
            for j in range(width):
                m2[j, i] = m1x[j]
    space = rng.randint(0, 1)
    s = 0
    while s < width - cd:
        a = int(cd * np.exp(fac * rng.random()))
        if space == 1:
            space = 0
        else:
            m1y = _ls_randmask(height, cd, fac1d, rng)
            for i in range(s, min(s + a, width)):
                for j in range(height):
                    m2[i, j] = min(m1y[j], m2[i, j])
            space = 1
        s += a
    for i in range(s, width):
        if space == 0:
            space = 1
        else:
            for j in range(height):
                m2[i, j] = min(m1y[j], m2[i, j])
    return 1 - m2 if field_type == "DF" else m2


def pattern_to_bits(m):
    """An M[x][y'] array -> mask.bin bytes (pixel (i = x, j = ndivy-1-y'),
    the inverse of the reader linear.py:203-215)."""
    m = np.asarray(m)
    return np.packbits(np.flipud(m.T).astype(np.uint8).reshape(-1)).tobytes()

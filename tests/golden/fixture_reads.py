"""The masks as the program revision that wrote the reference fixtures read them.

The reference's fixtures m3d/inputxx.csv and abbe/emint.csv (written by
cnn/intensity/m3d/m3d.cpp:468-558 and cnn/intensity/abbe/intensity.cpp:390-408
for the committed mask.bin) are not what the committed sources compute from
that mask: they miss it by 0.26-3.6 % of the row maxima and 2.5 % of the image
peak.  They are reproduced to their six significant digits (<= 3.6e-6 of each
row's maximum, 8.6e-7 of the image peak) when every read of a mask array in
y is one row off, i.e. pattern row p is taken from array row NDIVY - p instead
of NDIVY - 1 - p (the y flip of header.h:1157-1159 and m3d.cpp:287-291 with
the "- 1" missing).  Pattern row 0 then reads the row past the array's end:

* mask2d (the clip, `new int[NDIVSQ]`, m3d.cpp:83, intensity.cpp:111): zeros,
  i.e. an absorber row, in the EM solve and in m3d.cpp's thin mask alike;
* mask2d0 (the blank masks, m3d.cpp:84, allocated right after mask2d): the
  16 MiB array is a glibc mmap chunk of 16 MiB + 4 KiB whose data start 16
  bytes into the mapping, so the row past its end is 1020 zero ints (the rest
  of the last page), then the chunk header of mask2d's mapping, which lies
  directly above (prev_size = 0, 0; size = 0x01001002 | IS_MMAPPED in its low
  int, 0 in its high int), then the first 1024 pixels of mask2d itself.  A
  nonzero int reads as vacuum (header.h:1161-1164).

This module builds those four masks explicitly, so that the reference's own
outputs can pin this build's EM solve, M3D fit and Abbe image end to end.  It
is test infrastructure: the product reads masks as the committed sources do.
"""
import numpy as np

N_TAIL_ZEROS = 1020  # (4096 - 16) / 4: the last page of mask2d0's mapping after its data
HDR_NONZERO = 1022   # the low int of mask2d's mmap chunk size word


def _row_off(a, past_end):
    """Array rows as read with pattern row p <- array row NDIVY - p: row j of
    the result (mask2d layout, index NDIVX * j + i) holds row j + 1, and the
    last row holds what lies past the array's end."""
    out = np.empty_like(a)
    out[:-1] = a[1:]
    out[-1] = past_end
    return out


def fixture_reads(raw):
    """raw: (NDIVY, NDIVX) 0/1 array, raw[j, i] = mask2d[NDIVX * j + i] of mask.bin.

    Returns (em, thin, blank_vc, blank_ab), each (NDIVY, NDIVX) int8 in the same
    layout: the clip of the EM solves, the clip of m3d.cpp's thin mask, and the
    vacuum / absorber blank masks of the two narrow solves (m3d.cpp:214-239)."""
    raw = np.asarray(raw, dtype=np.int8)
    ny, nx = raw.shape
    em = _row_off(raw, 0)
    past0 = np.zeros(nx, dtype=np.int8)
    past0[HDR_NONZERO] = 1
    past0[N_TAIL_ZEROS + 4:] = raw[0, :nx - N_TAIL_ZEROS - 4]
    blank_vc = _row_off(np.ones_like(raw), past0)
    blank_ab = _row_off(np.zeros_like(raw), past0)
    return em, em.copy(), blank_vc, blank_ab

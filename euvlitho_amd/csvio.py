"""The reference's text file contract: inputlm.csv, inputxx.csv and the image
CSVs (emint.csv, emintY.csv, intft.csv, intlinear.csv, intsocs.csv).

Layouts:
  inputlm.csv   one line "l,m,x" per pupil point, ip-major (m3d.cpp:160-195);
                x = 1 when the point has at least 8 lattice sources.
  inputxx.csv   6 comma-terminated rows per clip (m3d.cpp:468-558):
                Re a0, Im a0 over every pupil point; Re ax, Im ax, Re ay, Im ay
                over the points with x = 1 only.  Several clips append.
  image CSVs    one row per y index from the top (y = ny-1 first), nx values
                per row, each followed by ',' (intensity.cpp:390-407,
                linear.py:245-258, socs.py:227-234).  Readers drop the empty
                field after the last comma (linear.py:233-235).

Number formats: "repr" (default) writes Python's shortest round-trip form
(17 significant digits at most, exact on reading back; what linear.py and
socs.py write); "g" writes C++ ostream's default 6-significant-digit form
(what m3d.cpp and intensity.cpp write), for byte diffs against reference
outputs.
"""
import numpy as np

FORMATS = ("repr", "g")


def _fmt(v, fmt):
    if fmt == "repr":
        return repr(float(v))
    if fmt == "g":
        return "%g" % v
    raise ValueError(f"unknown number format {fmt!r} (use one of {FORMATS})")


def _row(values, fmt):
    return "".join(_fmt(v, fmt) + "," for v in values) + "\n"


def _split(line):
    f = line.rstrip("\r\n").split(",")
    return [x for x in f if x != ""]


# ---------------------------------------------------------------- inputlm.csv
def write_inputlm(path, linput, minput, xinput):
    """xinput: the 0/1 flag (1 = at least 8 lattice sources)."""
    flag = np.asarray(xinput).astype(int)
    with open(path, "w") as f:
        for l, m, s in zip(np.asarray(linput), np.asarray(minput), flag):
            f.write(f"{int(l)},{int(m)},{int(s)}\n")


def read_inputlm(path):
    """(linput, minput, xinput) int arrays."""
    rows = [list(map(int, _split(line))) for line in open(path) if line.strip()]
    a = np.array(rows, dtype=np.int64).reshape(-1, 3)
    return a[:, 0], a[:, 1], a[:, 2]


# ---------------------------------------------------------------- inputxx.csv
def write_inputxx(path, a0, ax, ay, xinput, fmt="repr", append=False):
    """a0, ax, ay: complex, one value per pupil point (ax, ay are written only
    where xinput = 1); or ax, ay already reduced to those points."""
    a0 = np.asarray(a0)
    sel = np.asarray(xinput).astype(bool)
    ax, ay = np.asarray(ax), np.asarray(ay)
    if len(ax) == len(a0):
        ax, ay = ax[sel], ay[sel]
    if len(ax) != sel.sum() or len(ay) != sel.sum():
        raise ValueError("ax / ay do not match the pupil points with x = 1")
    with open(path, "a" if append else "w") as f:
        for v in (a0.real, a0.imag, ax.real, ax.imag, ay.real, ay.imag):
            f.write(_row(v, fmt))


def read_inputxx(path):
    """List of clips, each (re0, im0, rex, imx, rey, imy) float arrays, in the
    argument order of linear() (linear.py:190)."""
    rows = [np.array(_split(line), dtype=np.float64) for line in open(path) if line.strip()]
    if len(rows) % 6:
        raise ValueError(f"{path}: {len(rows)} rows, not a multiple of 6")
    return [tuple(rows[i:i + 6]) for i in range(0, len(rows), 6)]


def expand_xy(ax, xinput):
    """ax over the x = 1 points -> one value per pupil point (0 elsewhere),
    the axxx fill of linear.py:107-117 / socs.py:66-80 (xinput: the 0/1 flag)."""
    sel = np.asarray(xinput).astype(bool)
    out = np.zeros(len(sel), dtype=np.result_type(np.asarray(ax).dtype, np.complex128))
    if len(ax) != sel.sum():
        raise ValueError(f"{len(ax)} slope values for {int(sel.sum())} pupil points with x = 1")
    out[sel] = ax
    return out


# ------------------------------------------------------------------ images
def write_image(path, img, fmt="repr"):
    """img indexed [x][y] (nx x ny)."""
    img = np.real(np.asarray(img))
    nx, ny = img.shape
    with open(path, "w") as f:
        for j in range(ny - 1, -1, -1):
            f.write(_row(img[:, j], fmt))


def read_image(path):
    """Image CSV -> float array indexed [x][y]."""
    import gzip
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt") as f:
        rows = [np.array(_split(line), dtype=np.float64) for line in f if line.strip()]
    a = np.array(rows)
    return a[::-1].T.copy()

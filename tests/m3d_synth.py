"""Synthetic inputs for pinning the M3D least-squares fit (m3d.cpp:426-467).

The fit is affine in the wide amplitudes Ax: the thin-mask term fampxx
(m3d.cpp:329-356) enters as a constant.  So fit(Ax) - fit(0) is the fit of
f = Ax / phasexx alone, and these helpers build Ax from chosen per-pupil-point
(a0, ax, ay) so that f = a0 + ax * ls + ay * ms holds exactly over every valid
source point (ls = is - lsmaxX + lp/2, ms = js - lsmaxY + mp/2).
"""
import numpy as np


def valid_sources(lp, mp, lsx, lsy, nsx, nsy, in_na):
    """(is, js, ls, ms) of the source points that enter the fit of pupil
    point (lp, mp) (the two NA tests of m3d.cpp:434-437)."""
    out = []
    for is_ in range(nsx):
        for js in range(nsy):
            l_, m_ = is_ - lsx, js - lsy
            if in_na(l_, m_) and in_na(lp + l_, mp + m_):
                out.append((is_, js, l_ + lp / 2.0, m_ + mp / 2.0))
    return out


def identifiable(lin, min_, snum, lsx, lsy, nsx, nsy, in_na):
    """Per pupil point, which of (ax, ay) the fit determines: both where
    snum >= 8 and the sources span two dimensions; one where they lie on a
    single row (ms constant: ax) or column (ls constant: ay); none where
    snum < 8 (the mean)."""
    out = np.zeros((len(lin), 2), dtype=bool)
    for q, (lp, mp, sn) in enumerate(zip(lin, min_, snum)):
        if sn < 8:
            continue
        v = np.array([(ls, ms) for _, _, ls, ms in valid_sources(int(lp), int(mp), lsx, lsy, nsx, nsy, in_na)])
        out[q] = (np.ptp(v[:, 0]) > 0, np.ptp(v[:, 1]) > 0)
    return out


def linear_amplitudes(lin, min_, lidx, midx, lsx, lsy, nsx, nsy, in_na, phase, a0, ax, ay):
    """Ax[is, js, n] = phase[is, js] (a0 + ax ls + ay ms) of the pupil point
    whose order l = lp + is - lsx, m = mp + js - lsy is mode n.  Each
    (is, js, n) belongs to at most one pupil point, so the data are exact."""
    pos = {(int(l), int(m)): i for i, (l, m) in enumerate(zip(lidx, midx))}
    Ax = np.zeros((nsx, nsy, len(lidx)), dtype=np.complex128)
    for q, (lp, mp) in enumerate(zip(lin, min_)):
        for is_, js, ls, ms in valid_sources(int(lp), int(mp), lsx, lsy, nsx, nsy, in_na):
            n = pos[(int(lp) + is_ - lsx, int(mp) + js - lsy)]
            Ax[is_, js, n] = phase[is_, js] * (a0[q] + ax[q] * ls + ay[q] * ms)
    return Ax


def lstsq_fit(lin, min_, snum, lidx, midx, lsx, lsy, nsx, nsy, in_na, phase, Ax):
    """The fit of f = Ax / phase by numpy.linalg.lstsq on the design matrix
    [1, ls, ms] where snum >= 8, and the plain mean elsewhere."""
    pos = {(int(l), int(m)): i for i, (l, m) in enumerate(zip(lidx, midx))}
    out = np.zeros((3, len(lin)), dtype=np.complex128)
    for q, (lp, mp, sn) in enumerate(zip(lin, min_, snum)):
        rows, f = [], []
        for is_, js, ls, ms in valid_sources(int(lp), int(mp), lsx, lsy, nsx, nsy, in_na):
            n = pos[(int(lp) + is_ - lsx, int(mp) + js - lsy)]
            rows.append((1.0, ls, ms))
            f.append(Ax[is_, js, n] / phase[is_, js])
        if sn >= 8:
            out[:, q] = np.linalg.lstsq(np.array(rows), np.array(f), rcond=None)[0]
        else:
            out[0, q] = np.mean(f)
    return out

"""Order sets, pupil points, source lattice and derived optics constants.

Restates cnn/intensity/abbe/intensity.cpp:44-195,467-502,
cnn/intensity/m3d/m3d.cpp:44-195 and elitho/const.py:95-111.
A problem is a plain dict of raw parameters (see ``problems.py``).
"""
import math

import numpy as np

PI = math.atan(1.0) * 4.0  # intensity.cpp:54


def derived(p):
    """Derived integers and wave numbers (intensity.cpp:51-130, m3d.cpp:44-82)."""
    d = {}
    NDIVX, NDIVY, MX, MY = p["NDIVX"], p["NDIVY"], p["MX"], p["MY"]
    lam, NA = p["wavelength"], p["NA"]
    dx, dy = float(NDIVX), float(NDIVY)
    k = 2.0 * PI / lam
    phi0 = 90.0 - p["azimuth"]
    kx0 = k * math.sin(PI / 180.0 * p["theta0"]) * math.cos(PI / 180.0 * phi0)
    ky0 = k * math.sin(PI / 180.0 * p["theta0"]) * math.sin(PI / 180.0 * phi0)
    d.update(dx=dx, dy=dy, k=k, kx0=kx0, ky0=ky0)
    # intensity.cpp:119-126 (C truncation toward zero of positive values)
    d["lsmaxX"] = int(NA * dx / float(MX) / lam + 1)
    d["lsmaxY"] = int(NA * dy / float(MY) / lam + 1)
    d["lpmaxX"] = int(NA * dx / float(MX) * 2 / lam + 0.0001)
    d["lpmaxY"] = int(NA * dy / float(MY) * 2 / lam + 0.0001)
    d["nsourceX"] = 2 * d["lsmaxX"] + 1
    d["nsourceY"] = 2 * d["lsmaxY"] + 1
    d["noutX"] = 2 * d["lpmaxX"] + 1
    d["noutY"] = 2 * d["lpmaxY"] + 1
    # intensity.cpp:106-109: delta = 1 nm so the Fourier grid equals the mask grid
    d["FDIVX"] = int(dx / 1.0 + 0.000001)
    d["FDIVY"] = int(dy / 1.0 + 0.000001)
    return d


def order_set(p, kind, cutoff=None, K=None):
    """Diffraction-order list (lindex, mindex) and LMAX, MMAX.

    kind 'diamond' : intensity.cpp:138-164 / m3d.cpp:131-157 (rounded diamond)
    kind 'ellipse' : m3d.cpp:101-129 (narrow set)
    kind 'rect'    : a +-K x +-K rectangle (LMAX = MMAX = K); the reference
                     loop order (l outer, m inner) is kept.
    """
    if kind == "rect":
        LMAX = MMAX = int(K) if not isinstance(K, tuple) else int(K[0])
        if isinstance(K, tuple):
            MMAX = int(K[1])
    else:
        cutx = p["NA"] / p["MX"] * cutoff
        cuty = p["NA"] / p["MY"] * cutoff
        LMAX = int(cutx * float(p["NDIVX"]) / p["wavelength"])
        MMAX = int(cuty * float(p["NDIVY"]) / p["wavelength"])
    lidx, midx = [], []
    for i in range(2 * LMAX + 1):
        ii = i - LMAX
        for j in range(2 * MMAX + 1):
            jj = j - MMAX
            if kind == "diamond":
                ok = (abs(ii) / (LMAX + 0.01) + 1.0) * (abs(jj) / (MMAX + 0.01) + 1.0) <= 2.0
            elif kind == "ellipse":
                ok = (ii / (LMAX + 0.01)) ** 2.0 + (jj / (MMAX + 0.01)) ** 2.0 <= 1.0
            else:
                ok = True
            if ok:
                lidx.append(ii)
                midx.append(jj)
    return np.array(lidx, dtype=np.int64), np.array(midx, dtype=np.int64), LMAX, MMAX


def in_na(ls, ms, p, d):
    """Lattice incidence inside the NA (header.h:1398, intensity.cpp:262)."""
    return (ls * p["MX"] / d["dx"]) ** 2 + (ms * p["MY"] / d["dy"]) ** 2 <= (p["NA"] / p["wavelength"]) ** 2


def pupil_points(p, d):
    """Pupil points with at least one overlapping source (m3d.cpp:160-195).

    Returns linput, minput, snum (the C++ xinput is snum >= 8; the Python
    pupil.py:25-40 keeps the count itself)."""
    lin, min_, snum = [], [], []
    for ip in range(d["noutX"]):
        for jp in range(d["noutY"]):
            s = 0
            for is_ in range(d["nsourceX"]):
                for js in range(d["nsourceY"]):
                    if in_na(is_ - d["lsmaxX"], js - d["lsmaxY"], p, d) and in_na(
                        ip - d["lpmaxX"] + is_ - d["lsmaxX"], jp - d["lpmaxY"] + js - d["lsmaxY"], p, d
                    ):
                        s += 1
            if s > 0:
                lin.append(ip - d["lpmaxX"])
                min_.append(jp - d["lpmaxY"])
                snum.append(s)
    return np.array(lin, dtype=np.int64), np.array(min_, dtype=np.int64), np.array(snum, dtype=np.int64)


def ndivs_rule(p, d, sigmadiv=0.5):
    """Source divisions per axis (intensity.cpp:90-91)."""
    return max(1, int(180.0 / PI * p["wavelength"] / d["dx"] / sigmadiv))


def source_lattice(p, d, ndivs, stype, sigma1, sigma2, openangle):
    """Source points per division (intensity.cpp:467-502).

    stype 0 conventional, 1 annular, 2 X dipole, 3 Y dipole.  stype may also
    be a tuple of types: a point is kept when any of them accepts it (the
    union used for dipole+annular sources; each member keeps its own sigmas
    when given as (type, sigma1, sigma2, openangle) tuples).
    Returns {(nsx, nsy): (l0s array, m0s array)}."""
    k, dx, dy = d["k"], d["dx"], d["dy"]
    MX, MY, NA = p["MX"], p["MY"], p["NA"]
    members = []
    if isinstance(stype, (tuple, list)):
        for m in stype:
            members.append(m if isinstance(m, tuple) else (m, sigma1, sigma2, openangle))
    else:
        members.append((stype, sigma1, sigma2, openangle))
    s1max = max(m[1] for m in members)
    dkx, dky = 2.0 * PI / dx, 2.0 * PI / dy
    l0max = int(k * NA / MX * s1max / dkx + 1)
    m0max = int(k * NA / MY * s1max / dky + 1)
    out = {}
    for nsx in range(ndivs):
        for nsy in range(ndivs):
            ls_, ms_ = [], []
            for l in range(-l0max, l0max + 1):
                for m in range(-m0max, m0max + 1):
                    skx = l * dkx + 2.0 * PI / dx * nsx / float(ndivs)
                    sky = m * dky + 2.0 * PI / dy * nsy / float(ndivs)
                    skxo, skyo = skx * MX, sky * MY
                    r2 = skxo * skxo + skyo * skyo
                    r = math.sqrt(r2)
                    ok = False
                    for (t, s1, s2, oa) in members:
                        if t == 0:
                            ok |= r2 <= (k * NA * s1) ** 2
                        elif t == 1:
                            ok |= (r <= k * NA * s1) and (r >= k * NA * s2)
                        elif t == 2:
                            ok |= (r <= k * NA * s1) and (r >= k * NA * s2) and (
                                abs(skyo) <= abs(skxo) * math.tan(PI * oa / 180.0 / 2.0))
                        elif t == 3:
                            ok |= (r <= k * NA * s1) and (r >= k * NA * s2) and (
                                abs(skxo) <= abs(skyo) * math.tan(PI * oa / 180.0 / 2.0))
                    if ok:
                        ls_.append(l)
                        ms_.append(m)
            out[(nsx, nsy)] = (np.array(ls_, dtype=np.int64), np.array(ms_, dtype=np.int64))
    return out


def mesh_source(p, d, stype, sigma1, sigma2, openangle, mesh_deg=0.2):
    """0.2-degree source mesh of linear.py:46-97 / eigenfunction.py:6-56.

    Returns dsx, dsy (wave-number offsets of each source point)."""
    k, MX, MY, NA = d["k"], p["MX"], p["MY"], p["NA"]
    kmesh = k * mesh_deg * (PI / 180.0)
    l0max = int(k * NA / MX * sigma1 / kmesh) + 1
    m0max = int(k * NA / MY * sigma1 / kmesh) + 1
    dsx, dsy = [], []
    for l in range(-l0max, l0max + 1):
        for m in range(-m0max, m0max + 1):
            skx, sky = l * kmesh, m * kmesh
            skxo, skyo = skx * MX, sky * MY
            r = np.sqrt(skxo ** 2 + skyo ** 2)
            if stype == 0:
                ok = (skxo ** 2 + skyo ** 2) <= (k * NA * sigma1) ** 2
            elif stype == 1:
                ok = k * NA * sigma2 <= r <= k * NA * sigma1
            elif stype == 2:
                ok = (k * NA * sigma2 <= r <= k * NA * sigma1) and abs(skyo) <= abs(skxo) * np.tan(
                    PI * openangle / 180.0 / 2.0)
            else:
                ok = (k * NA * sigma2 <= r <= k * NA * sigma1) and abs(skxo) <= abs(skyo) * np.tan(
                    PI * openangle / 180.0 / 2.0)
            if ok:
                dsx.append(skx)
                dsy.append(sky)
    return np.array(dsx), np.array(dsy)

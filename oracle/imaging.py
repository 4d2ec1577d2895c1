"""Images from M3D parameters, restated from the reference Python drivers.

thin_linear_images : cnn/intensity/linear/linear.py:10-188 (int_abbe, int_linear)
reflect_amplitude  : cnn/intensity/linear/reflectivity.py:4-17
tcc_kernels        : cnn/intensity/stccsocs/eigenfunction.py:5-179
socs_image         : cnn/intensity/stccsocs/socs.py:12-154
The 0.2-degree source mesh is geometry.mesh_source (linear.py:46-97).
"""
import numpy as np

from .abbe import blank_amplitudes, rotate
from .em import pattern_xy
from .geometry import PI, mesh_source, pupil_points


def reflect_amplitude(p, d):
    """ampvc, ampab, phase0 of the blank masks at the central incidence (reflectivity.py:4-17)."""
    vc, ab = blank_amplitudes(p, d)
    ampvc = vc[d["lsmaxX"], d["lsmaxY"]]
    ampab = ab[d["lsmaxX"], d["lsmaxY"]]
    return ampvc, ampab, ampvc / abs(ampvc)


def pattern_spectrum(mask2d, p, d, ampvc, ampab):
    """fmask(ip, jp) = fft2(pattern)[l, m] / (NDIVX NDIVY) on the nout window (linear.py:120-128),
    pattern = mask (ampvc - ampab) + ampab, y flipped, [x][y] (linear.py:209-216)."""
    pat = pattern_xy(mask2d, p["NDIVX"], p["NDIVY"], p["NDIVX"], p["NDIVY"], ampab, ampvc)
    F = np.fft.fft2(pat) / (p["NDIVX"] * p["NDIVY"])
    lx = np.arange(-d["lpmaxX"], d["lpmaxX"] + 1)
    my = np.arange(-d["lpmaxY"], d["lpmaxY"] + 1)
    return F[np.ix_(lx % p["NDIVX"], my % p["NDIVY"])]


def thin_amplitudes(fmask, p, d, phase0):
    """fampxx = fmask phaseX phaseY / phase0 (linear.py:130-143)."""
    k, z0 = d["k"], p["z0"]
    kxp = 2.0 * np.pi * np.arange(-d["lpmaxX"], d["lpmaxX"] + 1) / d["dx"]
    kyp = 2.0 * np.pi * np.arange(-d["lpmaxY"], d["lpmaxY"] + 1) / d["dy"]
    phx = np.exp(-1j * (d["kx0"] * kxp + kxp * kxp / 2.0) / k * z0)
    phy = np.exp(-1j * (d["ky0"] * kyp + kyp * kyp / 2.0) / k * z0)
    return fmask * phx[:, None] * phy[None, :] / phase0


def _field(p, d, A, lp, mp, ls, ms, co):
    """high_na_electro_field(0, 0, A, 0, lp, mp, ls, ms) of electro_field.py:162-189."""
    k = d["k"]
    kxn = 2 * PI / d["dx"] * ls + 2 * PI * lp / d["dx"]
    kyn = 2 * PI / d["dy"] * ms + 2 * PI * mp / d["dy"]
    p2 = p["MX"] ** 2 * kxn ** 2 + p["MY"] ** 2 * kyn ** 2
    if (p["NA"] * k * co) ** 2 <= p2 <= (p["NA"] * k) ** 2:
        R = rotate(k, p["MX"], p["MY"], kxn, kyn, d["kx0"], d["ky0"])
        return 1j * k * R[:, 0] * A
    return np.zeros(3, dtype=np.complex128)


def abbe_sum(p, d, E, dsx, dsy, lin, min_, nx, ny, z):
    """int_abbe of linear.py:10-43 (E: [S, ncut, 3])."""
    k = d["k"]
    img = np.zeros((nx, ny))
    for s in range(len(dsx)):
        F = np.zeros((3, nx, ny), dtype=np.complex128)
        for n in range(len(lin)):
            kxn = dsx[s] + 2.0 * np.pi * lin[n] / d["dx"]
            kyn = dsy[s] + 2.0 * np.pi * min_[n] / d["dy"]
            p2 = p["MX"] ** 2 * kxn ** 2 + p["MY"] ** 2 * kyn ** 2
            if p2 <= (p["NA"] * k) ** 2:
                ph = np.exp(1j * ((kxn + d["kx0"]) ** 2 + (kyn + d["ky0"]) ** 2) / (2 * k) * p["z0"]
                            + 1j * p2 / (2 * k) * z)
                F[:, lin[n] % nx, min_[n] % ny] = E[s, n] * ph
        f = np.fft.ifft2(F, axes=(1, 2)) * nx * ny
        img += (np.abs(f) ** 2).sum(axis=0) / len(dsx)
    return img


def thin_linear_images(p, d, fampxx, a0, ax, ay, dsx, dsy, nx, ny, z=0.0, co=0.0):
    """(intlinear, intft) of int_linear, linear.py:45-188.  a0 (ncut); ax, ay (ncut,
    zero where the pupil point has fewer than 8 sources)."""
    k = d["k"]
    lin, min_, _ = pupil_points(p, d)
    lpx, lpy = d["lpmaxX"], d["lpmaxY"]
    S, ncut = len(dsx), len(lin)
    Ef = np.zeros((S, ncut, 3), dtype=np.complex128)
    El = np.zeros_like(Ef)
    for s in range(S):
        kx = d["kx0"] + dsx[s]
        ls = dsx[s] / (2.0 * np.pi / d["dx"])
        ms = dsy[s] / (2.0 * np.pi / d["dy"])
        norm = np.sqrt(k * k - kx * kx)
        for i in range(ncut):
            lp, mp = lin[i], min_[i]
            f = fampxx[lp + lpx, mp + lpy]
            Afx = f / norm
            Ax = (f + a0[i] + ax[i] * (ls + lp / 2.0) + ay[i] * (ms + mp / 2.0)) / norm
            Ef[s, i] = _field(p, d, Afx, lp, mp, ls, ms, co)
            El[s, i] = _field(p, d, Ax, lp, mp, ls, ms, co)
    intft = abbe_sum(p, d, Ef, dsx, dsy, lin, min_, nx, ny, z)
    intlinear = abbe_sum(p, d, El, dsx, dsy, lin, min_, nx, ny, z)
    return intlinear, intft


def tcc_kernels(p, d, dsx, dsy, z=0.0):
    """(alpha, phi) for the three TCCs (S0, SX, SY) of eigenfunction.py:5-179;
    phi[m, n] = component n of the m-th kernel, kernels sorted by |alpha|
    descending (stable, as the reference bubble sort)."""
    k = d["k"]
    lin, min_, _ = pupil_points(p, d)
    S, ncut = len(dsx), len(lin)
    sx0, sy0 = d["kx0"], d["ky0"]
    pmax = (k * p["NA"]) ** 2
    Pc = np.zeros((3, ncut, S), dtype=np.complex128)
    for i in range(ncut):
        kx = 2 * PI / d["dx"] * lin[i]
        ky = 2 * PI / d["dy"] * min_[i]
        for s in range(S):
            ksx, ksy = kx + dsx[s], ky + dsy[s]
            p2 = p["MX"] ** 2 * ksx ** 2 + p["MY"] ** 2 * ksy ** 2
            if p2 > pmax:
                # outside the pupil the reference multiplies by int(p2 <= pmax) = 0;
                # rotate() is not evaluated there (at NA 0.55 it would give NaN)
                continue
            R = rotate(k, p["MX"], p["MY"], ksx, ksy, sx0, sy0)
            ph = np.exp(1j * ((ksx + sx0) ** 2 + (ksy + sy0) ** 2) / 2.0 / k * p["z0"]
                        + 1j * p2 / 2.0 / k * z)
            Pc[:, i, s] = ph * R[:, 0]
    w = k * k / (k * k - (sx0 + dsx) ** 2)
    out = []
    for wt in (w, dsx * w, dsy * w):
        T = sum(Pc[c] @ (np.conj(Pc[c]).T * wt[:, None]) for c in range(3)) / S
        al, ph = np.linalg.eigh(T)
        order = np.argsort(-np.abs(al), kind="stable")
        out += [al[order], ph[:, order].T.copy()]
    return out  # alpha0, phi0, alphax, phix, alphay, phiy


def socs_image(p, d, fampxx, a0, ax, ay, kern, nx, ny, nsocs=100, nsocsxy=20, nout=64):
    """intsocs of socs.py:12-154 (a0, ax, ay as for thin_linear_images)."""
    al0, ph0, alx, phx, aly, phy = kern
    lin, min_, _ = pupil_points(p, d)
    lpx, lpy = d["lpmaxX"], d["lpmaxY"]
    ncut = len(lin)
    f = fampxx[lin + lpx, min_ + lpy]
    A0 = f + a0
    Ax_ = ax * d["dx"] / 2 / np.pi
    Ay_ = ay * d["dy"] / 2 / np.pi
    Axy = ax * lin / 2 + ay * min_ / 2
    px, py = lin % nout, min_ % nout

    def field(vals):
        F = np.zeros((nout, nout), dtype=np.complex128)
        F[px, py] = vals
        return np.fft.ifft2(F) * nout * nout

    small = np.zeros((nout, nout))
    for m in range(min(nsocs, ncut)):
        f0 = field(A0 * ph0[m])
        small += al0[m] * np.abs(f0) ** 2
        if m < nsocsxy:
            f0x, f0y = field(A0 * phx[m]), field(A0 * phy[m])
            fx, fy, fxy = field(Ax_ * phx[m]), field(Ay_ * phy[m]), field(Axy * ph0[m])
            small += (2 * alx[m] * (f0x.real * fx.real + f0x.imag * fx.imag)
                      + 2 * aly[m] * (f0y.real * fy.real + f0y.imag * fy.imag)
                      + 2 * al0[m] * (f0.real * fxy.real + f0.imag * fxy.imag))
    Fs = np.fft.fft2(small) / nout / nout
    big = np.zeros((nx, ny), dtype=np.complex128)
    h = nout // 2
    for i in range(-h, h):
        for j in range(-h, h):
            big[i, j] = Fs[i, j]
    return np.real(np.fft.ifft2(big) * nx * ny)

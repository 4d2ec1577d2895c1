"""Rigorous EM solve (ampS) restated with NumPy/LAPACK in FP64.

Follows cnn/include/header.h step by step:
  exponential/fourier   header.h:1621-1648
  maskamp               header.h:1147-1206
  multilayerS           header.h:5-363
  absorberS0            header.h:754-1145   (eigenpairs by LAPACK zgeev via numpy)
  ampS                  header.h:1208-1466
matinv (header.h:1468-1557) is an LU solve here; matproduct (header.h:1559-1619)
is the BLAS product.  Only NABS = 1 (one absorber layer) is restated, which is
what every reference driver uses (intensity.cpp:85, m3d.cpp:97).
"""
import numpy as np

from .geometry import PI, in_na
from .problems import layer_list


def cexp_table(fdiv):
    """exponential(): cexp[i] = exp(-2 pi i i/FDIV), i = 0..FDIV (header.h:1637-1648)."""
    i = np.arange(fdiv + 1, dtype=np.float64)
    return np.exp(-2.0 * PI * 1j * i / float(fdiv))


def _dft_matrix(fdiv, orders):
    """fourier() as a matrix: F[i, o] = cexp[j(i, l_o)] with the reference's
    index rule for negative l (header.h:1621-1635); divided by FDIV."""
    cexp = cexp_table(fdiv)
    i = np.arange(fdiv, dtype=np.int64)[:, None]
    l = np.asarray(orders, dtype=np.int64)[None, :]
    j = np.where(l >= 0, (i * l) % fdiv, fdiv - ((-i * l) % fdiv))
    return cexp[j] / float(fdiv)


def pattern_xy(mask2d, ndivx, ndivy, fdivx, fdivy, val0, val1):
    """pattern(i, j): x = i, y = j with the y flip of header.h:1157-1168."""
    meshx, meshy = fdivx // ndivx, fdivy // ndivy
    m = np.asarray(mask2d).reshape(ndivy, ndivx)  # [row jj][ii]
    jj = ndivy - 1 - np.arange(fdivy) // meshy
    ii = np.arange(fdivx) // meshx
    bits = m[jj[None, :], ii[:, None]]  # [i, j]
    return np.where(bits == 0, val0, val1).astype(np.complex128)


def maskamp(mask2d, ndivx, ndivy, fdivx, fdivy, LMAX, MMAX, ampta, ampvc):
    """Fourier coefficients famp(l + 2LMAX, m + 2MMAX), |l| <= 2LMAX, |m| <= 2MMAX
    (header.h:1147-1206): y transform first, then x."""
    pat = pattern_xy(mask2d, ndivx, ndivy, fdivx, fdivy, ampta, ampvc)
    Fy = _dft_matrix(fdivy, np.arange(-2 * MMAX, 2 * MMAX + 1))
    Fx = _dft_matrix(fdivx, np.arange(-2 * LMAX, 2 * LMAX + 1))
    ftmp = pat @ Fy  # [i, m]
    return Fx.T @ ftmp  # [l, m]


def layer_tables(mask2d, p, d, LMAX, MMAX, n=0):
    """eps, sigma(=1/eps), eta, zeta tables of one absorber layer (header.h:1236-1268)."""
    ea = p["eabs"][n]
    args = (mask2d, p["NDIVX"], p["NDIVY"], d["FDIVX"], d["FDIVY"], LMAX, MMAX)
    eps = maskamp(*args, ea, 1.0)
    sigma = maskamp(*args, 1.0 / ea, 1.0)
    leps = maskamp(*args, np.log(ea), 0.0)
    zl = 1j * 2.0 * PI * np.arange(-2 * LMAX, 2 * LMAX + 1) / d["dx"]
    zm = 1j * 2.0 * PI * np.arange(-2 * MMAX, 2 * MMAX + 1) / d["dy"]
    eta = zl[:, None] * leps
    zeta = zm[None, :] * leps
    return eps, sigma, eta, zeta


def multilayer(polar, k, kxplus, kyplus, kxy2, stack):
    """Per-mode diagonal T-matrix cascade (header.h:5-363).

    Returns URUU, URUB (diagonals) at the top of the Ru cap."""
    kt = kxplus if polar == "X" else kyplus
    L = layer_list(stack)
    al = [np.sqrt(k * k * e - kxy2 + 0j) for (e, _) in L]
    f = [k - kt * kt / k / e for (e, _) in L]
    U = np.ones_like(kxy2, dtype=np.complex128)
    B = np.zeros_like(U)
    for j in range(1, len(L)):
        cj = f[j - 1] / f[j]           # (k - kt^2/k/e_below)/(k - kt^2/k/e_cur)
        r = al[j - 1] / al[j]           # al_below / al_cur
        g = np.exp(1j * al[j] * L[j][1])
        ul = 0.5 * (cj + r) / g
        ur = 0.5 * (cj - r) / g
        bl = 0.5 * (cj - r) * g
        br = 0.5 * (cj + r) * g
        U, B = ul * U + ur * B, bl * U + br * B
    return U, B


def toeplitz(tab, lidx, midx, LMAX, MMAX):
    """T(i, j) = tab(l_i - l_j + 2LMAX, m_i - m_j + 2MMAX) (header.h:806-820)."""
    return tab[(lidx[:, None] - lidx[None, :]) + 2 * LMAX, (midx[:, None] - midx[None, :]) + 2 * MMAX]


def amp_s(polar, mask2d, p, d, lidx, midx, LMAX, MMAX, sx0, sy0, eig=None, return_fg=False):
    """ampS (header.h:1208-1466) for one polarisation and one incidence offset.

    Returns Ax[nsourceX, nsourceY, N] (zero where the lattice incidence is
    outside the NA).  ``eig`` may replace numpy.linalg.eig (used by the
    eigenvector-invariance pin)."""
    k = d["k"]
    stack = p["stack"]
    N = len(lidx)
    eps, sigma, eta, zeta = layer_tables(mask2d, p, d, LMAX, MMAX)
    kxplus = sx0 + 2 * PI * lidx / d["dx"]
    kyplus = sy0 + 2 * PI * midx / d["dy"]
    kxy2 = kxplus ** 2 + kyplus ** 2
    klm = np.sqrt(k * k - kxy2).astype(np.complex128)  # header.h:1283 (real sqrt)
    URUU, URUB = multilayer(polar, k, kxplus, kyplus, kxy2, stack)
    eru = stack["n_ru"] ** 2
    alru = np.sqrt(k * k * eru - kxy2 + 0j)
    kt = kxplus if polar == "X" else kyplus
    Bru = 1j * k - 1j / k / eru * kt ** 2                 # header.h:1294-1297
    # absorberS0 (header.h:754-1145)
    D = k * k * toeplitz(eps, lidx, midx, LMAX, MMAX)
    if polar == "X":
        D = D - 1j * toeplitz(eta, lidx, midx, LMAX, MMAX) * kxplus[None, :]
    else:
        D = D - 1j * toeplitz(zeta, lidx, midx, LMAX, MMAX) * kyplus[None, :]
    D[np.arange(N), np.arange(N)] -= kxy2
    w, V = (eig or np.linalg.eig)(D)                       # header.h:834-899
    al1 = np.sqrt(w)                                      # header.h:918-921
    Cjp = np.linalg.inv(V)                                # header.h:923-975
    Sig = toeplitz(sigma, lidx, midx, LMAX, MMAX) * kt[None, :]
    B1 = 1j * (k * V - (kt / k)[:, None] * (Sig @ V))     # header.h:979-1052
    Cj = np.linalg.solve(B1, np.diag(Bru))                # header.h:1053-1055
    dab = p["dabs"][0]
    gam = np.exp(1j * al1 * dab)
    A = 0.5 * Cj
    Bm = 0.5 * (alru[None, :] / al1[:, None]) * Cjp
    T1UL = (A + Bm) / gam[:, None]
    T1UR = (A - Bm) / gam[:, None]
    T1BL = (A - Bm) * gam[:, None]
    T1BR = (A + Bm) * gam[:, None]
    U1U = T1UL * URUU[None, :] + T1UR * URUB[None, :]     # header.h:1085-1144
    U1B = T1BL * URUU[None, :] + T1BR * URUB[None, :]
    # top solve (header.h:1355-1391)
    T0L = (klm[:, None] + al1[None, :]) * V
    T0R = (klm[:, None] - al1[None, :]) * V
    U0 = T0L @ U1U + T0R @ U1B
    U1 = np.linalg.solve(U0.T, (U1U - U1B).T).T           # (U1U - U1B) U0^-1
    FG = (al1[None, :] * V / klm[:, None]) @ U1
    # per-incidence amplitudes (header.h:1394-1465)
    Ax = np.zeros((d["nsourceX"], d["nsourceY"], N), dtype=np.complex128)
    pos = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(lidx, midx))}
    for ls in range(-d["lsmaxX"], d["lsmaxX"] + 1):
        for ms in range(-d["lsmaxY"], d["lsmaxY"] + 1):
            if not in_na(ls, ms, p, d):
                continue
            kx = sx0 + ls * 2.0 * PI / d["dx"]
            ky = sy0 + ms * 2.0 * PI / d["dy"]
            kz = np.sqrt(k * k - kx * kx - ky * ky)
            col = np.zeros(N, dtype=np.complex128)
            i = pos.get((ls, ms))
            if i is not None:
                col = -FG[:, i] * (2.0 * kz)
                col[i] += 1.0
            Ax[ls + d["lsmaxX"], ms + d["lsmaxY"]] = col
    if return_fg:
        return Ax, FG
    return Ax

"""The TCCEE image of elitho/stcc.py, restated (TEST INFRASTRUCTURE ONLY).

Follows cnn/intensity/stccsocs/elitho/stcc.py:12-431 with the elitho
constants (const.py) passed in as a dict ``c`` and the reference's quirks
kept:

* the propagation phase divides by k z0 (stcc.py:121-127);
* the pupil wave vector of the field terms is kx0 + 2 pi l / dx / 2
  (stcc.py:196-197);
* the Y branch reads the M3D arrays always (``type == 1`` compares the
  builtin, stcc.py:239);
* TCCEE is XDIV x XDIV in both axes (stcc.py:300).

``tcc_matrices_literal`` / ``calc_tccee_literal`` are the reference's loops
(small problems); ``tcc_matrices`` / ``calc_tccee`` the same sums vectorised,
pinned to the loops in tests/test_stcc.py.
"""
import numpy as np


def constants(ndiv=2048, mx=4, my=4, na=0.33, wavelength=13.5, theta0=-6.0, azimuth=0.0, mesh=0.5,
              sigma1=0.9, sigma2=0.55, openangle=90.0, illumination="annular", z0=102.0):
    """The derived constants of elitho/const.py:8-130 (non-high-NA branch by
    default)."""
    c = dict(NDIVX=ndiv, NDIVY=ndiv, MX=mx, MY=my, NA=na, wavelength=wavelength, mesh=mesh, sigma1=sigma1,
             sigma2=sigma2, openangle=openangle, illumination=illumination, z0=z0)
    c["dx"], c["dy"] = ndiv, ndiv
    c["XDIV"] = ndiv // mx
    c["k"] = 2.0 * np.pi / wavelength
    phi0 = 90.0 - azimuth
    c["kx0"] = c["k"] * np.sin(np.deg2rad(theta0)) * np.cos(np.deg2rad(phi0))
    c["ky0"] = c["k"] * np.sin(np.deg2rad(theta0)) * np.sin(np.deg2rad(phi0))
    c["lsmaxX"] = int(na * c["dx"] / mx / wavelength + 1)
    c["lsmaxY"] = int(na * c["dy"] / my / wavelength + 1)
    c["lpmaxX"] = int(na * c["dx"] / mx * 2 / wavelength + 0.0001)
    c["lpmaxY"] = int(na * c["dy"] / my * 2 / wavelength + 0.0001)
    c["nsourceX"], c["nsourceY"] = 2 * c["lsmaxX"] + 1, 2 * c["lsmaxY"] + 1
    c["noutX"], c["noutY"] = 2 * c["lpmaxX"] + 1, 2 * c["lpmaxY"] + 1
    return c


def pupil_coordinates(c):
    """pupil.py:5-44 (offsets 0): pupil points with at least one lattice
    source inside the NA whose shifted point is inside too."""
    lin, mn, cnt = [], [], []
    na2 = (c["NA"] / c["wavelength"]) ** 2
    for ip in range(c["noutX"]):
        for jp in range(c["noutY"]):
            n = 0
            for i in range(c["nsourceX"]):
                for j in range(c["nsourceY"]):
                    sx, sy = i - c["lsmaxX"], j - c["lsmaxY"]
                    if ((sx * c["MX"] / c["dx"]) ** 2 + (sy * c["MY"] / c["dy"]) ** 2 <= na2 and
                            ((ip - c["lpmaxX"] + sx) * c["MX"] / c["dx"]) ** 2
                            + ((jp - c["lpmaxY"] + sy) * c["MY"] / c["dy"]) ** 2 <= na2):
                        n += 1
            if n > 0:
                lin.append(ip - c["lpmaxX"])
                mn.append(jp - c["lpmaxY"])
                cnt.append(n)
    return np.array(lin), np.array(mn), np.array(cnt)


def uniform_k_source(c):
    """source.py:157-199: (dkx, dky) of the mesh points inside the source."""
    k = c["k"]
    kmesh = k * c["mesh"] * (np.pi / 180.0)
    skangx = k * c["NA"] / c["MX"] * c["sigma1"]
    skangy = k * c["NA"] / c["MY"] * c["sigma1"]
    l0max, m0max = int(skangx / kmesh + 1), int(skangy / kmesh + 1)
    L, M = np.meshgrid(np.arange(-l0max, l0max + 1), np.arange(-m0max, m0max + 1), indexing="ij")
    skx, sky = L * kmesh, M * kmesh
    skxo, skyo = skx * c["MX"], sky * c["MY"]
    r = np.sqrt(skxo ** 2 + skyo ** 2)
    if c["illumination"] == "circular":
        sel = skxo ** 2 + skyo ** 2 <= (k * c["NA"] * c["sigma1"]) ** 2
    elif c["illumination"] == "annular":
        sel = (k * c["NA"] * c["sigma2"] <= r) & (r <= k * c["NA"] * c["sigma1"])
    else:
        raise ValueError(c["illumination"])
    return skx[sel].tolist(), sky[sel].tolist()


def _phase(c, ksx, ksy):
    return np.exp(1j * ((ksx + c["kx0"]) ** 2 + (ksy + c["ky0"]) ** 2) / (2 * c["k"]) * c["z0"])


def tcc_matrices_literal(c, lin, mn, src):
    """stcc.py:12-103, loop for loop."""
    n = len(lin)
    T = {key: np.zeros((n, n), dtype=np.complex128) for key in ("XS0", "XSX", "XSY", "YS0", "YSX", "YSY")}
    pmax = (c["k"] * c["NA"]) ** 2
    sdiv = len(src[0])
    for i in range(n):
        kx = 2 * np.pi / c["dx"] * lin[i]
        ky = 2 * np.pi / c["dy"] * mn[i]
        for j in range(i + 1):
            kxp = 2 * np.pi / c["dx"] * lin[j]
            kyp = 2 * np.pi / c["dy"] * mn[j]
            s = dict.fromkeys(T, 0j)
            for sx, sy in zip(*src):
                ksx, ksy, ksxp, ksyp = kx + sx, ky + sy, kxp + sx, kyp + sy
                if (c["MX"] ** 2 * ksx ** 2 + c["MY"] ** 2 * ksy ** 2) <= pmax and (
                        c["MX"] ** 2 * ksxp ** 2 + c["MY"] ** 2 * ksyp ** 2) <= pmax:
                    pp = _phase(c, ksx, ksy) * np.conj(_phase(c, ksxp, ksyp))
                    dxn = c["k"] ** 2 - (c["kx0"] + sx) ** 2
                    dyn = c["k"] ** 2 - (c["ky0"] + sy) ** 2
                    s["XS0"] += pp / dxn
                    s["YS0"] += pp / dyn
                    s["XSX"] += sx * pp / dxn
                    s["YSX"] += sx * pp / dyn
                    s["XSY"] += sy * pp / dxn
                    s["YSY"] += sy * pp / dyn
            for key in T:
                T[key][i, j] = s[key] / sdiv
                T[key][j, i] = np.conj(T[key][i, j])
    return T


def tcc_matrices(c, lin, mn, src):
    """The sums of tcc_matrices_literal as P diag(w) P^H / SDIV."""
    sx, sy = np.asarray(src[0]), np.asarray(src[1])
    kx = 2 * np.pi / c["dx"] * np.asarray(lin)[:, None] + sx[None, :]
    ky = 2 * np.pi / c["dy"] * np.asarray(mn)[:, None] + sy[None, :]
    inside = (c["MX"] ** 2 * kx ** 2 + c["MY"] ** 2 * ky ** 2) <= (c["k"] * c["NA"]) ** 2
    P = np.where(inside, _phase(c, kx, ky), 0)
    dxn = c["k"] ** 2 - (c["kx0"] + sx) ** 2
    dyn = c["k"] ** 2 - (c["ky0"] + sy) ** 2
    w = {"XS0": 1 / dxn, "XSX": sx / dxn, "XSY": sy / dxn, "YS0": 1 / dyn, "YSX": sx / dyn, "YSY": sy / dyn}
    return {key: (P * v[None, :]) @ P.conj().T / len(sx) for key, v in w.items()}


def fampxx(c, fft_mask_window, amp_absorber, amp_vacuum, phasexx):
    """shift_center + propagation + / phasexx (stcc.py:106-128, 415-418) from
    the forward-normalised FFT of the 0/1 mask on the window |l| <= lpmaxX,
    |m| <= lpmaxY (index [l + lpmaxX][m + lpmaxY]): the FFT of hfpattern is
    (ab - vc) times it plus vc at l = m = 0."""
    f = (amp_absorber - amp_vacuum) * fft_mask_window
    f[c["lpmaxX"], c["lpmaxY"]] += amp_vacuum
    ip, jp = np.meshgrid(np.arange(c["noutX"]), np.arange(c["noutY"]), indexing="ij")
    kxp = 2.0 * np.pi * (ip - c["lpmaxX"]) / c["dx"]
    kyp = 2.0 * np.pi * (jp - c["lpmaxY"]) / c["dy"]
    ph = np.exp(-1j * (c["kx0"] * kxp + kxp ** 2 / 2 + c["ky0"] * kyp + kyp ** 2 / 2) / (c["k"] * c["z0"]))
    return f * ph / phasexx


def electric_field(c, polar, lin, mn, famp, a0xx=None, axxx=None, ayxx=None):
    """stcc.py:166-295: twelve per-pupil-point field terms (dict)."""
    k, zi = c["k"], 1j
    lin, mn = np.asarray(lin), np.asarray(mn)
    kxp = c["kx0"] + 2 * np.pi * lin / c["dx"] / 2.0
    kyp = c["ky0"] + 2 * np.pi * mn / c["dy"] / 2.0
    kzp = -np.sqrt(k * k - kxp * kxp - kyp * kyp + 0j)
    ip, jp = lin + c["lpmaxX"], mn + c["lpmaxY"]
    lp, mp = ip - c["lpmaxX"], jp - c["lpmaxY"]
    gx, gy = c["dx"] / (2 * np.pi), c["dy"] / (2 * np.pi)
    ef = {}
    if polar == "X":
        if a0xx is None:
            A, dA, eA = famp[ip, jp], 0.0, 0.0
        else:
            A, dA, eA = famp[ip, jp] + a0xx[ip, jp], axxx[ip, jp] * gx, ayxx[ip, jp] * gy
        ef["Exs0"] = zi * k * A - zi / k * kxp * kxp * A
        ef["Exsx"] = -2 * zi / k * kxp * A + zi * k * dA - zi / k * kxp * kxp * dA
        ef["Exsy"] = zi * k * eA - zi / k * kxp * kxp * eA
        ef["Eys0"] = -zi / k * kxp * kyp * A
        ef["Eysx"] = -zi / k * kyp * A - zi / k * kxp * kyp * dA
        ef["Eysy"] = -zi / k * kxp * A - zi / k * kxp * kyp * eA
        ef["Ezs0"] = -zi / k * kxp * kzp * A
        ef["Ezsx"] = -zi / k * kzp * A - zi / k * kxp * kzp * dA
        ef["Ezsy"] = -zi / k * kxp * kzp * eA
    else:
        A, dA, eA = famp[ip, jp] + a0xx[ip, jp], axxx[ip, jp] * gx, ayxx[ip, jp] * gy
        ef["Exs0"] = -zi / k * kxp * kyp * A
        ef["Exsx"] = -zi / k * kyp * A - zi / k * kxp * kyp * dA
        ef["Exsy"] = -zi / k * kxp * A - zi / k * kxp * kyp * eA
        ef["Eys0"] = zi / k * A - zi / k * kyp * kyp * A
        ef["Eysx"] = zi / k * dA - zi / k * kyp * kyp * dA
        ef["Eysy"] = -2 * zi / k * kyp * A + zi / k * eA - zi / k * kyp * kyp * eA
        ef["Ezs0"] = -zi / k * kyp * kzp * A
        ef["Ezsx"] = -zi / k * kyp * kzp * dA
        ef["Ezsy"] = -zi / k * kzp * A - zi / k * kyp * kzp * eA
    for comp in "xyz":
        ef[f"E{comp}sxy"] = ef[f"E{comp}sx"] / gx * lp / 2 + ef[f"E{comp}sy"] / gy * mp / 2
    return ef


def _pair_terms(ef, a, b):
    """sum over the components of E_a(i) conj(E_b(j)), [i][j]."""
    return sum(np.outer(ef[f"E{c}{a}"], np.conj(ef[f"E{c}{b}"])) for c in "xyz")


def calc_tccee(c, polar, lin, mn, ef, T):
    """stcc.py:298-391 vectorised: the summand of every (i, j) binned at
    ((l_i - l_j) mod XDIV, (m_i - m_j) mod XDIV)."""
    p = "X" if polar == "X" else "Y"
    S = (T[f"{p}S0"] * _pair_terms(ef, "s0", "s0") + 2.0 * T[f"{p}SX"] * _pair_terms(ef, "s0", "sx")
         + 2.0 * T[f"{p}SY"] * _pair_terms(ef, "s0", "sy") + 2.0 * T[f"{p}S0"] * _pair_terms(ef, "s0", "sxy"))
    X = c["XDIV"]
    lin, mn = np.asarray(lin), np.asarray(mn)
    px = (lin[:, None] - lin[None, :] + X) % X
    py = (mn[:, None] - mn[None, :] + X) % X
    out = np.zeros((X, X), dtype=np.complex128)
    np.add.at(out, (px.ravel(), py.ravel()), S.ravel())
    return out


def calc_tccee_literal(c, polar, lin, mn, ef, T):
    """stcc.py:298-391, loop for loop."""
    p = "X" if polar == "X" else "Y"
    X = c["XDIV"]
    out = np.zeros((X, X), dtype=np.complex128)
    n = len(lin)
    dot = lambda a, b, i, j: sum(ef[f"E{q}{a}"][i] * np.conj(ef[f"E{q}{b}"][j]) for q in "xyz")
    for i in range(n):
        for j in range(n):
            px = (lin[i] - lin[j] + X) % X
            py = (mn[i] - mn[j] + X) % X
            out[px, py] += T[f"{p}S0"][i, j] * dot("s0", "s0", i, j)
            out[px, py] += 2.0 * T[f"{p}SX"][i, j] * dot("s0", "sx", i, j)
            out[px, py] += 2.0 * T[f"{p}SY"][i, j] * dot("s0", "sy", i, j)
            out[px, py] += 2.0 * T[f"{p}S0"][i, j] * dot("s0", "sxy", i, j)
    return out


def intensity(c, polar, fft_mask_window, amp_absorber, amp_vacuum, phasexx, a0xx=None, axxx=None, ayxx=None,
              literal=False):
    """stcc.py:394-431 from the mask's forward FFT window: the image
    ifft2(TCCEE, norm="forward").real, [x][y] on XDIV x XDIV."""
    lin, mn, _ = pupil_coordinates(c)
    src = uniform_k_source(c)
    famp = fampxx(c, fft_mask_window, amp_absorber, amp_vacuum, phasexx)
    ef = electric_field(c, polar, lin, mn, famp, a0xx, axxx, ayxx)
    T = (tcc_matrices_literal if literal else tcc_matrices)(c, lin, mn, src)
    tccee = (calc_tccee_literal if literal else calc_tccee)(c, polar, lin, mn, ef, T)
    return np.fft.ifft2(tccee, norm="forward").real


def mask_fft_window(c, mask):
    """np.fft.fft2(mask, norm="forward") sampled on the window, index
    [l + lpmaxX][m + lpmaxY] (shift_center, stcc.py:106-114)."""
    F = np.fft.fft2(np.asarray(mask, dtype=np.float64), norm="forward")
    l = (np.arange(c["noutX"]) - c["lpmaxX"]) % mask.shape[0]
    m = (np.arange(c["noutY"]) - c["lpmaxY"]) % mask.shape[1]
    return F[np.ix_(l, m)]

"""Rigorous Abbe image and M3D extraction restated from the reference drivers.

abbe_image : cnn/intensity/abbe/intensity.cpp:219-408 (+ rotate :423-465)
m3d_params : cnn/intensity/m3d/m3d.cpp:211-467
"""
import numpy as np

from . import em
from .geometry import PI, in_na, pupil_points, order_set


def rotate(k, MX, MY, px, py, sx0, sy0):
    """3x2 polarisation rotation (intensity.cpp:423-465)."""
    s0 = np.array([sx0, sy0, -np.sqrt(k * k - sx0 * sx0 - sy0 * sy0)])
    p = np.array([px, py, 0.0])
    pp = np.array([MX * px, MY * py, -np.sqrt(k * k - MX * px * MX * px - MY * py * MY * py)])
    ez = np.array([0.0, 0.0, 1.0])
    if (px * px + py * py) > 0.00001:
        es = np.cross(p, s0)
        es = es / np.linalg.norm(es)
        esp = np.cross(pp, ez)
        esp = esp / np.linalg.norm(esp)
        ps0 = (p + s0) / np.linalg.norm(p + s0)
        em_ = np.cross(es, ps0)
        emp = np.cross(esp, pp) / k
    else:
        es = np.cross(ez, s0)
        es = es / np.linalg.norm(es)
        esp = -es
        em_ = np.cross(es, s0) / k
        emp = np.cross(esp, -ez)
    return np.outer(esp, es[:2]) + np.outer(emp, em_[:2])


def division_fields(p, d, Ax, divoff, l0s, m0s, lin, min_, lidx, midx, z, polar="X"):
    """Pupil spectra of one division: returns list over sources of (px, py, fx, fy, fz)
    entries restating intensity.cpp:248-368 (X: Ax = amp/sqrt(k^2-kx^2) on column 0
    of rotate(); Y: Ay = amp/sqrt(k^2-ky^2) on column 1, intensity.cpp:295-305)."""
    col = 0 if polar == "X" else 1
    k, dx, dy = d["k"], d["dx"], d["dy"]
    MX, MY = p["MX"], p["MY"]
    sx0, sy0, fxo, fyo = divoff
    kx0, ky0 = d["kx0"], d["ky0"]
    pos = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(lidx, midx))}
    out = []
    for l0, m0 in zip(l0s, m0s):
        kin = sx0 + 2.0 * PI / dx * l0 if col == 0 else sy0 + 2.0 * PI / dy * m0
        ls, ms = l0 + d["lsmaxX"], m0 + d["lsmaxY"]
        rows = []
        for li, mi in zip(lin, min_):
            kxn = 2.0 * PI / dx * fxo + 2.0 * PI / dx * l0 + 2.0 * PI * li / dx
            kyn = 2.0 * PI / dy * fyo + 2.0 * PI / dy * m0 + 2.0 * PI * mi / dy
            if not (MX * MX * kxn * kxn + MY * MY * kyn * kyn) <= (p["NA"] * k) ** 2:
                continue
            amp = -1000.0 + 0j
            n = pos.get((int(l0 + li), int(m0 + mi)))
            if n is not None and 0 <= ls < d["nsourceX"] and 0 <= ms < d["nsourceY"]:
                amp = Ax[ls, ms, n]
            A = amp / np.sqrt(k * k - kin * kin)
            R = rotate(k, MX, MY, kxn, kyn, kx0, ky0)
            E = 1j * k * R[:, col] * A
            phase = np.exp(1j * ((kxn + kx0) ** 2 + (kyn + ky0) ** 2) / 2.0 / k * p["z0"]
                           + 1j * (MX * MX * kxn * kxn + MY * MY * kyn * kyn) / 2.0 / k * z)
            rows.append((int(li), int(mi), E * phase))
        out.append(rows)
    return out


def abbe_image(p, d, mask2d, lidx, midx, LMAX, MMAX, divisions, nx, ny, z=0.0, amps=None, polar="X"):
    """Mean over all source points of |Ex|^2+|Ey|^2+|Ez|^2 on an nx x ny grid
    (intensity.cpp:219-408).  divisions: list of ((sx0, sy0, fx, fy), l0s, m0s).
    amps: optional precomputed Ax per division."""
    lin, min_, _ = pupil_points(p, d)
    img = np.zeros((nx, ny))
    total = 0
    for q, (off, l0s, m0s) in enumerate(divisions):
        Ax = amps[q] if amps is not None else em.amp_s(polar, mask2d, p, d, lidx, midx, LMAX, MMAX, off[0], off[1])
        for rows in division_fields(p, d, Ax, off, l0s, m0s, lin, min_, lidx, midx, z, polar):
            F = np.zeros((3, nx, ny), dtype=np.complex128)
            for li, mi, f in rows:
                F[:, li % nx, mi % ny] = f
            fld = np.fft.ifft2(F, axes=(1, 2)) * nx * ny  # unnormalised inverse (cufftExecZ2Z INVERSE)
            img += (np.abs(fld) ** 2).sum(axis=0)
            total += 1
    return img / total


def blank_amplitudes(p, d, narrow=("ellipse", 1.5), masks=None):
    """vcxx, abxx on the narrow order set (m3d.cpp:101-129, 211-258).  In-NA
    lattice incidences that the truncated ellipse misses (small pitches) are
    appended to the set: the reference reads an unset entry there, and a
    uniform mask couples no modes, so the other amplitudes are unchanged.
    masks: optional (vacuum, absorber) 0/1 masks in place of the uniform ones."""
    li, mi, L, M = order_set(p, narrow[0], cutoff=narrow[1])
    have = set(zip(li.tolist(), mi.tolist()))
    extra = [(ls, ms) for ls in range(-d["lsmaxX"], d["lsmaxX"] + 1) for ms in range(-d["lsmaxY"], d["lsmaxY"] + 1)
             if in_na(ls, ms, p, d) and (ls, ms) not in have]
    if extra:
        el, em_ = np.array(extra, dtype=np.int64).T
        li, mi = np.concatenate([li, el]), np.concatenate([mi, em_])
        L, M = max(L, int(np.abs(el).max())), max(M, int(np.abs(em_).max()))
    n = p["NDIVX"] * p["NDIVY"]
    m_vc, m_ab = (np.ones(n, np.int64), np.zeros(n, np.int64)) if masks is None else masks
    Avc = em.amp_s("X", np.asarray(m_vc).reshape(-1), p, d, li, mi, L, M, d["kx0"], d["ky0"])
    Aab = em.amp_s("X", np.asarray(m_ab).reshape(-1), p, d, li, mi, L, M, d["kx0"], d["ky0"])
    vc = np.zeros((d["nsourceX"], d["nsourceY"]), dtype=np.complex128)
    ab = np.zeros_like(vc)
    pos = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(li, mi))}
    for ls in range(-d["lsmaxX"], d["lsmaxX"] + 1):
        for ms in range(-d["lsmaxY"], d["lsmaxY"] + 1):
            if in_na(ls, ms, p, d) and (ls, ms) in pos:
                vc[ls + d["lsmaxX"], ms + d["lsmaxY"]] = Avc[ls + d["lsmaxX"], ms + d["lsmaxY"], pos[(ls, ms)]]
                ab[ls + d["lsmaxX"], ms + d["lsmaxY"]] = Aab[ls + d["lsmaxX"], ms + d["lsmaxY"], pos[(ls, ms)]]
    return vc, ab


def thin_spectrum(mask2d, p, d):
    """fmask(ip, jp) of m3d.cpp:287-327 (indicator, y flipped), window lpmax."""
    pat = em.pattern_xy(mask2d, p["NDIVX"], p["NDIVY"], d["FDIVX"], d["FDIVY"], 0.0, 1.0)
    Fy = em._dft_matrix(d["FDIVY"], np.arange(-d["lpmaxY"], d["lpmaxY"] + 1))
    Fx = em._dft_matrix(d["FDIVX"], np.arange(-d["lpmaxX"], d["lpmaxX"] + 1))
    return Fx.T @ (pat @ Fy)


def m3d_params(p, d, mask2d, lidx, midx, Ax, vcxx, abxx):
    """a0 (ncut), ax, ay (ncut, zero where snum < 8) from m3d.cpp:329-467.
    Ax: wide amplitudes at the central incidence (nsourceX, nsourceY, N)."""
    k = d["k"]
    lsx, lsy, lpx, lpy = d["lsmaxX"], d["lsmaxY"], d["lpmaxX"], d["lpmaxY"]
    nsx, nsy = d["nsourceX"], d["nsourceY"]
    fmask = thin_spectrum(mask2d, p, d)
    lin, min_, snum = pupil_points(p, d)
    pos = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(lidx, midx))}
    kxs, kys = d["kx0"], d["ky0"]
    phasexx = vcxx / np.abs(np.where(vcxx == 0, 1, vcxx))
    a0 = np.zeros(len(lin), dtype=np.complex128)
    ax = np.zeros_like(a0)
    ay = np.zeros_like(a0)
    for n, (li, mi, sn) in enumerate(zip(lin, min_, snum)):
        kxp = 2.0 * PI * li / d["dx"]
        kyp = 2.0 * PI * mi / d["dy"]
        ph = np.exp(-1j * (kxs * kxp + kxp * kxp / 2.0 + kys * kyp + kyp * kyp / 2.0) / k * p["z0"])
        fc = fmask[li + lpx, mi + lpy] * ph * (vcxx[lsx, lsy] - abxx[lsx, lsy])
        if li == 0 and mi == 0:
            fc += abxx[lsx, lsy]
        fc = fc / phasexx[lsx, lsy]
        f1 = fx = fy = 0j
        c1 = cx = cy = cx2 = cy2 = cxy = 0.0
        for is_ in range(nsx):
            for js in range(nsy):
                ls_, ms_ = is_ - lsx, js - lsy
                if not (in_na(ls_, ms_, p, d) and in_na(li + ls_, mi + ms_, p, d)):
                    continue
                nn = pos.get((int(li + ls_), int(mi + ms_)))
                am = Ax[is_, js, nn] if nn is not None else 0j
                f = am / phasexx[is_, js] - fc
                ls = ls_ + li / 2.0
                ms = ms_ + mi / 2.0
                f1 += f
                fx += ls * f
                fy += ms * f
                c1 += 1
                cx += ls
                cy += ms
                cx2 += ls * ls
                cy2 += ms * ms
                cxy += ls * ms
        cd = c1 * cx2 * cy2 + 2.0 * cx * cxy * cy - c1 * cxy * cxy - cx2 * cy * cy - cy2 * cx * cx
        if sn >= 8 and cd == 0.0:
            # deviation from m3d.cpp:451-457, which divides by cd = 0 when the sources lie on one
            # row or column: fit the slope along the direction that varies, the other slope is 0
            vx, vy = c1 * cx2 - cx * cx, c1 * cy2 - cy * cy
            if vx > 0:
                ax[n] = (c1 * fx - cx * f1) / vx
                a0[n] = (f1 - ax[n] * cx) / c1
            elif vy > 0:
                ay[n] = (c1 * fy - cy * f1) / vy
                a0[n] = (f1 - ay[n] * cy) / c1
            else:
                a0[n] = f1 / c1
        elif sn >= 8:
            a0[n] = ((cx2 * cy2 - cxy * cxy) * f1 + (-cx * cy2 + cy * cxy) * fx + (cx * cxy - cy * cx2) * fy) / cd
            ax[n] = ((-cx * cy2 + cy * cxy) * f1 + (c1 * cy2 - cy * cy) * fx + (-c1 * cxy + cx * cy) * fy) / cd
            ay[n] = ((cx * cxy - cy * cx2) * f1 + (-c1 * cxy + cx * cy) * fx + (c1 * cx2 - cx * cx) * fy) / cd
        else:
            a0[n] = f1 / c1
    return a0, ax, ay

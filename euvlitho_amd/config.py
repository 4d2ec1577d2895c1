"""Typed configuration of one simulation: optics, stack, order set, source, image grid.

The derived integers follow the reference drivers (cnn/intensity/abbe/intensity.cpp:44-195,
cnn/intensity/m3d/m3d.cpp:44-195, elitho/const.py:95-111); the material table defaults
to cnn/include/header.h:1217-1230.  Geometry is computed vectorised with NumPy on the
host (integer bookkeeping, not part of the GPU hot path).
"""
from dataclasses import dataclass, field
from functools import cached_property
import math

import numpy as np

PI = math.atan(1.0) * 4.0


@dataclass(frozen=True)
class Stack:
    """Mo/Si multilayer with interdiffusion layers (header.h:1217-1230)."""
    n_mo: complex = 0.9238 + 0.006435j
    n_si: complex = 0.999 + 0.001826j
    n_ru: complex = 0.8863 + 0.01706j
    n_mosi2: complex = 0.9693 + 0.004333j
    n_rusi: complex = 0.9099 + 0.01547j
    n_sio2: complex = 0.978 + 0.01083j
    d_mo: float = 2.052
    d_mosi: float = 1.661
    d_si: float = 2.283
    d_simo: float = 1.045
    d_siru: float = 0.8
    d_ru: float = 2.5
    n_pairs: int = 40

    def layers(self):
        """(eps, thickness) from the substrate (index 0) to the Ru cap (last),
        in the cascade order of header.h:294-362."""
        e = lambda n: complex(n) * complex(n)
        out = [(e(self.n_sio2), 0.0), (e(self.n_mo), self.d_mo)]
        for i in reversed(range(self.n_pairs)):
            if i < self.n_pairs - 1:
                out.append((e(self.n_mo), self.d_mo))
            out.append((e(self.n_mosi2), self.d_mosi))
            out.append((e(self.n_si), self.d_si))
            out.append((e(self.n_mosi2), self.d_simo) if i > 0 else (e(self.n_rusi), self.d_siru))
        out.append((e(self.n_ru), self.d_ru))
        return out


REFERENCE_STACK = Stack()
# 40 x Mo/Si 2.8/4.2 nm with sharp interfaces, Ru cap 2.5 nm, SiO2 substrate.
HIGH_NA_STACK = Stack(n_mo=0.924 + 0.0064j, n_si=0.999 + 0.0018j, d_mo=2.8, d_si=4.2,
                      d_mosi=0.0, d_simo=0.0, d_siru=0.0)


@dataclass(frozen=True)
class Source:
    """Illumination: kind 0 conventional, 1 annular, 2 X dipole, 3 Y dipole
    (intensity.cpp:68-74).  ``extra`` adds further (kind, sigma_out, sigma_in,
    open_angle) members whose union forms the source (e.g. dipole + annular)."""
    kind: int = 1
    sigma_out: float = 0.9
    sigma_in: float = 0.55
    open_angle: float = 90.0
    extra: tuple = ()
    ndivs: int = 0  # 0: reference rule max(1, int(180/pi lambda/dx/0.5)) (intensity.cpp:90-91)
    # > 0: keep exactly this many of the rule's lattice points (all divisions
    # pooled), spread evenly in azimuth; the explicit point list of a workload
    # whose point count is given (e.g. 61 or 200)
    count: int = 0

    def members(self):
        return ((self.kind, self.sigma_out, self.sigma_in, self.open_angle),) + tuple(self.extra)


@dataclass(frozen=True)
class Config:
    ndivx: int = 2048
    ndivy: int = 2048
    mx: int = 4
    my: int = 4
    na: float = 0.33
    wavelength: float = 13.5
    theta0: float = -6.0
    azimuth: float = 0.0
    n_abs: complex = 0.9567 + 0.0343j
    d_abs: float = 60.0
    z0: float = 60.0 + 42.0
    stack: Stack = REFERENCE_STACK
    orders: tuple = ("diamond", 6.0)        # ('diamond'|'ellipse', cutoff) or ('rect', K)
    narrow_orders: tuple = ("ellipse", 1.5)  # blank-mask solves of m3d.cpp:101-129
    source: Source = Source()
    image: tuple = (0, 0)                    # (nx, ny); 0 = ndivx/mx (intensity.cpp:53)
    obscuration: float = 0.0                 # central obscuration (elitho const.co)

    # ---- derived scalars (intensity.cpp:51-130) ----
    @property
    def dx(self):
        return float(self.ndivx)

    @property
    def dy(self):
        return float(self.ndivy)

    @property
    def k(self):
        return 2.0 * PI / self.wavelength

    @property
    def kx0(self):
        return self.k * math.sin(PI / 180.0 * self.theta0) * math.cos(PI / 180.0 * (90.0 - self.azimuth))

    @property
    def ky0(self):
        return self.k * math.sin(PI / 180.0 * self.theta0) * math.sin(PI / 180.0 * (90.0 - self.azimuth))

    @property
    def eps_abs(self):
        return complex(self.n_abs) ** 2

    @property
    def lsmax(self):
        return (int(self.na * self.dx / float(self.mx) / self.wavelength + 1),
                int(self.na * self.dy / float(self.my) / self.wavelength + 1))

    @property
    def lpmax(self):
        return (int(self.na * self.dx / float(self.mx) * 2 / self.wavelength + 0.0001),
                int(self.na * self.dy / float(self.my) * 2 / self.wavelength + 0.0001))

    @property
    def nsource(self):
        a, b = self.lsmax
        return 2 * a + 1, 2 * b + 1

    @property
    def image_shape(self):
        nx, ny = self.image
        return (nx or self.ndivx // self.mx, ny or self.ndivx // self.mx)

    @property
    def ndivs(self):
        if self.source.ndivs:
            return self.source.ndivs
        return max(1, int(180.0 / PI * self.wavelength / self.dx / 0.5))

    def in_na(self, l, m):
        l = np.asarray(l, dtype=np.float64)
        m = np.asarray(m, dtype=np.float64)
        return (l * self.mx / self.dx) ** 2 + (m * self.my / self.dy) ** 2 <= (self.na / self.wavelength) ** 2

    # ---- order sets ----
    def order_set(self, spec=None):
        """(lindex, mindex, LMAX, MMAX); l-major enumeration as intensity.cpp:148-164."""
        kind, par = spec or self.orders
        if kind == "rect":
            L, M = (par, par) if np.isscalar(par) else par
        else:
            L = int(self.na / self.mx * par * self.dx / self.wavelength)
            M = int(self.na / self.my * par * self.dy / self.wavelength)
        l, m = np.meshgrid(np.arange(-L, L + 1), np.arange(-M, M + 1), indexing="ij")
        if kind == "diamond":
            keep = (np.abs(l) / (L + 0.01) + 1.0) * (np.abs(m) / (M + 0.01) + 1.0) <= 2.0
        elif kind == "ellipse":
            keep = (l / (L + 0.01)) ** 2.0 + (m / (M + 0.01)) ** 2.0 <= 1.0
        elif kind == "rect":
            keep = np.ones_like(l, dtype=bool)
        else:
            raise ValueError(f"unknown order set {kind!r}")
        return l[keep].astype(np.int32), m[keep].astype(np.int32), int(L), int(M)

    def blank_order_set(self):
        """Order set of the blank-mask solves: the narrow set of m3d.cpp:101-129
        plus every lattice incidence inside the NA that it misses.  The blank
        amplitude of incidence (ls, ms) is read at mode (ls, ms)
        (m3d.cpp:240-258); at small pitches the truncated ellipse can miss an
        in-NA incidence (256 nm at NA 0.55: (2, 1)), where the reference reads
        an unset entry.  A uniform mask couples no modes, so the added modes
        leave every other amplitude unchanged."""
        li, mi, L, M = self.order_set(self.narrow_orders)
        have = set(zip(li.tolist(), mi.tolist()))
        lsx, lsy = self.lsmax
        extra = [(ls, ms) for ls in range(-lsx, lsx + 1) for ms in range(-lsy, lsy + 1)
                 if self.in_na(ls, ms) and (ls, ms) not in have]
        if not extra:
            return li, mi, L, M
        el, em = np.array(extra, dtype=np.int32).T
        return (np.concatenate([li, el]), np.concatenate([mi, em]), max(L, int(np.abs(el).max())),
                max(M, int(np.abs(em).max())))

    @staticmethod
    def mode_map(lidx, midx, L, M):
        mp = np.full((2 * L + 1) * (2 * M + 1), -1, dtype=np.int32)
        mp[(lidx + L) * (2 * M + 1) + (midx + M)] = np.arange(len(lidx), dtype=np.int32)
        return mp

    # ---- pupil points (intensity.cpp:166-195, m3d.cpp:160-195) ----
    @cached_property
    def pupil(self):
        """(linput, minput, snum): pupil points with at least one overlapping
        lattice source, ip-major as the reference loops; snum = number of sources."""
        (lsx, lsy), (lpx, lpy) = self.lsmax, self.lpmax
        ls, ms = np.meshgrid(np.arange(-lsx, lsx + 1), np.arange(-lsy, lsy + 1), indexing="ij")
        ok_s = self.in_na(ls, ms)
        ls, ms = ls[ok_s], ms[ok_s]
        lp, mp = np.meshgrid(np.arange(-lpx, lpx + 1), np.arange(-lpy, lpy + 1), indexing="ij")
        lp, mp = lp.ravel(), mp.ravel()
        cnt = self.in_na(lp[:, None] + ls[None, :], mp[:, None] + ms[None, :]).sum(axis=1)
        keep = cnt > 0
        return lp[keep].astype(np.int32), mp[keep].astype(np.int32), cnt[keep].astype(np.int32)

    # ---- source lattice (intensity.cpp:467-502) ----
    def source_points(self):
        """(nsx, nsy, l0s array, m0s array) per division; computed once per
        configuration (the arrays are read-only)."""
        return self._source_points

    @cached_property
    def _source_points(self):
        out = self._lattice_points()
        for _, _, l0, m0 in out:
            l0.flags.writeable = False
            m0.flags.writeable = False
        return tuple(out)

    def _lattice_points(self):
        nd = self.ndivs
        k, NA = self.k, self.na
        mem = self.source.members()
        s1 = max(m[1] for m in mem)
        dkx, dky = 2.0 * PI / self.dx, 2.0 * PI / self.dy
        l0max = int(k * NA / self.mx * s1 / dkx + 1)
        m0max = int(k * NA / self.my * s1 / dky + 1)
        l, m = np.meshgrid(np.arange(-l0max, l0max + 1), np.arange(-m0max, m0max + 1), indexing="ij")
        l, m = l.ravel(), m.ravel()
        out = []
        for nsx in range(nd):
            for nsy in range(nd):
                skx = l * dkx + 2.0 * PI / self.dx * nsx / float(nd)
                sky = m * dky + 2.0 * PI / self.dy * nsy / float(nd)
                sxo, syo = skx * self.mx, sky * self.my
                r2 = sxo * sxo + syo * syo
                r = np.sqrt(r2)
                ok = np.zeros(len(l), dtype=bool)
                for (t, so, si, oa) in mem:
                    ring = (r <= k * NA * so) & (r >= k * NA * si)
                    if t == 0:
                        ok |= r2 <= (k * NA * so) ** 2
                    elif t == 1:
                        ok |= ring
                    elif t == 2:
                        ok |= ring & (np.abs(syo) <= np.abs(sxo) * math.tan(PI * oa / 180.0 / 2.0))
                    elif t == 3:
                        ok |= ring & (np.abs(sxo) <= np.abs(syo) * math.tan(PI * oa / 180.0 / 2.0))
                out.append((nsx, nsy, l[ok].astype(np.int32), m[ok].astype(np.int32)))
        if self.source.count > 0:
            out = self._thin(out, self.source.count)
        return out

    def _thin(self, divs, count):
        """Exactly ``count`` of the pooled points: sorted by azimuth (then
        radius) of their pupil position, every (total/count)-th is kept."""
        nd = self.ndivs
        pts = []
        for q, (nsx, nsy, l0, m0) in enumerate(divs):
            sx = (l0 + nsx / float(nd)) * self.mx / self.dx
            sy = (m0 + nsy / float(nd)) * self.my / self.dy
            for a, b, x, y in zip(l0, m0, sx, sy):
                pts.append((math.atan2(y, x), x * x + y * y, q, int(a), int(b)))
        if count > len(pts):
            raise ValueError(f"source rule gives {len(pts)} points, fewer than the {count} asked for")
        pts.sort()
        keep = [pts[int((i + 0.5) * len(pts) / count)] for i in range(count)]
        out = []
        for q, (nsx, nsy, _, _) in enumerate(divs):
            sel = sorted((a, b) for _, _, qq, a, b in keep if qq == q)
            l0 = np.array([a for a, _ in sel], dtype=np.int32)
            m0 = np.array([b for _, b in sel], dtype=np.int32)
            out.append((nsx, nsy, l0, m0))
        return out

    def mesh_source(self, mesh_deg=0.2):
        """(dsx, dsy) wave-number offsets of the source mesh of linear.py:46-97
        (eigenfunction.py:6-56): a square mesh of mesh_deg degrees, filtered by
        the source shape (union of its members)."""
        k, NA = self.k, self.na
        kmesh = k * mesh_deg * (PI / 180.0)
        mem = self.source.members()
        s1 = max(m[1] for m in mem)
        l0max = int(k * NA / self.mx * s1 / kmesh) + 1
        m0max = int(k * NA / self.my * s1 / kmesh) + 1
        l, m = np.meshgrid(np.arange(-l0max, l0max + 1), np.arange(-m0max, m0max + 1), indexing="ij")
        skx, sky = (l * kmesh).ravel(), (m * kmesh).ravel()
        sxo, syo = skx * self.mx, sky * self.my
        r = np.sqrt(sxo ** 2 + syo ** 2)
        ok = np.zeros(len(skx), dtype=bool)
        for (t, so, si, oa) in mem:
            ring = (k * NA * si <= r) & (r <= k * NA * so)
            if t == 0:
                ok |= (sxo ** 2 + syo ** 2) <= (k * NA * so) ** 2
            elif t == 1:
                ok |= ring
            elif t == 2:
                ok |= ring & (np.abs(syo) <= np.abs(sxo) * np.tan(PI * oa / 180.0 / 2.0))
            elif t == 3:
                ok |= ring & (np.abs(sxo) <= np.abs(syo) * np.tan(PI * oa / 180.0 / 2.0))
        return skx[ok], sky[ok]

    def offsets(self):
        """Incidence offsets (sx0, sy0, fx, fy) per division (intensity.cpp:234-238)."""
        nd = self.ndivs
        out = []
        for nsx in range(nd):
            for nsy in range(nd):
                out.append((2.0 * PI / self.dx * nsx / nd + self.kx0, 2.0 * PI / self.dy * nsy / nd + self.ky0,
                            nsx / float(nd), nsy / float(nd)))
        return np.array(out, dtype=np.float64)


def reference_default():
    """The reference drivers' hard-coded problem (intensity.cpp:44-87, m3d.cpp:44-111)."""
    return Config()


def high_na(ndiv, K, source, ndivs=0, image=(0, 0)):
    """NA 0.55 anamorphic 4x/8x, CRAO 5.355 deg, 60 nm absorber n = 0.947+0.031i on
    40 x Mo/Si 2.8/4.2 nm with a 2.5 nm Ru cap; rectangular +-K orders."""
    return Config(ndivx=ndiv, ndivy=ndiv, mx=4, my=8, na=0.55, theta0=-5.355, n_abs=0.947 + 0.031j,
                  stack=HIGH_NA_STACK, orders=("rect", K), source=Source(**{**source, "ndivs": ndivs}),
                  image=image)


ANNULAR = dict(kind=1, sigma_out=0.8, sigma_in=0.3)
DIPOLE_ANNULAR = dict(kind=2, sigma_out=0.9, sigma_in=0.5, open_angle=70.0, extra=((1, 0.8, 0.7, 90.0),))
# annular 0.3-0.8 at 61 points; X dipole 0.5-0.9 / 70 deg + annular 0.7-0.8 at 200 points
ANNULAR_C1 = dict(ANNULAR, count=61)
DIPOLE_ANNULAR_C2 = dict(DIPOLE_ANNULAR, count=200)


def workload_mix():
    """The problems of the uneven-clip workloads (batched and served forms):
    pitch 256-1024 nm, order truncation +-8..+-20 growing with the pitch,
    annular or dipole + annular sources of 61-200 points."""
    return [
        high_na(256, 8, dict(ANNULAR, count=61), ndivs=4, image=(64, 64)),
        high_na(512, 12, dict(ANNULAR, count=80), ndivs=2, image=(128, 128)),
        high_na(512, 12, dict(DIPOLE_ANNULAR, count=100), ndivs=3, image=(128, 128)),
        high_na(768, 16, dict(ANNULAR, count=150), ndivs=2, image=(192, 192)),
        preset("c2"),
        high_na(1024, 20, dict(ANNULAR, count=61), ndivs=1, image=(256, 256)),
    ]


def preset(name):
    """Named problems: 'reference' (the reference drivers' own), 'c1' (512 nm
    clip, +-10 orders, annular 0.3-0.8 at 61 points of the 2 x 2 division
    lattice, 256 x 256 image) and 'c2' (1024 nm clip, +-20 orders, X dipole
    0.5-0.9 / 70 deg + annular 0.7-0.8 at 200 points of the 2 x 2 division
    lattice, 256 x 256 image)."""
    if name == "reference":
        return reference_default()
    if name == "c1":
        return high_na(512, 10, ANNULAR_C1, ndivs=2, image=(256, 256))
    if name == "c2":
        return high_na(1024, 20, DIPOLE_ANNULAR_C2, ndivs=2, image=(256, 256))
    raise ValueError(f"unknown preset {name!r} (reference, c1, c2)")

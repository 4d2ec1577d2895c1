"""The alternative STCC image of the Python elitho package: TCCEE
(cnn/intensity/stccsocs/elitho/stcc.py:12-431) on the GPU.

``intensity(mask, polar, a0xx, axxx, ayxx)`` keeps the reference's call: the
mask is the package's array M[x][y'] (0/1, as LinePattern returns it), the
optional M3D arrays are (noutX, noutY) complex.  The pupil points
(pupil.py:5-66) and the uniform k-source (source.py:157-207) are built on the
host from the constants; the mask spectrum, the field terms, the three TCCs
(batched ZGEMMs over the source points), the TCCEE gather and the image
transform run in libeuvl (``euvl_stcc_image``).

Constants follow elitho/const.py (the non-high-NA branch: NA 0.33, 4x, 2048 nm,
mesh 0.5 deg, annular 0.55-0.9).  The zero-order amplitudes (amp_absorber,
amp_vacuum, phasexx of diffraction_amplitude.zero_order_amplitude) come from
this library's blank-mask solve of the same optics, or are passed in.
"""
import ctypes
import functools
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib


@dataclass(frozen=True)
class StccConfig:
    ndiv: int = 2048
    mx: int = 4
    my: int = 4
    na: float = 0.33
    wavelength: float = 13.5
    theta0: float = -6.0
    azimuth: float = 0.0
    mesh: float = 0.5
    sigma1: float = 0.9
    sigma2: float = 0.55
    illumination: str = "annular"   # or "circular"
    z0: float = 102.0               # dabst + 42

    @property
    def k(self):
        return 2.0 * np.pi / self.wavelength

    @property
    def kx0(self):
        return self.k * np.sin(np.deg2rad(self.theta0)) * np.cos(np.deg2rad(90.0 - self.azimuth))

    @property
    def ky0(self):
        return self.k * np.sin(np.deg2rad(self.theta0)) * np.sin(np.deg2rad(90.0 - self.azimuth))

    @property
    def xdiv(self):
        return self.ndiv // self.mx

    @property
    def lsmax(self):
        return (int(self.na * self.ndiv / self.mx / self.wavelength + 1),
                int(self.na * self.ndiv / self.my / self.wavelength + 1))

    @property
    def lpmax(self):
        return (int(self.na * self.ndiv / self.mx * 2 / self.wavelength + 0.0001),
                int(self.na * self.ndiv / self.my * 2 / self.wavelength + 0.0001))

    def pupil(self):
        """pupil.py:24-44 (offsets 0): (l, m) of the points with at least one
        lattice source inside the NA whose shifted point is inside too,
        ip-major."""
        return _pupil(self).copy()

    def source(self):
        """source.py:157-199: (sx, sy) of the mesh points inside the source."""
        return _source(self).copy()


@functools.lru_cache(maxsize=16)
def _pupil(cfg):
    lsx, lsy = cfg.lsmax
    lpx, lpy = cfg.lpmax
    na2 = (cfg.na / cfg.wavelength) ** 2
    sx = np.arange(-lsx, lsx + 1)[None, None, :, None]
    sy = np.arange(-lsy, lsy + 1)[None, None, None, :]
    ip = np.arange(2 * lpx + 1)[:, None, None, None]
    jp = np.arange(2 * lpy + 1)[None, :, None, None]
    src_ok = (sx * cfg.mx / cfg.ndiv) ** 2 + (sy * cfg.my / cfg.ndiv) ** 2 <= na2
    ok = ((ip - lpx + sx) * cfg.mx / cfg.ndiv) ** 2 + ((jp - lpy + sy) * cfg.my / cfg.ndiv) ** 2 <= na2
    hit = np.any(src_ok & ok, axis=(2, 3))
    I, J = np.nonzero(hit)  # ip-major, as the reference's loops
    return np.stack([I - lpx, J - lpy], axis=1).astype(np.int32)


@functools.lru_cache(maxsize=16)
def _source(cfg):
    """source.py:157-199: (sx, sy) of the mesh points inside the source."""
    k = cfg.k
    kmesh = k * cfg.mesh * (np.pi / 180.0)
    l0max = int(k * cfg.na / cfg.mx * cfg.sigma1 / kmesh + 1)
    m0max = int(k * cfg.na / cfg.my * cfg.sigma1 / kmesh + 1)
    L, M = np.meshgrid(np.arange(-l0max, l0max + 1), np.arange(-m0max, m0max + 1), indexing="ij")
    skx, sky = L * kmesh, M * kmesh
    skxo, skyo = skx * cfg.mx, sky * cfg.my
    if cfg.illumination == "circular":
        sel = skxo ** 2 + skyo ** 2 <= (k * cfg.na * cfg.sigma1) ** 2
    elif cfg.illumination == "annular":
        r = np.sqrt(skxo ** 2 + skyo ** 2)
        sel = (k * cfg.na * cfg.sigma2 <= r) & (r <= k * cfg.na * cfg.sigma1)
    else:
        raise ValueError("Invalid illumination type")
    return np.stack([skx[sel], sky[sel]], axis=1).astype(np.float64)


def zero_order_amplitude(cfg: StccConfig = StccConfig(), device="cuda:0"):
    """(amp_absorber, amp_vacuum, phasexx) of the blank-mask solve at the
    central incidence (diffraction_amplitude.py:156-169), from this library's
    EM solve of the matching reference optics."""
    from .config import reference_default
    from .engine import Engine
    if cfg != StccConfig():
        raise ValueError("zero-order amplitudes are solved for the reference optics only; pass them explicitly")
    eng = Engine(reference_default(), device)
    vc, ab, ph = eng.reflect_amplitude()
    eng.close()
    return ab, vc, ph


def _bits_of(mask, ndiv):
    """The package's mask array M[x][y'] -> mask.bin bytes (pixel (x, ndiv-1-y')):
    mask.bin row ndiv-1-y' packs M[x][y'] over x, so the bits are packed down
    the first axis (eight x rows at a time, contiguous) and only the packed
    array, an eighth of the size, is transposed."""
    m = np.asarray(mask)
    if m.shape != (ndiv, ndiv) or ndiv % 8:
        raise ValueError(f"mask must be {ndiv} x {ndiv} with ndiv a multiple of 8")
    u = (m != 0).view(np.uint8).reshape(ndiv // 8, 8, ndiv)
    packed = np.zeros((ndiv // 8, ndiv), dtype=np.uint8)
    for bit in range(8):
        packed |= u[:, bit] << np.uint8(7 - bit)
    return np.ascontiguousarray(packed[:, ::-1].T).reshape(-1)


def intensity(mask, polar="X", a0xx=None, axxx=None, ayxx=None, cfg: StccConfig = StccConfig(), amplitudes=None,
              device="cuda:0", handle=None, timing=None):
    """stcc.py:394-431: the TCCEE image [x][y] (xdiv x xdiv float64 host
    array).  polar 'X' or 'Y' (elitho PolarizationDirection); the M3D arrays
    are all given or all None ('Y' needs them, as the reference's branch
    reads them unconditionally).  amplitudes = (amp_absorber, amp_vacuum,
    phasexx) overrides the blank-mask solve.  timing (a dict) receives the
    stream time in ms of the library calls ("device_ms")."""
    polar = getattr(polar, "name", polar)
    if polar not in ("X", "Y"):
        raise ValueError("polar must be 'X' or 'Y'")
    ab, vc, ph = amplitudes if amplitudes is not None else zero_order_amplitude(cfg, device)
    lpx, lpy = cfg.lpmax
    dev = torch.device(device)
    bits = torch.from_numpy(_bits_of(mask, cfg.ndiv)).to(dev)
    fm = torch.empty((2 * lpx + 1, 2 * lpy + 1), dtype=torch.complex128, device=dev)
    if handle is None:
        # the per-thread handle of this package: its device workspace (the
        # three TCCs, ~170 MB at the reference size) is kept between calls
        from .cnn import _handle
        lib, h = _handle(dev)
    else:
        lib, h = _lib.load(), handle
        _lib.check(lib.euvl_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)),
                   "euvl_set_stream")
    if timing is not None:
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
    _lib.check(lib.euvl_mask_spectrum(h, _lib.ptr(bits), cfg.ndiv, cfg.ndiv, 1, lpx, lpy, _lib.ptr(fm)),
               "euvl_mask_spectrum")
    m3d = [None if a is None else torch.as_tensor(np.asarray(a, dtype=np.complex128)).to(dev).contiguous()
           for a in (a0xx, axxx, ayxx)]
    if any(a is not None and tuple(a.shape) != tuple(fm.shape) for a in m3d):
        raise ValueError(f"M3D arrays must be {tuple(fm.shape)}")
    pupil = np.ascontiguousarray(cfg.pupil())
    src = np.ascontiguousarray(cfg.source())
    prm = _lib.StccParams(cfg.k, cfg.kx0, cfg.ky0, cfg.z0, float(cfg.ndiv), float(cfg.ndiv), cfg.na, cfg.mx,
                          cfg.my, lpx, lpy, cfg.xdiv)
    img = torch.empty((cfg.xdiv, cfg.xdiv), dtype=torch.float64, device=dev)
    cplx = lambda z: (ctypes.c_double * 2)(complex(z).real, complex(z).imag)
    p = lambda t: _lib.ptr(t) if t is not None else None
    _lib.check(lib.euvl_stcc_image(h, polar.encode(), ctypes.byref(prm), _lib.ptr(fm), cplx(ab), cplx(vc),
                                   cplx(ph), p(m3d[0]), p(m3d[1]), p(m3d[2]), len(pupil), pupil.ctypes.data,
                                   len(src), src.ctypes.data, _lib.ptr(img)), "euvl_stcc_image")
    if timing is not None:
        ev[1].record()
        ev[1].synchronize()
        timing["device_ms"] = ev[0].elapsed_time(ev[1])
    return img.cpu().numpy()

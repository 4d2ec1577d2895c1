"""Python callables with the reference's names, argument order and return
shapes, running on libeuvl:

  reflect_amplitude()                          reflectivity.py:4-17
  eigenfunction(z)                             eigenfunction.py:5-179
  linear(z, maskname, re0, im0, rex, imx, rey, imy)  linear.py:190-225
  socs(z, maskname, a0, ax, ay)                socs.py:156-204

plus the two C++ drivers as functions:

  m3d(maskname, ndata)                         m3d.cpp main (inputxx.csv rows)
  abbe(maskname, polar, z)                     intensity.cpp main (emint.csv)

The reference keeps the reflectivity scalars and the SOCS kernels as files in
the working directory (reflectivity/reflectivity.txt read back with eval,
eigenfunction/eigenfunctions.pkl); here they are cached in memory per
Session, keyed by what they depend on, and stay on the device.

The optics come from the Session's Config (default: the reference drivers'
problem, config.reference_default()).  Images are float64 arrays indexed
[x][y]; M3D slopes ax, ay are given and returned over the pupil points with
x = 1 only, as in inputxx.csv.  Every call needs the GPU library; there is no
CPU path.
"""
import numpy as np

from . import csvio
from .config import Config, reference_default
from .masks import read_mask_bin

_session = None


class Session:
    def __init__(self, cfg: Config = None, device="cuda:0"):
        self.cfg = cfg if cfg is not None else reference_default()
        self.device = device
        self._engine = None

    @property
    def engine(self):
        if self._engine is None:
            from .engine import Engine
            self._engine = Engine(self.cfg, self.device)
        return self._engine

    @property
    def xinput(self):
        """0/1 flag per pupil point: at least 8 lattice sources (inputlm.csv x)."""
        return (self.cfg.pupil[2] >= 8).astype(np.int64)

    def _bits(self, maskname, sample=0):
        m = read_mask_bin(maskname, self.cfg.ndivx, self.cfg.ndivy, sample)
        return self.engine.upload_mask(m)

    # -------------------------------------------------------------- reference names
    def reflect_amplitude(self):
        """(ampvc, ampab, phase0) complex scalars."""
        return self.engine.reflect_amplitude()

    def eigenfunction(self, z, nk=None, mesh_deg=0.2):
        """(alphaXs0, phiXs0, alphaXsx, phiXsx, alphaXsy, phiXsy): eigenvalues
        sorted by |alpha| descending and kernels phi[m, n] (kernel m, pupil
        point n); nk = number kept (default: all)."""
        ncut = len(self.cfg.pupil[0])
        alpha, phi = self.engine.eigenfunction(float(z), ncut if nk is None else int(nk), mesh_deg)
        a, p = alpha.cpu().numpy(), phi.cpu().numpy()
        return a[0], p[0], a[1], p[1], a[2], p[2]

    def linear(self, z, maskname, re0, im0, rex, imx, rey, imy):
        """(intlinear, intft)."""
        xi = self.xinput
        a0 = np.asarray(re0) + 1j * np.asarray(im0)
        ax = csvio.expand_xy(np.asarray(rex) + 1j * np.asarray(imx), xi)
        ay = csvio.expand_xy(np.asarray(rey) + 1j * np.asarray(imy), xi)
        il, it = self.engine.linear(self._bits(maskname), a0, ax, ay, z=float(z))
        return il.cpu().numpy(), it.cpu().numpy()

    def socs(self, z, maskname, a0, ax, ay, nsocs=100, nsocsxy=20, nout=64):
        """intsocs (socs.py defaults: 100 kernels, 20 with the slope terms, nout 64)."""
        xi = self.xinput
        img = self.engine.socs(self._bits(maskname), np.asarray(a0), csvio.expand_xy(np.asarray(ax), xi),
                               csvio.expand_xy(np.asarray(ay), xi), nsocs=nsocs, nsocsxy=nsocsxy, nout=nout,
                               z=float(z))
        return img.cpu().numpy()

    # -------------------------------------------------------------- C++ drivers
    def m3d(self, maskname="mask.bin", ndata=1):
        """Per clip (a0, ax, ay): a0 over every pupil point, ax / ay over the
        points with x = 1 (the six rows of inputxx.csv, as complex)."""
        sel = self.xinput.astype(bool)
        out = []
        for s in range(ndata):
            a0, ax, ay = (t.cpu().numpy() for t in self.engine.m3d(self._bits(maskname, s)))
            out.append((a0, ax[sel], ay[sel]))
        return out

    def abbe(self, maskname="mask.bin", polar="X", z=0.0, sample=0):
        """Rigorous Abbe image (emint.csv for 'X', emintY.csv for 'Y'); z may be
        a sequence (through-focus stack, one EM solve)."""
        img = self.engine.abbe(self._bits(maskname, sample), z=z, polar=polar)
        return img.cpu().numpy()


def session(cfg: Config = None):
    """The module-level Session used by the free functions (created on first
    use; pass a Config to replace it)."""
    global _session
    if cfg is not None or _session is None:
        _session = Session(cfg)
    return _session


def reflect_amplitude():
    return session().reflect_amplitude()


def eigenfunction(z):
    return session().eigenfunction(z)


def linear(z, maskname, re0, im0, rex, imx, rey, imy):
    return session().linear(z, maskname, re0, im0, rex, imx, rey, imy)


def socs(z, maskname, a0, ax, ay):
    return session().socs(z, maskname, a0, ax, ay)

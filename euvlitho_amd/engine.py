"""Host layer over libeuvl: the EM solve, M3D extraction and rigorous Abbe image
of one clip, mirroring the reference drivers

  ampS            cnn/include/header.h:1208-1466       -> Engine.amp_s
  m3d.cpp main    cnn/intensity/m3d/m3d.cpp:28-569      -> Engine.m3d
  intensity.cpp   cnn/intensity/abbe/intensity.cpp:35-421 -> Engine.abbe
  reflectivity.py cnn/intensity/linear/reflectivity.py:4-17 -> Engine.reflect_amplitude
  linear.py       cnn/intensity/linear/linear.py:10-225   -> Engine.linear
  eigenfunction.py cnn/intensity/stccsocs/eigenfunction.py:5-179 -> Engine.eigenfunction
  socs.py         cnn/intensity/stccsocs/socs.py:12-204   -> Engine.socs

Device memory and streams come from PyTorch (ROCm); all arithmetic runs in
libeuvl's HIP kernels.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .config import Config
from .masks import pack_bits


class Resident:
    """Mask-independent state kept on the device (or host) across clips and
    shared between engines and streams: blank amplitudes and SOCS kernels per
    configuration.  Thread-safe; an entry is complete (its stream synchronised)
    before it is published."""

    def __init__(self):
        import threading
        self._d = {}
        self._lock = threading.Lock()

    def get(self, key):
        with self._lock:
            return self._d.get(key)

    def put(self, key, value):
        with self._lock:
            return self._d.setdefault(key, value)

    def __len__(self):
        return len(self._d)


class Engine:
    def __init__(self, cfg: Config, device="cuda:0", precision=_lib.PREC_FP64, handle=None, resident=None):
        """precision: euvl_set_precision mask (0 = the FP64 drop-in; _lib.PREC_MIXED =
        the coupling and top-solve products on FP32 MFMA and the imaging in FP32, the
        matrix functions and the multilayer cascade in FP64; _lib.PREC_FAST = every
        product of the EM solve on FP32 MFMA too).
        handle: an existing euvl handle to run on (not owned: close() leaves it);
        resident: a Resident store shared between engines for the mask-independent
        state (blank amplitudes, SOCS kernels)."""
        if not torch.cuda.is_available():
            raise _lib.EuvlError("a ROCm GPU is required (torch.cuda.is_available() is False)")
        self.cfg = cfg
        self.device = torch.device(device)
        self.lib = _lib.load()
        self._owns = handle is None
        if handle is None:
            with torch.cuda.device(self.device):
                handle = ctypes.c_void_p()
                _lib.check(self.lib.euvl_create(ctypes.byref(handle)), "euvl_create")
        self.h = handle
        self.precision = int(precision)
        _lib.check(self.lib.euvl_set_precision(self.h, self.precision), "euvl_set_precision")
        self.resident = resident if resident is not None else Resident()
        self._layers = cfg.stack.layers()
        eps = np.array([[complex(e).real, complex(e).imag] for e, _ in self._layers], dtype=np.float64).ravel()
        thk = np.array([d for _, d in self._layers], dtype=np.float64)
        self._eps, self._thk = eps, thk
        lsx, lsy = cfg.lsmax
        ea = cfg.eps_abs
        self.params = _lib.EmParams(
            cfg.ndivx, cfg.ndivy, cfg.mx, cfg.my, cfg.na, cfg.wavelength, cfg.k, cfg.dx, cfg.dy, lsx, lsy,
            ea.real, ea.imag, cfg.d_abs, len(self._layers),
            eps.ctypes.data_as(_lib.P_double), thk.ctypes.data_as(_lib.P_double))
        self.lidx, self.midx, self.L, self.M = cfg.order_set()
        self.modemap = Config.mode_map(self.lidx, self.midx, self.L, self.M)
        self._modemap_c = _lib.arr_i(self.modemap)
        self._abbe_tab = None  # host tables of the Abbe image (built on first use)
        self.last_info = None

    def _abbe_tables(self):
        """(n_off, offsets, n_src, (division, l0, m0) per source point, n_pupil,
        pupil (l, m)) as ctypes-ready arrays, built once per engine."""
        if self._abbe_tab is None:
            cfg = self.cfg
            offs = cfg.offsets()
            src = []
            for o, (_, _, l0, m0) in enumerate(cfg.source_points()):
                src += [(o, int(a), int(b)) for a, b in zip(l0, m0)]
            src = np.array(src, dtype=np.int32).reshape(-1, 3)
            lin, min_, _ = cfg.pupil
            self._abbe_tab = (len(offs), _lib.arr_d(offs), len(src), _lib.arr_i(src), len(lin),
                              _lib.arr_i(np.stack([lin, min_], axis=1)))
        return self._abbe_tab

    def close(self):
        if self.h and self._owns:
            self.lib.euvl_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self):
        """Time the GEMM kernel variants on this problem's main product shape
        (N x N x N, batched over the source offsets) in FP64 and, for the
        precision form, in FP32 compute; the fastest is kept for that size
        bucket by every later GEMM of the process.  Returns
        {"fp64" | "fp32": (variant, [ms per variant])}."""
        n, b = len(self.lidx), len(self.cfg.offsets())
        out = {}
        for fp32 in ((0, 1) if self.precision else (0,)):
            v = ctypes.c_int()
            ms = (ctypes.c_double * 7)()
            self._bind_stream()
            _lib.check(self.lib.euvl_autotune_gemm(self.h, n, n, n, b, fp32, ctypes.byref(v), ms),
                       "euvl_autotune_gemm")
            out["fp32" if fp32 else "fp64"] = (v.value, list(ms))
        return out

    def _bind_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib.euvl_set_stream(self.h, ctypes.c_void_p(s)), "euvl_set_stream")

    # ------------------------------------------------------------------ inputs
    def upload_mask(self, mask2d):
        """mask2d (0/1 array of ndivx*ndivy, index ndivx*j + i) or packed mask.bin bytes -> device uint8."""
        if isinstance(mask2d, (bytes, bytearray)):
            raw = np.frombuffer(mask2d, dtype=np.uint8)
        else:
            raw = np.frombuffer(pack_bits(mask2d), dtype=np.uint8)
        return torch.from_numpy(raw.copy()).to(self.device)

    def mask_spectrum(self, bits, lw, mw, n_clips=1):
        """Indicator spectrum window (n_clips, 2lw+1, 2mw+1) complex128 (euvl_mask_spectrum).

        The last window computed is remembered with the mask tensor it came
        from (object and in-place version): the M3D fit and the images ask for
        the pupil window, which lies inside the EM solve's window of the same
        clip, and get a slice of it instead of a second transform."""
        last = getattr(self, "_spec_last", None)
        if last is not None:
            b, ver, n, lw0, mw0, spec = last
            if b is bits and ver == bits._version and n == n_clips and lw <= lw0 and mw <= mw0:
                return spec[:, lw0 - lw:lw0 + lw + 1, mw0 - mw:mw0 + mw + 1].contiguous()
        out = torch.empty((n_clips, 2 * lw + 1, 2 * mw + 1), dtype=torch.complex128, device=self.device)
        self._bind_stream()
        _lib.check(self.lib.euvl_mask_spectrum(self.h, _lib.ptr(bits), self.cfg.ndivx, self.cfg.ndivy, n_clips,
                                               lw, mw, _lib.ptr(out)), "euvl_mask_spectrum")
        self._spec_last = (bits, bits._version, n_clips, lw, mw, out)
        return out

    # ---------------------------------------------------------------- EM solve
    def amp_s(self, bits, polar="X", offsets=None, orders=None):
        """Ax[o, ls, ms, n] for the incidence offsets (sx0, sy0) (default: the
        source divisions of the config).  ``orders``: (lidx, midx, L, M)."""
        cfg = self.cfg
        lidx, midx, L, M = orders if orders is not None else (self.lidx, self.midx, self.L, self.M)
        if offsets is None:
            offsets = cfg.offsets()[:, :2]
        offsets = np.asarray(offsets, dtype=np.float64).reshape(-1, 2)
        spec = self.mask_spectrum(bits, 2 * L, 2 * M)
        nsx, nsy = cfg.nsource
        n_off = len(offsets)
        Ax = torch.empty((n_off, nsx, nsy, len(lidx)), dtype=torch.complex128, device=self.device)
        li, lp = _lib.arr_i(lidx)
        mi, mp = _lib.arr_i(midx)
        sx, sxp = _lib.arr_d(offsets[:, 0])
        sy, syp = _lib.arr_d(offsets[:, 1])
        info = np.zeros(4 * n_off, dtype=np.int32)
        self._bind_stream()
        _lib.check(self.lib.euvl_amp_s(self.h, polar.encode(), ctypes.byref(self.params), lp, mp, len(lidx), L, M,
                                       _lib.ptr(spec), n_off, sxp, syp, _lib.ptr(Ax),
                                       info.ctypes.data_as(_lib.P_int)), "euvl_amp_s")
        self.last_info = info.reshape(n_off, 4)
        return Ax

    def blank_amplitudes(self, polar="X", masks=None):
        """vcxx, abxx: blank vacuum / absorber amplitudes on the narrow order set at
        the central offset (m3d.cpp:211-258); kept in the resident store per
        configuration and polarisation.  masks: optional (vacuum, absorber)
        0/1 arrays of ndivx*ndivy to solve in place of the uniform blanks (not
        cached)."""
        key = ("blank", self.cfg, polar, self.precision)
        if masks is None:
            got = self.resident.get(key)
            if got is not None:
                return got
        cfg = self.cfg
        orders = cfg.blank_order_set()
        n = cfg.ndivx * cfg.ndivy
        m_vc, m_ab = (np.ones(n, np.int8), np.zeros(n, np.int8)) if masks is None else masks
        vc = self.amp_s(self.upload_mask(np.asarray(m_vc).reshape(-1)), polar, [(cfg.kx0, cfg.ky0)], orders)[0]
        ab = self.amp_s(self.upload_mask(np.asarray(m_ab).reshape(-1)), polar, [(cfg.kx0, cfg.ky0)], orders)[0]
        lidx, midx, L, M = orders
        mm = Config.mode_map(lidx, midx, L, M)
        nsx, nsy = cfg.nsource
        lsx, lsy = cfg.lsmax
        vcxx = np.zeros((nsx, nsy), dtype=np.complex128)
        abxx = np.zeros((nsx, nsy), dtype=np.complex128)
        vch, abh = vc.cpu().numpy(), ab.cpu().numpy()
        for is_ in range(nsx):
            for js in range(nsy):
                ls, ms = is_ - lsx, js - lsy
                if cfg.in_na(ls, ms) and abs(ls) <= L and abs(ms) <= M:
                    n_ = mm[(ls + L) * (2 * M + 1) + (ms + M)]
                    if n_ >= 0:
                        vcxx[is_, js] = vch[is_, js, n_]
                        abxx[is_, js] = abh[is_, js, n_]
        if masks is not None:
            return vcxx, abxx
        return self.resident.put(key, (vcxx, abxx))

    # --------------------------------------------------------------- M3D params
    def m3d(self, bits, Ax0=None, blanks=None):
        """(a0, ax, ay) of one clip (m3d.cpp main); ax, ay only where xinput = 1.
        Ax0: wide amplitudes at the central offset (computed when None);
        blanks: (vcxx, abxx) in place of the resident blank amplitudes."""
        cfg = self.cfg
        vcxx, abxx = self.blank_amplitudes() if blanks is None else blanks
        if Ax0 is None:
            Ax0 = self.amp_s(bits, "X", [(cfg.kx0, cfg.ky0)])[0]
        lpx, lpy = cfg.lpmax
        fmask = self.mask_spectrum(bits, lpx, lpy)[0]
        lin, min_, snum = cfg.pupil
        pup = np.stack([lin, min_, (snum >= 8).astype(np.int32)], axis=1)
        ncut = len(lin)
        a0 = torch.empty(ncut, dtype=torch.complex128, device=self.device)
        ax = torch.empty_like(a0)
        ay = torch.empty_like(a0)
        mmp = self._modemap_c[1]
        pu, pup_p = _lib.arr_i(pup)
        vc, vcp = _lib.arr_d(np.stack([vcxx.real, vcxx.imag], -1))
        ab, abp = _lib.arr_d(np.stack([abxx.real, abxx.imag], -1))
        self._bind_stream()
        _lib.check(self.lib.euvl_m3d_fit(self.h, ctypes.byref(self.params), _lib.ptr(fmask), lpx, lpy,
                                         _lib.ptr(Ax0.contiguous()), len(self.lidx), self.L, self.M, mmp, vcp, abp,
                                         cfg.kx0, cfg.ky0, cfg.z0, ncut, pup_p, _lib.ptr(a0), _lib.ptr(ax),
                                         _lib.ptr(ay)), "euvl_m3d_fit")
        return a0, ax, ay

    # ----------------------------------------------------------- Abbe (EM amps)
    def abbe(self, bits, z=0.0, Ax=None, polar="X"):
        """Rigorous Abbe image (nx, ny) of one clip (intensity.cpp main): polar
        'X' gives emint.csv, 'Y' emintY.csv.  z: defocus, or a sequence of
        defocus values (the z loop of intensity.cpp:219) -> (nz, nx, ny); the
        EM amplitudes do not depend on z and are solved once."""
        cfg = self.cfg
        if polar not in ("X", "Y"):
            raise ValueError("polar must be 'X' or 'Y'")
        if Ax is None:
            Ax = self.amp_s(bits, polar)
        if np.ndim(z) > 0:
            return torch.stack([self.abbe(bits, float(zz), Ax, polar) for zz in z])
        n_off, (_, ofp), n_src, (_, srp), n_pup, (_, pup_p) = self._abbe_tables()
        nx, ny = cfg.image_shape
        img = torch.zeros((nx, ny), dtype=torch.float64, device=self.device)
        mmp = self._modemap_c[1]
        self._bind_stream()
        _lib.check(self.lib.euvl_abbe_image(self.h, polar.encode(), ctypes.byref(self.params), _lib.ptr(Ax),
                                            len(self.lidx), self.L,
                                            self.M, mmp, n_off, ofp, n_src, srp, n_pup, pup_p, cfg.kx0,
                                            cfg.ky0, cfg.z0, float(z), nx, ny, 1.0 / max(n_src, 1),
                                            _lib.ptr(img)), "euvl_abbe_image")
        return img

    # ------------------------------------------------ images from M3D params
    def reflect_amplitude(self):
        """(ampvc, ampab, phase0): blank amplitudes at the central incidence
        (reflectivity.py:4-17); device-resident engine cache, no file."""
        vcxx, abxx = self.blank_amplitudes()
        lsx, lsy = self.cfg.lsmax
        ampvc, ampab = complex(vcxx[lsx, lsy]), complex(abxx[lsx, lsy])
        return ampvc, ampab, ampvc / abs(ampvc)

    def _m3d_inputs(self, a0, ax, ay):
        ncut = len(self.cfg.pupil[0])
        out = []
        for v in (a0, ax, ay):
            if torch.is_tensor(v) and v.device == self.device and v.dtype == torch.complex128:
                t = v.reshape(-1)  # already on the device: no round trip through the host
            else:
                t = torch.as_tensor(np.asarray(v.cpu() if torch.is_tensor(v) else v), dtype=torch.complex128)
            if t.numel() != ncut:
                raise ValueError(f"M3D vector has {t.numel()} values, the pupil has {ncut} points")
            out.append(t.to(self.device).contiguous())
        return out

    def linear(self, bits, a0, ax, ay, z=0.0, mesh_deg=0.2):
        """(intlinear, intft) images (linear.py:190-225): Abbe imaging of the
        M3D-linear model and of the thin mask on the source mesh."""
        cfg = self.cfg
        ampvc, ampab, ph0 = self.reflect_amplitude()
        lpx, lpy = cfg.lpmax
        fmask = self.mask_spectrum(bits, lpx, lpy)[0]
        a0, ax, ay = self._m3d_inputs(a0, ax, ay)
        lin, min_, _ = cfg.pupil
        pu, pup = _lib.arr_i(np.stack([lin, min_], axis=1))
        dsx, dsy = cfg.mesh_source(mesh_deg)
        sr, srp = _lib.arr_d(np.stack([dsx, dsy], axis=1))
        nx, ny = cfg.image_shape
        il = torch.empty((nx, ny), dtype=torch.float64, device=self.device)
        it = torch.empty_like(il)
        cv = lambda z_: (ctypes.c_double * 2)(z_.real, z_.imag)
        self._bind_stream()
        _lib.check(self.lib.euvl_linear_images(self.h, ctypes.byref(self.params), _lib.ptr(fmask), lpx, lpy, cv(ampvc),
                                               cv(ampab), cv(ph0), _lib.ptr(a0), _lib.ptr(ax), _lib.ptr(ay), len(lin),
                                               pup, len(dsx), srp, cfg.kx0, cfg.ky0, cfg.z0, float(z),
                                               float(cfg.obscuration), nx, ny, _lib.ptr(il), _lib.ptr(it)),
                   "euvl_linear_images")
        return il, it

    def eigenfunction(self, z=0.0, nk=100, mesh_deg=0.2):
        """STCC kernels (eigenfunction.py:5-179) kept on the device:
        (alpha [3, nk], phi [3, nk, ncut]); cached per (z, nk, mesh).
        last_sweeps: the Jacobi sweeps of the full decomposition, or minus the
        subspace iterations when the leading pairs came from those (nk <= 24)."""
        key = ("kernels", self.cfg, float(z), int(nk), float(mesh_deg))
        got = self.resident.get(key)
        if got is not None:
            return got
        cfg = self.cfg
        lin, min_, _ = cfg.pupil
        nk = min(nk, len(lin))
        pu, pup = _lib.arr_i(np.stack([lin, min_], axis=1))
        dsx, dsy = cfg.mesh_source(mesh_deg)
        sr, srp = _lib.arr_d(np.stack([dsx, dsy], axis=1))
        alpha = torch.empty((3, nk), dtype=torch.float64, device=self.device)
        phi = torch.empty((3, nk, len(lin)), dtype=torch.complex128, device=self.device)
        sweeps = ctypes.c_int()
        self._bind_stream()
        _lib.check(self.lib.euvl_tcc_kernels(self.h, ctypes.byref(self.params), len(lin), pup, len(dsx), srp, cfg.kx0,
                                             cfg.ky0, cfg.z0, float(z), nk, _lib.ptr(alpha), _lib.ptr(phi),
                                             ctypes.byref(sweeps)), "euvl_tcc_kernels")
        self.last_sweeps = sweeps.value
        torch.cuda.current_stream(self.device).synchronize()  # other streams may read them next
        return self.resident.put(key, (alpha, phi))

    def socs(self, bits, a0, ax, ay, nsocs=100, nsocsxy=20, nout=64, z=0.0, mesh_deg=0.2):
        """STCC-SOCS image (socs.py:156-204) on the config's image grid."""
        cfg = self.cfg
        alpha, phi = self.eigenfunction(z, max(nsocs, 1), mesh_deg)
        nk = alpha.shape[1]
        nsocs = min(nsocs, nk)
        nsocsxy = min(nsocsxy, nsocs)  # the xy terms ride inside the first nsocs kernels (socs.py:100-140)
        ampvc, ampab, ph0 = self.reflect_amplitude()
        lpx, lpy = cfg.lpmax
        fmask = self.mask_spectrum(bits, lpx, lpy)[0]
        a0, ax, ay = self._m3d_inputs(a0, ax, ay)
        lin, min_, _ = cfg.pupil
        pu, pup = _lib.arr_i(np.stack([lin, min_], axis=1))
        nx, ny = cfg.image_shape
        img = torch.empty((nx, ny), dtype=torch.float64, device=self.device)
        cv = lambda z_: (ctypes.c_double * 2)(z_.real, z_.imag)
        self._bind_stream()
        _lib.check(self.lib.euvl_socs_image(self.h, ctypes.byref(self.params), _lib.ptr(fmask), lpx, lpy, cv(ampvc),
                                            cv(ampab), cv(ph0), _lib.ptr(a0), _lib.ptr(ax), _lib.ptr(ay), len(lin), pup,
                                            _lib.ptr(alpha), _lib.ptr(phi), nk, nsocs, nsocsxy, nout, cfg.kx0, cfg.ky0,
                                            cfg.z0, nx, ny, _lib.ptr(img)), "euvl_socs_image")
        return img

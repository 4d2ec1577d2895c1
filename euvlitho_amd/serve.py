"""Served form: a continuous-batching loop over a stream of clip requests.

Each request names a problem (Config: pitch, order truncation, source) and a
mask.  The server owns S slots; a slot is a HIP stream with its own libeuvl
handle (device workspace) and per-configuration engines bound to that
handle.  Worker threads, one per slot, take requests as they arrive: a
request is admitted when a slot is idle and its device-memory footprint fits
the HBM budget (the slot's workspace may have to grow), so requests that
arrive while others run start at once on another stream and their kernels
overlap on the device.  Requests that fit may pass an older one that does
not (admission by footprint), but the oldest waiting request that fits is
always taken first.

Mask-independent state is resident and shared by every slot (one Resident
store): the blank-mask amplitudes per configuration and the SOCS kernels.
Per request the server computes the M3D parameters, the rigorous Abbe image
and the SOCS image (the per-clip work of the FP64 drop-in, in the precision
mode chosen for the server).

Results are independent of the mix: every request runs the same kernels on
its own stream and workspace, and nothing a request reads is written by
another, so a clip's images are identical whether it runs alone or among
others.
"""
import ctypes
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .config import Config, workload_mix
from .engine import Engine, Resident

MIB = 1 << 20


@dataclass
class ClipRequest:
    rid: int
    cfg: Config
    mask: np.ndarray          # 0/1 per pixel (index ndivx*j + i), or packed mask.bin bytes
    arrival: float = 0.0      # seconds after the start of the stream


@dataclass
class ClipResult:
    rid: int
    slot: int
    arrival: float
    start: float
    finish: float
    a0: np.ndarray = field(repr=False)
    image: np.ndarray = field(repr=False)
    socs: np.ndarray = field(repr=False)

    @property
    def latency(self):
        return self.finish - self.arrival


def served_configs():
    """The problem mix of the served stream (config.workload_mix): pitch
    256-1024 nm with order truncation +-8..+-20 growing with the pitch,
    annular or dipole+annular sources of 61-200 points."""
    return workload_mix()


def footprint(cfg: Config):
    """Device bytes one request of this configuration needs beyond the
    resident state: the EM workspace (13 N x N complex matrices per offset
    plus vectors and the incidence columns) or the image workspace, whichever
    is larger, plus the amplitude and image tensors."""
    lidx, _, _, _ = cfg.order_set()
    N, B = len(lidx), len(cfg.offsets())
    lsx, lsy = cfg.lsmax
    ls, ms = np.meshgrid(np.arange(-lsx, lsx + 1), np.arange(-lsy, lsy + 1), indexing="ij")
    ninc = int(cfg.in_na(ls, ms).sum())
    mat, vec = 16 * N * N * B, 16 * N * B
    em = 13 * mat + 8 * vec + 2 * 16 * N * ninc * B + 4 * N * B + 2 * MIB
    imaging = 272 * MIB
    nx, ny = cfg.image_shape
    tensors = 16 * B * (2 * lsx + 1) * (2 * lsy + 1) * N + 3 * 8 * nx * ny
    total = max(em, imaging) + tensors
    return -(-total // MIB) * MIB  # whole MiB (the library rounds reservations up to 256 bytes)


class _Slot:
    def __init__(self, idx, device, precision, resident):
        self.idx = idx
        self.device = device
        self.lib = _lib.load()
        with torch.cuda.device(device):
            self.stream = torch.cuda.Stream(device)  # default priority
            h = ctypes.c_void_p()
            _lib.check(self.lib.euvl_create(ctypes.byref(h)), "euvl_create")
        self.h = h
        self.precision = precision
        self.resident = resident
        self.engines = {}
        self.reserved = 0

    def engine(self, cfg):
        eng = self.engines.get(cfg)
        if eng is None:
            eng = Engine(cfg, self.device, self.precision, handle=self.h, resident=self.resident)
            self.engines[cfg] = eng
        return eng

    def workspace(self):
        return int(self.lib.euvl_workspace_bytes(self.h))

    def close(self):
        for e in self.engines.values():
            e.close()
        if self.h:
            self.lib.euvl_destroy(self.h)
            self.h = None


class Server:
    def __init__(self, n_slots=4, precision=_lib.PREC_MIXED, hbm_budget=None, device="cuda:0",
                 socs_kernels=24, socs_xy=20):
        if not torch.cuda.is_available():
            raise _lib.EuvlError("a ROCm GPU is required")
        self.device = torch.device(device)
        if hbm_budget is None:
            free, total = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.5 * free)
        self.budget = int(hbm_budget)
        self.resident = Resident()
        self.slots = [_Slot(i, self.device, precision, self.resident) for i in range(n_slots)]
        self.precision = precision
        self.socs_kernels, self.socs_xy = socs_kernels, socs_xy
        self.peak_reserved = 0

    # ------------------------------------------------------------ residency
    def warm(self, cfgs):
        """Build the resident state of every configuration (blank amplitudes,
        SOCS kernels) on slot 0, and grow each slot's workspace to the largest
        footprint that the budget allows, so that serving does not allocate."""
        s0 = self.slots[0]
        with torch.cuda.stream(s0.stream):
            for cfg in cfgs:
                eng = s0.engine(cfg)
                eng.blank_amplitudes()
                eng.eigenfunction(0.0, self.socs_kernels)
        s0.stream.synchronize()
        need = max(footprint(c) for c in cfgs)
        for s in self.slots:
            if self._reserved_total() - s.reserved + need <= self.budget:
                _lib.check(self.lib_reserve(s, need), "euvl_reserve_workspace")

    def lib_reserve(self, slot, nbytes):
        rc = slot.lib.euvl_reserve_workspace(slot.h, ctypes.c_int64(int(nbytes)))
        slot.reserved = slot.workspace()
        self.peak_reserved = max(self.peak_reserved, self._reserved_total())
        return rc

    def _reserved_total(self):
        return sum(s.reserved for s in self.slots)

    # ------------------------------------------------------------ one clip
    def process(self, slot, req):
        """M3D parameters, Abbe image and SOCS image of one request on one slot
        (runs on the slot's stream; returns host arrays)."""
        with torch.cuda.stream(slot.stream):
            eng = slot.engine(req.cfg)
            bits = eng.upload_mask(req.mask)
            Ax = eng.amp_s(bits, "X")
            a0, ax, ay = eng.m3d(bits, Ax0=Ax[0])
            img = eng.abbe(bits, Ax=Ax)
            socs = eng.socs(bits, a0, ax, ay, nsocs=self.socs_kernels, nsocsxy=self.socs_xy)
            out = (a0.cpu().numpy(), img.cpu().numpy(), socs.cpu().numpy())
        slot.stream.synchronize()
        return out

    # ------------------------------------------------------------ the stream
    def run(self, requests):
        """Serve `requests` (arrival times in seconds from now); returns the
        results in request order.  Exceptions of a worker are re-raised."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        need = {r.rid: footprint(r.cfg) for r in reqs}
        if any(n > self.budget for n in need.values()):
            raise _lib.EuvlError("a request needs more device memory than the HBM budget")
        pending = list(reqs)
        results, errors = {}, []
        cv = threading.Condition()
        t0 = time.perf_counter()
        state = {"done": 0}

        def pick(slot, now):
            # oldest arrived request whose footprint fits this slot under the budget
            for i, r in enumerate(pending):
                if r.arrival > now:
                    break
                grow = max(0, need[r.rid] - slot.reserved)
                if self._reserved_total() + grow <= self.budget:
                    return pending.pop(i), grow
            return None, 0

        def worker(slot):
            try:
                while True:
                    with cv:
                        while True:
                            if state["done"] == len(reqs) or errors:
                                return
                            now = time.perf_counter() - t0
                            req, grow = pick(slot, now)
                            if req is not None:
                                if grow:
                                    slot.reserved += grow  # claim the growth before releasing the lock
                                    self.peak_reserved = max(self.peak_reserved, self._reserved_total())
                                break
                            nxt = [r.arrival for r in pending if r.arrival > now]
                            cv.wait(timeout=min(0.05, max(1e-4, (min(nxt) - now) if nxt else 0.05)))
                    if grow:
                        _lib.check(slot.lib.euvl_reserve_workspace(slot.h, ctypes.c_int64(slot.reserved)),
                                   "euvl_reserve_workspace")
                    start = time.perf_counter() - t0
                    a0, img, socs = self.process(slot, req)
                    finish = time.perf_counter() - t0
                    with cv:
                        slot.reserved = max(slot.reserved, slot.workspace())
                        results[req.rid] = ClipResult(req.rid, slot.idx, req.arrival, start, finish, a0, img, socs)
                        state["done"] += 1
                        cv.notify_all()
            except BaseException as e:  # surface it to run()
                with cv:
                    errors.append(e)
                    cv.notify_all()

        threads = [threading.Thread(target=worker, args=(s,), daemon=True) for s in self.slots]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        self.wall = time.perf_counter() - t0
        return [results[r.rid] for r in requests]

    def close(self):
        for s in self.slots:
            s.close()


def poisson_stream(n, rate, seed=0, cfgs=None):
    """n seeded requests: configuration uniform over `cfgs` (default
    served_configs()), Manhattan mask of vacuum fraction U(0.3, 0.7), Poisson
    arrivals at `rate` per second."""
    from .masks import manhattan
    cfgs = cfgs or served_configs()
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    out = []
    for i in range(n):
        cfg = cfgs[int(rng.integers(len(cfgs)))]
        dens = float(rng.uniform(0.3, 0.7))
        out.append(ClipRequest(i, cfg, manhattan(cfg.ndivx, cfg.ndivy, dens, seed * 100003 + i), float(t[i])))
    return out

"""Served form: a continuous-batching loop over a stream of clip requests.

Each request names a problem (Config: pitch, order truncation, source) and a
mask.  The server owns S slots; a slot is a HIP stream with its own libeuvl
handle (device workspace), per-configuration engines and batched plans
(libeuvl plans for 1..B clips of each problem) bound to that handle.  Worker
threads, one per slot, take requests as they arrive: when a slot is idle it
takes the oldest arrived request whose batch fits the HBM budget, together
with the later arrived requests of the same problem (up to the problem's
batch size), and runs them as one batched plan run -- mask spectra, EM
solves batched over clips x source divisions, M3D fits and Abbe images --
then the SOCS image of each clip.  Requests that arrive while others run are
admitted as slots and memory free up, and their kernels overlap on the
device with those of the other slots.  Requests that fit may pass an older
one that does not (admission by footprint), but the oldest waiting request
that fits is always taken first.

Mask-independent state is resident and shared by every slot (one Resident
store): the blank-mask amplitudes per configuration, the SOCS kernels and
the decision of the fixed-decision EM solve (its doubling count,
calibrated once per problem at warm-up on seeded masks with a
factor-2 margin on the norm bound, and checked on the device for every clip;
a flagged batch is run again with host decisions).

Results are independent of the mix: every clip of a batch runs the same
kernels with the same decisions as when it runs alone, each batch entry is
computed by its own tiles, and nothing a request reads is written by another,
so a clip's images agree whether it runs alone or among others.
"""
import ctypes
import gc
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
    cfg: Config
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


def batch_size(cfg: Config, max_batch=4):
    """Clips of one problem batched in one plan run of the served form."""
    from .batch import clips_per_run
    return clips_per_run(cfg, max_batch)


def footprint(cfg: Config, n_clips=1):
    """Device bytes one batch of n_clips requests of this configuration needs
    beyond the resident state: the EM workspace (13 N x N complex matrices per
    clip and offset plus vectors, the incidence columns, the coupling
    product's FFT spectra and the mask spectra)
    or the image workspace, whichever is larger, plus the amplitude and image
    tensors."""
    lidx, _, L, M = cfg.order_set()
    N, B = len(lidx), len(cfg.offsets()) * n_clips
    lsx, lsy = cfg.lsmax
    ls, ms = np.meshgrid(np.arange(-lsx, lsx + 1), np.arange(-lsy, lsy + 1), indexing="ij")
    ninc = int(cfg.in_na(ls, ms).sum())
    mat, vec = 16 * N * N * B, 16 * N * B
    lpx, lpy = cfg.lpmax
    spectra = 16 * n_clips * ((4 * L + 1) * (4 * M + 1) + (2 * lpx + 1) * (2 * lpy + 1))
    em = 13 * mat + 8 * vec + 3 * 16 * N * ninc * B + 4 * N * B + 16 * 8192 * B + 2 * MIB
    imaging = 272 * MIB
    nx, ny = cfg.image_shape
    tensors = 16 * B * (2 * lsx + 1) * (2 * lsy + 1) * N + 3 * 8 * nx * ny * n_clips
    total = max(em, imaging) + spectra + tensors
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
        self.runs = {}  # (cfg, n) -> _Run of a plan for n clips
        self.graphs = {}  # (cfg, n) -> HIP graph of that run in fixed mode (euvl_graph), when captured
        self.recaptures = 0
        self.reserved = 0

    def bind(self):
        _lib.check(self.lib.euvl_set_stream(self.h, ctypes.c_void_p(self.stream.cuda_stream)), "euvl_set_stream")

    def engine(self, cfg):
        eng = self.engines.get(cfg)
        if eng is None:
            eng = Engine(cfg, self.device, self.precision, handle=self.h, resident=self.resident)
            self.engines[cfg] = eng
        return eng

    def run_for(self, cfg, n, decisions):
        """The plan run (plan + device buffers) of n clips of cfg on this slot."""
        from .batch import Plan, _Run
        r = self.runs.get((cfg, n))
        if r is None:
            with torch.cuda.stream(self.stream):
                self.bind()
                plan = Plan(self.engine(cfg), n)
                _lib.check(self.lib.euvl_plan_set_decisions(plan.pl, int(decisions)), "euvl_plan_set_decisions")
                nbytes = cfg.ndivx * cfg.ndivy // 8
                bits = torch.empty(n * nbytes, dtype=torch.uint8, device=self.device)
                flags = torch.zeros(n * plan.n_off, dtype=torch.int32, device=self.device)
                r = _Run(plan, list(range(n)), bits, self.device, flags)
            self.runs[(cfg, n)] = r
        return r

    def workspace(self):
        return int(self.lib.euvl_workspace_bytes(self.h))

    def reserve_plans(self):
        """Grow this slot's workspace and LU scratch to what every plan of the
        slot needs (graphs are captured after this, so that they stay valid)."""
        plans = [r.plan.pl for r in self.runs.values()]
        arr = (ctypes.c_void_p * len(plans))(*plans)
        with torch.cuda.stream(self.stream):
            self.bind()
            _lib.check(self.lib.euvl_plan_reserve(self.h, len(plans), arr), "euvl_plan_reserve")

    def capture(self, key):
        """Capture the fixed-decision run `key` = (cfg, n) of this slot as a HIP
        graph on the slot's stream (the workspace must be reserved for it)."""
        r = self.runs[key]
        with torch.cuda.stream(self.stream):
            self.bind()
            _lib.check(self.lib.euvl_graph_begin(self.h), "euvl_graph_begin")
            g = ctypes.c_void_p()
            try:
                r.launch(fixed=True)
            finally:
                rc = self.lib.euvl_graph_end(self.h, ctypes.byref(g))
            _lib.check(rc, "euvl_graph_end")
        self.graphs[key] = g

    def drop_graphs(self):
        for g in self.graphs.values():
            self.lib.euvl_graph_destroy(g)
        self.graphs = {}

    def launch_fixed(self, key, run):
        """Enqueue the fixed-decision run on the slot's stream: its graph when
        one was captured (captured again first if the workspace has moved since),
        else the plan's kernels one by one."""
        g = self.graphs.get(key)
        if g is None:
            run.launch(fixed=True)
            return
        self.bind()
        rc = self.lib.euvl_graph_launch(g, self.h)
        if rc == _lib.ERR_STALE:  # every graph of the slot addresses the old workspace
            self.drop_graphs()
            self.recaptures += 1
            self.reserve_plans()
            self.capture(key)
            rc = self.lib.euvl_graph_launch(self.graphs[key], self.h)
        _lib.check(rc, "euvl_graph_launch")

    def close(self):
        self.drop_graphs()
        for r in self.runs.values():
            r.plan.close()
        self.runs = {}
        for e in self.engines.values():
            e.close()
        if self.h:
            self.lib.euvl_destroy(self.h)
            self.h = None


class Server:
    def __init__(self, n_slots=4, precision=_lib.PREC_MIXED, hbm_budget=None, device="cuda:0",
                 socs_kernels=24, socs_xy=20, max_batch=4, graphs=True):
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
        self.max_batch = max_batch
        self.use_graphs = graphs
        self.peak_reserved = 0
        self.reruns = 0
        self.batches = []  # sizes of the batches run by the last run()

    # ------------------------------------------------------------ residency
    def decisions(self, cfg):
        """Doubling count of the fixed-decision EM solve of cfg: calibrated once (resident), on seeded masks of vacuum fraction
        0.3, 0.5 and 0.7 in adaptive runs, with a factor-2 margin on the norm
        bound behind the doubling count."""
        key = ("decisions", cfg, self.precision)
        got = self.resident.get(key)
        if got is not None:
            return got
        from .batch import Plan
        from .masks import manhattan, pack_bits
        s0 = self.slots[0]
        with torch.cuda.stream(s0.stream):
            s0.bind()
            plan = Plan(s0.engine(cfg), 1, alpha_margin=2.0)
            try:
                nb = cfg.ndivx * cfg.ndivy // 8
                from .batch import _Run
                r = _Run(plan, [0], torch.empty(nb, dtype=torch.uint8, device=self.device), self.device,
                         torch.zeros(plan.n_off, dtype=torch.int32, device=self.device))
                for k, dens in enumerate((0.3, 0.5, 0.7)):
                    raw = pack_bits(manhattan(cfg.ndivx, cfg.ndivy, dens, 7919 + k))
                    r.bits.copy_(torch.from_numpy(np.frombuffer(raw, dtype=np.uint8).copy()))
                    r.launch(fixed=False)
                dec = plan.decisions()
            finally:
                plan.close()
        s0.stream.synchronize()
        return self.resident.put(key, dec)

    def batch_size(self, cfg):
        return batch_size(cfg, self.max_batch)

    def warm(self, cfgs):
        """Build the resident state of every configuration (blank amplitudes,
        SOCS kernels, EM decisions) on slot 0, make every slot's plans, and
        grow each slot's workspace to the largest batch footprint that the
        budget allows, so that serving does not allocate."""
        s0 = self.slots[0]
        with torch.cuda.stream(s0.stream):
            for cfg in cfgs:
                eng = s0.engine(cfg)
                eng.blank_amplitudes()
                eng.eigenfunction(0.0, self.socs_kernels)
        s0.stream.synchronize()
        for cfg in cfgs:
            dec = self.decisions(cfg)
            for s in self.slots:
                for n in range(1, self.batch_size(cfg) + 1):
                    s.run_for(cfg, n, dec)
        need = max(footprint(c, self.batch_size(c)) for c in cfgs)
        for s in self.slots:
            if self._reserved_total() - s.reserved + need <= self.budget:
                _lib.check(self.lib_reserve(s, need), "euvl_reserve_workspace")
                if self.use_graphs:
                    # every plan of the slot fits the reservation (the footprint covers the plans), so
                    # its fixed runs are captured once here and replayed by process()
                    s.reserve_plans()
                    s.reserved = max(s.reserved, s.workspace())
                    self.peak_reserved = max(self.peak_reserved, self._reserved_total())
                    for key in s.runs:
                        if key not in s.graphs:
                            s.capture(key)
        torch.cuda.synchronize(self.device)

    def lib_reserve(self, slot, nbytes):
        rc = slot.lib.euvl_reserve_workspace(slot.h, ctypes.c_int64(int(nbytes)))
        slot.reserved = slot.workspace()
        self.peak_reserved = max(self.peak_reserved, self._reserved_total())
        return rc

    def _reserved_total(self):
        return sum(s.reserved for s in self.slots)

    # ------------------------------------------------------------ one batch
    def process(self, slot, reqs):
        """M3D parameters, Abbe images and SOCS images of a batch of requests of
        one problem on one slot (one batched plan run, then the SOCS image of
        each clip; runs on the slot's stream; returns host arrays per request)."""
        from .masks import pack_bits
        cfg = reqs[0].cfg
        run = slot.run_for(cfg, len(reqs), self.decisions(cfg))
        raw = b"".join(r.mask if isinstance(r.mask, (bytes, bytearray)) else pack_bits(r.mask) for r in reqs)
        nb = cfg.ndivx * cfg.ndivy // 8
        with torch.cuda.stream(slot.stream):
            slot.bind()
            run.bits.copy_(torch.from_numpy(np.frombuffer(raw, dtype=np.uint8).copy()))
            slot.launch_fixed((cfg, len(reqs)), run)
            if int(run.flags.max().item()) != 0:  # a device-side check failed: host decisions for this batch
                self.reruns += 1
                run.launch(fixed=False)
            eng = slot.engine(cfg)
            out = []
            for k in range(len(reqs)):
                bits = run.bits[k * nb:(k + 1) * nb]
                socs = eng.socs(bits, run.a0[k], run.ax[k], run.ay[k], nsocs=self.socs_kernels,
                                nsocsxy=self.socs_xy)
                out.append((run.a0[k].cpu().numpy(), run.img[k].cpu().numpy(), socs.cpu().numpy()))
        slot.stream.synchronize()
        return out

    # ------------------------------------------------------------ the stream
    def run(self, requests):
        """Serve `requests` (arrival times in seconds from now); returns the
        results in request order.  Exceptions of a worker are re-raised."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for c in {r.cfg for r in reqs}:
            if footprint(c) > self.budget:
                raise _lib.EuvlError("a request needs more device memory than the HBM budget")
            self.decisions(c)
        # Objects made so far (the set-up, torch's own) go to the permanent generation, so that a
        # full collection during the stream does not stall a worker for tens of milliseconds.
        gc.collect()
        gc.freeze()
        pending = list(reqs)
        results, errors = {}, []
        self.batches = []
        cv = threading.Condition()
        t0 = time.perf_counter()
        state = {"done": 0}

        def pick(slot, now):
            # the oldest arrived request whose batch fits this slot under the budget, with the
            # later arrived requests of its problem (up to the batch size), largest batch that fits;
            # pending is in arrival order, so the arrived requests are a prefix of it
            arrived = 0
            while arrived < len(pending) and pending[arrived].arrival <= now:
                arrived += 1
            for i in range(arrived):
                r = pending[i]
                same = [j for j in range(i, arrived) if pending[j].cfg == r.cfg]
                for n in range(min(len(same), self.batch_size(r.cfg)), 0, -1):
                    grow = max(0, footprint(r.cfg, n) - slot.reserved)
                    if self._reserved_total() + grow <= self.budget:
                        take = same[:n]
                        batch = [pending[j] for j in take]
                        for j in reversed(take):
                            pending.pop(j)
                        return batch, grow
            return None, 0

        def worker(slot):
            try:
                while True:
                    with cv:
                        while True:
                            if state["done"] == len(reqs) or errors:
                                return
                            now = time.perf_counter() - t0
                            batch, grow = pick(slot, now)
                            if batch is not None:
                                if grow:
                                    slot.reserved += grow  # claim the growth before releasing the lock
                                    self.peak_reserved = max(self.peak_reserved, self._reserved_total())
                                break
                            nxt = next((r.arrival for r in pending if r.arrival > now), None)
                            cv.wait(timeout=min(0.05, max(1e-4, (nxt - now) if nxt is not None else 0.05)))
                    if grow:
                        _lib.check(slot.lib.euvl_reserve_workspace(slot.h, ctypes.c_int64(slot.reserved)),
                                   "euvl_reserve_workspace")
                    start = time.perf_counter() - t0
                    outs = self.process(slot, batch)
                    finish = time.perf_counter() - t0
                    with cv:
                        slot.reserved = max(slot.reserved, slot.workspace())
                        self.batches.append(len(batch))
                        for req, (a0, img, socs) in zip(batch, outs):
                            results[req.rid] = ClipResult(req.rid, req.cfg, slot.idx, req.arrival, start, finish,
                                                          a0, img, socs)
                        state["done"] += len(batch)
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

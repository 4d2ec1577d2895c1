"""Batched form: many clips of uneven pitch and order truncation in one
captured HIP graph.

The reference runs its drivers once per mask (the ndata loop of
cnn/intensity/m3d/m3d.cpp:197-558, and intensity.cpp:35-421 per mask file).
Here the clips are grouped by problem (pitch, order set, source), each group
is cut into runs of a few clips whose EM products are batched over clips x
source divisions (libeuvl plans: euvl_plan_create / euvl_plan_run), and the
runs of every group are captured once into one HIP graph that is replayed
with one launch.  Decisions the eager path takes on the host (doubling count,
pivoted fallback of the top solve) are calibrated by an
adaptive run of the loaded clips and then checked on the device; a clip whose
check fails is run again adaptively after the graph.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from .config import Config
from .engine import Engine, Resident
from .masks import pack_bits

# GEMM tiles of 64 x 64 per batched product that fill the GPU several times over
TARGET_TILES = 1536
MAX_CLIPS_PER_RUN = 16
# flag bits of a fixed-decision run (euvl_plan_run)
FLAG_NAMES = {1: "doubling bound", 4: "top-solve residual", 8: "zero pivot"}


def clips_per_run(cfg: Config, n_clips: int):
    """Clips batched in one run of a problem: enough batch entries (clips x
    divisions) for about TARGET_TILES 64 x 64 tiles per N x N product."""
    n = len(cfg.order_set()[0])
    per_entry = math.ceil(n / 64) ** 2
    b = len(cfg.offsets())
    c = max(1, math.ceil(TARGET_TILES / (per_entry * b)))
    return int(min(c, MAX_CLIPS_PER_RUN, n_clips))


class _Run:
    """One plan run: its plan, its clips and its device buffers (inputs and
    outputs at fixed addresses, so that a captured graph can be replayed)."""

    def __init__(self, plan, clip_ids, bits, dev, flags):
        self.plan = plan
        self.clip_ids = clip_ids
        c = len(clip_ids)
        self.bits = bits
        self.Ax = torch.empty((c,) + plan.ax_shape, dtype=torch.complex128, device=dev)
        self.a0 = torch.empty((c, plan.ncut), dtype=torch.complex128, device=dev)
        self.ax = torch.empty_like(self.a0)
        self.ay = torch.empty_like(self.a0)
        self.img = torch.empty((c,) + plan.cfg.image_shape, dtype=torch.float64, device=dev)
        self.flags = flags  # a slice of the engine's flag array (c x divisions entries)

    def launch(self, fixed):
        p = self.plan
        _lib.check(p.lib.euvl_plan_run(p.pl, _lib.ptr(self.bits), _lib.ptr(self.Ax), _lib.ptr(self.a0),
                                       _lib.ptr(self.ax), _lib.ptr(self.ay), _lib.ptr(self.img),
                                       _lib.ptr(self.flags), int(fixed)), "euvl_plan_run")


class Plan:
    """A libeuvl plan of one problem for `n_clips` clips per run."""

    def __init__(self, eng: Engine, n_clips: int, polar="X", z=0.0, alpha_margin=1.0):
        cfg = eng.cfg
        self.cfg, self.lib, self.n_clips = cfg, eng.lib, n_clips
        vcxx, abxx = eng.blank_amplitudes(polar)
        lidx, midx, L, M = eng.lidx, eng.midx, eng.L, eng.M
        offs = cfg.offsets()
        self.n_off = len(offs)
        src = []
        for o, (_, _, l0, m0) in enumerate(cfg.source_points()):
            src += [(o, int(a), int(b)) for a, b in zip(l0, m0)]
        lin, min_, snum = cfg.pupil
        self.ncut = len(lin)
        nsx, nsy = cfg.nsource
        self.ax_shape = (self.n_off, nsx, nsy, len(lidx))
        lpx, lpy = cfg.lpmax
        nx, ny = cfg.image_shape
        self._keep = [_lib.arr_i(lidx), _lib.arr_i(midx), _lib.arr_d(offs),
                      _lib.arr_i(np.stack([lin, min_, (snum >= 8).astype(np.int32)], axis=1)),
                      _lib.arr_d(np.stack([vcxx.real, vcxx.imag], -1)), _lib.arr_d(np.stack([abxx.real, abxx.imag], -1)),
                      _lib.arr_i(np.array(src, dtype=np.int32).reshape(-1, 3))]
        (_, li), (_, mi), (_, of), (_, pu), (_, vc), (_, ab), (_, sr) = self._keep
        d = _lib.PlanDesc(polar.encode(), ctypes.pointer(eng.params), li, mi, len(lidx), L, M, self.n_off, of,
                          n_clips, cfg.kx0, cfg.ky0, cfg.z0, self.ncut, pu, lpx, lpy, vc, ab, len(src), sr, float(z),
                          nx, ny, 1.0 / max(len(src), 1), float(alpha_margin))
        self.pl = ctypes.c_void_p()
        _lib.check(self.lib.euvl_plan_create(eng.h, ctypes.byref(d), ctypes.byref(self.pl)), "euvl_plan_create")

    def decisions(self):
        """The plan's doubling count of the matrix functions (fixed-mode runs)."""
        d = ctypes.c_int()
        _lib.check(self.lib.euvl_plan_decisions(self.pl, ctypes.byref(d)), "euvl_plan_decisions")
        return d.value

    def close(self):
        if self.pl:
            self.lib.euvl_plan_destroy(self.pl)
        self.pl = None


class BatchedEngine:
    """Runs a list of (config, mask) clips as batched plan runs inside one
    captured HIP graph on a stream and handle of its own.

    load(clips) makes the plans, uploads the masks (they stay resident in
    HBM), runs every clip once adaptively (each plan keeps the largest
    doubling count its clips needed) and captures the graph;
    launch() replays the graph; results() synchronises, reruns flagged clips
    adaptively and returns per-clip (a0, ax, ay, image) as device tensors."""

    def __init__(self, device="cuda:0", precision=_lib.PREC_FP64, graph=True, resident=None):
        if not torch.cuda.is_available():
            raise _lib.EuvlError("a ROCm GPU is required (torch.cuda.is_available() is False)")
        self.device = torch.device(device)
        self.lib = _lib.load()
        self.precision = int(precision)
        self.use_graph = graph
        self.stream = torch.cuda.Stream(self.device)  # default priority; never the null stream
        self.h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.euvl_create(ctypes.byref(self.h)), "euvl_create")
        _lib.check(self.lib.euvl_set_stream(self.h, ctypes.c_void_p(self.stream.cuda_stream)), "euvl_set_stream")
        _lib.check(self.lib.euvl_set_precision(self.h, self.precision), "euvl_set_precision")
        self.resident = resident if resident is not None else Resident()
        self.engines, self.plans, self.runs = {}, [], []
        self.graph = None
        self.n_clips = 0
        self.reruns = 0

    def _engine(self, cfg):
        if cfg not in self.engines:
            self.engines[cfg] = Engine(cfg, self.device, precision=self.precision, handle=self.h,
                                       resident=self.resident)
        return self.engines[cfg]

    def load(self, clips):
        """clips: list of (Config, mask) with mask a 0/1 array (ndivx*ndivy,
        index ndivx*j + i) or packed mask.bin bytes."""
        self.close_graph()
        self.n_clips = len(clips)
        groups = {}
        for i, (cfg, _) in enumerate(clips):
            groups.setdefault(cfg, []).append(i)
        self._bind()
        with torch.cuda.stream(self.stream):
            for cfg in groups:  # blank amplitudes first: eager solves that may grow the workspace
                self._engine(cfg).blank_amplitudes()
            plans = {}
            for cfg, ids in groups.items():
                c = clips_per_run(cfg, len(ids))
                for s in range(0, len(ids), c):
                    part = ids[s:s + c]
                    key = (cfg, len(part))
                    if key not in plans:
                        plans[key] = Plan(self._engine(cfg), len(part))
                        self.plans.append(plans[key])
                    raw = b"".join(m if isinstance(m, (bytes, bytearray)) else pack_bits(m)
                                   for m in (clips[i][1] for i in part))
                    bits = torch.from_numpy(np.frombuffer(raw, dtype=np.uint8).copy()).to(self.device)
                    self.runs.append((plans[key], part, bits))
            n_flags = sum(len(part) * p.n_off for p, part, _ in self.runs)
            self.all_flags = torch.zeros(n_flags, dtype=torch.int32, device=self.device)
            runs, at = [], 0
            for p, part, bits in self.runs:
                k = len(part) * p.n_off
                runs.append(_Run(p, part, bits, self.device, self.all_flags[at:at + k]))
                at += k
            self.runs = runs
            self._bind()
            arr = (ctypes.c_void_p * len(self.plans))(*[p.pl for p in self.plans])
            _lib.check(self.lib.euvl_plan_reserve(self.h, len(self.plans), arr), "euvl_plan_reserve")
            for r in self.runs:  # every run once adaptive: each plan keeps the largest decisions
                r.launch(fixed=False)
            if self.use_graph:
                _lib.check(self.lib.euvl_graph_begin(self.h), "euvl_graph_begin")
                try:
                    for r in self.runs:
                        r.launch(fixed=True)
                finally:
                    g = ctypes.c_void_p()
                    rc = self.lib.euvl_graph_end(self.h, ctypes.byref(g))
                _lib.check(rc, "euvl_graph_end")
                self.graph = g
        self.stream.synchronize()

    def _bind(self):
        _lib.check(self.lib.euvl_set_stream(self.h, ctypes.c_void_p(self.stream.cuda_stream)), "euvl_set_stream")

    def graph_nodes(self):
        n = ctypes.c_int64()
        _lib.check(self.lib.euvl_graph_nodes(self.graph, ctypes.byref(n)), "euvl_graph_nodes")
        return n.value

    def launch(self):
        """Enqueue one pass over every loaded clip (one graph launch)."""
        self._bind()
        if self.use_graph:
            _lib.check(self.lib.euvl_graph_launch(self.graph, self.h), "euvl_graph_launch")
        else:
            with torch.cuda.stream(self.stream):
                for r in self.runs:
                    r.launch(fixed=True)

    def results(self):
        """Synchronise, rerun flagged runs adaptively, and return per clip
        (a0, ax, ay, image) views of the device outputs.  One read-back of the
        flag array when nothing is flagged."""
        self.stream.synchronize()
        out = [None] * self.n_clips
        any_flag = int(self.all_flags.max().item()) != 0
        for r in self.runs:
            if any_flag and int(r.flags.max().item()) != 0:
                self.reruns += 1
                self._bind()
                with torch.cuda.stream(self.stream):
                    r.launch(fixed=False)
                self.stream.synchronize()
            for k, i in enumerate(r.clip_ids):
                out[i] = (r.a0[k], r.ax[k], r.ay[k], r.img[k])
        return out

    def flags(self):
        """Per run: the OR of its entries' flag bits after the last launch."""
        self.stream.synchronize()
        return [int(np.bitwise_or.reduce(r.flags.cpu().numpy())) for r in self.runs]

    def close_graph(self):
        if self.graph is not None:
            self.lib.euvl_graph_destroy(self.graph)
        self.graph = None
        for p in self.plans:
            p.close()
        self.plans, self.runs = [], []

    def close(self):
        self.stream.synchronize()
        self.close_graph()
        for e in self.engines.values():
            e.close()
        self.engines = {}
        if self.h:
            self.lib.euvl_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Benchmark: EM + Abbe mask clips/s on one MI355X.

Workload (one step = one clip, FP64): a seeded 1024 x 1024 nm Manhattan clip
(1 nm pixels, rectangles 16-64 nm, vacuum fraction 0.5), NA 0.55, 4x/8x,
CRAO 5.355 deg, 60 nm absorber n = 0.947+0.031i on 40 x Mo/Si 2.8/4.2 nm with
a 2.5 nm Ru cap, orders +-20 x +-20 (N = 1681 modes), X-dipole (sigma 0.5-0.9,
70 deg) + annular (sigma 0.7-0.8) source: 200 points of the 2 x 2 lattice
divisions (4 EM solves per clip), outputs: M3D parameters, the
rigorous Abbe image (256 x 256) and the STCC-SOCS image from the M3D
parameters (24 kernels, 20 of them with the x/y slope terms, nout 64).  The
SOCS kernels (TCC eigendecomposition on the 0.2 degree source mesh) are
mask-independent: built once, resident on the device, timed separately.

Prints one JSON line: metric, value = clips/s of the plain FP64 form,
roofline of the dominant kernel (the FP64 MFMA complex GEMM), cpu_baseline =
the FP64 NumPy/LAPACK oracle on clips of the same workload, and under
"forms" the precision form timed on the same clips with its error against
the FP64 form, the batched and served forms over the uneven-clip mix, and
the reference-default m3d.cpp path (one mask: the N = 8705 solve, the two
N = 1009 blank solves and the fit).

Roofline accounting of the EM solve: the algorithmic work is the
reference algorithm's 164 N^3 flops per (clip, polarisation, offset) solve
(5 ZGEMM + 3 inverses + one ZGEEV with right vectors at 100 N^3); achieved =
that work over the summed duration of the GEMM launches of one clip (HIP
events on the library's stream).  The GEMM kernel's own executed rate
(8mnk per complex product) is reported beside it.
"""
import argparse
import gc
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
EM_FLOPS_PER_N3 = 164.0  # reference algorithm per (clip, polarisation, offset) solve
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix
SOCS_KERNELS, SOCS_XY = 24, 20


def copy_calibration(mib=512, reps=3):
    """Plain device-to-device copies of a known size: the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    x = torch.ones(mib << 17, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": mib << 20, "reps": reps}), flush=True)


def run_served(n_req, slots, precision=None, loads=(0.8,), graphs=True):
    """Served form: requests of six problems (pitch 256-1024 nm, orders +-8..+-20,
    61-200 source points), mixed precision, `slots` streams with admission by
    device-memory footprint; arrived requests of one problem share a batched
    plan run.  A burst of 24 requests measures the capacity; the Poisson
    stream is then offered at loads[0] (80 %) of it, and at each further load
    (reported under "at_load", same seed) for the latency-throughput curve."""
    from euvlitho_amd import _lib
    from euvlitho_amd.serve import Server, poisson_stream, served_configs
    cfgs = served_configs()
    prec = _lib.PREC_MIXED if precision is None else precision
    srv = Server(n_slots=slots, precision=prec, graphs=graphs)
    try:
        t = time.perf_counter()
        srv.warm(cfgs)
        warm_s = time.perf_counter() - t
        burst = srv.run(poisson_stream(24, 0.0, seed=1, cfgs=cfgs))
        cap = len(burst) / srv.wall
        curve = {}
        for load in loads[1:]:
            r2 = srv.run(poisson_stream(n_req, load * cap, seed=2, cfgs=cfgs))
            l2 = np.array([r.latency for r in r2])
            curve["%.2f" % load] = {"value": len(r2) / srv.wall, "offered_rate": load * cap,
                                    "p50_ms": 1000.0 * float(np.percentile(l2, 50)),
                                    "p99_ms": 1000.0 * float(np.percentile(l2, 99)),
                                    "p99_over_p50": float(np.percentile(l2, 99) / np.percentile(l2, 50)),
                                    "mean_batch": float(np.mean(srv.batches)) if srv.batches else 0.0}
        rate = loads[0] * cap
        res = srv.run(poisson_stream(n_req, rate, seed=2, cfgs=cfgs))
        lat = np.array([r.latency for r in res])
        svc = np.array([r.finish - r.start for r in res])
        wait = np.array([r.start - r.arrival for r in res])
        p_wait, w99 = erlang_c_wait(slots, cap, rate)
        return {"value": len(res) / srv.wall, "unit": "clips/s", "requests": len(res),
                "offered_rate": rate, "burst_capacity": cap, "slots": slots,
                "p50_ms": 1000.0 * float(np.percentile(lat, 50)), "p99_ms": 1000.0 * float(np.percentile(lat, 99)),
                "service_p50_ms": 1000.0 * float(np.percentile(svc, 50)),
                "wait_p99_ms": 1000.0 * float(np.percentile(wait, 99)),
                "queue_model": {"model": "M/M/c (Erlang C), c = slots, service rate burst_capacity / c, "
                                         "arrivals at offered_rate",
                                "p_wait": p_wait, "wait_p99_ms": 1000.0 * w99},
                "dtype": ("f32 MFMA for the coupling / top-solve products and the imaging; f64 matrix functions, "
                          "multilayer and storage" if prec == _lib.PREC_MIXED else
                          "f32 MFMA for every EM product and the imaging; f64 multilayer and storage"),
                "per_problem_ms": {name: [round(1000.0 * float(np.percentile(l, 50)), 1),
                                          round(1000.0 * float(np.percentile(l, 99)), 1)]
                                   for name, l in per_problem(res, cfgs).items()},
                "resident_build_s": warm_s, "peak_reserved_gib": srv.peak_reserved / 2**30,
                "batches": len(srv.batches), "mean_batch": float(np.mean(srv.batches)) if srv.batches else 0.0,
                "adaptive_reruns": srv.reruns, "p99_over_p50": float(np.percentile(lat, 99) / np.percentile(lat, 50)),
                "graphs_captured": sum(len(s.graphs) for s in srv.slots),
                "graph_recaptures": sum(s.recaptures for s in srv.slots),
                "load": loads[0], "at_load": curve,
                "problems": ["%dnm K%d %s %dpts" % (c.ndivx, c.orders[1], "ann" if c.source.kind == 1 else "dip+ann",
                                                       sum(len(x[2]) for x in c.source_points())) for c in cfgs]}
    finally:
        srv.close()


def erlang_c_wait(slots, capacity, rate, q=0.99):
    """Waiting time of an M/M/c queue (Erlang C) with c = slots servers of rate
    capacity / c each and Poisson arrivals at `rate`: the probability that a
    request waits, and the q-quantile of the wait (seconds), which is
    ln(P_wait / (1 - q)) / (c mu - lambda) when P_wait > 1 - q, else 0.  It is
    the tail that first-come-first-served admission on `slots` streams cannot
    avoid at that load."""
    c, mu = int(slots), capacity / slots
    a = rate / mu
    if a >= c:
        return 1.0, float("inf")
    terms = [a ** k / math.factorial(k) for k in range(c)]
    top = a ** c / math.factorial(c) * c / (c - a)
    p_wait = top / (sum(terms) + top)
    tq = math.log(p_wait / (1.0 - q)) / (c * mu - rate) if p_wait > 1.0 - q else 0.0
    return p_wait, tq


def per_problem(res, cfgs):
    """Latencies (s) of the served results grouped by problem."""
    from euvlitho_amd.serve import poisson_stream  # noqa: F401
    out = {}
    for r in res:
        c = r.cfg
        name = "%dnm K%d %dpts" % (c.ndivx, c.orders[1], sum(len(x[2]) for x in c.source_points()))
        out.setdefault(name, []).append(r.latency)
    return out


def measured_traffic(form="fp64"):
    """HBM bytes per GEMM launch of this workload (form "fp64", "mixed" or "fast"),
    measured with rocprofv3 PMC counters (FETCH_SIZE, WRITE_SIZE, scaled by a
    plain copy of known size) and kept as a summary under profiles/; empty
    when absent."""
    path = os.path.join(ROOT, "profiles", f"zgemm_mfma_c2_{form}_traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def workload_config():
    from euvlitho_amd.config import preset
    return preset("c2")


def cpu_clip(cfg, m):
    """Oracle (NumPy + LAPACK, FP64) on one whole clip: the EM solves of every
    source division, the Abbe accumulation over every source point and the
    M3D fit; returns (seconds, (em, abbe, fit) seconds).  The blank solves are
    done before, untimed, as the GPU path keeps them resident; the SOCS image
    (24 + 5 x 20 small IFFTs, well under 1 % of the clip on the CPU) is left
    out."""
    from oracle import abbe as OA, em as OE, geometry as G, problems as OP
    p = OP.hna(cfg.ndivx, cfg.orders[1], None)
    d = G.derived(p)
    li, mi, L, M = cfg.order_set()
    li, mi = li.astype(np.int64), mi.astype(np.int64)
    off = cfg.offsets()
    divs = cfg.source_points()
    nx, ny = cfg.image_shape
    key = (cfg.ndivx, cfg.orders)
    if key not in _BLANK:
        _BLANK[key] = OA.blank_amplitudes(p, d)
    vc, ab = _BLANK[key]
    t0 = time.perf_counter()
    amps = [OE.amp_s("X", m, p, d, li, mi, L, M, o[0], o[1]) for o in off]
    t1 = time.perf_counter()
    OA.abbe_image(p, d, m, li, mi, L, M, [(tuple(o), l0, m0) for o, (_, _, l0, m0) in zip(off, divs)], nx, ny,
                  amps=amps)
    t2 = time.perf_counter()
    OA.m3d_params(p, d, m, li, mi, amps[0], vc, ab)
    t3 = time.perf_counter()
    return t3 - t0, (t1 - t0, t2 - t1, t3 - t2)


_BLANK = {}


def cpu_baseline(cfg, masks, threads, reps=3):
    """The oracle on one whole clip of the workload, and on `reps` clips
    (median) of the 512 nm / +-10-order / 61-point problem beside it."""
    from euvlitho_amd.config import preset
    from euvlitho_amd.masks import manhattan
    t, (em, ab, fit) = cpu_clip(cfg, masks[0])
    c1 = preset("c1")
    t1 = [cpu_clip(c1, manhattan(c1.ndivx, c1.ndivy, 0.5, s))[0] for s in range(reps)]
    n_src = sum(len(x[2]) for x in cfg.source_points())
    return dict(value=1.0 / t, unit="clips/s", cores=threads, kind="port",
                sample=f"one whole clip (seed 0), {t:.1f} s: oracle EM solves of {len(cfg.offsets())} divisions "
                       f"({em:.1f} s, N={len(cfg.order_set()[0])}), Abbe over {n_src} source points ({ab:.1f} s), "
                       f"M3D fit ({fit:.1f} s); LAPACK threads = cores",
                c1={"value": 1.0 / float(np.median(t1)), "unit": "clips/s", "runs_s": t1,
                    "sample": f"median of {reps} clips (seeds 0-{reps - 1}): 512 nm, +-10 orders (N=441), 61 annular "
                              "points, 4 divisions, Abbe 256x256"})


def batched_clips(n_clips=64, seed=0):
    """The uneven-clip workload: n_clips seeded Manhattan clips, each of a
    problem drawn from the mix (pitch 256-1024 nm, orders +-8..+-20, 61-200
    source points) with a vacuum fraction drawn from U(0.3, 0.7)."""
    from euvlitho_amd.config import workload_mix
    from euvlitho_amd.masks import manhattan
    cfgs = workload_mix()
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_clips):
        cfg = cfgs[int(rng.integers(len(cfgs)))]
        out.append((cfg, manhattan(cfg.ndivx, cfg.ndivy, float(rng.uniform(0.3, 0.7)), int(rng.integers(1 << 30)))))
    return out


def run_batched(n_clips=64, reps=5, precision=None, with_eager=True):
    """Batched form: the uneven-clip workload in batched plan runs captured
    into one HIP graph (one launch per pass over all clips), FP64 unless a
    precision mask is given.  Roofline: 164 N^3 per (clip, division) solve over
    the pass's wall time.  Beside it: the same clips through the eager
    per-clip path (Engine, host decisions), and the largest difference
    between the two (M3D parameters of max|a0|, images of peak)."""
    import torch
    from euvlitho_amd import _lib
    from euvlitho_amd.batch import BatchedEngine
    from euvlitho_amd.engine import Engine
    prec = _lib.PREC_FP64 if precision is None else precision
    clips = batched_clips(n_clips)
    be = BatchedEngine(precision=prec)
    t0 = time.perf_counter()
    be.load(clips)
    setup_s = time.perf_counter() - t0
    be.launch()
    out = be.results()
    gc.collect()
    gc.freeze()  # set-up objects out of the collector's full passes (a ~40 ms host stall otherwise)
    walls = []
    for _ in range(reps):
        t0 = time.perf_counter()
        be.launch()
        be.results()  # synchronises, reads the flags once, reruns flagged clips
        walls.append(time.perf_counter() - t0)
    out = [[t.cpu().numpy() for t in r] for r in be.results()]
    flops = sum(EM_FLOPS_PER_N3 * float(len(c.order_set()[0])) ** 3 * len(c.offsets()) for c, _ in clips)
    wall = float(np.median(walls))
    res = {"value": n_clips / wall, "unit": "clips/s", "clips": n_clips, "ms_per_pass": 1000 * wall,
           "p50_pass_ms": 1000 * float(np.percentile(walls, 50)), "p99_pass_ms": 1000 * float(np.percentile(walls, 99)),
           "runs": len(be.runs), "plans": len(be.plans), "graph_nodes": be.graph_nodes(), "graph_launches_per_pass": 1,
           "adaptive_reruns": be.reruns, "setup_s": setup_s,
           "decisions": {f"{p.cfg.ndivx}nm K{p.cfg.orders[1]} x{p.n_clips}": p.decisions() for p in be.plans},
           "problems": {f"{c.ndivx}nm K{c.orders[1]} {sum(len(x[2]) for x in c.source_points())}pts "
                        f"{len(c.offsets())}div": sum(1 for cc, _ in clips if cc is c) for c in dict.fromkeys(
                            cc for cc, _ in clips)},
           "roofline": {"bound": "mfma", "achieved": flops / wall / 1e12, "peak": FP64_PEAK_TFLOPS if not prec
                        else FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": flops / wall / 1e12 / (FP64_PEAK_TFLOPS if not prec else FP32_PEAK_TFLOPS),
                        "flops_counted": "164 N^3 per (clip, division) solve over the whole pass (wall)"}}
    be.close()
    if with_eager:
        engs = {}
        for c, _ in clips:
            if c not in engs:
                engs[c] = Engine(c, precision=prec)
                engs[c].blank_amplitudes()
        bits = [engs[c].upload_mask(m) for c, m in clips]
        torch.cuda.synchronize()

        def eager():
            r = []
            for (c, _), b in zip(clips, bits):
                e = engs[c]
                Ax = e.amp_s(b, "X")
                a0, ax, ay = e.m3d(b, Ax0=Ax[0])
                r.append((a0, ax, ay, e.abbe(b, Ax=Ax)))
            torch.cuda.synchronize()
            return r

        eager()
        t0 = time.perf_counter()
        ref = eager()
        te = time.perf_counter() - t0
        em, ei = 0.0, 0.0
        for g, r in zip(out, ref):
            r = [t.cpu().numpy() for t in r]
            s = np.abs(r[0]).max()
            em = max(em, max(float(np.abs(g[i] - r[i]).max() / s) for i in range(3)))
            ei = max(ei, float(np.abs(g[3] - r[3]).max() / r[3].max()))
        res["eager_per_clip"] = {"value": n_clips / te, "unit": "clips/s", "ms_per_pass": 1000 * te}
        res["err_vs_eager"] = {"m3d_of_max_a0": em, "image_of_peak": ei}
        for e in engs.values():
            e.close()
    return res


def ref_default_leg(reps=3):
    """The reference's own m3d.cpp path on its mask (tests/golden/ref_mask.bin,
    2048 nm, NA 0.33, 4x/4x): the two blank solves on the narrow ellipse set
    (N = 1009), the wide solve on the diamond set (N = 8705) and the M3D fit,
    per mask, FP64 -- the run that the reference's README quotes at about
    one minute per mask.  Roofline: 164 N^3 per solve over the GEMM time of
    the run (HIP events), and over its wall time."""
    import ctypes
    import torch
    from euvlitho_amd import _lib
    from euvlitho_amd.config import reference_default
    from euvlitho_amd.engine import Engine, Resident
    cfg = reference_default()
    eng = Engine(cfg)
    with open(os.path.join(ROOT, "tests", "golden", "ref_mask.bin"), "rb") as f:
        bits = eng.upload_mask(f.read())

    def run():
        eng.resident = Resident()  # the reference solves the blank masks for every mask
        Ax = eng.amp_s(bits, "X")
        return eng.m3d(bits, Ax0=Ax[0])

    run()
    torch.cuda.synchronize()
    walls = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
    stats = (ctypes.c_double * 8)()
    _lib.check(eng.lib.euvl_gemm_stats(eng.h, 1, None), "euvl_gemm_stats")
    run()
    _lib.check(eng.lib.euvl_gemm_stats(eng.h, 0, stats), "euvl_gemm_stats")
    N, Nb = len(eng.lidx), len(cfg.blank_order_set()[0])
    eng.close()
    wall = float(np.median(walls))
    flops = EM_FLOPS_PER_N3 * (float(N) ** 3 + 2.0 * float(Nb) ** 3)
    ach = flops / (stats[1] * 1e-3) / 1e12
    return {"value": 1.0 / wall, "unit": "masks/s", "s_per_mask": wall, "runs_s": walls,
            "config": f"reference default: 2048 nm mask (tests/golden/ref_mask.bin), N={N} wide solve + 2 blank "
                      f"solves N={Nb} + M3D fit, X polarisation, FP64",
            "published": "~1 min per mask (reference README.md:32, hardware not stated)",
            "roofline": {"bound": "mfma", "kernel": "zgemm_mfma_kernel", "achieved": ach,
                         "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
                         "flops_counted": "164 N^3 per solve (3 solves)", "gemm_ms_per_mask": stats[1],
                         "gemm_launches": int(stats[0]), "whole_run_tflops": flops / wall / 1e12}}


def mask_producer(n_masks=64, ndiv=2048, reps=5, with_cpu=True):
    """The line/space mask program (mask.cpp) on the GPU: n_masks 2048 x 2048
    clips per call (host draws + one raster launch).  Raster roofline: the
    kernel writes n_masks * ndiv^2 / 8 bytes and reads one pattern byte per
    pixel from a pool of ~(ndiv/56) patterns per clip that stays in L2, so the
    algorithmic HBM bytes are the packed output plus the per-clip column
    table and pool.  CPU baseline: the reference program built from its own
    sources (oracle/_ref/mask.out: one clip plus its maskimage.csv, as the
    program always writes both), where it has been built."""
    import ctypes
    import subprocess
    import tempfile
    import torch
    from euvlitho_amd import _lib
    from euvlitho_amd.masks import line_space_bits, line_space_plan
    lib = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(lib.euvl_create(ctypes.byref(h)), "euvl_create")
    out = line_space_bits(ndiv, ndiv, n_masks, 1, handle=h)  # warm-up (workspace)
    torch.cuda.synchronize()
    walls, kern = [], []
    ms = ctypes.c_double()
    for r in range(reps):
        t0 = time.perf_counter()
        out = line_space_bits(ndiv, ndiv, n_masks, 1 + r, handle=h)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
        _lib.check(lib.euvl_mask_raster_ms(h, ctypes.byref(ms)), "euvl_mask_raster_ms")
        kern.append(ms.value * 1e-3)
    lib.euvl_destroy(h)
    lays = line_space_plan(ndiv, ndiv, 1, 1)
    pool = sum(len(p) for p in lays[0].patterns[0])  # bytes of the clip's pattern pool
    algo = n_masks * (ndiv * ndiv // 8 + 4 * ndiv + pool)
    wall, dt = float(np.median(walls)), float(np.median(kern))
    res = {"value": n_masks / wall, "unit": "masks/s", "config": f"{n_masks} clips of {ndiv}x{ndiv} px per call "
           f"(cd 56, gap 80: the reference mask program's clip)", "ms_per_call": 1000.0 * wall,
           "kernel_ms_per_call": 1000.0 * dt,
           "roofline": {"bound": "hbm", "kernel": "mask_raster_words_kernel", "achieved": algo / dt / 1e9,
                        "peak": 8000.0, "unit": "GB/s", "frac": algo / dt / 1e9 / 8000.0,
                        "algorithmic_bytes_per_launch": algo,
                        "note": "the call's rate is set by the host draws and the layout upload"}}
    ref = os.path.join(ROOT, "oracle", "_ref", "mask.out")
    if with_cpu and os.path.exists(ref):
        with tempfile.TemporaryDirectory() as td:
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                subprocess.run([ref], cwd=td, check=True, timeout=120)
                ts.append(time.perf_counter() - t0)
        res["cpu_baseline"] = {"value": 1.0 / float(np.median(ts)), "unit": "masks/s", "cores": 1,
                               "kind": "reference", "sample": "the reference mask program (one 2048 x 2048 clip "
                               "and its maskimage.csv), median of 3 runs"}
    return res


def cnn_predictor(batch=64, reps=3, with_cpu=True):
    """The CNN M3D predictor (predict.py): six networks of the reference's
    architecture with seeded random FP32 weights (the trained checkpoints are
    not distributed), on `batch` line/space clips (2048 x 2048 px, compmask to
    512 x 512).  Conv roofline: the algorithmic conv flops (2 * H * W * cin *
    cout * 9 per block and sample) over the conv blocks' time (HIP events on
    the stream); FC roofline: fc1 + fc2 weight bytes (read once per batch)
    over the FC time.  CPU baseline: the reference's model class on torch CPU
    (FP32, all host threads) for one clip."""
    import torch
    from euvlitho_amd.cnn import CNNmodel, CNNmodels, CHANNELS, compmask
    from euvlitho_amd.masks import line_space_bits
    g = torch.Generator(device="cuda:0")
    g.manual_seed(0)

    def params(n_out):
        p = {}
        for i in range(1, 6):
            cin, cout = CHANNELS[i - 1], CHANNELS[i]
            bd = 1.0 / (cin * 9) ** 0.5
            p[f"conv{i}.weight"] = (torch.rand((cout, cin, 3, 3), generator=g, device="cuda:0") * 2 - 1) * bd
            p[f"conv{i}.bias"] = (torch.rand(cout, generator=g, device="cuda:0") * 2 - 1) * bd
            p[f"norm{i}.weight"] = torch.ones(cout, device="cuda:0")
            p[f"norm{i}.bias"] = torch.zeros(cout, device="cuda:0")
            p[f"norm{i}.running_mean"] = torch.zeros(cout, device="cuda:0")
            p[f"norm{i}.running_var"] = torch.ones(cout, device="cuda:0")
        for name, (i, o) in (("fc1", (65536, 4096)), ("fc2", (4096, n_out))):
            bd = 1.0 / i ** 0.5
            p[f"{name}.weight"] = (torch.rand((o, i), generator=g, device="cuda:0") * 2 - 1) * bd
            p[f"{name}.bias"] = (torch.rand(o, generator=g, device="cuda:0") * 2 - 1) * bd
        return p

    nets = [CNNmodel(params(1901)), CNNmodel(params(1901))] + [CNNmodels(params(1749)) for _ in range(4)]
    bits = line_space_bits(2048, 2048, batch, 1)
    torch.cuda.synchronize()

    def run(timing=None):
        x = compmask(bits)
        return [n.forward(x, timing) for n in nets]

    run()
    torch.cuda.synchronize()
    walls, conv_ms, fc_ms = [], [], []
    for _ in range(reps):
        tm = {}
        t0 = time.perf_counter()
        run(tm)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
        conv_ms.append(sum(e[0].elapsed_time(e[1]) for e in tm["events"]))
        fc_ms.append(sum(e[1].elapsed_time(e[2]) for e in tm["events"]))
    wall, tc, tf = float(np.median(walls)), float(np.median(conv_ms)), float(np.median(fc_ms))
    conv_flops, hw = 0.0, 512
    for i in range(5):
        conv_flops += 2.0 * hw * hw * CHANNELS[i] * CHANNELS[i + 1] * 9
        hw //= 2
    conv_flops *= 6 * batch
    fc_bytes = 4.0 * (2 * (65536 * 4096 + 4096 * 1901) + 4 * (65536 * 4096 + 4096 * 1749))
    res = {"value": batch / wall, "unit": "clips/s", "config": f"6 networks (re0 im0 rex imx rey imy), batch "
           f"{batch} clips of 2048x2048 px -> 512x512, FP32, seeded random weights", "ms_per_batch": 1000 * wall,
           "conv_ms": tc, "fc_ms": tf,
           "roofline": {"bound": "mfma", "kernel": "conv_block_kernel", "achieved": conv_flops / (tc * 1e-3) / 1e12,
                        "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": conv_flops / (tc * 1e-3) / 1e12 / FP32_PEAK_TFLOPS},
           "fc_roofline": {"bound": "hbm", "kernel": "fc_splitk_kernel (fc1, fc2)", "achieved": fc_bytes / (tf * 1e-3) / 1e9,
                           "peak": 8000.0, "unit": "GB/s", "frac": fc_bytes / (tf * 1e-3) / 1e9 / 8000.0}}
    del nets
    torch.cuda.empty_cache()
    if with_cpu:
        from oracle import cnn as OC
        rng = np.random.default_rng(0)
        threads = len(os.sched_getaffinity(0))
        torch.set_num_threads(threads)
        x = torch.from_numpy(rng.random((1, 1, 512, 512), dtype=np.float32))
        ts = []
        for n_out in (1901, 1749):
            m = OC.torch_model(OC.random_params(n_out, 1, nin=512, dtype=np.float32), n_out, 512, "float32")
            with torch.no_grad():
                t0 = time.perf_counter()
                m(x)
                ts.append(time.perf_counter() - t0)
            del m
        per_clip = 2 * ts[0] + 4 * ts[1]
        res["cpu_baseline"] = {"value": 1.0 / per_clip, "unit": "clips/s", "cores": threads, "kind": "port",
                               "sample": "one clip through the six networks: the reference's model class on torch "
                                         f"CPU FP32 ({per_clip:.2f} s per clip)"}
    return res


def stcc_leg(reps=5, with_cpu=True):
    """The TCCEE image of elitho/stcc.py at the reference constants (2048 nm,
    NA 0.33, 1901 pupil points, the 0.5-degree annular source), X
    polarisation with M3D arrays: images/s over the whole call (mask spectrum,
    field terms, three TCCs, TCCEE, transform).  Roofline: the TCC ZGEMMs
    (8 ncut^2 n_src flops each, three per image), the call's dominant
    arithmetic, over the call's time.  CPU baseline: the oracle (NumPy,
    vectorised) on the same inputs."""
    import torch
    from euvlitho_amd.stcc import StccConfig, intensity
    cfg = StccConfig()
    rng = np.random.default_rng(0)
    mask = (rng.random((cfg.ndiv, cfg.ndiv)) > 0.5).astype(int)
    lpx, lpy = cfg.lpmax
    shape = (2 * lpx + 1, 2 * lpy + 1)
    extra = [0.05 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(3)]
    amps = (0.21 - 0.13j, 0.83 + 0.31j, np.exp(0.37j))
    intensity(mask, "X", *extra, cfg=cfg, amplitudes=amps)
    torch.cuda.synchronize()
    ts, dts = [], []
    for _ in range(reps):
        tm = {}
        t0 = time.perf_counter()
        intensity(mask, "X", *extra, cfg=cfg, amplitudes=amps, timing=tm)
        ts.append(time.perf_counter() - t0)
        dts.append(tm["device_ms"] * 1e-3)
    t, dt = float(np.median(ts)), float(np.median(dts))
    ncut, ns = len(cfg.pupil()), len(cfg.source())
    flops = 3 * 8.0 * ncut * ncut * ns
    res = {"value": 1.0 / t, "unit": "images/s", "config": f"TCCEE X with M3D arrays, 2048 nm, {ncut} pupil "
           f"points, {ns} source points, {cfg.xdiv}x{cfg.xdiv} image", "ms_per_image": 1000 * t,
           "device_ms_per_image": 1000 * dt,
           "roofline": {"bound": "mfma", "kernel": "zgemm_mfma_kernel (TCC)", "achieved": flops / dt / 1e12,
                        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": flops / dt / 1e12 / FP64_PEAK_TFLOPS,
                        "note": "TCC flops over the stream time of all the image's kernels and uploads"}}
    if with_cpu:
        from oracle import stcc as OS
        c = OS.constants()
        t0 = time.perf_counter()
        OS.intensity(c, "X", OS.mask_fft_window(c, mask), *amps, *extra)
        tc = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": 1.0 / tc, "unit": "images/s", "cores": len(os.sched_getaffinity(0)),
                               "kind": "port", "sample": f"one image, the oracle (NumPy, vectorised sums): {tc:.1f} s"}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="time the FP64 form only (profiling runs)")
    ap.add_argument("--served", type=int, default=2000,
                    help="requests of the served Poisson stream (0 = skip; 2000 = the full stream)")
    ap.add_argument("--slots", type=int, default=4, help="streams of the served form")
    ap.add_argument("--served-only", action="store_true", help="only run the served form (mixed precision)")
    ap.add_argument("--served-eager", action="store_true",
                    help="served form with eager plan runs instead of replayed graphs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the plain copy that scales the HBM byte counters")
    ap.add_argument("--no-mask-producer", action="store_true", help="skip the mask producer leg")
    ap.add_argument("--no-cnn", action="store_true", help="skip the CNN predictor leg")
    ap.add_argument("--cnn-only", action="store_true", help="only run the CNN predictor leg")
    ap.add_argument("--no-stcc", action="store_true", help="skip the TCCEE image leg")
    ap.add_argument("--no-ref-default", action="store_true", help="skip the reference-default m3d.cpp leg")
    ap.add_argument("--batched", type=int, default=64, help="clips of the batched form (0 = skip)")
    ap.add_argument("--batched-only", action="store_true", help="only run the batched form")
    ap.add_argument("--only-form", choices=["fp64", "mixed", "fast"],
                    help="profiling runs: time one precision form on the workload and print its GEMM launch figures")
    ap.add_argument("--mask-producer-only", action="store_true", help="only run the mask producer leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); clips are independent, so N GPUs "
              f"would each run this workload on their own share of the clip stream.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return
    if args.mask_producer_only:
        print(json.dumps({"mask_producer": mask_producer(with_cpu=not args.no_cpu_baseline)}), flush=True)
        return
    if args.served_only:
        print(json.dumps({"served": run_served(args.served, args.slots, loads=(0.8, 0.6),
                                               graphs=not args.served_eager)}), flush=True)
        return
    if args.batched_only:
        print(json.dumps({"batched": run_batched(args.batched)}), flush=True)
        return
    if args.cnn_only:
        print(json.dumps({"cnn_predictor": cnn_predictor(with_cpu=not args.no_cpu_baseline)}), flush=True)
        return
    import ctypes
    import torch
    from euvlitho_amd.engine import Engine
    from euvlitho_amd.masks import manhattan
    from euvlitho_amd import _lib

    cfg = workload_config()
    n = args.warmup + args.steps
    masks = [manhattan(cfg.ndivx, cfg.ndivy, 0.5, seed) for seed in range(n)]

    def run_form(precision):
        """Time one form: warmup, then `steps` clips each bracketed by a device
        synchronisation; then one more clip with per-GEMM HIP events."""
        eng = Engine(cfg, "cuda:0", precision=precision)
        tuned = eng.autotune()  # GEMM tile choice for this size bucket
        bits = [eng.upload_mask(m) for m in masks]
        eng.blank_amplitudes()  # mask-independent, resident for every clip
        torch.cuda.synchronize()
        t_k = time.perf_counter()
        eng.eigenfunction(0.0, SOCS_KERNELS)  # mask-independent, resident for every clip
        torch.cuda.synchronize()
        tcc_ms = 1000.0 * (time.perf_counter() - t_k)

        def step(b):
            Ax = eng.amp_s(b, "X")
            a0, ax, ay = eng.m3d(b, Ax0=Ax[0])
            img = eng.abbe(b, Ax=Ax)
            socs = eng.socs(b, a0, ax, ay, nsocs=SOCS_KERNELS, nsocsxy=SOCS_XY)
            return Ax, a0, img, socs

        for i in range(args.warmup):
            step(bits[i])
        torch.cuda.synchronize()
        # The set-up's objects (about 170k, mostly torch's) go to the permanent generation: a full
        # collection over them otherwise stalls the host for ~40 ms inside one step.
        gc.collect()
        gc.freeze()
        lat = []
        t0 = time.perf_counter()
        for i in range(args.warmup, n):
            ts = time.perf_counter()
            step(bits[i])
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - ts)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # GEMM roofline: HIP events on the library's stream around every GEMM launch of one extra clip
        # (all launches, then those whose products ran on FP32 MFMA)
        stats = (ctypes.c_double * 8)()
        _lib.check(eng.lib.euvl_gemm_stats(eng.h, 1, None), "euvl_gemm_stats")
        outs = [t.cpu().numpy() for t in step(bits[args.warmup])]
        _lib.check(eng.lib.euvl_gemm_stats(eng.h, 2, stats), "euvl_gemm_stats")
        info = eng.last_info  # per offset: LU path, doublings, (reserved), incidences
        em = {"lu_path": [int(x) for x in info[:, 0]], "doublings": int(info[:, 1].max()),
              "incidences": int(info[:, 3].max())}
        eng.close()
        return dict(lat=lat, wall=wall, tcc_ms=tcc_ms, outs=outs, launches=stats[0], gemm_ms=stats[1],
                    gemm_flops=stats[2], gemm_bytes=stats[3], f32_launches=stats[4], f32_ms=stats[5],
                    f32_flops=stats[6], f32_bytes=stats[7], tuned=tuned, em=em)

    if args.only_form:
        r = run_form({"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED, "fast": _lib.PREC_FAST}[args.only_form])
        print(json.dumps({"form": args.only_form, "clips_per_s": args.steps / r["wall"],
                          "launches_per_clip": int(r["launches"]), "f32_launches_per_clip": int(r["f32_launches"]),
                          "gemm_ms_per_clip": r["gemm_ms"], "f32_gemm_ms_per_clip": r["f32_ms"],
                          "gemm_flops_8mnk": r["gemm_flops"], "f32_flops_8mnk": r["f32_flops"],
                          "algorithmic_bytes_per_launch": r["gemm_bytes"] / max(r["launches"], 1),
                          "f32_algorithmic_bytes_per_launch": r["f32_bytes"] / max(r["f32_launches"], 1),
                          "step_ms": [round(1000.0 * t, 2) for t in r["lat"]]}),
              flush=True)
        return
    f64 = run_form(_lib.PREC_FP64)
    mix = None if args.no_mixed else run_form(_lib.PREC_MIXED)
    fast = None if args.no_mixed else run_form(_lib.PREC_FAST)
    ms = 1000.0 * f64["wall"] / args.steps
    launches, gemm_ms, gemm_flops, gemm_bytes = f64["launches"], f64["gemm_ms"], f64["gemm_flops"], f64["gemm_bytes"]
    executed = gemm_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    n_solves = len(cfg.offsets())  # X polarisation, one solve per source division
    N = len(cfg.order_set()[0])
    algo_flops = EM_FLOPS_PER_N3 * float(N) ** 3 * n_solves
    achieved = algo_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    traffic = measured_traffic()
    lat = f64["lat"]

    def form_line(r, dtype, traffic_key):
        """A precision form against the FP64 form on the same clip: rate,
        latency, errors per quantity, and the GEMM launches split by the
        precision their products ran in."""
        (A64, a064, I64, S64), (A32, a032, I32, S32) = f64["outs"], r["outs"]
        f32_ms, f64_ms = r["f32_ms"], r["gemm_ms"] - r["f32_ms"]
        f32_l, f64_l = r["f32_launches"], r["launches"] - r["f32_launches"]
        f64_flops = r["gemm_flops"] - r["f32_flops"]
        t = measured_traffic(traffic_key)
        return {
            "value": args.steps / r["wall"], "unit": "clips/s", "ms_per_step": 1000.0 * r["wall"] / args.steps,
            "p50_ms": 1000.0 * float(np.percentile(r["lat"], 50)),
            "p99_ms": 1000.0 * float(np.percentile(r["lat"], 99)),
            "dtype": dtype, "em_solve": r["em"],
            "err_vs_fp64": {"abbe_image_of_peak": float(np.abs(I32 - I64).max() / I64.max()),
                            "socs_image_of_peak": float(np.abs(S32 - S64).max() / np.abs(S64).max()),
                            "amplitudes_of_max": float(np.abs(A32 - A64).max() / np.abs(A64).max()),
                            "a0_of_max": float(np.abs(a032 - a064).max() / np.abs(a064).max())},
            "roofline": {"bound": "mfma", "kernel": "zgemm_mfma_kernel",
                         "achieved": algo_flops / (r["gemm_ms"] * 1e-3) / 1e12, "unit": "TFLOP/s",
                         "flops_counted": "164 N^3 per solve over the clip's GEMM time (FP64 and FP32 launches)",
                         "peak": FP64_PEAK_TFLOPS if f64_ms >= f32_ms else FP32_PEAK_TFLOPS,
                         "frac": algo_flops / (r["gemm_ms"] * 1e-3) / 1e12 /
                                 (FP64_PEAK_TFLOPS if f64_ms >= f32_ms else FP32_PEAK_TFLOPS),
                         "fp32_launches": {"launches_per_clip": int(f32_l), "ms_per_clip": f32_ms,
                                           "executed_8mnk_tflops": r["f32_flops"] / max(f32_ms, 1e-9) / 1e9,
                                           "avg_launch_us": 1000.0 * f32_ms / max(f32_l, 1),
                                           "algorithmic_bytes_per_launch": r["f32_bytes"] / max(f32_l, 1),
                                           "traffic": t.get("traffic_bytes_per_launch"),
                                           "traffic_source": t.get("source")},
                         "fp64_launches": {"launches_per_clip": int(f64_l), "ms_per_clip": f64_ms,
                                           "executed_8mnk_tflops": f64_flops / max(f64_ms, 1e-9) / 1e9},
                         "autotuned_variant": r["tuned"].get("fp32", (None,))[0],
                         "complex_form": "3M (3 real MFMAs per complex product)"},
        }

    mixed = None
    if mix is not None:
        mixed = form_line(mix, "f32 MFMA for the coupling / top-solve products (FP32 accumulation) and the imaging "
                               "(x pass on FP32 MFMA, Kahan-compensated source sums); f64 matrix functions, "
                               "multilayer cascade and storage", "mixed")
        mixed["fast_variant"] = form_line(fast, "f32 MFMA for every EM product (matrix functions too) and the "
                                                "imaging; f64 multilayer and storage", "fast")
    tcc_ms = f64["tcc_ms"]

    line = {
        "metric": "EM+Abbe mask clips/s at <=1e-3 intensity err vs FP64; p50/p99 clip latency",
        "value": args.steps / f64["wall"],
        "unit": "clips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "p50_ms": 1000.0 * float(np.percentile(lat, 50)),
        "p99_ms": 1000.0 * float(np.percentile(lat, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Manhattan clips)",
        "config": {"workload": "plain FP64 drop-in, 1024x1024 nm clip, +-20 orders (N=1681), dipole+annular "
                               "200 source pts on 2x2 divisions (4 EM solves), M3D params + Abbe 256x256 + "
                               "24-kernel SOCS 256x256",
                   "form": "fp64", "clips": args.steps, "socs_kernel_build_ms": tcc_ms, "em_solve": f64["em"]},
        "roofline": {"bound": "mfma", "kernel": "zgemm_mfma_kernel", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "flops_counted": "reference algorithm: 164 N^3 per (clip, polarisation, offset) solve "
                                      f"x {n_solves} solves = {algo_flops:.3e} per clip, over the clip's GEMM time",
                     "frac_note": "the reference algorithm's work per GEMM time: above 1 because this route needs "
                                  "fewer flops than it (per solve 8 N x N x N GEMM products, the products by the "
                                  "absorber operator and the coupling product as 2D FFT correlations, a block "
                                  "Gauss-Jordan solve; no eigendecomposition) and runs part of its work in the FFT "
                                  "kernels; the MFMA units' own utilisation is mfma_issue_frac, the reference "
                                  "algorithm's work over the whole clip time clip_reference_tflops",
                     "clip_reference_tflops": algo_flops / (ms * 1e-3) / 1e12,
                     "route_flops_per_clip_8mnk": gemm_flops,
                     "algorithmic_flops_per_launch": algo_flops / max(launches, 1),
                     "traffic": traffic.get("traffic_bytes_per_launch"),
                     "traffic_source": traffic.get("source"),
                     "algorithmic_bytes_per_launch": gemm_bytes / max(launches, 1),
                     "traffic_over_algorithmic": (traffic["traffic_bytes_per_launch"] / (gemm_bytes / launches)
                                                  if traffic.get("traffic_bytes_per_launch") and launches else None),
                     "launches_per_clip": int(launches),
                     "avg_launch_us": 1000.0 * gemm_ms / max(launches, 1),
                     "gemm_ms_per_clip": gemm_ms, "gemm_share_of_step": gemm_ms / ms,
                     "autotuned_variant": f64["tuned"]["fp64"][0],
                     # the route's own GEMMs: 8mnk per complex product, of which the 3M form issues 6mnk
                     "gemm_executed_8mnk_tflops": executed,
                     "gemm_executed_8mnk_frac": executed / FP64_PEAK_TFLOPS,
                     "complex_form": "3M (3 real MFMAs per complex product)",
                     "mfma_issue_frac": 0.75 * executed / FP64_PEAK_TFLOPS},
        "forms": {"mixed": mixed} if mixed else {},
    }
    if args.batched > 0:
        line["forms"]["batched"] = run_batched(args.batched)
    if args.served > 0:
        line["forms"]["served"] = run_served(args.served, args.slots, loads=(0.8, 0.6),
                                             graphs=not args.served_eager)
        line["forms"]["served_fast"] = run_served(args.served, args.slots, _lib.PREC_FAST,
                                                  graphs=not args.served_eager)
    if not args.no_mask_producer:
        line["mask_producer"] = mask_producer(with_cpu=not args.no_cpu_baseline)
    if not args.no_cnn:
        line["cnn_predictor"] = cnn_predictor(with_cpu=not args.no_cpu_baseline)
    if not args.no_stcc:
        line["tccee_image"] = stcc_leg(with_cpu=not args.no_cpu_baseline)
    if not args.no_ref_default:
        line["reference_default_m3d"] = ref_default_leg()
    if not args.no_cpu_baseline:
        threads = len(os.sched_getaffinity(0))
        line["cpu_baseline"] = cpu_baseline(cfg, masks, threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

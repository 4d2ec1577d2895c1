"""Benchmark: EM + Abbe mask clips/s on one MI355X.

Workload (one step = one clip, FP64): a seeded 1024 x 1024 nm Manhattan clip
(1 nm pixels, rectangles 16-64 nm, vacuum fraction 0.5), NA 0.55, 4x/8x,
CRAO 5.355 deg, 60 nm absorber n = 0.947+0.031i on 40 x Mo/Si 2.8/4.2 nm with
a 2.5 nm Ru cap, orders +-20 x +-20 (N = 1681 modes), X-dipole (sigma 0.5-0.9,
60 deg) + annular (sigma 0.7-0.8) source on 2 x 2 lattice divisions
(190 source points, 4 EM solves per clip), outputs: M3D parameters and the
rigorous Abbe image (256 x 256).

Prints one JSON line (metric, value = clips/s, roofline of the dominant
kernel = the FP64 MFMA complex GEMM, cpu_baseline = the FP64 NumPy/LAPACK
oracle on a bounded sample of the same clip).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)


def workload_config():
    from euvlitho_amd.config import high_na
    src = dict(kind=2, sigma_out=0.9, sigma_in=0.5, open_angle=60.0, extra=((1, 0.8, 0.7, 90.0),))
    return high_na(1024, 20, src, ndivs=2, image=(256, 256))


def cpu_baseline(cfg, masks, threads):
    """Oracle (NumPy + LAPACK zgeev/zgesv, FP64) on a bounded sample: the EM solve
    of one source division plus its Abbe accumulation and the clip's M3D fit;
    the clip time is extrapolated to all divisions."""
    from oracle import abbe as OA, em as OE, geometry as G, problems as OP
    p = OP.hna(cfg.ndivx, cfg.orders[1], None)
    d = G.derived(p)
    li, mi, L, M = cfg.order_set()
    li, mi = li.astype(np.int64), mi.astype(np.int64)
    m = masks[0]
    off = cfg.offsets()
    divs = cfg.source_points()
    t0 = time.perf_counter()
    Ax = OE.amp_s("X", m, p, d, li, mi, L, M, off[0][0], off[0][1])
    t1 = time.perf_counter()
    _, _, l0, m0 = divs[0]
    nx, ny = cfg.image_shape
    OA.abbe_image(p, d, m, li, mi, L, M, [(tuple(off[0]), l0, m0)], nx, ny, amps=[Ax])
    t2 = time.perf_counter()
    vc, ab = OA.blank_amplitudes(p, d)
    t3 = time.perf_counter()
    OA.m3d_params(p, d, m, li, mi, Ax, vc, ab)
    t4 = time.perf_counter()
    n_off = len(off)
    n_src = sum(len(x[2]) for x in divs)
    t_clip = n_off * (t1 - t0) + (t2 - t1) * n_src / max(len(l0), 1) + (t4 - t3)
    return dict(value=1.0 / t_clip, unit="clips/s", cores=threads, kind="port",
                sample=f"oracle EM solve of 1 of {n_off} divisions ({t1 - t0:.1f} s, N={len(li)}), Abbe of its "
                       f"{len(l0)} of {n_src} source points ({t2 - t1:.1f} s), M3D fit ({t4 - t3:.1f} s); "
                       f"clip time extrapolated to {t_clip:.1f} s; blank solves cached as on the GPU")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); clips are independent, so N GPUs "
              f"would each run this workload on their own share of the clip stream.", file=sys.stderr)
        sys.exit(2)

    import torch
    from euvlitho_amd.engine import Engine
    from euvlitho_amd.masks import manhattan
    from euvlitho_amd import _lib

    cfg = workload_config()
    n = args.warmup + args.steps
    masks = [manhattan(cfg.ndivx, cfg.ndivy, 0.5, seed) for seed in range(n)]
    eng = Engine(cfg, "cuda:0")
    bits = [eng.upload_mask(m) for m in masks]
    eng.blank_amplitudes()  # mask-independent, resident for every clip
    torch.cuda.synchronize()

    def step(b):
        Ax = eng.amp_s(b, "X")
        a0, ax, ay = eng.m3d(b, Ax0=Ax[0])
        img = eng.abbe(b, Ax=Ax)
        return a0, img

    for i in range(args.warmup):
        step(bits[i])
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for i in range(args.warmup, n):
        ts = time.perf_counter()
        step(bits[i])
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = 1000.0 * wall / args.steps

    # GEMM roofline: HIP events around every GEMM launch of one extra clip
    import ctypes
    stats = (ctypes.c_double * 3)()
    _lib.check(eng.lib.euvl_gemm_stats(eng.h, 1, None), "euvl_gemm_stats")
    step(bits[0])
    _lib.check(eng.lib.euvl_gemm_stats(eng.h, 0, stats), "euvl_gemm_stats")
    launches, gemm_ms, gemm_flops = stats[0], stats[1], stats[2]
    achieved = gemm_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0

    line = {
        "metric": "EM+Abbe mask clips/s at <=1e-3 intensity err vs FP64; p50/p99 clip latency",
        "value": args.steps / wall,
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
                               "190 source pts on 2x2 divisions (4 EM solves), M3D params + Abbe 256x256",
                   "form": "fp64", "clips": args.steps},
        "roofline": {"bound": "mfma", "kernel": "zgemm_mfma_kernel", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": None, "launches_per_clip": int(launches),
                     "gemm_ms_per_clip": gemm_ms, "gemm_share_of_step": gemm_ms / ms},
    }
    if not args.no_cpu_baseline:
        threads = len(os.sched_getaffinity(0))
        line["cpu_baseline"] = cpu_baseline(cfg, masks, threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

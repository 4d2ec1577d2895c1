"""Batched form on the GPU: clips of uneven pitch, order truncation and
source in batched plan runs captured into one HIP graph, against the eager
per-clip path (Engine) on the same masks.  Tolerances (FP64): M3D parameters
1e-9 of the largest |a0|, images 1e-10 of the peak; graph replays agree with
the first to 1e-14."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)


def mix(n_per):
    from euvlitho_amd.config import workload_mix
    from euvlitho_amd.masks import manhattan
    cfgs = workload_mix()[:3]
    rng = np.random.default_rng(3)
    clips = []
    for cfg, n in zip(cfgs, n_per):
        for _ in range(n):
            dens = float(rng.uniform(0.3, 0.7))
            clips.append((cfg, manhattan(cfg.ndivx, cfg.ndivy, dens, int(rng.integers(1 << 30)))))
    return clips


def eager(clips):
    from euvlitho_amd.engine import Engine
    engs, out = {}, []
    for cfg, m in clips:
        if cfg not in engs:
            engs[cfg] = Engine(cfg)
        e = engs[cfg]
        b = e.upload_mask(m)
        Ax = e.amp_s(b, "X")
        a0, ax, ay = e.m3d(b, Ax0=Ax[0])
        img = e.abbe(b, Ax=Ax)
        out.append([t.cpu().numpy() for t in (a0, ax, ay, img)])
    return out


def compare(got, ref):
    worst = [0.0, 0.0]
    for g, r in zip(got, ref):
        g = [t.cpu().numpy() for t in g]
        s = np.abs(r[0]).max()
        worst[0] = max(worst[0], max(np.abs(g[i] - r[i]).max() / s for i in range(3)))
        worst[1] = max(worst[1], np.abs(g[3] - r[3]).max() / r[3].max())
    return worst


def test_batched_graph_equals_per_clip():
    from euvlitho_amd.batch import BatchedEngine
    clips = mix([5, 3, 2])  # 256 nm runs of 4 + 1 clips, 512 nm runs of 3 and 2
    say("batched: 10 clips of 3 problems")
    ref = eager(clips)
    be = BatchedEngine()
    try:
        be.load(clips)
        say(f"  runs {len(be.runs)}, plans {len(be.plans)}, graph nodes {be.graph_nodes()}, "
            f"decisions {[p.decisions() for p in be.plans]}")
        assert len(be.runs) == 4 and len(be.plans) == 4
        be.launch()
        first = [[t.cpu().numpy().copy() for t in r] for r in be.results()]
        assert be.reruns == 0 and all(f == 0 for f in be.flags())
        e = compare(be.results(), ref)
        say(f"  graph vs per clip: M3D {e[0]:.2e} of max|a0|, image {e[1]:.2e} of peak")
        assert e[0] < 1e-9 and e[1] < 1e-10
        # replays, with eager work of the program and synchronisations between them
        for k in range(3):
            torch.ones(1 << 20, device="cuda").sum().item()
            be.launch()
            again = [[t.cpu().numpy() for t in r] for r in be.results()]
            d = max(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300) for x, y in zip(again, first)
                    for a, b in zip(x, y))
            say(f"  replay {k}: max rel difference to the first {d:.1e}")
            assert d < 1e-14
    finally:
        be.close()


def test_batched_flags_rerun_adaptively():
    """A decision fixed too low (no doublings at all) is caught by the
    device-side check: the run's flags carry the doubling bit, and the
    results are those of the adaptive rerun."""
    from euvlitho_amd.batch import BatchedEngine
    from euvlitho_amd import _lib
    clips = mix([2, 0, 0])
    ref = eager(clips)
    be = BatchedEngine(graph=False)
    try:
        be.load(clips)
        for p in be.plans:
            _lib.check(p.lib.euvl_plan_set_decisions(p.pl, 0), "euvl_plan_set_decisions")
        be.launch()
        fl = be.flags()
        say(f"  flags with 0 doublings: {fl}")
        assert all(f & 1 for f in fl)
        got = be.results()
        assert be.reruns == len(be.runs)
        e = compare(got, ref)
        say(f"  after the adaptive rerun: M3D {e[0]:.2e}, image {e[1]:.2e}")
        assert e[0] < 1e-9 and e[1] < 1e-10
    finally:
        be.close()

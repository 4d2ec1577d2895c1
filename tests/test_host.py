"""Host-side logic (CPU): the C ABI library loads and exports what
include/euvl.h declares; the typed config reproduces the reference geometry
(checked against the oracle restatement and the reference fixture)."""
import ctypes
import os
import re

import numpy as np
import pytest

from euvlitho_amd import _lib
from euvlitho_amd.config import Config, high_na, reference_default
from euvlitho_amd.masks import manhattan, pack_bits, unpack_bits
from oracle import geometry as G, problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    with open(os.path.join(ROOT, "include", "euvl.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(euvl_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libeuvl.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert "euvl_amp_s" in syms and "euvl_zgemm_batched" in syms
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.EXPORTS)
    _lib.load()  # prototypes bind


def test_reference_default_geometry_matches_oracle(golden):
    cfg = reference_default()
    p = P.REF_DEFAULT
    d = G.derived(p)
    assert cfg.lsmax == (d["lsmaxX"], d["lsmaxY"]) and cfg.lpmax == (d["lpmaxX"], d["lpmaxY"])
    li, mi, L, M = cfg.order_set()
    oli, omi, oL, oM = G.order_set(p, "diamond", cutoff=6.0)
    assert (L, M) == (oL, oM) and np.array_equal(li, oli) and np.array_equal(mi, omi)
    lin, min_, snum = cfg.pupil
    ref = np.loadtxt(os.path.join(golden, "ref_inputlm.csv"), delimiter=",", dtype=np.int64)
    assert np.array_equal(lin, ref[:, 0]) and np.array_equal(min_, ref[:, 1])
    assert np.array_equal(snum >= 8, ref[:, 2] == 1)
    (_, _, l0, m0), = cfg.source_points()
    src = G.source_lattice(p, d, 1, 1, 0.9, 0.55, 90.0)[(0, 0)]
    assert np.array_equal(l0, src[0]) and np.array_equal(m0, src[1])
    assert abs(cfg.kx0 - d["kx0"]) == 0 and abs(cfg.ky0 - d["ky0"]) == 0


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_high_na_sources_match_oracle(nd):
    cfg = high_na(512, 10, dict(kind=2, sigma_out=0.9, sigma_in=0.5, open_angle=60.0,
                                extra=((1, 0.8, 0.7, 90.0),)), ndivs=nd)
    p = P.hna(512, 10, None)
    d = G.derived(p)
    ref = G.source_lattice(p, d, nd, ((2, 0.9, 0.5, 60.0), (1, 0.8, 0.7, 90.0)), None, None, None)
    for nsx, nsy, l0, m0 in cfg.source_points():
        assert np.array_equal(l0, ref[(nsx, nsy)][0]) and np.array_equal(m0, ref[(nsx, nsy)][1])


def test_stack_layers_match_oracle():
    from euvlitho_amd.config import HIGH_NA_STACK, REFERENCE_STACK
    for st, ost in ((REFERENCE_STACK, P.REF_STACK), (HIGH_NA_STACK, P.HNA_STACK)):
        a = st.layers()
        b = P.layer_list(ost)
        assert len(a) == len(b)
        for (e1, d1), (e2, d2) in zip(a, b):
            assert abs(e1 - e2) == 0 and d1 == d2


def test_mask_pack_roundtrip_and_density():
    m = manhattan(256, 128, 0.5, 9)
    assert 0.5 <= m.mean() < 0.56
    assert np.array_equal(unpack_bits(pack_bits(m), 256, 128), m)
    m2 = manhattan(256, 128, 0.5, 9)
    assert np.array_equal(m, m2)


def test_engine_requires_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from euvlitho_amd.engine import Engine
    with pytest.raises(_lib.EuvlError):
        Engine(reference_default())


def test_served_stream_and_footprint():
    from euvlitho_amd.serve import footprint, poisson_stream, served_configs
    cfgs = served_configs()
    a = poisson_stream(50, 10.0, seed=7, cfgs=cfgs)
    b = poisson_stream(50, 10.0, seed=7, cfgs=cfgs)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert all(np.array_equal(x.mask, y.mask) for x, y in zip(a, b))
    assert all(a[i].arrival < a[i + 1].arrival for i in range(49))
    assert abs(a[-1].arrival / 50 - 0.1) < 0.05  # mean gap 1 / rate
    assert {r.cfg for r in a} <= set(cfgs)
    for r in a[:10]:
        dens = r.mask.mean()
        assert 0.29 <= dens <= 0.72
    # the EM workspace dominates the footprint: 13 N x N complex matrices per offset
    c2 = cfgs[4]
    n, off = len(c2.order_set()[0]), len(c2.offsets())
    assert footprint(c2) >= 13 * 16 * n * n * off
    assert footprint(c2) < 1.2 * 13 * 16 * n * n * off + 2**30
    # source point counts of the mix are in 61..200
    pts = [sum(len(x[2]) for x in c.source_points()) for c in cfgs]
    assert all(61 <= p <= 200 for p in pts)


@pytest.mark.parametrize("name,count,n_off", [("c1", 61, 4), ("c2", 200, 4)])
def test_presets_have_the_workload_point_counts(name, count, n_off):
    """Presets carry the workloads' source point counts exactly (61 annular
    0.3-0.8 at 512 nm; 200 dipole + annular at 1024 nm), as a deterministic
    subset of the reference lattice rule's points, spread over all 2 x 2
    divisions."""
    from euvlitho_amd.config import preset
    import dataclasses
    cfg = preset(name)
    divs = cfg.source_points()
    assert len(divs) == n_off == len(cfg.offsets())
    assert sum(len(d[2]) for d in divs) == count
    assert all(len(d[2]) > 0 for d in divs)
    full = dataclasses.replace(cfg, source=dataclasses.replace(cfg.source, count=0)).source_points()
    for (a, b, l0, m0), (_, _, fl, fm) in zip(divs, full):
        assert set(zip(l0.tolist(), m0.tolist())) <= set(zip(fl.tolist(), fm.tolist()))
    assert [len(d[2]) for d in preset(name).source_points()] == [len(d[2]) for d in divs]


def test_bench_defines_every_leg():
    """bench.py imports without a GPU and defines every leg main() calls."""
    import ast
    import bench
    tree = ast.parse(open(os.path.join(ROOT, "bench.py")).read())
    main = next(n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name == "main")
    called = {n.func.id for n in ast.walk(main) if isinstance(n, ast.Call) and isinstance(n.func, ast.Name)}
    local = {n.name for n in ast.walk(main) if isinstance(n, ast.FunctionDef)}
    local |= {a.asname or a.name for n in ast.walk(main) if isinstance(n, ast.ImportFrom) for a in n.names}
    import builtins
    missing = [c for c in called if not hasattr(bench, c) and not hasattr(builtins, c) and c not in local]
    assert not missing, missing


def test_erlang_c_queue_model():
    """The served leg's queue model: with one server it is M/M/1 (the wait
    exceeds t with probability rho exp(-(mu - lambda) t)); with c servers
    P_wait matches the Erlang C value of a textbook case (c = 2, a = 1: 1/3),
    and a load below the quantile's mass gives no wait."""
    import math
    import bench
    p, t = bench.erlang_c_wait(1, 10.0, 8.0)
    assert abs(p - 0.8) < 1e-12 and abs(t - math.log(0.8 / 0.01) / 2.0) < 1e-12
    p, t = bench.erlang_c_wait(2, 2.0, 1.0)
    assert abs(p - 1.0 / 3.0) < 1e-12 and abs(t - math.log((1.0 / 3.0) / 0.01) / 1.0) < 1e-12
    p, t = bench.erlang_c_wait(4, 100.0, 1.0)
    assert p < 0.01 and t == 0.0
    assert bench.erlang_c_wait(4, 10.0, 12.0)[1] == float("inf")

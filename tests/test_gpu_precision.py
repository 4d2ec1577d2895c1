"""GPU: the precision form (FP32 MFMA products, FP64 storage and FP64
everywhere else) against the FP64 drop-in on the same inputs.

Bars: an FP32-compute GEMM within 2e-5 of the largest entry of the FP64
product (K up to 2100, FP32 accumulation); for a whole clip, images within
1e-3 of their peak of the FP64 images (the bar the precision form is held to)
and amplitudes within 1e-3 of the largest amplitude.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.fail("GPU test without a GPU")
    from euvlitho_amd import _lib
    return _lib.load()


@pytest.mark.parametrize("m,n,k,batch", [(17, 33, 5, 2), (64, 64, 64, 1), (100, 70, 130, 3), (300, 257, 2100, 1),
                                         (1681, 12, 1681, 2)])
def test_zgemm_fp32_compute(lib, m, n, k, batch):
    from euvlitho_amd import _lib
    say(f"FP32-compute zgemm m={m} n={n} k={k} batch={batch}")
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((batch, m, k)) + 1j * rng.standard_normal((batch, m, k))
    B = rng.standard_normal((batch, k, n)) + 1j * rng.standard_normal((batch, k, n))
    C0 = rng.standard_normal((batch, m, n)) + 1j * rng.standard_normal((batch, m, n))
    dA, dB, dC = (torch.from_numpy(x).cuda() for x in (A, B, C0))
    h = ctypes.c_void_p()
    assert lib.euvl_create(ctypes.byref(h)) == 0
    try:
        al, be, ga = (ctypes.c_double * 2)(0.5, -1.0), (ctypes.c_double * 2)(0.25, 0.5), (ctypes.c_double * 2)(2, 0)
        ref = (0.5 - 1j) * A @ B + (0.25 + 0.5j) * C0 + 2.0 * np.eye(m, n)[None]
        got = {}
        assert lib.euvl_gemm_stats(h, 1, None) == 0  # record the three launches below
        # the 3M form (default) and the 4M form, both with FP32 products
        for name, mask in (("3m", _lib.PREC_FP32_GEMM), ("4m", _lib.PREC_FP32_GEMM | _lib.PREC_GEMM_4M)):
            assert lib.euvl_set_precision(h, mask) == 0
            dC.copy_(torch.from_numpy(C0))
            rc = lib.euvl_zgemm_batched(h, m, n, k, al, ctypes.c_void_p(dA.data_ptr()), k, m * k,
                                        ctypes.c_void_p(dB.data_ptr()), n, k * n, be, ctypes.c_void_p(dC.data_ptr()),
                                        n, m * n, ga, batch)
            assert rc == 0, lib.euvl_last_error().decode()
            torch.cuda.synchronize()
            got[name] = dC.cpu().numpy()
            e = np.abs(got[name] - ref).max() / np.abs(ref).max()
            say(f"  {name}: max err / max |C| = {e:.2e}")
            assert e < 2e-5
            assert e > 1e-12  # it really ran in FP32
        assert np.abs(got["3m"] - got["4m"]).max() > 0.0  # two different kernels ran
        # the mask selects: FP64 again after resetting it
        assert lib.euvl_set_precision(h, 0) == 0
        dC.copy_(torch.from_numpy(C0))
        rc = lib.euvl_zgemm_batched(h, m, n, k, al, ctypes.c_void_p(dA.data_ptr()), k, m * k,
                                    ctypes.c_void_p(dB.data_ptr()), n, k * n, be, ctypes.c_void_p(dC.data_ptr()), n,
                                    m * n, ga, batch)
        assert rc == 0
        torch.cuda.synchronize()
        assert np.abs(dC.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-13
        # launch records: all three, then the two whose products ran on FP32 MFMA
        st = (ctypes.c_double * 8)()
        assert lib.euvl_gemm_stats(h, 2, st) == 0
        fl = 8.0 * m * n * k * batch
        assert (st[0], st[2], st[4], st[6]) == (3.0, 3 * fl, 2.0, 2 * fl)
        assert st[1] > st[5] > 0.0
        assert lib.euvl_gemm_stats(h, 0, st) == 0 and st[0] == 0.0  # recording was off
        assert lib.euvl_set_precision(h, 64) != 0  # unknown bit
    finally:
        lib.euvl_destroy(h)


@pytest.mark.parametrize("m,n,k,batch", [(17, 33, 5, 2), (100, 70, 130, 3), (300, 257, 2100, 1), (1681, 12, 1681, 2),
                                         (200, 300, 400, 5)])
def test_zgemm_3m_form(lib, m, n, k, batch):
    """FP64 products in the 3-multiplication form (the default) and in the
    4-multiplication form (EUVL_PREC_GEMM_4M): both within the FP64 GEMM bar
    of the exact product (1e-13 of its largest entry per sqrt(k)), and not
    bit-identical to the 4M form (it really ran), on operands whose real and
    imaginary parts differ in scale (the case where 3M's cancellation in the
    imaginary part is largest)."""
    from euvlitho_amd import _lib
    say(f"3M zgemm m={m} n={n} k={k} batch={batch}")
    rng = np.random.default_rng(7 * m + n + k)
    A = 3.0 * rng.standard_normal((batch, m, k)) + 0.01j * rng.standard_normal((batch, m, k))
    B = rng.standard_normal((batch, k, n)) - 2j * rng.standard_normal((batch, k, n))
    C0 = rng.standard_normal((batch, m, n)) + 1j * rng.standard_normal((batch, m, n))
    ref = (0.5 - 1j) * A @ B + (0.25 + 0.5j) * C0 + 2.0 * np.eye(m, n)[None]
    h = ctypes.c_void_p()
    assert lib.euvl_create(ctypes.byref(h)) == 0
    out = {}
    try:
        for name, mask in (("4m", _lib.PREC_GEMM_4M), ("3m", _lib.PREC_FP64)):
            dA, dB, dC = (torch.from_numpy(x).cuda() for x in (A, B, C0))
            assert lib.euvl_set_precision(h, mask) == 0
            al, be, ga = (ctypes.c_double * 2)(0.5, -1.0), (ctypes.c_double * 2)(0.25, 0.5), (ctypes.c_double * 2)(2, 0)
            rc = lib.euvl_zgemm_batched(h, m, n, k, al, ctypes.c_void_p(dA.data_ptr()), k, m * k,
                                        ctypes.c_void_p(dB.data_ptr()), n, k * n, be, ctypes.c_void_p(dC.data_ptr()),
                                        n, m * n, ga, batch)
            assert rc == 0, lib.euvl_last_error().decode()
            torch.cuda.synchronize()
            out[name] = dC.cpu().numpy()
    finally:
        lib.euvl_destroy(h)
    tol = 1e-13 * np.abs(ref).max() * np.sqrt(k)
    for name, got in out.items():
        e = np.abs(got - ref).max()
        say(f"  {name}: max err {e:.2e} (bar {tol:.2e}); imag part err {np.abs(got.imag - ref.imag).max():.2e}")
        assert e < tol
    assert np.abs(out["3m"] - out["4m"]).max() > 0.0


def test_3m_form_clip_vs_4m(lib):
    """512 nm clip through the EM solve, M3D fit and Abbe image with FP64
    products in the 3M form (the default) against the 4M form: amplitudes, a0
    and the image agree to 1e-9 of their largest value (the drop-in's FP64
    bar against the reference is 1e-9 for amplitudes, 1e-6 of peak for
    intensity)."""
    from euvlitho_amd import _lib
    from euvlitho_amd.config import preset
    from euvlitho_amd.engine import Engine
    from euvlitho_amd.masks import manhattan
    cfg = preset("c1")
    m = manhattan(cfg.ndivx, cfg.ndivy, 0.5, 43)
    out = {}
    for name, prec in (("4m", _lib.PREC_GEMM_4M), ("3m", _lib.PREC_FP64)):
        eng = Engine(cfg, precision=prec)
        bits = eng.upload_mask(m)
        Ax = eng.amp_s(bits, "X")
        a0, ax, ay = eng.m3d(bits, Ax0=Ax[0])
        img = eng.abbe(bits, Ax=Ax)
        out[name] = [t.cpu().numpy() for t in (Ax, a0, img)]
        eng.close()
    (A4, a04, I4), (A3, a03, I3) = out["4m"], out["3m"]
    ea = np.abs(A3 - A4).max() / np.abs(A4).max()
    e0 = np.abs(a03 - a04).max() / np.abs(a04).max()
    ei = np.abs(I3 - I4).max() / I4.max()
    say(f"  3M vs 4M: amplitudes {ea:.2e}, a0 {e0:.2e}, Abbe image {ei:.2e}")
    assert ea < 1e-9 and e0 < 1e-9 and ei < 1e-9


def test_precision_form_clip_vs_fp64(lib):
    """512 nm clip, +-10 orders, 4 offsets: amplitudes, M3D parameters, Abbe
    and SOCS images of the mixed form against the FP64 drop-in."""
    from euvlitho_amd import _lib
    from euvlitho_amd.config import preset
    from euvlitho_amd.engine import Engine
    from euvlitho_amd.masks import manhattan
    cfg = preset("c1")
    m = manhattan(cfg.ndivx, cfg.ndivy, 0.5, 41)
    out = {}
    for name, prec in (("fp64", _lib.PREC_FP64), ("mixed", _lib.PREC_MIXED)):
        say(f"clip in form {name}")
        eng = Engine(cfg, precision=prec)
        bits = eng.upload_mask(m)
        Ax = eng.amp_s(bits, "X")
        a0, ax, ay = eng.m3d(bits, Ax0=Ax[0])
        img = eng.abbe(bits, Ax=Ax)
        socs = eng.socs(bits, a0, ax, ay, nsocs=24, nsocsxy=20)
        out[name] = [t.cpu().numpy() for t in (Ax, a0, img, socs)]
        eng.close()
    (A64, a064, I64, S64), (A32, a032, I32, S32) = out["fp64"], out["mixed"]
    ea = np.abs(A32 - A64).max() / np.abs(A64).max()
    e0 = np.abs(a032 - a064).max() / np.abs(a064).max()
    ei = np.abs(I32 - I64).max() / I64.max()
    es = np.abs(S32 - S64).max() / np.abs(S64).max()
    say(f"  amplitudes {ea:.2e}, a0 {e0:.2e}, Abbe image {ei:.2e}, SOCS image {es:.2e} (of max / peak)")
    assert ea < 1e-3 and e0 < 1e-3
    assert ei < 1e-3 and es < 1e-3
    assert ei > 0.0  # the mixed form is a different computation


def test_autotune_keeps_results(lib):
    """The autotuner picks a valid variant per size bucket; results after
    tuning equal the untuned ones (every variant sums k in the same order)."""
    from euvlitho_amd.config import preset
    from euvlitho_amd.engine import Engine
    from euvlitho_amd.masks import manhattan
    from euvlitho_amd import _lib
    cfg = preset("c1")
    m = manhattan(cfg.ndivx, cfg.ndivy, 0.5, 51)
    eng = Engine(cfg, precision=_lib.PREC_MIXED)
    bits = eng.upload_mask(m)
    before = eng.amp_s(bits, "X").cpu().numpy()
    tuned = eng.autotune()
    say(f"  tuned {tuned}")
    for key in ("fp64", "fp32"):
        v, ms = tuned[key]
        timed = [t for t in ms if t > 0]
        # the 3M forms (the default): FP64 has variants 0-6 (5 is the 32x32 tile,
        # timed on this small grid), FP32 0-4 and 6 (5 is not one of its variants)
        assert len(ms) == 7
        if key == "fp64":
            assert len(timed) == 7
        else:
            assert ms[5] == -1.0 and len(timed) == 6 and v != 5
        # the fastest, or a larger tile within 1 % of it
        assert ms[v] <= 1.01 * min(timed)
    after = eng.amp_s(bits, "X").cpu().numpy()
    e = np.abs(after - before).max() / np.abs(before).max()
    say(f"  amplitudes after tuning vs before: {e:.1e}")
    assert e < 1e-12


@pytest.mark.parametrize("m,kn,batch", [(200, 48, 2), (1681, 64, 1), (130, 20, 3), (200, 32, 2), (300, 100, 1)])
def test_zgemm_in_place_rows(lib, m, kn, batch):
    """C = A B with C the same storage as A (k = n, as the LU's panel products
    use it): the launcher keeps or picks a tile that spans all n columns, so
    every workgroup reads only the rows it overwrites, and edge tiles are not
    moved onto rows that another workgroup writes.  Small grids otherwise take
    the 64 x 32 tile: n = 32 keeps it, n = 48 and 64 move to a 64-column tile,
    n = 100 to a 128-column one.  Checked in the FP64 forms (3M, 4M) and the
    FP32-compute 3M form, to 1e-13 (FP64) and 1e-5 (FP32) of the exact
    product."""
    from euvlitho_amd import _lib
    say(f"in-place zgemm m={m} k=n={kn} batch={batch}")
    rng = np.random.default_rng(m + kn)
    A = rng.standard_normal((batch, m, kn)) + 1j * rng.standard_normal((batch, m, kn))
    B = rng.standard_normal((batch, kn, kn)) + 1j * rng.standard_normal((batch, kn, kn))
    ref = A @ B
    h = ctypes.c_void_p()
    assert lib.euvl_create(ctypes.byref(h)) == 0
    try:
        for mask, bar in ((_lib.PREC_FP64, 1e-13), (_lib.PREC_GEMM_4M, 1e-13), (_lib.PREC_FP32_GEMM, 1e-5)):
            dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
            assert lib.euvl_set_precision(h, mask) == 0
            one, zero = (ctypes.c_double * 2)(1.0, 0.0), (ctypes.c_double * 2)(0.0, 0.0)
            rc = lib.euvl_zgemm_batched(h, m, kn, kn, one, ctypes.c_void_p(dA.data_ptr()), kn, m * kn,
                                        ctypes.c_void_p(dB.data_ptr()), kn, kn * kn, zero,
                                        ctypes.c_void_p(dA.data_ptr()), kn, m * kn, None, batch)
            assert rc == 0, lib.euvl_last_error().decode()
            torch.cuda.synchronize()
            e = np.abs(dA.cpu().numpy() - ref).max() / np.abs(ref).max()
            say(f"  prec {mask}: rel err {e:.2e}")
            assert e < bar
    finally:
        lib.euvl_destroy(h)


def test_zgemm_in_place_too_wide_is_refused(lib):
    """An in-place product wider than every tile (C on A with n = 200) would
    race between workgroups: it is refused with EUVL_ERR_ARG and C is left
    untouched."""
    m, kn = 96, 200
    rng = np.random.default_rng(7)
    A = rng.standard_normal((m, kn)) + 1j * rng.standard_normal((m, kn))
    B = rng.standard_normal((kn, kn)) + 1j * rng.standard_normal((kn, kn))
    h = ctypes.c_void_p()
    assert lib.euvl_create(ctypes.byref(h)) == 0
    try:
        dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
        one, zero = (ctypes.c_double * 2)(1.0, 0.0), (ctypes.c_double * 2)(0.0, 0.0)
        rc = lib.euvl_zgemm_batched(h, m, kn, kn, one, ctypes.c_void_p(dA.data_ptr()), kn, 0,
                                    ctypes.c_void_p(dB.data_ptr()), kn, 0, zero,
                                    ctypes.c_void_p(dA.data_ptr()), kn, 0, None, 1)
        assert rc == -1 and b"overlaps" in lib.euvl_last_error()
        torch.cuda.synchronize()
        assert np.array_equal(dA.cpu().numpy(), A)
    finally:
        lib.euvl_destroy(h)

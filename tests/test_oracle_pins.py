"""Pins of the FP64 oracle against the reference's fixtures, textbook closed
forms and exact invariants (CPU only)."""
import gzip
import math
import os

import numpy as np
import pytest

from oracle import abbe as OA
from oracle import em, geometry as G, maskio, problems as P


def small(ndiv=512, K=6, sigma=(0.8, 0.3)):
    p = P.hna(ndiv, K, dict(type=1, sigma1=sigma[0], sigma2=sigma[1], openangle=90.0))
    return p, G.derived(p)


def rand_mask(ndivx, ndivy, seed, density=0.5):
    rng = np.random.default_rng(seed)
    m = np.zeros((ndivy, ndivx), np.int64)
    while m.mean() < density:
        w, h = rng.integers(16, 65, 2)
        x0, y0 = rng.integers(0, ndivx), rng.integers(0, ndivy)
        m[np.ix_((y0 + np.arange(h)) % ndivy, (x0 + np.arange(w)) % ndivx)] = 1
    return m.ravel()


# ---------------------------------------------------------------- fixtures
def test_inputlm_exact(golden):
    """Pupil enumeration at the reference default equals m3d/inputlm.csv."""
    p = P.REF_DEFAULT
    d = G.derived(p)
    lin, min_, snum = G.pupil_points(p, d)
    ref = np.loadtxt(os.path.join(golden, "ref_inputlm.csv"), delimiter=",", dtype=np.int64)
    assert ref.shape == (1901, 3)
    assert np.array_equal(ref[:, 0], lin) and np.array_equal(ref[:, 1], min_)
    assert np.array_equal(ref[:, 2], (snum >= 8).astype(np.int64))
    rows = maskio.read_inputxx(os.path.join(golden, "ref_inputxx.csv"))
    assert [len(r) for r in rows] == [1901, 1901, 1749, 1749, 1749, 1749]


def test_reference_order_counts():
    p = P.REF_DEFAULT
    li, mi, L, M = G.order_set(p, "diamond", cutoff=6.0)
    assert (L, M, len(li)) == (75, 75, 8705)
    li, mi, L, M = G.order_set(p, "ellipse", cutoff=1.5)
    assert (L, M, len(li)) == (18, 18, 1009)
    d = G.derived(p)
    nd = G.ndivs_rule(p, d)
    src = G.source_lattice(p, d, nd, 1, 0.9, 0.55, 90.0)
    assert nd == 1 and len(src[(0, 0)][0]) == 256


def test_mask_bin_roundtrip(golden):
    with open(os.path.join(golden, "ref_mask.bin"), "rb") as f:
        buf = f.read()
    m = maskio.unpack(buf, 2048, 2048)
    assert m.mean() == 0.63946533203125
    assert maskio.pack(m) == buf
    # bit order: MSB first
    assert maskio.pack(np.array([1, 0, 0, 0, 0, 0, 0, 1])) == bytes([0x81])


# ------------------------------------------------------------- mask spectrum
def test_maskamp_is_textbook_dft():
    """maskamp equals numpy.fft.fft2 of the y-flipped pattern on the window."""
    nx, ny, L, M = 64, 32, 5, 3
    m = rand_mask(nx, ny, 3)
    ea = (0.947 + 0.031j) ** 2
    tab = em.maskamp(m, nx, ny, nx, ny, L, M, ea, 1.0)
    img = np.where(m.reshape(ny, nx) == 0, ea, 1.0)  # [row jj][x]
    pat = img[::-1, :].T  # [x][y] with y counted from the bottom row
    F = np.fft.fft2(pat) / (nx * ny)
    ref = np.array([[F[l % nx, mm % ny] for mm in range(-2 * M, 2 * M + 1)] for l in range(-2 * L, 2 * L + 1)])
    assert np.abs(tab - ref).max() < 1e-15


def test_tables_affine_in_indicator():
    nx, ny, L, M = 64, 64, 4, 4
    m = rand_mask(nx, ny, 4)
    ea = (0.947 + 0.031j) ** 2
    mh = em.maskamp(m, nx, ny, nx, ny, L, M, 0.0, 1.0)
    dlt = np.zeros_like(mh)
    dlt[2 * L, 2 * M] = 1.0
    assert np.abs(em.maskamp(m, nx, ny, nx, ny, L, M, ea, 1.0) - (ea * dlt + (1 - ea) * mh)).max() < 1e-14
    assert np.abs(em.maskamp(m, nx, ny, nx, ny, L, M, 1 / ea, 1.0) - (dlt / ea + (1 - 1 / ea) * mh)).max() < 1e-14
    assert np.abs(em.maskamp(m, nx, ny, nx, ny, L, M, np.log(ea), 0.0) - np.log(ea) * (dlt - mh)).max() < 1e-14


# ------------------------------------------------------------------ multilayer
def parratt(k, kt, kxy2, layers, f_weight):
    """Textbook recursion on the up/down ratio X = B/U, bottom to top, with
    admittance al/f (f = k - kt^2/(k eps) when f_weight else k)."""
    X = 0.0 + 0j
    e0 = layers[0][0]
    al_b = np.sqrt(k * k * e0 - kxy2 + 0j)
    y_b = al_b / ((k - kt * kt / (k * e0)) if f_weight else k)
    for e, dth in layers[1:]:
        al = np.sqrt(k * k * e - kxy2 + 0j)
        y = al / ((k - kt * kt / (k * e)) if f_weight else k)
        r = (y - y_b) / (y + y_b)
        X = (r + X) / (1 + r * X)        # at the bottom of the new layer
        X = X * np.exp(2j * al * dth)    # carried to its top
        y_b = y
    return X


@pytest.mark.parametrize("stack", ["ref", "hna"])
def test_multilayer_vs_parratt(stack):
    st = P.REF_STACK if stack == "ref" else P.HNA_STACK
    k = 2 * math.pi / 13.5
    kx = np.array([0.0, 0.01, -0.05, 0.08])
    ky = np.array([-0.0486, 0.02, 0.07, -0.1])
    kxy2 = kx ** 2 + ky ** 2
    U, B = em.multilayer("X", k, kx, ky, kxy2, st)
    layers = P.layer_list(st)
    for i in range(len(kx)):
        X = parratt(k, kx[i], kxy2[i], layers, True)
        assert abs(B[i] / U[i] - X) < 1e-12
    # EUV multilayer reflectance near normal incidence (inside the Ru cap)
    R = abs(B[0] / U[0]) ** 2
    assert 0.6 < R < 0.8


def test_zero_thickness_interlayers_compose():
    """T(a,c) T(c,b) = T(a,b) when layer c has zero thickness."""
    k = 2 * math.pi / 13.5
    kx, ky = np.array([0.03]), np.array([-0.05])
    st = dict(P.HNA_STACK)
    U1, B1 = em.multilayer("Y", k, kx, ky, kx ** 2 + ky ** 2, st)
    st2 = dict(st, n_mosi2=1.3 + 0.2j, n_rusi=0.7 + 0.1j)  # materials of zero-thickness layers are irrelevant
    U2, B2 = em.multilayer("Y", k, kx, ky, kx ** 2 + ky ** 2, st2)
    assert abs(B1[0] / U1[0] - B2[0] / U2[0]) < 1e-13


# ------------------------------------------------------------------- EM core
@pytest.mark.parametrize("blank", [1, 0])
def test_blank_mask_specular_is_airy(blank):
    """Blank vacuum (1) / absorber (0) mask: D is diagonal and the specular
    amplitude of an l = 0 incidence (kx = 0, TE) equals the Airy reflection
    coefficient of vacuum | layer(60 nm) | multilayer."""
    p, d = small(256, 3)
    li, mi, L, M = G.order_set(p, "rect", K=3)
    n = p["NDIVX"] * p["NDIVY"]
    Ax = em.amp_s("X", np.full(n, blank), p, d, li, mi, L, M, 0.0, d["ky0"])
    k = d["k"]
    e_layer = 1.0 if blank else p["eabs"][0]
    layers = P.layer_list(p["stack"]) + [(e_layer, p["dabs"][0]), (1.0, 0.0)]
    checked = 0
    for ms in range(-d["lsmaxY"], d["lsmaxY"] + 1):
        if not G.in_na(0, ms, p, d):
            continue
        ky = d["ky0"] + 2 * math.pi * ms / d["dy"]
        X = parratt(k, 0.0, ky * ky, layers[:-1], False)
        # top interface vacuum/layer, no propagation in the vacuum half-space
        al = np.sqrt(k * k * e_layer - ky * ky + 0j)
        kz = math.sqrt(k * k - ky * ky)
        r = (kz - al) / (kz + al)
        ref = (r + X) / (1 + r * X)
        i = int(np.flatnonzero((li == 0) & (mi == ms))[0])
        got = Ax[d["lsmaxX"], ms + d["lsmaxY"], i]
        assert abs(got - ref) < 1e-11 * abs(ref)
        off = np.delete(Ax[d["lsmaxX"], ms + d["lsmaxY"]], i)
        assert np.abs(off).max() < 1e-13
        checked += 1
    assert checked >= 1


def test_eigenvector_scaling_and_order_invariance():
    p, d = small(256, 4)
    li, mi, L, M = G.order_set(p, "rect", K=4)
    m = rand_mask(256, 256, 7)
    A0 = em.amp_s("X", m, p, d, li, mi, L, M, d["kx0"], d["ky0"])
    rng = np.random.default_rng(1)

    def eig_scrambled(D):
        w, V = np.linalg.eig(D)
        perm = rng.permutation(len(w))
        s = rng.normal(size=len(w)) + 1j * rng.normal(size=len(w))
        return w[perm], V[:, perm] * s[perm][None, :]

    A1 = em.amp_s("X", m, p, d, li, mi, L, M, d["kx0"], d["ky0"], eig=eig_scrambled)
    assert np.abs(A1 - A0).max() < 1e-11 * np.abs(A0).max()


def _m3d_small(m, p, d, K):
    li, mi, L, M = G.order_set(p, "rect", K=K)
    Ax = em.amp_s("X", m, p, d, li, mi, L, M, d["kx0"], d["ky0"])
    vc, ab = OA.blank_amplitudes(p, d)
    return OA.m3d_params(p, d, m, li, mi, Ax, vc, ab)


def test_integer_shift_theorem():
    """Shifting the mask by dx pixels multiplies every M3D parameter at pupil
    order lp by exp(-2 pi i lp dx / NDIVX)."""
    p, d = small(512, 6)
    m = rand_mask(512, 512, 11)
    sh = 37
    m2 = np.roll(m.reshape(512, 512), sh, axis=1).ravel()
    a = _m3d_small(m, p, d, 6)
    b = _m3d_small(m2, p, d, 6)
    lin, min_, _ = G.pupil_points(p, d)
    ph = np.exp(-2j * math.pi * lin * sh / 512)
    for x, y in zip(a, b):
        assert np.abs(y - x * ph).max() < 1e-10 * max(np.abs(x).max(), 1e-30)


def test_x_mirror_symmetry():
    """Mirroring the mask in x (i -> -i mod N) maps a0(l,m) -> a0(-l,m),
    ax -> -ax(-l,m), ay -> ay(-l,m) (incidence in the y-z plane)."""
    p, d = small(512, 6)
    p = dict(p, theta0=-5.355)
    m = rand_mask(512, 512, 12).reshape(512, 512)
    mm = m[:, (-np.arange(512)) % 512]
    a0, ax, ay = _m3d_small(m.ravel(), p, d, 6)
    b0, bx, by = _m3d_small(mm.ravel(), p, d, 6)
    lin, min_, _ = G.pupil_points(p, d)
    idx = {(int(l), int(q)): i for i, (l, q) in enumerate(zip(lin, min_))}
    j = np.array([idx[(-int(l), int(q))] for l, q in zip(lin, min_)])
    s = np.abs(a0).max()
    assert np.abs(b0 - a0[j]).max() < 1e-9 * s
    assert np.abs(bx + ax[j]).max() < 1e-9 * s
    assert np.abs(by - ay[j]).max() < 1e-9 * s


# ------------------------------------------------------------------- imaging
def _fit_setup(seed=5):
    from tests import m3d_synth as S
    p, d = small(512, 6)
    li, mi, L, M = G.order_set(p, "rect", K=6)
    lin, min_, snum = G.pupil_points(p, d)
    rng = np.random.default_rng(seed)
    nsx, nsy = d["nsourceX"], d["nsourceY"]
    vc = rng.normal(size=(nsx, nsy)) + 1j * rng.normal(size=(nsx, nsy))
    ab = 0.3 * (rng.normal(size=(nsx, nsy)) + 1j * rng.normal(size=(nsx, nsy)))
    m = rand_mask(512, 512, 17)
    in_na = lambda l_, m_: G.in_na(l_, m_, p, d)
    geo = (lin, min_, li, mi, d["lsmaxX"], d["lsmaxY"], nsx, nsy, in_na)
    return S, p, d, li, mi, lin, min_, snum, vc, ab, m, geo, rng


def test_m3d_fit_recovers_linear_data_exactly():
    """Ax built as phase (a0 + ax ls + ay ms) per pupil point: the fit returns
    (a0, ax, ay) where snum >= 8 (3x3 normal equations by Cramer's rule,
    m3d.cpp:451-457) and a0 where snum < 8 (the mean, m3d.cpp:459-464), to
    rounding.  fit(0) carries the thin-mask term and is subtracted."""
    S, p, d, li, mi, lin, min_, snum, vc, ab, m, geo, rng = _fit_setup()
    n = len(lin)
    big = snum >= 8
    assert big.any() and (~big).any()  # both branches are exercised
    # The valid sources of a pupil point are symmetric under the point
    # reflection that swaps the two NA disks, so sum(ls) = sum(ms) = 0 exactly
    # and the cofactor terms carrying cx or cy never reach an output; the
    # cross moment sum(ls ms) is not zero, so its terms are exercised.
    cxy = []
    for lp, mp in zip(lin, min_):
        v = np.array([s[2:] for s in S.valid_sources(int(lp), int(mp), *geo[4:])])
        assert v[:, 0].sum() == 0 and v[:, 1].sum() == 0
        cxy.append(abs((v[:, 0] * v[:, 1]).sum()))
    assert max(cxy) > 0
    a = rng.normal(size=n) + 1j * rng.normal(size=n)
    x = np.where(big, rng.normal(size=n) + 1j * rng.normal(size=n), 0)
    y = np.where(big, rng.normal(size=n) + 1j * rng.normal(size=n), 0)
    phase = vc / np.abs(vc)
    Ax = S.linear_amplitudes(*geo, phase, a, x, y)
    got = OA.m3d_params(p, d, m, li, mi, Ax, vc, ab)
    zero = OA.m3d_params(p, d, m, li, mi, np.zeros_like(Ax), vc, ab)
    for g, z, t in zip(got, zero, (a, x, y)):
        assert np.abs((g - z) - t).max() < 1e-12 * np.abs(t).max()
    assert np.abs(got[1][~big]).max() == 0 and np.abs(got[2][~big]).max() == 0


def test_m3d_fit_single_row_sources():
    """768 nm, +-16 orders: pupil points where snum >= 8 sources lie on one
    row of the lattice make the reference's 3 x 3 system singular
    (m3d.cpp:451-457 divides by zero there).  The fit takes the slope along
    the row with a0 and sets the other to 0: exact recovery of data without
    the unidentifiable slope, and every output finite."""
    from tests import m3d_synth as S
    p = P.hna(768, 16, dict(type=1, sigma1=0.8, sigma2=0.3, openangle=90.0))
    d = G.derived(p)
    li, mi, L, M = G.order_set(p, "rect", K=16)
    lin, min_, snum = G.pupil_points(p, d)
    nsx, nsy = d["nsourceX"], d["nsourceY"]
    in_na = lambda l_, m_: G.in_na(l_, m_, p, d)
    ok = S.identifiable(lin, min_, snum, d["lsmaxX"], d["lsmaxY"], nsx, nsy, in_na)
    big = snum >= 8
    single = big & (ok.sum(axis=1) == 1)
    assert single.sum() >= 8
    rng = np.random.default_rng(8)
    n = len(lin)
    a = rng.normal(size=n) + 1j * rng.normal(size=n)
    x = np.where(ok[:, 0], rng.normal(size=n) + 1j * rng.normal(size=n), 0)
    y = np.where(ok[:, 1], rng.normal(size=n) + 1j * rng.normal(size=n), 0)
    vc = rng.normal(size=(nsx, nsy)) + 1j * rng.normal(size=(nsx, nsy))
    ab = 0.3 * vc
    m = rand_mask(768, 768, 5)
    Ax = S.linear_amplitudes(lin, min_, li, mi, d["lsmaxX"], d["lsmaxY"], nsx, nsy, in_na, vc / np.abs(vc), a, x, y)
    got = OA.m3d_params(p, d, m, li, mi, Ax, vc, ab)
    zero = OA.m3d_params(p, d, m, li, mi, np.zeros_like(Ax), vc, ab)
    for g, z, t in zip(got, zero, (a, x, y)):
        assert np.isfinite(g).all()
        assert np.abs((g - z) - t).max() < 1e-12 * np.abs(t).max()


def test_m3d_fit_equals_lstsq():
    """On random amplitudes the fit equals numpy.linalg.lstsq on [1, ls, ms]
    (snum >= 8) and the mean (snum < 8)."""
    S, p, d, li, mi, lin, min_, snum, vc, ab, m, geo, rng = _fit_setup(6)
    nsx, nsy = d["nsourceX"], d["nsourceY"]
    Ax = rng.normal(size=(nsx, nsy, len(li))) + 1j * rng.normal(size=(nsx, nsy, len(li)))
    phase = vc / np.abs(vc)
    got = OA.m3d_params(p, d, m, li, mi, Ax, vc, ab)
    zero = OA.m3d_params(p, d, m, li, mi, np.zeros_like(Ax), vc, ab)
    want = S.lstsq_fit(lin, min_, snum, li, mi, d["lsmaxX"], d["lsmaxY"], nsx, nsy,
                       lambda l_, m_: G.in_na(l_, m_, p, d), phase, Ax)
    for g, z, t in zip(got, zero, want):
        assert np.abs((g - z) - t).max() < 1e-12 * np.abs(want[0]).max()


def test_abbe_grid_refinement_subsamples():
    p, d = small(256, 4)
    li, mi, L, M = G.order_set(p, "rect", K=4)
    m = rand_mask(256, 256, 5)
    nd = G.ndivs_rule(p, d)
    srcs = G.source_lattice(p, d, 1, 1, 0.8, 0.3, 90.0)
    divs = [((d["kx0"], d["ky0"], 0.0, 0.0),) + srcs[(0, 0)]]
    Ax = [em.amp_s("X", m, p, d, li, mi, L, M, d["kx0"], d["ky0"])]
    i1 = OA.abbe_image(p, d, m, li, mi, L, M, divs, 64, 64, amps=Ax)
    i2 = OA.abbe_image(p, d, m, li, mi, L, M, divs, 128, 128, amps=Ax)
    assert nd >= 1
    assert np.abs(i2[::2, ::2] - i1).max() < 1e-12 * i1.max()


def test_abbe_blank_mask_uniform():
    """Blank vacuum mask: every source point sees a single plane wave, so the
    image is flat."""
    p, d = small(256, 3)
    li, mi, L, M = G.order_set(p, "rect", K=3)
    n = 256 * 256
    srcs = G.source_lattice(p, d, 1, 1, 0.8, 0.3, 90.0)
    divs = [((d["kx0"], d["ky0"], 0.0, 0.0),) + srcs[(0, 0)]]
    img = OA.abbe_image(p, d, np.ones(n, np.int64), li, mi, L, M, divs, 32, 32)
    assert img.std() < 1e-12 * img.mean()


# ------------------------------------------------------------ end to end (slow)
@pytest.mark.slow
def test_reference_default_end_to_end(golden, tmp_path):
    """Oracle at the reference-default optics vs m3d/inputxx.csv and
    abbe/emint.csv (one N = 8705 eigenproblem; ~15 min on 6 cores), at the
    2e-5 bar that the fixtures' 6 significant digits support, on the masks as
    the fixtures' program revision read them (tests/golden/fixture_reads.py)."""
    from tests.golden.ref_default_e2e import run
    rep = run(golden, out=str(tmp_path / "ref_default_e2e.json"))
    assert all(e < 2e-5 for e in rep["inputxx_rel_per_row"])
    assert rep["emint_max_rel"] < 2e-5


# ----------------------------------------------- images from M3D parameters
def _imaging_setup(ndiv=256, K=3):
    from oracle import imaging as OI
    p, d = small(ndiv, K)
    m = rand_mask(ndiv, ndiv, 21)
    ampvc, ampab, ph0 = OI.reflect_amplitude(p, d)
    fmask = OI.pattern_spectrum(m, p, d, ampvc, ampab)
    famp = OI.thin_amplitudes(fmask, p, d, ph0)
    dsx, dsy = G.mesh_source(p, d, 1, 0.8, 0.3, 90.0, mesh_deg=1.0)
    return OI, p, d, m, famp, dsx, dsy


def test_pattern_spectrum_is_affine_in_indicator():
    OI, p, d, m, famp, dsx, dsy = _imaging_setup()
    ampvc, ampab, _ = OI.reflect_amplitude(p, d)
    fm = OI.pattern_spectrum(m, p, d, ampvc, ampab)
    mh = em.maskamp(m, p["NDIVX"], p["NDIVY"], p["NDIVX"], p["NDIVY"], 0, 0, 0.0, 1.0)  # mean only
    ind = OI.pattern_spectrum(m, p, d, 1.0, 0.0)
    dl = np.zeros_like(ind)
    dl[d["lpmaxX"], d["lpmaxY"]] = 1.0
    assert abs(ind[d["lpmaxX"], d["lpmaxY"]] - mh[0, 0]) < 1e-15
    assert np.abs(fm - (ampab * dl + (ampvc - ampab) * ind)).max() < 1e-15


def test_linear_image_reduces_to_thin_mask():
    """a0 = ax = ay = 0 => intlinear == intft (linear.py:166-169)."""
    OI, p, d, m, famp, dsx, dsy = _imaging_setup()
    lin, _, _ = G.pupil_points(p, d)
    z = np.zeros(len(lin), complex)
    il, ift = OI.thin_linear_images(p, d, famp, z, z, z, dsx, dsy, 64, 64)
    assert np.abs(il - ift).max() == 0.0
    assert ift.max() > 0


def test_socs_all_kernels_equals_abbe():
    """Hopkins/SOCS identity: with every kernel, ax = ay = 0 and an alias-free
    grid, the SOCS image equals the Abbe image of the same mesh source."""
    OI, p, d, m, famp, dsx, dsy = _imaging_setup()
    lin, _, _ = G.pupil_points(p, d)
    rng = np.random.default_rng(3)
    a0 = 0.01 * (rng.normal(size=len(lin)) + 1j * rng.normal(size=len(lin)))
    z = np.zeros(len(lin), complex)
    il, _ = OI.thin_linear_images(p, d, famp, a0, z, z, dsx, dsy, 64, 64)
    kern = OI.tcc_kernels(p, d, dsx, dsy)
    for T in (kern[0], kern[2], kern[4]):
        assert np.all(np.diff(np.abs(T)) <= 0)  # sorted by |alpha| descending
    img = OI.socs_image(p, d, famp, a0, z, z, kern, 64, 64, nsocs=len(lin), nsocsxy=0)
    assert np.abs(img - il).max() < 1e-12 * il.max()


def test_socs_linear_terms_are_first_order_exact():
    """The STCC terms (ax, ay) reproduce the linear-model Abbe image to first
    order: with every kernel, SOCS - Abbe scales as eps^2 when ax, ay -> eps ax, eps ay."""
    OI, p, d, m, famp, dsx, dsy = _imaging_setup()
    lin, _, snum = G.pupil_points(p, d)
    rng = np.random.default_rng(4)
    n = len(lin)
    a0 = 0.01 * (rng.normal(size=n) + 1j * rng.normal(size=n))
    ax0 = 0.02 * (rng.normal(size=n) + 1j * rng.normal(size=n))
    ay0 = 0.02 * (rng.normal(size=n) + 1j * rng.normal(size=n))
    kern = OI.tcc_kernels(p, d, dsx, dsy)
    diffs = []
    for eps in (0.01, 0.02):
        il, _ = OI.thin_linear_images(p, d, famp, a0, eps * ax0, eps * ay0, dsx, dsy, 64, 64)
        img = OI.socs_image(p, d, famp, a0, eps * ax0, eps * ay0, kern, 64, 64, nsocs=n, nsocsxy=n)
        diffs.append(np.abs(img - il).max())
    assert diffs[0] > 0 and abs(diffs[1] / diffs[0] - 4.0) < 0.05

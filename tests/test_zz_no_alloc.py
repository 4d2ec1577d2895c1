"""GPU: the allocation-free mask spectrum (euvl_mask_spectrum_noalloc) against
euvl_mask_spectrum through the C ABI, on two calls in a row (the second with
a smaller input, so the fixed scratch holds the first call's contents), to
1e-14 of the largest coefficient (the y sums may be split differently); a
request above the scratch maximum is refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def lib_h():
    if not torch.cuda.is_available():
        pytest.fail("GPU test without a GPU")
    from euvlitho_amd import _lib
    lib = _lib.load()
    torch.cuda.init()
    h = ctypes.c_void_p()
    assert lib.euvl_create(ctypes.byref(h)) == 0
    yield lib, h
    lib.euvl_destroy(h)


def spectra(lib, h, nx, ny, nc, lw, mw, seed):
    from euvlitho_amd.masks import manhattan, pack_bits
    raw = b"".join(pack_bits(manhattan(nx, ny, 0.5, seed + c)) for c in range(nc))
    bits = torch.from_numpy(np.frombuffer(raw, dtype=np.uint8).copy()).cuda()
    shape = (nc, 2 * lw + 1, 2 * mw + 1)
    a = torch.zeros(shape, dtype=torch.complex128, device="cuda")
    b = torch.full(shape, np.nan, dtype=torch.complex128, device="cuda")
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    assert lib.euvl_mask_spectrum(h, P(bits), nx, ny, nc, lw, mw, P(a)) == 0, lib.euvl_last_error()
    rc = lib.euvl_mask_spectrum_noalloc(h, P(bits), nx, ny, nc, lw, mw, P(b))
    assert rc == 0, lib.euvl_last_error()
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def test_noalloc_equals_allocating_entry_point(lib_h):
    lib, h = lib_h
    say(f"scratch maximum {lib.euvl_mask_spectrum_noalloc_max() / 2**20:.0f} MiB")
    for (nx, ny, nc, lw, mw, seed) in [(1024, 1024, 2, 40, 40, 1), (512, 256, 1, 12, 9, 7)]:
        say(f"  {nc} clips of {nx} x {ny}, window {2 * lw + 1} x {2 * mw + 1}")
        a, b = spectra(lib, h, nx, ny, nc, lw, mw, seed)
        assert np.isfinite(b).all()
        e = np.abs(a - b).max() / np.abs(a).max()
        say(f"    max difference / max |coefficient| = {e:.1e}")
        assert e <= 1e-14


def test_noalloc_refuses_above_the_maximum(lib_h):
    lib, h = lib_h
    nbytes = 2048 * 2048 // 8 * 8
    bits = torch.zeros(nbytes, dtype=torch.uint8, device="cuda")
    out = torch.zeros((8, 301, 301), dtype=torch.complex128, device="cuda")
    rc = lib.euvl_mask_spectrum_noalloc(h, ctypes.c_void_p(bits.data_ptr()), 2048, 2048, 8, 150, 150,
                                        ctypes.c_void_p(out.data_ptr()))
    assert rc == -1 and b"maximum" in lib.euvl_last_error()

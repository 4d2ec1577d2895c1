"""CPU: the CNN predictor's oracle pinned to independent implementations of
the same layers (torch.nn on the CPU, as the reference's model is built; an
FFT circular correlation; the reference's own float16 compmask procedure)."""
import numpy as np
import pytest

from oracle import cnn as OC

torch = pytest.importorskip("torch")


def test_compmask_block_mean_is_the_reference_procedure():
    """compmask.py's float16 rowmat . mask0 . colmat == the block means (all
    values are multiples of 1/f^2, exact in float16)."""
    rng = np.random.default_rng(3)
    for ndiv, nout in ((64, 16), (256, 64), (128, 128)):
        bits = np.packbits(rng.integers(0, 2, ndiv * ndiv).astype(np.uint8)).tobytes()
        lit = OC.compmask_literal(bits, ndiv, nout).astype(np.float64)
        np.testing.assert_array_equal(OC.compmask(bits, ndiv, nout), lit)


def test_circular_conv_matches_fft_correlation():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((3, 16, 24))
    w = rng.standard_normal((5, 3, 3, 3))
    b = rng.standard_normal(5)
    got = OC.conv3x3_circular(x, w, b)
    # cross-correlation with the 3 x 3 kernel centred at (1, 1), periodic: via FFT
    want = np.zeros_like(got)
    for o in range(5):
        acc = np.zeros((16, 24), dtype=complex)
        for c in range(3):
            k = np.zeros((16, 24))
            for dy in range(3):
                for dx in range(3):
                    k[(dy - 1) % 16, (dx - 1) % 24] = w[o, c, dy, dx]
            acc += np.fft.ifft2(np.fft.fft2(x[c]) * np.conj(np.fft.fft2(k)))
        want[o] = acc.real + b[o]
    np.testing.assert_allclose(got, want, atol=1e-12)


@pytest.mark.parametrize("nin,n_out", [(64, 1901), (32, 1749)])
def test_oracle_forward_matches_torch_nn(nin, n_out):
    """The oracle's network == torch.nn's (circular padding, cross-correlation,
    ReLU before the pool, BatchNorm eval formula, flatten order, Linear
    orientation), float64."""
    p = OC.random_params(n_out, 7, nin=nin)
    x = np.random.default_rng(8).random((2, 1, nin, nin))
    want = OC.torch_model(p, n_out, nin)(torch.from_numpy(x)).detach().numpy()
    got = OC.forward(p, x)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())

"""Pins of the CZT oracle (CPU): Bluestein against the defining sum, the
DFT as a special case, the zoomed FFT against a brute-force band-limited
interpolation, the CZT propagation against the FFT propagation on the same
grid."""
import numpy as np
import pytest

from oracle import czt as oc
from oracle import propagation as op


@pytest.mark.parametrize("n,m", [(16, 16), (37, 5), (8, 29), (128, 64)])
def test_bluestein_matches_direct_sum(n, m):
    rng = np.random.default_rng(n + m)
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    a = np.exp(-1j * 0.37)
    w = np.exp(1j * 2 * np.pi / (3.3 * m))
    np.testing.assert_allclose(oc.czt(x, m, a, w), oc.czt_direct(x, m, a, w), atol=1e-9 * n)


def test_czt_with_dft_parameters_is_the_fft():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((4, 50)) + 1j * rng.standard_normal((4, 50))
    y = oc.czt_direct(x, 50, 1.0, np.exp(-2j * np.pi / 50))
    np.testing.assert_allclose(y, np.fft.fft(x, axis=-1), atol=1e-10)


def test_zoomed_fft_full_range_is_the_fft():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((12, 20)) + 1j * rng.standard_normal((12, 20))
    # k from 0 to 2 pi without the end point: the ordinary DFT with e^{+i}
    y = oc.zoomed_fft(x, 0.0, 2 * np.pi, (12, 20), include_end=False)
    np.testing.assert_allclose(y, np.conj(np.fft.fft2(np.conj(x))), atol=1e-10)


def test_czt_propagation_on_the_input_grid():
    """Output grid = input grid, even m = N, dx = 1 (so alpha = N dx^2 = N,
    the centred inverse DFT's 1/N per axis).  The reference's modulation
    uses N' = (m - 1) // 2 = N/2 - 1 instead of the centred index N/2, which
    leaves a ramp exp(i 2 pi r / T) per axis: removing it gives the plain ASM
    propagation exactly."""
    rng = np.random.default_rng(2)
    n, dx = 32, 1.0
    u = np.exp(2j * np.pi * rng.random((n, n)))
    out, _, _ = oc.czt_propagate(u, (dx, dx), 0.532, 10.0, 1.33, dx, (n, n))
    r = (np.arange(n) - n / 2) * dx
    ramp = np.exp(2j * np.pi * r / (n * dx))
    ref = op.propagate(u, (dx, dx), 0.532, 10.0, 1.33, 0)
    got = out * np.conj(ramp[:, None] * ramp[None, :])
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-10

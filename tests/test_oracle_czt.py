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


@pytest.mark.parametrize("shape", [(16, 16), (15, 12), (20, 9)])
@pytest.mark.parametrize("odx,shift", [((0.2, 0.35), (1.0, -0.5)), ((0.3, 0.3), (0.0, 0.0))])
def test_scaled_convolution_fft_form_matches_double_sum(shape, odx, shift):
    """The reference's FFT-convolution steps (propagation.py:444-458) against
    the scaled convolution written as an explicit double sum."""
    rng = np.random.default_rng(sum(shape))
    u = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    H = op.asm_propagator(shape, (0.3, 0.25), 0.532, 5.0, 1.33)
    a = oc.yu_kernel_propagate(u, H, (0.3, 0.25), odx, shift)
    b = oc.yu_kernel_propagate(u, H, (0.3, 0.25), odx, shift, direct=True)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


@pytest.mark.parametrize("n", [8, 16, 24])
def test_scaled_convolution_is_a_windowed_dft(n):
    """For an even length and the grid at the origin, the chirps combine into
    the DFT kernel exp(2 pi i x_j f_k) times the linear phase of the "same"
    window's one-sample offset: with a = j - n/2, b = k - n/2, c = dx df,
    y_j = (dx / alpha) e^{-i pi c} sum_{0 <= j + n/2 - 1 - k < n} B_k
          e^{2 pi i c (a b + a - b)}."""
    rng = np.random.default_rng(n)
    B = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dx_in, out_dx = 0.3, 0.45
    df = 1.0 / (n * dx_in)
    alpha = out_dx / df
    f = (np.arange(n) - n // 2) * df
    x = (np.arange(n) - n / 2) * out_dx
    got = oc.yu_axis_direct(B[None, :], alpha, f, x, out_dx, -1)[0]
    c = out_dx * df
    ref = np.zeros(n, complex)
    for j in range(n):
        for k in range(n):
            if 0 <= j + n // 2 - 1 - k < n:
                a, b = j - n / 2, k - n / 2
                ref[j] += B[k] * np.exp(2j * np.pi * c * (a * b + a - b))
    ref *= out_dx / alpha * np.exp(-1j * np.pi * c)
    np.testing.assert_allclose(got, ref, atol=1e-12 * np.abs(ref).max())


def test_high_na_lens_matches_direct_sum():
    """high_na_ff_lens (lenses.py:98-156) against the defocus factor and the
    zoomed DFT written out as sums of exp(i j k_p) over each axis."""
    rng = np.random.default_rng(3)
    n, dx, lam, nm, NA, f = 12, 0.2, 0.532, 1.33, 0.8, 40.0
    u = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    got, odx = oc.high_na_ff_lens(u, (dx, dx), lam, f, nm, NA, output_shape=(7, 9), output_dx=0.15)
    fr = (np.arange(n) - n // 2) / (n * dx)
    s2 = (fr[:, None] ** 2 + fr[None, :] ** 2) * (lam / nm) ** 2
    sz = np.sqrt(np.maximum(1 - NA**2 * s2, 0))
    d = np.zeros_like(u)
    ok = sz > 0
    d[ok] = np.exp(1j * 2 * np.pi * nm / lam * sz[ok] * f) / sz[ok]
    zoom = 2 * NA * 7 * 0.15 / ((n - 1) * lam)
    ky = -zoom * np.pi + np.arange(7) * (2 * zoom * np.pi / 6)
    kx = -zoom * np.pi + np.arange(9) * (2 * zoom * np.pi / 8)
    j = np.arange(n)
    Ey = np.exp(1j * np.outer(ky, j))  # (7, n)
    Ex = np.exp(1j * np.outer(j, kx))  # (n, 9)
    ref = Ey @ (u * d) @ Ex
    np.testing.assert_allclose(got, ref, atol=1e-10 * np.abs(ref).max())
    np.testing.assert_allclose(odx, [0.15, 0.15])

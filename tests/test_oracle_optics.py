"""Pins of the optical-FFT / transform / SAS oracle (CPU): the reference
tests' closed forms and invariants, and an analytic Gaussian."""
import numpy as np
import pytest

from oracle import optics as oo
from oracle import propagation as op
from oracle import sources as os_
from tests.conftest import rel_l2

D, Z, LAM, N_MED = 40.0, 100.0, 0.532, 1.33


def _square(shape, width, dx):
    g = op.grid(shape, dx)
    u = ((np.abs(g[..., 0]) <= width / 2) & (np.abs(g[..., 1]) <= width / 2)).astype(np.complex128)
    return u / np.sqrt(np.sum(np.abs(u) ** 2) * dx[0] * dx[1])


@pytest.mark.parametrize(("shape", "pad"), [((256, 256), (512, 512)), ((1024, 256), (256, 512))])
def test_transform_square_aperture(shape, pad):
    """tests/test_propagate.py:36-52."""
    dx = (D / shape[0], D / shape[1])
    out, du = oo.transform_propagate(_square(shape, D, dx), dx, LAM, Z, N_MED, pad)
    I_num = np.abs(out) ** 2
    I_ana = np.abs(op.square_aperture_field(op.grid(shape, du), Z, D, LAM, N_MED)) ** 2
    assert np.mean((I_ana - I_num) ** 2) / np.mean(I_ana**2) < 2e-2


@pytest.mark.parametrize("shape", [(256, 256), (1024, 256)])
def test_sas_square_aperture(shape):
    """tests/test_propagate.py:60-76."""
    dx = (D / shape[0], D / shape[1])
    out, du = oo.transform_propagate_sas(_square(shape, D, dx), dx, LAM, Z, N_MED)
    I_num = np.abs(out) ** 2
    I_ana = np.abs(op.square_aperture_field(op.grid(shape, du), Z, D, LAM, N_MED)) ** 2
    assert np.mean((I_ana - I_num) ** 2) / np.mean(I_ana**2) < 2e-2


def test_transform_round_trip():
    """tests/test_propagate.py:53-57: z then -z at pad 0."""
    shape = (256, 256)
    dx = (D / 256, D / 256)
    u = _square(shape, D / 10, dx)
    out, du = oo.transform_propagate(u, dx, LAM, Z, N_MED, 0)
    back, d2 = oo.transform_propagate(out, du, LAM, -Z, N_MED, 0)
    np.testing.assert_allclose(d2, dx, rtol=1e-12)
    np.testing.assert_allclose(back, u, rtol=2e-5, atol=1e-8)


def test_ff_lens_twice_and_inverse():
    """tests/test_lenses.py:6-22."""
    u = os_.objective_point_source((512, 512), (0.3, 0.3), 0.532, 0.0, f=10.0, n=1.0, NA=0.8)
    u2, d2 = oo.ff_lens(u, (0.3, 0.3), 0.532, 10.0, 1.0)
    u3, d3 = oo.ff_lens(u2, d2, 0.532, 10.0, 1.0)
    u2b, _ = oo.ff_lens(u3, d3, 0.532, 10.0, 1.0, inverse=True)
    np.testing.assert_allclose(np.abs(u2) ** 2, np.abs(u2b) ** 2, atol=1e-5)
    assert u3[256, 256] != 0.0


def test_optical_fft_gaussian_closed_form():
    """The lens FT of exp(-|r|^2/w^2) is -i/L2 * pi w^2 exp(-pi^2 w^2 |x'|^2/L2^2)."""
    shape, dx, w, f, n = (256, 256), (0.25, 0.25), 4.0, 300.0, 1.0
    g = op.grid(shape, dx)
    u = np.exp(-(g[..., 0] ** 2 + g[..., 1] ** 2) / w**2)
    out, du = oo.optical_fft(u, dx, LAM, f, n)
    L2 = LAM * f / n
    gp = op.grid(shape, du)
    expect = -1j / L2 * np.pi * w**2 * np.exp(-np.pi**2 * w**2 * (gp[..., 0] ** 2 + gp[..., 1] ** 2) / L2**2)
    assert rel_l2(out, expect) < 1e-6
    # z < 0 inverts z > 0 up to a global sign: the reference's norms give
    # (-i)(-i) = -1 (convenience.py:43-58), invisible in intensities
    back, d2 = oo.optical_fft(out, du, LAM, -f, n)
    np.testing.assert_allclose(d2, dx, rtol=1e-12)
    assert rel_l2(back, -u) < 1e-12


@pytest.mark.parametrize("z", [30.0, -30.0])
def test_ref32_close_to_fp64(z):
    shape, dx = (128, 96), (0.3, 0.25)
    rng = np.random.default_rng(0)
    u = np.exp(2j * np.pi * rng.random(shape))
    a, _ = oo.transform_propagate(u, dx, LAM, z, N_MED, 16)
    b, _ = oo.transform_propagate(u, dx, LAM, z, N_MED, 16, precision="ref32")
    assert 0 < rel_l2(b, a) < 1e-4
    a, _ = oo.transform_propagate_sas(u, dx, LAM, z, N_MED)
    b, _ = oo.transform_propagate_sas(u, dx, LAM, z, N_MED, precision="ref32")
    assert 0 < rel_l2(b, a) < 1e-3


def test_point_source_power_and_pupil():
    """tests/test_sources.py:73-79: power normalised to 1; pupil support."""
    u = os_.objective_point_source((256, 256), (0.3, 0.3), 0.532, [-1.0, 0.0, 2.0], 100.0, 1.33, 0.8)
    assert u.shape == (3, 256, 256)
    np.testing.assert_allclose(os_.power(u, (0.3, 0.3))[:, 0, 0], 1.0, rtol=1e-12)
    D = 2 * 100.0 * 0.8 / 1.33
    g = op.grid((256, 256), (0.3, 0.3))
    outside = g[..., 0] ** 2 + g[..., 1] ** 2 > (D / 2) ** 2
    assert np.all(u[:, outside] == 0) and np.all(np.abs(u[:, ~outside]) > 0)

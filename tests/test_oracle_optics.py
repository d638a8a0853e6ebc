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


# ---- SAS against the exact angular-spectrum integral ------------------------------
def _exact_as_gaussian(r, z, w, lam, n, fmax, m=10001):
    """Angular-spectrum field of u0 = exp(-|r|^2 / w^2) at distance z, by the
    Hankel transform of its closed-form spectrum pi w^2 exp(-pi^2 w^2 rho^2):
    u(r) = 2 pi int rho U(rho) exp(i 2 pi |z| sqrt((n/lam)^2 - rho^2)) J0(2 pi rho r) drho
    (conjugated phase for z < 0), Simpson's rule on [0, fmax]."""
    from scipy.integrate import simpson
    from scipy.special import j0

    rho = np.linspace(0.0, fmax, m)
    U = np.pi * w**2 * np.exp(-((np.pi * w * rho) ** 2))
    H = np.exp(2j * np.pi * abs(z) * np.sqrt((n / lam) ** 2 - rho**2 + 0j))
    if z < 0:
        H = np.conj(H)
    return 2 * np.pi * simpson(j0(2 * np.pi * np.outer(r, rho)) * (rho * U * H)[None, :], x=rho, axis=1)


def _sas_vs_exact(n_px, dx, w, z, lam, n):
    """rel. L2 of transform_propagate_sas of a sampled Gaussian against the
    exact angular-spectrum field on the SAS output grid.  The reference's
    Fresnel transform leaves out the constant exp(i k |z|) / i (conjugated for
    z < 0), so the SAS output times that constant is the physical field."""
    y = (np.arange(n_px) - n_px / 2) * dx
    u = np.exp(-(y[:, None] ** 2 + y[None, :] ** 2) / w**2)
    out, du = oo.transform_propagate_sas(u, (dx, dx), lam, z, n)
    yo = (np.arange(n_px) - n_px / 2) * np.atleast_1d(du)[0]
    R = np.sqrt(yo[:, None] ** 2 + yo[None, :] ** 2)
    radii, inv = np.unique(np.round(R, 12), return_inverse=True)
    exact = _exact_as_gaussian(radii, z, w, lam, n, min(1 / (2 * dx), 12 / (np.pi * w)))
    exact = exact[inv].reshape(R.shape)
    c = np.exp(2j * np.pi * n * abs(z) / lam) / 1j
    c = c if z >= 0 else np.conj(c)
    return rel_l2(out * c, exact)


# (pixels, dx, waist, z, lambda, n): paraxial and wide-angle forward cases,
# and a back-propagation whose output grid resolves the beam
_SAS_CASES = [(64, 0.5, 2.0, 100.0, 0.532, 1.0), (64, 0.25, 1.0, 40.0, 0.532, 1.33),
              (64, 0.5, 3.0, -100.0, 0.532, 1.0)]


@pytest.mark.parametrize("case", _SAS_CASES)
def test_sas_matches_exact_angular_spectrum(case):
    assert _sas_vs_exact(*case) < 1e-8


def test_sas_exact_pin_fails_on_planted_mistakes(monkeypatch):
    """The same comparison with the precompensation's sign flipped
    (exp(-i kz (H_AS - H_Fr))) or left out (plain Fresnel transform) fails."""
    orig = oo.sas_precompensation
    monkeypatch.setattr(oo, "sas_precompensation", lambda *a, **k: np.conj(orig(*a, **k)))
    assert min(_sas_vs_exact(*c) for c in _SAS_CASES[:2]) > 1e-2
    monkeypatch.setattr(oo, "sas_precompensation", lambda *a, **k: np.ones_like(orig(*a, **k)))
    assert min(_sas_vs_exact(*c) for c in _SAS_CASES[:2]) > 5e-3


def _sas_window_error(prec):
    """Bins where the precompensation's window W differs from the band
    |s| <= s* per axis, s* the root of s^2 (2 + 1/t^2) = 1 with
    t = extent/|z| + |s| (propagation.py:85-103: L = 2 * extent, t = L/2/|z| + |s|),
    found by bisection; anisotropic grids, both signs of z."""
    from scipy.optimize import brentq

    bad = 0
    for shape, sp_, lam, n, z in (((96, 128), (0.2, 0.15), 0.532, 1.33, 30.0),
                                  ((128, 64), (0.25, 0.3), 0.45, 1.0, -12.0),
                                  ((64, 64), (0.18, 0.18), 0.65, 1.33, 80.0)):
        dH = prec(shape, sp_, lam, z, n)
        keep = []
        for nn, d in zip(shape, sp_):
            a = nn * d / abs(z)
            s_star = brentq(lambda s: s * s * (2 + 1 / (a + s) ** 2) - 1, 1e-12, 1.0)
            keep.append(np.abs(lam * np.fft.fftfreq(nn, d) / n) <= s_star)
        want = keep[0][:, None] & keep[1][None, :]
        assert 0 < want.sum() < want.size
        bad += int(np.sum((np.abs(dH) > 0) != want))
    return bad


def test_sas_window_edge():
    assert _sas_window_error(oo.sas_precompensation) == 0
    # planted mistakes: L = extent instead of 2 * extent (same as doubling z
    # in t), and the spacings of the two axes exchanged
    assert _sas_window_error(lambda sh, sp_, lam, z, n: oo.sas_precompensation(sh, sp_, lam, 2 * z, n)) > 0
    assert _sas_window_error(
        lambda sh, sp_, lam, z, n: oo.sas_precompensation(sh, sp_[::-1], lam, z, n)) > 0


def test_sas_window_keeps_only_frequencies_that_stay_inside_the_padding():
    """Physical sufficiency of W: the precompensation moves frequency s by
    |z| |s| (1/sqrt(1 - s^2) - 1) sideways (its group delay); every kept
    frequency moves less than the padded extent L/2."""
    for shape, sp_, z in (((128, 128), (0.15, 0.15), 20.0), ((96, 64), (0.2, 0.18), -9.0)):
        dH = oo.sas_precompensation(shape, sp_, 0.532, z, 1.33)
        for ax, (nn, d) in enumerate(zip(shape, sp_)):
            s = np.abs(0.532 * np.fft.fftfreq(nn, d) / 1.33)
            kept = (np.abs(dH) > 0).any(axis=1 - ax)
            shift = abs(z) * s[kept] * (1 / np.sqrt(1 - s[kept] ** 2) - 1)
            assert np.all(shift <= nn * d)
            assert not kept.all()  # the window is active in these cases

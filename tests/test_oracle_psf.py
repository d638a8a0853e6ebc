"""Pins of the 4f PSF oracle (CPU): finite differences for the hand-written
adjoint, Parseval for the lens transform, the reference's spacing rule."""
import numpy as np
import pytest

from oracle import psf as op
from oracle import sources as os_

LAM, N_IMM, F_OBJ, NA, F_TUBE = 0.532, 1.33, 8000.0, 0.6, 200000.0


def _setup(n=16, nz=2, seed=0):
    rng = np.random.default_rng(seed)
    dx = op.required_spacing(n, 6.5 * 32 / n, F_TUBE, N_IMM, LAM)
    phase = rng.uniform(-np.pi, np.pi, (n, n))
    z = np.linspace(-5.0, 5.0, nz)
    return phase, (dx, dx), z, rng


def test_gradient_matches_finite_differences():
    phase, sp, z, rng = _setup()
    target = rng.uniform(0, 1e-9, (len(z),) + phase.shape)
    args = (sp, LAM, z, F_OBJ, N_IMM, NA, F_TUBE)
    loss, g = op.psf4f_mse_grad(phase, target, *args)
    eps = 1e-6
    for (i, j) in [(8, 8), (5, 9), (10, 3), (7, 11)]:
        p1, p2 = phase.copy(), phase.copy()
        p1[i, j] += eps
        p2[i, j] -= eps
        l1 = np.mean((op.psf4f(p1, *args) - target) ** 2)
        l2 = np.mean((op.psf4f(p2, *args) - target) ** 2)
        fd = (l1 - l2) / (2 * eps)
        assert abs(fd - g[i, j]) <= 1e-6 * np.max(np.abs(g)) + 1e-7 * abs(fd), (i, j, fd, g[i, j])


def test_gradient_with_upstream_cotangent_is_linear():
    phase, sp, z, rng = _setup(n=20, nz=3, seed=1)
    args = (sp, LAM, z, F_OBJ, N_IMM, NA, F_TUBE)
    a = rng.standard_normal((3, 20, 20))
    b = rng.standard_normal((3, 20, 20))
    ga = op.psf4f_grad(phase, a, *args)
    gb = op.psf4f_grad(phase, b, *args)
    gab = op.psf4f_grad(phase, 2 * a - 3 * b, *args)
    np.testing.assert_allclose(gab, 2 * ga - 3 * gb, rtol=1e-10, atol=1e-12 * np.abs(gab).max())
    # and matches a finite difference of <dL/dI, I>
    eps = 1e-6
    p1, p2 = phase.copy(), phase.copy()
    p1[10, 10] += eps
    p2[10, 10] -= eps
    fd = (np.sum(a * op.psf4f(p1, *args)) - np.sum(a * op.psf4f(p2, *args))) / (2 * eps)
    assert abs(fd - ga[10, 10]) <= 1e-6 * abs(fd) + 1e-9 * np.abs(ga).max()


def test_power_is_conserved_by_the_tube_lens():
    """Parseval for the lens norm -i dx^2 / L^2: sum I du^2 = power = 1."""
    phase, sp, z, _ = _setup(n=64, nz=3)
    out, du = op.psf4f_field(phase, sp, LAM, z, F_OBJ, N_IMM, NA, F_TUBE)
    np.testing.assert_allclose(np.sum(np.abs(out) ** 2, axis=(-2, -1)) * du[0] * du[1], 1.0,
                               rtol=1e-12)


def test_output_spacing_is_the_camera_pitch():
    """compute_required_spacing inverts the lens' sampling change."""
    n, pitch = 1280, 6.5
    dx = op.required_spacing(n, pitch, F_TUBE, N_IMM, LAM)
    np.testing.assert_allclose(dx, 9.615384615, rtol=1e-9)
    _, du = op.psf4f_field(np.zeros((32, 32)), (dx, dx), LAM, 0.0, F_OBJ, N_IMM, NA, F_TUBE)
    np.testing.assert_allclose(du, (F_TUBE * LAM / (N_IMM * 32 * dx),) * 2, rtol=1e-12)


def test_in_focus_flat_phase_is_symmetric_and_peaked():
    """A flat mask at z = 0 gives a centro-symmetric PSF peaked at the centre."""
    n = 64
    dx = op.required_spacing(n, 6.5 * 20, F_TUBE, N_IMM, LAM)
    I = op.psf4f(np.zeros((n, n)), (dx, dx), LAM, 0.0, F_OBJ, N_IMM, NA, F_TUBE)[0]
    assert np.unravel_index(np.argmax(I), I.shape) == (n // 2, n // 2)
    np.testing.assert_allclose(I[1:, 1:], I[1:, 1:][::-1, ::-1], rtol=1e-9, atol=1e-12 * I.max())


def test_point_source_defocus_phase():
    """objective_point_source: the defocus phase is -pi (z/f) r^2 / L^2."""
    n, dx, z = 64, 50.0, 3.0
    u = os_.objective_point_source((n, n), (dx, dx), LAM, z, F_OBJ, N_IMM, NA)
    u0 = os_.objective_point_source((n, n), (dx, dx), LAM, 0.0, F_OBJ, N_IMM, NA)
    g = (np.arange(n) - n / 2) * dx
    r2 = g[:, None] ** 2 + g[None, :] ** 2
    inside = u0 != 0
    L2 = LAM * F_OBJ / N_IMM
    np.testing.assert_allclose(np.angle(u[inside] / u0[inside]),
                               np.angle(np.exp(-1j * np.pi * (z / F_OBJ) * r2[inside] / L2)),
                               atol=1e-12)


def test_chromatic_psf_reduces_to_monochromatic():
    # one wavelength, and the same wavelength listed twice with any weights:
    # the chromatic composition (per-wavelength spacing and modulation, the
    # whole field normalised, density-weighted intensity) gives the
    # monochromatic PSF exactly
    rng = np.random.default_rng(0)
    n, pitch, ft = 64, 6.5 * 20, 200000.0
    phase = rng.uniform(-np.pi, np.pi, (n, n))
    z = np.array([-2.0, 1.5])
    dx = op.required_spacing(n, pitch, ft, 1.33, 0.532)
    mono = op.psf4f(phase, (dx, dx), 0.532, z, 8000.0, 1.33, 0.6, ft)
    one = op.psf4f_chromatic(phase, pitch, [0.532], [1.0], z, 8000.0, 1.33, 0.6, ft)
    two = op.psf4f_chromatic(phase, pitch, [0.532, 0.532], [0.3, 0.9], z, 8000.0, 1.33, 0.6, ft)
    np.testing.assert_allclose(one, mono, rtol=1e-12, atol=1e-12 * mono.max())
    np.testing.assert_allclose(two, mono, rtol=1e-12, atol=1e-12 * mono.max())


def test_chromatic_psf_power_and_weights():
    # the normalised chromatic source has power P in the reference's sense at
    # every wavelength's area, and the lens conserves each wavelength's
    # energy: sum_k w_k dx_k^2 sum|u_k|^2 = ... the output intensity summed
    # over the plane times the output pixel area equals sum_k w_k P_k with
    # P_k = P S_k / S (S_k = sum |u_k|^2 before the lens, S = sum_k w_k S_k)
    rng = np.random.default_rng(1)
    n, pitch, ft, P = 96, 6.5 * 10, 200000.0, 2.0
    lam, wts = np.array([0.5, 0.6]), np.array([1.0, 3.0])
    phase = rng.uniform(-np.pi, np.pi, (n, n))
    I = op.psf4f_chromatic(phase, pitch, lam, wts, 0.0, 8000.0, 1.33, 0.6, ft, power=P)[0]
    w = wts / wts.sum()
    from oracle.sources import objective_point_source

    dxs = [op.required_spacing(n, pitch, ft, 1.33, l) for l in lam]
    Sk = [np.sum(np.abs(objective_point_source((n, n), (d, d), l, 0.0, 8000.0, 1.33, 0.6, None)) ** 2)
          for l, d in zip(lam, dxs)]
    S = np.dot(w, Sk)
    # per wavelength, after normalisation, sum |u_k|^2 dx_k^2 = P S_k / S
    expect = sum(wk * P * sk / S for wk, sk in zip(w, Sk))
    assert abs(I.sum() * pitch**2 / expect - 1.0) < 1e-9

"""Pins of the propagation oracle (CPU).

Each check is independent of the oracle's own formula text: closed forms
from the reference's tests, the physical plane-wave eigenfunction, round
trips and brute-force DFTs.
"""
import numpy as np
import pytest

from oracle import propagation as op
from tests.conftest import rel_l2

# constants of tests/test_propagate.py:10-14
D, W_SMALL, Z, LAM, N_MED = 40.0, 4.0, 100.0, 0.532, 1.33


def _square_aperture(shape, width, dx):
    g = op.grid(shape, dx)
    return ((np.abs(g[..., 0]) <= width / 2) & (np.abs(g[..., 1]) <= width / 2)).astype(np.complex128)


@pytest.mark.parametrize("method", ["transfer", "asm"])
@pytest.mark.parametrize(("shape", "pad"), [((256, 256), (512, 512)), ((1024, 256), (256, 512))])
def test_square_aperture_closed_form(method, shape, pad):
    """tests/test_propagate.py:79-171: intensity rel. MSE < 2e-2 vs Fresnel integrals."""
    dx = (D / shape[0], D / shape[1])
    u = _square_aperture(shape, D, dx)
    u = u / np.sqrt(np.sum(np.abs(u) ** 2) * dx[0] * dx[1])  # plane_wave(power=1)
    out = op.propagate(u, dx, LAM, Z, N_MED, pad, method=method, mode="same")
    I_num = np.abs(out) ** 2
    I_ana = np.abs(op.square_aperture_field(op.grid(shape, dx), Z, D, LAM, N_MED)) ** 2
    err = np.mean((I_ana - I_num) ** 2) / np.mean(I_ana**2)
    assert err < 2e-2


@pytest.mark.parametrize("method", ["transfer", "asm"])
def test_round_trip(method):
    """z then -z returns the field (tests/test_propagate.py:53-57, 127-149)."""
    shape = (256, 256)
    dx = (D / 256, D / 256)
    u = _square_aperture(shape, W_SMALL, dx)
    # as in the reference test, the asm round trip uses remove_evanescent=True
    # (evanescent components decay in both directions)
    kw = {"remove_evanescent": True} if method == "asm" else {}
    out = op.propagate(u, dx, LAM, Z, N_MED, 0, method=method, **kw)
    back = op.propagate(out, dx, LAM, -Z, N_MED, 0, method=method, **kw)
    np.testing.assert_allclose(back, u, rtol=2e-5, atol=1e-8)


@pytest.mark.parametrize("precision", ["fp64", "ref32"])
def test_asm_plane_wave_eigen(precision):
    """A sampled plane wave at bin (ky, kx) is multiplied by the physical
    angular-spectrum phase exp(i 2 pi z sqrt((n/lam)^2 - fy^2 - fx^2))."""
    ny, nx, dy, dx = 64, 48, 0.3, 0.25
    rng = np.random.default_rng(0)
    for _ in range(5):
        ky, kx = rng.integers(-ny // 2, ny // 2), rng.integers(-nx // 2, nx // 2)
        y = np.arange(ny)[:, None]
        x = np.arange(nx)[None, :]
        u = np.exp(2j * np.pi * (ky * y / ny + kx * x / nx))
        fy, fx = ky / (ny * dy), kx / (nx * dx)
        kz2 = (N_MED / LAM) ** 2 - fy**2 - fx**2
        expect = np.exp(1j * 2 * np.pi * 7.5 * np.sqrt(kz2 + 0j)) * u
        H = op.asm_propagator((ny, nx), (dy, dx), LAM, 7.5, N_MED, precision=precision)
        out = op.kernel_propagate(u, H)
        if precision == "fp64":
            assert rel_l2(out, expect) < 1e-12
        else:
            # float32 error bound of the reference's order: the phase
            # A*sqrt(kern) with kern = 1 - s^2 rounded to ~2 ulp(1), plus the
            # rounding of the phase itself
            A = 2 * np.pi * 7.5 * N_MED / LAM
            kern = kz2 * (LAM / N_MED) ** 2
            eps = float(np.finfo(np.float32).eps)
            bound = A * 3 * eps / (2 * np.sqrt(kern)) + 4 * A * np.sqrt(kern) * eps + 1e-6
            assert rel_l2(out, expect) < bound


def test_transfer_plane_wave_eigen():
    ny, nx, d = 32, 32, 0.5
    ky, kx = 3, -5
    y = np.arange(ny)[:, None]
    x = np.arange(nx)[None, :]
    u = np.exp(2j * np.pi * (ky * y / ny + kx * x / nx))
    f2 = (ky / (ny * d)) ** 2 + (kx / (nx * d)) ** 2
    expect = np.exp(-1j * np.pi * LAM / N_MED * 12.0 * f2) * u
    out = op.kernel_propagate(u, op.transfer_propagator((ny, nx), (d, d), LAM, 12.0, N_MED))
    assert rel_l2(out, expect) < 1e-12


def test_kernel_propagate_brute_force_dft():
    rng = np.random.default_rng(1)
    n = 8
    u = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    F = np.exp(-2j * np.pi * np.outer(np.arange(n), np.arange(n)) / n)
    Fi = np.conj(F) / n
    expect = Fi @ ((F @ u @ F.T) * H) @ Fi.T
    assert rel_l2(op.kernel_propagate(u, H), expect) < 1e-12


def test_parseval_unit_modulus():
    rng = np.random.default_rng(2)
    u = np.exp(2j * np.pi * rng.random((128, 128)))
    out = op.propagate(u, (0.3, 0.3), LAM, 50.0, N_MED, 0, method="transfer", mode="full")
    assert abs(np.sum(np.abs(out) ** 2) / np.sum(np.abs(u) ** 2) - 1) < 1e-12


def test_ref32_matches_fp64_to_float32_phase_rounding():
    """The ref32 propagator differs from FP64 only by float32 phase rounding:
    |dphase| <= ~ulp(phase) everywhere (phase reaches ~1.5e3 rad here)."""
    shape, dx = (256, 256), (0.3, 0.3)
    H64 = op.asm_propagator(shape, dx, LAM, 100.0, N_MED)
    H32 = op.asm_propagator(shape, dx, LAM, 100.0, N_MED, precision="ref32")
    dphi = np.abs(np.angle(H32 * np.conj(H64)))
    phase_max = 2 * np.pi * 100.0 * N_MED / LAM
    assert dphi.max() < 8 * np.spacing(np.float32(phase_max))
    assert dphi.max() > 0  # the rounding is really there


def test_negative_z_is_conjugate():
    shape, dx = (64, 64), (0.2, 0.2)
    for prec in ("fp64", "ref32"):
        Hp = op.asm_propagator(shape, dx, LAM, 30.0, N_MED, precision=prec)
        Hm = op.asm_propagator(shape, dx, LAM, -30.0, N_MED, precision=prec)
        np.testing.assert_array_equal(Hm, np.conj(Hp))


def test_evanescent_decay():
    """Small dx puts bins beyond n/lam: |H| = exp(-2 pi z sqrt(f^2-(n/lam)^2))."""
    ny = nx = 32
    d = 0.1
    H = op.asm_propagator((ny, nx), (d, d), LAM, 0.5, N_MED)
    fy = np.fft.fftfreq(ny, d)[:, None]
    fx = np.fft.fftfreq(nx, d)[None, :]
    ev = fy**2 + fx**2 - (N_MED / LAM) ** 2
    m = ev > 0
    assert m.any()
    np.testing.assert_allclose(np.abs(H[m]), np.exp(-2 * np.pi * 0.5 * np.sqrt(ev[m])), rtol=1e-10)
    Hr = op.asm_propagator((ny, nx), (d, d), LAM, 0.5, N_MED, remove_evanescent=True)
    # the reference's remove_evanescent sets delay = 0 there: unit modulus, zero phase
    np.testing.assert_allclose(Hr[m], 1.0, atol=1e-12)


def test_bandlimit_mask_is_subset():
    shape, dx = (128, 128), (0.5, 0.5)
    H = op.asm_propagator(shape, dx, LAM, 2000.0, N_MED)
    Hb = op.asm_propagator(shape, dx, LAM, 2000.0, N_MED, bandlimit=True)
    kept = np.abs(Hb) > 0
    assert 0 < kept.sum() < kept.size
    np.testing.assert_allclose(Hb[kept], H[kept])
    # the limit u_limit = 1/(lam sqrt((2 df z)^2 + 1)) of Matsushima & Shimobaba (2009), on-axis
    df = 1 / (shape[0] * dx[0])
    ulim = 1 / (LAM * np.sqrt((2 * df * 2000.0) ** 2 + 1))
    f = np.fft.fftfreq(shape[0], dx[0])
    np.testing.assert_array_equal(kept[:, 0], np.abs(f) <= ulim)

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


# ---- off-axis options: kykx, shift_yx and the off-axis band limit ------------------
# Each pin is a property the mathematics fixes (a shift theorem, a roll, the
# case table of the paper the band limit comes from).  Every pin is also run
# against a copy of the oracle with the sign flipped, and must fail there.

def _bandlimited_field(shape, radius, seed):
    """A field whose spectrum is supported on |bin| <= radius around DC, so
    that it can be modulated by a bin-frequency carrier without wrapping."""
    ny, nx = shape
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(ny) * ny
    kx = np.fft.fftfreq(nx) * nx
    keep = (ky[:, None] ** 2 + kx[None, :] ** 2) <= radius**2
    spec = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * keep
    return np.fft.ifft2(spec)


def _modulation_identity_error(prop, method, precision="fp64"):
    """Shift theorem of the frequency grid: with c a bin frequency and v
    band-limited, propagating v * exp(i 2 pi c.r) on axis equals
    exp(i 2 pi c.r) times v propagated with H evaluated at f + c, which the
    reference writes as kykx = -c (H uses f - kykx, propagation.py:483-487,
    :538-542).  Returns the worst rel. L2 over a few carriers and distances."""
    ny, nx, dy, dx = 64, 80, 0.3, 0.25
    v = _bandlimited_field((ny, nx), 9, 3)
    y = np.arange(ny)[:, None]
    x = np.arange(nx)[None, :]
    worst = 0.0
    for (my, mx), z in (((7, -11), 12.0), ((-13, 5), 40.0), ((3, 17), -25.0)):
        carrier = np.exp(2j * np.pi * (my * y / ny + mx * x / nx))
        c = (my / (ny * dy), mx / (nx * dx))
        H0 = prop((ny, nx), (dy, dx), LAM, z, N_MED, (0.0, 0.0), precision)
        Hc = prop((ny, nx), (dy, dx), LAM, z, N_MED, (-c[0], -c[1]), precision)
        lhs = op.kernel_propagate(v * carrier, H0)
        rhs = carrier * op.kernel_propagate(v, Hc)
        worst = max(worst, rel_l2(lhs, rhs))
    return worst


def _asm(shape, spacing, lam, z, n, kykx, precision):
    return op.asm_propagator(shape, spacing, lam, z, n, kykx=kykx, precision=precision)


def _transfer(shape, spacing, lam, z, n, kykx, precision):
    return op.transfer_propagator(shape, spacing, lam, z, n, kykx=kykx, precision=precision)


def _flip_kykx(fn):
    """The oracle with the sign of kykx flipped (f + kykx instead of f - kykx)."""
    return lambda shape, spacing, lam, z, n, kykx, precision: fn(
        shape, spacing, lam, z, n, (-kykx[0], -kykx[1]), precision)


@pytest.mark.parametrize(("prop", "method"), [(_asm, "asm"), (_transfer, "transfer")])
def test_kykx_shift_theorem(prop, method):
    assert _modulation_identity_error(prop, method) < 1e-12
    # planted mistake: the sign of kykx flipped in the oracle
    assert _modulation_identity_error(_flip_kykx(prop), method) > 1e-2


@pytest.mark.parametrize(("prop", "method"), [(_asm, "asm"), (_transfer, "transfer")])
def test_kykx_ref32_within_float32_rounding(prop, method):
    """The reference-order float32 propagator with kykx differs from FP64 by
    the rounding of a float32 phase only (max phase ~ 2 pi |z| n / lam)."""
    shape, sp = (96, 64), (0.3, 0.25)
    kykx = (0.37, -0.81)
    H64 = prop(shape, sp, LAM, 40.0, N_MED, kykx, "fp64")
    H32 = prop(shape, sp, LAM, 40.0, N_MED, kykx, "ref32")
    fy = np.fft.fftfreq(shape[0], sp[0])[:, None] - kykx[0]
    fx = np.fft.fftfreq(shape[1], sp[1])[None, :] - kykx[1]
    # bins well inside the propagating band: near the cutoff sqrt(1 - s^2)
    # amplifies the float32 rounding of 1 - s^2 without bound
    inside = (LAM / N_MED) ** 2 * (fy**2 + fx**2) < 0.9
    dphi = np.abs(np.angle(H32 * np.conj(H64)))[inside]
    A = 2 * np.pi * 40.0 * N_MED / LAM
    eps = float(np.finfo(np.float32).eps)
    phase_max = max(A, np.pi * LAM / N_MED * 40.0 * float(np.max(fy**2 + fx**2)))
    bound = A * 3 * eps / (2 * np.sqrt(0.1)) + 16 * float(np.spacing(np.float32(phase_max)))
    assert dphi.max() < bound
    Hflip = _flip_kykx(prop)(shape, sp, LAM, 40.0, N_MED, kykx, "ref32")
    assert np.abs(np.angle(Hflip * np.conj(H64)))[inside].max() > 1e-2


def test_kykx_tilted_plane_wave_closed_form():
    """A plane wave at bin frequency f0 is an eigenfunction: its factor is the
    physical exp(i 2 pi z sqrt((n/lam)^2 - |f0 - kykx|^2))."""
    ny, nx, dy, dx = 48, 64, 0.3, 0.25
    kykx = (0.21, -0.33)
    for (ky, kx), z in (((5, -9), 7.5), ((-11, 4), -3.0)):
        y = np.arange(ny)[:, None]
        x = np.arange(nx)[None, :]
        u = np.exp(2j * np.pi * (ky * y / ny + kx * x / nx))
        f0 = (ky / (ny * dy), kx / (nx * dx))
        kz = np.sqrt((N_MED / LAM) ** 2 - (f0[0] - kykx[0]) ** 2 - (f0[1] - kykx[1]) ** 2 + 0j)
        fac = np.exp(2j * np.pi * abs(z) * kz)
        fac = fac if z >= 0 else np.conj(fac)
        H = op.asm_propagator((ny, nx), (dy, dx), LAM, z, N_MED, kykx=kykx)
        assert rel_l2(op.kernel_propagate(u, H), fac * u) < 1e-12


def _shift_roll_error(shift_sign, precision):
    """shift_yx = (my dy, mx dx) moves the output window by whole pixels:
    the output equals the unshifted output rolled by (-my, -mx)
    (2 pi f.shift_yx is added to the phase, propagation.py:551)."""
    ny, nx, dy, dx = 64, 48, 0.3, 0.25
    rng = np.random.default_rng(4)
    u = np.exp(2j * np.pi * rng.random((ny, nx)))
    worst = 0.0
    for my, mx in ((3, -5), (-7, 2), (0, 11)):
        s = (my * dy, mx * dx)
        a = op.propagate(u, (dy, dx), LAM, 20.0, N_MED, 0, method="asm", precision=precision,
                         fft_precision="fp64")
        b = op.propagate(u, (dy, dx), LAM, 20.0, N_MED, 0, method="asm", precision=precision,
                         fft_precision="fp64", shift_yx=(shift_sign * s[0], shift_sign * s[1]))
        worst = max(worst, rel_l2(b, np.roll(a, (-my, -mx), axis=(0, 1))))
    return worst


def test_shift_yx_is_a_roll():
    assert _shift_roll_error(+1, "fp64") < 1e-12
    # float32 phase 2 pi f.s rounds at ~ulp(2 pi * 0.5/d * 11 d)
    assert _shift_roll_error(+1, "ref32") < 2e-5
    # planted mistake: the sign of the shift phase flipped
    assert _shift_roll_error(-1, "fp64") > 0.5


def _matsushima_table1(s, extent, lam, z):
    """Band of the shifted ASM, Matsushima (2010) Table 1, case by case.
    u_lim(+/-) = 1 / (lam sqrt((s +/- S)^-2 z^2 + 1)) with S = 1/(2 du) the
    half extent of the (padded) sampling window, du = 1/extent.  The cases are
    selected by comparing s with +/- extent, as the reference does
    (propagation.py:567-575).  Returns (centre u0, half width)."""
    S = extent / 2
    up = 1 / (lam * np.sqrt((s + S) ** -2 * z**2 + 1))
    un = 1 / (lam * np.sqrt((s - S) ** -2 * z**2 + 1))
    if s > extent:
        return (up + un) / 2, (up - un) / 2
    if s < -extent:
        return -(up + un) / 2, (un - up) / 2
    return (up - un) / 2, (up + un) / 2


def _band_limit_error(shift_of):
    """Count of bins where the oracle's band-limit mask differs from the
    Table 1 band, over shifts in all three regimes of both axes."""
    shape, sp = (96, 128), (0.4, 0.3)
    bad = 0
    for sy, sx in ((0.0, 0.0), (5.0, -9.0), (60.0, -70.0), (-45.0, 20.0), (12.0, 90.0)):
        for z in (40.0, 150.0, -400.0):
            mask = op._band_limit(shape, sp, LAM, z, shift_of((sy, sx)), "fp64")
            want = []
            for nn, d, s in ((shape[0], sp[0], sy), (shape[1], sp[1], sx)):
                u0, half = _matsushima_table1(s, nn * d, LAM, z)
                want.append(np.abs(np.fft.fftfreq(nn, d) - u0) <= half)
            bad += int(np.sum((want[0][:, None] & want[1][None, :]) != mask))
    return bad


def test_off_axis_band_limit_table1():
    assert _band_limit_error(lambda s: s) == 0
    # planted mistake: the mask of the mirrored shift
    assert _band_limit_error(lambda s: (-s[0], -s[1])) > 0


def test_off_axis_band_limit_mirror_symmetry():
    """shift -> -shift mirrors the band (u0 -> -u0, same width)."""
    shape, sp = (64, 64), (0.4, 0.4)
    for s in ((30.0, -8.0), (70.0, 40.0)):
        m = op._band_limit(shape, sp, LAM, 300.0, s, "fp64")
        mm = op._band_limit(shape, sp, LAM, 300.0, (-s[0], -s[1]), "fp64")
        # bin k <-> bin -k (index (N - k) % N)
        mirror = np.roll(mm[::-1, ::-1], (1, 1), axis=(0, 1))
        np.testing.assert_array_equal(m, mirror)


def test_off_axis_band_limit_ref32_matches_fp64():
    shape, sp = (96, 128), (0.4, 0.3)
    for s in ((5.0, -9.0), (60.0, -70.0)):
        m64 = op._band_limit(shape, sp, LAM, 400.0, s, "fp64")
        m32 = op._band_limit(shape, sp, LAM, 400.0, s, "ref32")
        # bins within float32 rounding of the band edge may differ
        assert np.sum(m64 != m32) <= 4

"""Pins of the multislice oracle (CPU): finite differences for the
hand-written adjoint, a plane wave through a uniform slab (closed form),
the stack/final-field consistency."""
import numpy as np
import pytest

from oracle import multislice as om
from oracle import propagation as op
from oracle.multislice import Multislice, sample_factor

LAM, NM, T = 0.532, 1.33, 0.5


def _case(h=16, w=12, D=3, pad=4, seed=0, NA=None, bandlimit=False):
    rng = np.random.default_rng(seed)
    m = Multislice((h, w), (0.3, 0.25), LAM, NM, T, D, pad, NA=NA, kykx=(0.1, -0.2),
                   bandlimit=bandlimit)
    u = np.exp(2j * np.pi * rng.random((h, w))) * (1 + 0.1 * rng.random((h, w)))
    dn = 0.05 * rng.standard_normal((D, h, w))
    ab = 0.01 * rng.random((D, h, w))
    return m, u, dn, ab, rng


@pytest.mark.parametrize("kind", ["upstream", "mse", "complex"])
@pytest.mark.parametrize("NA", [None, 0.9])
def test_gradients_match_finite_differences(kind, NA):
    m, u, dn, ab, rng = _case(NA=NA)
    out = m.forward(u, dn, ab)
    if kind == "upstream":
        cot = rng.standard_normal(out.shape)
        loss = lambda d, a: np.sum(cot * np.abs(m.forward(u, d, a)) ** 2)
    elif kind == "mse":
        cot = np.abs(out) ** 2 * rng.uniform(0.5, 1.5, out.shape)
        loss = lambda d, a: np.mean((np.abs(m.forward(u, d, a)) ** 2 - cot) ** 2)
    else:
        cot = rng.standard_normal(out.shape) + 1j * rng.standard_normal(out.shape)
        loss = lambda d, a: np.sum(np.real(np.conj(cot) * m.forward(u, d, a)))
    gdn, gab = m.grad(u, dn, cot, ab, kind)
    eps = 1e-6
    for (i, y, x) in [(0, 8, 6), (1, 3, 9), (2, 12, 2)]:
        for which, g in (("dn", gdn), ("ab", gab)):
            d1, d2, a1, a2 = dn.copy(), dn.copy(), ab.copy(), ab.copy()
            if which == "dn":
                d1[i, y, x] += eps
                d2[i, y, x] -= eps
            else:
                a1[i, y, x] += eps
                a2[i, y, x] -= eps
            fd = (loss(d1, a1) - loss(d2, a2)) / (2 * eps)
            assert abs(fd - g[i, y, x]) <= 1e-6 * np.abs(g).max() + 1e-6 * abs(fd), (which, i, fd)


def test_plane_wave_through_uniform_slab():
    """No padding, k = 0: every slice multiplies by exp(i 2 pi t n / lam) (the
    propagator at f = 0) and exp(i 2 pi dn t / lam); the final back-propagation
    by D t / 2 multiplies by exp(-i 2 pi (D t / 2) n / lam)."""
    D, dn0 = 7, 0.02
    m = Multislice((32, 32), (0.3, 0.3), LAM, NM, T, D, 0)
    u = np.ones((32, 32), complex)
    out = m.forward(u, np.full((D, 32, 32), dn0))
    phase = 2 * np.pi / LAM * (D * T * NM + D * T * dn0 - D * T / 2 * NM)
    np.testing.assert_allclose(out, np.exp(1j * phase), rtol=1e-12)


def test_stack_last_plane_is_the_field_after_the_last_slice():
    m, u, dn, ab, _ = _case()
    stack = m.forward(u, dn, ab, return_stack=True)
    fs, _ = m.fields(u, dn, ab)
    assert stack.shape == (3, 16, 12)
    np.testing.assert_allclose(stack[-1], fs[-1][4:20, 4:16], rtol=1e-12)


def test_ref32_sample_factor_is_close_to_fp64():
    rng = np.random.default_rng(1)
    dn = 0.01 * rng.standard_normal((64, 64))
    ab = 0.001 * rng.random((64, 64))
    a = sample_factor(dn, ab, T, LAM)
    b = sample_factor(dn, ab, T, LAM, "ref32")
    assert 0 < np.max(np.abs(a - b)) < 1e-6


def test_fluorescent_multislice_without_fluorescence_is_a_z_stack():
    """No emitters and no index change: the backward sweep visits the field
    propagated to each slice depth, so slice j holds |u propagated by j t|^2
    (the plain angular-spectrum propagation of the oracle, padded grid)."""
    rng = np.random.default_rng(1)
    H = W = 24
    pad, D, t = 4, 5, 1.5
    u = np.exp(2j * np.pi * rng.random((H, W)))
    zero = np.zeros((D, H, W))
    phases = rng.uniform(0, 2 * np.pi, (1, D, H, W))
    got = om.fluorescent_multislice(u, zero, zero, (0.3, 0.3), 0.532, 1.33, t, pad, phases)
    up = np.pad(u, pad)
    for j in range(D):
        Hj = op.asm_propagator(up.shape, (0.3, 0.3), 0.532, j * t, 1.33) if j else 1.0
        ref = np.abs(np.fft.ifft2(np.fft.fft2(up) * Hj)[pad:-pad, pad:-pad]) ** 2
        np.testing.assert_allclose(got[j], ref, rtol=1e-9, atol=1e-12)


def test_fluorescent_multislice_point_emitter_is_linear_and_phase_blind():
    """With no incident field, one emitter of amplitude a at slice k is a
    point source added before slice k's step: slice j sees |a|^2 |impulse
    propagated by (j - k) t|^2 (back-propagated for j < k) whatever its
    random phase (the grid has no evanescent bins, so the +t and -t steps
    compose to one propagation)."""
    H = W = 16
    pad, D, t, k = 4, 4, 2.0, 1
    fl = np.zeros((D, H, W))
    fl[k, 7, 9] = 3.0
    dn = np.zeros((D, H, W))
    u0 = np.zeros((H, W), complex)
    for phase in (0.0, 2.1):
        phases = np.full((1, D, H, W), phase)
        got = om.fluorescent_multislice(u0, fl, dn, (0.3, 0.3), 0.532, 1.33, t, pad, phases)
        src = np.zeros((H + 2 * pad, W + 2 * pad), complex)
        src[7 + pad, 9 + pad] = 3.0
        for j in range(D):
            # D - k forward steps of t from the emitter, D - j backward steps
            z = (j - k) * t
            Hz = op.asm_propagator(src.shape, (0.3, 0.3), 0.532, z, 1.33) if z else 1.0
            ref = np.abs(np.fft.ifft2(np.fft.fft2(src) * Hz)[pad:-pad, pad:-pad]) ** 2
            np.testing.assert_allclose(got[j], ref, rtol=1e-8, atol=1e-12 * ref.max())

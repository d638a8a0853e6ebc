"""Pins of the multislice oracle (CPU): finite differences for the
hand-written adjoint, a plane wave through a uniform slab (closed form),
the stack/final-field consistency."""
import numpy as np
import pytest

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

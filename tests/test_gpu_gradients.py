"""Reverse-mode rules of the library's entry points against autodiff of the
same mathematics: torch autograd on the CPU in complex128 (an independent
differentiation of the reference formulas), for loss L = Re sum conj(c) out
(or a real loss of the output).  Bar: relative L2 <= 1e-5 for complex64."""
import numpy as np
import pytest
import torch

import chromatix_amd as cx
import chromatix_amd.functional as cf
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _rand_c(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def _grads_gpu(fn, inputs, c):
    """Gradients of Re sum conj(c) fn(*inputs) on the GPU."""
    ts = [torch.as_tensor(x).cuda().requires_grad_(True) for x in inputs]
    out = fn(*ts)
    loss = (out * torch.as_tensor(c).cuda().conj()).real.sum()
    return [g.cpu().numpy() for g in torch.autograd.grad(loss, ts)]


def _grads_cpu(fn, inputs, c):
    ts = [torch.as_tensor(x.astype(np.complex128 if np.iscomplexobj(x) else np.float64))
          .requires_grad_(True) for x in inputs]
    out = fn(*ts)
    loss = (out * torch.as_tensor(c.astype(np.complex128)).conj()).real.sum()
    return [g.numpy() for g in torch.autograd.grad(loss, ts)]


@pytest.mark.parametrize("chromatic", [False, True])
def test_phase_change_gradients(chromatic):
    rng = np.random.default_rng(1)
    wl = np.array([0.45, 0.532, 0.65])
    if chromatic:
        shape, spec = (2, 48, 40, 3), cx.Spectrum(wl, [1.0, 2.0, 1.0])
        mod = np.asarray(spec.spectral_modulation, np.float64)
    else:
        shape, spec, mod = (2, 48, 40), 0.532, np.ones(1)
    u, c = _rand_c(rng, shape), _rand_c(rng, shape)
    ph = rng.uniform(-np.pi, np.pi, (48, 40)).astype(np.float32)

    def gpu(u_, ph_):
        f = (cx.ChromaticScalarField(u_, 0.3, 0.0, spec) if chromatic
             else cx.ScalarField(u_, 0.3, 0.0, spec))
        return cf.phase_change(f, ph_).u

    def cpu(u_, ph_):
        m = torch.as_tensor(mod)
        e = torch.exp(1j * (ph_[..., None] * m if chromatic else ph_ * m[0]))
        return u_ * e

    gg = _grads_gpu(gpu, [u, ph], c)
    gc = _grads_cpu(cpu, [u, ph], c)
    for a, b in zip(gg, gc):
        assert rel_l2(a, b) < TOL


@pytest.mark.parametrize("chromatic", [False, True])
def test_thin_sample_gradients(chromatic):
    rng = np.random.default_rng(2)
    wl = np.array([0.45, 0.532, 0.65])
    shape = (2, 40, 36, 3) if chromatic else (2, 40, 36)
    u, c = _rand_c(rng, shape), _rand_c(rng, shape)
    dn = (0.05 * rng.standard_normal((40, 36))).astype(np.float32)
    ab = (0.01 * rng.random((40, 36))).astype(np.float32)
    t = 0.7

    def gpu(u_, dn_, ab_):
        f = (cx.ChromaticScalarField(u_, 0.3, 0.0, cx.Spectrum(wl)) if chromatic
             else cx.ScalarField(u_, 0.3, 0.0, 0.532))
        return cf.thin_sample(f, ab_, dn_, t).u

    def cpu(u_, dn_, ab_):
        lam = torch.as_tensor(wl) if chromatic else torch.tensor(0.532, dtype=torch.float64)
        if chromatic:
            return u_ * torch.exp(1j * 2 * np.pi * (dn_[..., None] + 1j * ab_[..., None]) * t / lam)
        return u_ * torch.exp(1j * 2 * np.pi * (dn_ + 1j * ab_) * t / lam)

    gg = _grads_gpu(gpu, [u, dn, ab], c)
    gc = _grads_cpu(cpu, [u, dn, ab], c)
    for a, b in zip(gg, gc):
        assert rel_l2(a, b) < TOL

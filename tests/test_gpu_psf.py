"""GPU parity of the 4f PSF path (cxm_psf4f_fwd / _bwd, cxm_point_source_c64,
cxm_phase_change_c64) against the FP64 oracle.  Bar: relative L2 <= 1e-5 in
both phase modes (the ref32 defocus phase differs from FP64 by float32
rounding of a phase of a few radians, ~1e-7)."""
import numpy as np
import pytest
import torch

import chromatix_amd as cx
import chromatix_amd.functional as cf
from chromatix_amd.systems import Optical4FSystemPSF
from oracle import psf as op
from oracle import sources as os_
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5
LAM, NIMM, FOBJ, NA, FTUBE, PITCH = 0.532, 1.33, 8000.0, 0.6, 200000.0, 6.5


def _case(n, nz, seed=1, pitch=PITCH):
    rng = np.random.default_rng(seed)
    phase = rng.uniform(-np.pi, np.pi, (n, n)).astype(np.float32)
    dx = op.required_spacing(n, pitch, FTUBE, NIMM, LAM)
    z = np.linspace(-5.0, 5.0, nz)
    return phase, dx, z, rng


@pytest.mark.parametrize("pm", ["ref32", "accurate"])
@pytest.mark.parametrize("n,nz,pitch", [(256, 3, 6.5 * 5), (160, 1, 6.5 * 8), (1280, 11, 6.5)])
def test_psf_forward(pm, n, nz, pitch):
    phase, dx, z, _ = _case(n, nz, pitch=pitch)
    z_arg = z if nz > 1 else float(z[0])
    ph = torch.as_tensor(phase).cuda()
    got = cf.psf_4f(ph, dx, LAM, z_arg, FOBJ, NIMM, NA, FTUBE, phase_mode=pm).cpu().numpy()
    ref = op.psf4f(phase.astype(np.float64), (dx, dx), LAM, z, FOBJ, NIMM, NA, FTUBE)
    if nz == 1:
        ref = ref[0]
    assert got.shape == ref.shape
    err = rel_l2(got, ref)
    print(f"psf fwd n={n} nz={nz} {pm}: rel L2 {err:.2e}")
    assert err < TOL


def test_psf_field_and_spacing():
    phase, dx, z, _ = _case(128, 2, pitch=6.5 * 10)
    f = cf.psf_4f(torch.as_tensor(phase).cuda(), dx, LAM, z, FOBJ, NIMM, NA, FTUBE, output="field")
    ref, du = op.psf4f_field(phase.astype(np.float64), (dx, dx), LAM, z, FOBJ, NIMM, NA, FTUBE)
    assert rel_l2(f.u.cpu().numpy(), ref) < TOL
    np.testing.assert_allclose(f.dx, du, rtol=1e-12)
    np.testing.assert_allclose(f.dx, 6.5 * 10, rtol=1e-9)


@pytest.mark.parametrize("pm", ["ref32", "accurate"])
@pytest.mark.parametrize("n,nz", [(256, 3), (1280, 11)])
def test_psf_mse_gradient(pm, n, nz):
    phase, dx, z, rng = _case(n, nz, pitch=6.5 * 1280 / n)
    args = ((dx, dx), LAM, z, FOBJ, NIMM, NA, FTUBE)
    I_ref = op.psf4f(phase.astype(np.float64), *args)
    target = (I_ref * rng.uniform(0.5, 1.5, I_ref.shape)).astype(np.float32)
    _, g_ref = op.psf4f_mse_grad(phase.astype(np.float64), target.astype(np.float64), *args)
    g = cf.psf_4f_mse_grad(torch.as_tensor(phase).cuda(), torch.as_tensor(target).cuda(), dx, LAM,
                           z, FOBJ, NIMM, NA, FTUBE, phase_mode=pm).cpu().numpy()
    err = rel_l2(g, g_ref)
    print(f"psf mse grad n={n} nz={nz} {pm}: rel L2 {err:.2e}")
    assert err < TOL


@pytest.mark.parametrize("n,nz", [(256, 3), (1280, 11), (96, 40)])
def test_psf_value_and_gradient(n, nz):
    """psf_4f_mse_grad(return_intensity=True): the gradient of the plain
    call bit for bit, and the intensity stack the adjoint stores on the way
    equal to psf_4f's bit for bit (same transforms of the same source) and to
    the FP64 oracle; 40 planes run as two library calls."""
    phase, dx, z, rng = _case(n, nz, pitch=6.5 * 1280 / n)
    args = ((dx, dx), LAM, z, FOBJ, NIMM, NA, FTUBE)
    I_ref = op.psf4f(phase.astype(np.float64), *args)
    target = torch.as_tensor((I_ref * rng.uniform(0.5, 1.5, I_ref.shape)).astype(np.float32)).cuda()
    ph = torch.as_tensor(phase).cuda()
    g0 = cf.psf_4f_mse_grad(ph, target, dx, LAM, z, FOBJ, NIMM, NA, FTUBE)
    g1, I1 = cf.psf_4f_mse_grad(ph, target, dx, LAM, z, FOBJ, NIMM, NA, FTUBE, return_intensity=True)
    I0 = cf.psf_4f(ph, dx, LAM, z, FOBJ, NIMM, NA, FTUBE)
    assert torch.equal(g0, g1)
    assert torch.equal(I0, I1)
    assert rel_l2(I1.cpu().numpy(), I_ref) < TOL


@pytest.mark.parametrize("pm", ["ref32", "accurate"])
@pytest.mark.parametrize("n,nz", [(512, 5), (1280, 11), (2048, 3)])
def test_dedicated_psf_passes_match_pipeline(pm, n, nz):
    """The table-driven dedicated row and phase-gradient passes
    (CXM_OPT_PSF_PASSES 0) against the operator pipeline (1): intensity,
    the MSE gradient and the autograd gradient, with a point-source offset
    and without power normalisation, to float32 rounding."""
    from chromatix_amd import _lib as L

    phase, dx, z, rng = _case(n, nz, pitch=6.5 * 1280 / n)
    ph = torch.as_tensor(phase).cuda()
    w = torch.as_tensor(rng.standard_normal((nz, n, n)).astype(np.float32)).cuda()
    res = []
    try:
        for opt in (0, 1):
            L.check(L.lib().cxm_set_option(L.CXM_OPT_PSF_PASSES, opt))
            I = cf.psf_4f(ph, dx, LAM, z, FOBJ, NIMM, NA, FTUBE, power=None, amplitude=2.0,
                          offset=(0.4, -0.7), phase_mode=pm)
            g, I2 = cf.psf_4f_mse_grad(ph, I * 1.1, dx, LAM, z, FOBJ, NIMM, NA, FTUBE, power=None,
                                       amplitude=2.0, offset=(0.4, -0.7), phase_mode=pm,
                                       return_intensity=True)
            phr = ph.clone().requires_grad_(True)
            (cf.psf_4f(phr, dx, LAM, z, FOBJ, NIMM, NA, FTUBE, phase_mode=pm) * w).sum().backward()
            res.append((I.cpu().numpy(), g.cpu().numpy(), I2.cpu().numpy(), phr.grad.cpu().numpy()))
    finally:
        L.check(L.lib().cxm_set_option(L.CXM_OPT_PSF_PASSES, 0))
    errs = [rel_l2(a, b) for a, b in zip(res[0], res[1])]
    print(f"dedicated vs pipeline psf n={n} nz={nz} {pm}: " + " ".join(f"{e:.2e}" for e in errs))
    assert max(errs) < 2e-6


@pytest.mark.parametrize("n,nz", [(1024, 11), (1280, 11), (1536, 7)])
def test_dedicated_adjoint_is_deterministic(n, nz):
    """The in-place column adjoint (forward column FFT, cotangent, inverse
    FFT in one workgroup) gives bit-identical gradients over repeated calls.
    When the compiler issued the first transform's last LDS reads after the
    barrier that lets the other waves write the second transform's first
    stage, a strip of some plane came out wrong in about 4 calls in 10 at
    these sizes; the barriers now keep every load on its side."""
    phase, dx, z, rng = _case(n, nz, pitch=6.5 * 1280 / n)
    ph = torch.as_tensor(phase).cuda()
    w = torch.as_tensor(rng.standard_normal((nz, n, n)).astype(np.float32)).cuda()
    grads = []
    for _ in range(8):
        phr = ph.clone().requires_grad_(True)
        (cf.psf_4f(phr, dx, LAM, z, FOBJ, NIMM, NA, FTUBE, phase_mode="ref32") * w).sum().backward()
        grads.append(phr.grad)
    same = [bool(torch.equal(g, grads[0])) for g in grads[1:]]
    print(f"dedicated adjoint {n}^2 x {nz}: {sum(same)}/{len(same)} repeats bit-identical")
    assert all(same)


def test_psf_more_planes_than_one_call():
    # 70 planes run as chunks of 32: intensity, field, autograd and the
    # MSE gradient (whose mean covers all planes) against the oracle
    n, nz = 96, 70
    phase, dx, z, rng = _case(n, nz, pitch=6.5 * 1280 / n)
    args = ((dx, dx), LAM, z, FOBJ, NIMM, NA, FTUBE)
    I_ref = op.psf4f(phase.astype(np.float64), *args)
    ph = torch.as_tensor(phase).cuda()
    I = cf.psf_4f(ph, dx, LAM, z, FOBJ, NIMM, NA, FTUBE, phase_mode="accurate")
    assert tuple(I.shape) == (nz, n, n)
    assert rel_l2(I.cpu().numpy(), I_ref) < TOL
    fld = cf.psf_4f(ph, dx, LAM, z, FOBJ, NIMM, NA, FTUBE, phase_mode="accurate", output="field")
    assert rel_l2(np.abs(fld.u.cpu().numpy()) ** 2, I_ref) < TOL
    target = (I_ref * rng.uniform(0.5, 1.5, I_ref.shape)).astype(np.float32)
    _, g_ref = op.psf4f_mse_grad(phase.astype(np.float64), target.astype(np.float64), *args)
    g = cf.psf_4f_mse_grad(ph, torch.as_tensor(target).cuda(), dx, LAM, z, FOBJ, NIMM, NA, FTUBE,
                           phase_mode="accurate").cpu().numpy()
    assert rel_l2(g, g_ref) < TOL
    phr = ph.clone().requires_grad_(True)
    I2 = cf.psf_4f(phr, dx, LAM, z, FOBJ, NIMM, NA, FTUBE, phase_mode="accurate")
    w = torch.as_tensor(rng.standard_normal(I2.shape).astype(np.float32)).cuda()
    (I2 * w).sum().backward()
    g_ref = op.psf4f_grad(phase.astype(np.float64), w.cpu().numpy().astype(np.float64), *args)
    assert rel_l2(phr.grad.cpu().numpy(), g_ref) < TOL


def test_psf_autograd_upstream():
    phase, dx, z, rng = _case(160, 4, pitch=6.5 * 1280 / 160)
    ph = torch.as_tensor(phase).cuda().requires_grad_(True)
    I = cf.psf_4f(ph, dx, LAM, z, FOBJ, NIMM, NA, FTUBE, phase_mode="accurate")
    w = torch.as_tensor(rng.standard_normal(I.shape).astype(np.float32)).cuda()
    (I * w).sum().backward()
    g_ref = op.psf4f_grad(phase.astype(np.float64), w.cpu().numpy().astype(np.float64),
                          (dx, dx), LAM, z, FOBJ, NIMM, NA, FTUBE)
    assert rel_l2(ph.grad.cpu().numpy(), g_ref) < TOL


def test_optical4f_system_psf():
    class Microscope:  # the attributes Optical4FSystemPSF reads
        f, n, NA, spectrum, padding_ratio = FOBJ, NIMM, NA, LAM, 0
    n = 128
    phase, _, z, _ = _case(n, 3, pitch=6.5 * 10)
    sysm = Optical4FSystemPSF((n, n), 6.5 * 10, FTUBE, torch.as_tensor(phase).cuda())
    field = sysm(Microscope(), z)
    dx = Optical4FSystemPSF.compute_required_spacing(n, 6.5 * 10, FTUBE, NIMM, LAM)
    ref, du = op.psf4f_field(phase.astype(np.float64), (dx, dx), LAM, z, FOBJ, NIMM, NA, FTUBE)
    assert rel_l2(field.u.cpu().numpy(), ref) < TOL
    np.testing.assert_allclose(field.dx, 6.5 * 10, rtol=1e-9)


@pytest.mark.parametrize("pm", ["ref32", "accurate"])
def test_objective_point_source(pm):
    n, dx = 200, 50.0
    z = np.array([-4.0, 0.0, 2.5])
    f = cf.objective_point_source((n, n), dx, LAM, z, FOBJ, NIMM, NA, offset=(30.0, -20.0),
                                  phase_mode=pm)
    ref = os_.objective_point_source((n, n), (dx, dx), LAM, z, FOBJ, NIMM, NA, offset=(30.0, -20.0))
    assert f.u.shape == (3, n, n)
    assert rel_l2(f.u.cpu().numpy(), ref) < TOL
    np.testing.assert_allclose(f.power.cpu().numpy().reshape(-1), 1.0, rtol=1e-5)
    single = cf.objective_point_source((n, n), dx, LAM, 1.0, FOBJ, NIMM, NA, power=None,
                                       amplitude=2.0)
    ref1 = os_.objective_point_source((n, n), (dx, dx), LAM, 1.0, FOBJ, NIMM, NA, power_=None,
                                      amplitude=2.0)
    assert single.u.shape == (n, n)
    assert rel_l2(single.u.cpu().numpy(), ref1) < TOL


def test_phase_change_mono_and_chromatic():
    rng = np.random.default_rng(7)
    u = (rng.standard_normal((2, 64, 48)) + 1j * rng.standard_normal((2, 64, 48))).astype(np.complex64)
    phi = rng.uniform(-np.pi, np.pi, (64, 48)).astype(np.float32)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.3, 0.0, LAM)
    out = cf.phase_change(f, torch.as_tensor(phi).cuda()).u.cpu().numpy()
    assert rel_l2(out, u * np.exp(1j * phi.astype(np.float64))) < 1e-6
    wl = np.array([0.45, 0.55, 0.65])
    uc = (rng.standard_normal((64, 48, 3)) + 1j * rng.standard_normal((64, 48, 3))).astype(np.complex64)
    fc = cx.ChromaticScalarField(torch.as_tensor(uc).cuda(), 0.3, 0.0, cx.Spectrum(wl))
    outc = cf.phase_change(fc, torch.as_tensor(phi).cuda()).u.cpu().numpy()
    ref = uc * np.exp(1j * phi[..., None].astype(np.float64) * (wl[0] / wl))
    assert rel_l2(outc, ref) < 1e-6


@pytest.mark.parametrize("lams,wts", [([0.532], [1.0]), ([0.45, 0.55, 0.65], [1.0, 2.0, 0.5])])
def test_optical4f_chromatic(lams, wts):
    """Optical4FSystemPSF with a chromatic spectrum: per-wavelength pupil
    spacing and phase modulation, the point sources normalised as one field,
    the density-weighted intensity (field output and fused intensity) against
    the FP64 oracle; the phase gradient against finite differences of the
    oracle."""
    rng = np.random.default_rng(5)
    n, pitch, ft = 128, 6.5 * 10, FTUBE
    phase = rng.uniform(-np.pi, np.pi, (n, n)).astype(np.float32)
    z = np.array([-3.0, 0.0, 2.0])

    class Mic:
        f, n, NA, padding_ratio = FOBJ, NIMM, NA, 0
        spectrum = (np.array(lams), np.array(wts))

    ph = torch.as_tensor(phase).cuda()
    sysp = Optical4FSystemPSF((n, n), pitch, ft, ph)
    ref = op.psf4f_chromatic(phase.astype(np.float64), pitch, lams, wts, z, FOBJ, NIMM, NA, ft)
    I = sysp.intensity(Mic(), z).cpu().numpy()
    assert rel_l2(I, ref) < TOL
    fld = sysp(Mic(), z)
    np.testing.assert_allclose(np.asarray(fld.dx, np.float64).reshape(-1, 2), pitch, rtol=1e-9)
    assert rel_l2(fld.intensity.cpu().numpy().reshape(ref.shape), ref) < TOL
    # gradient of <I, W> with respect to the phase at a few pixels
    W = rng.standard_normal(ref.shape)
    phr = torch.as_tensor(phase).cuda().requires_grad_(True)
    sysg = Optical4FSystemPSF((n, n), pitch, ft, phr)
    (sysg.intensity(Mic(), z) * torch.as_tensor(W.astype(np.float32)).cuda()).sum().backward()
    g = phr.grad.cpu().numpy()
    p64 = phase.astype(np.float64)
    c = (n // 2, n // 2)
    for (i, j) in [(c[0], c[1] + 3), (c[0] - 5, c[1] + 2), (c[0] + 7, c[1] - 6)]:
        e = np.zeros((n, n))
        e[i, j] = 1e-4
        fp = op.psf4f_chromatic(p64 + e, pitch, lams, wts, z, FOBJ, NIMM, NA, ft)
        fm = op.psf4f_chromatic(p64 - e, pitch, lams, wts, z, FOBJ, NIMM, NA, ft)
        fd = np.sum(W * (fp - fm)) / 2e-4
        assert abs(g[i, j] - fd) <= 2e-3 * np.max(np.abs(g)), (i, j, g[i, j], fd)

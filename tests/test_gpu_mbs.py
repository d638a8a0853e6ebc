"""GPU parity of the modified Born series iteration (cxm_mbs_*) against the
FP64 oracle: the same shift/scale and iteration count give the same iterate
(relative L2 <= 1e-5), and converged solves satisfy the vector Helmholtz
equation (FP64 residual evaluated on the host)."""
import numpy as np
import pytest
import torch

from chromatix_amd.experimental import modified_born_series as mbs
from oracle import mbs as om
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
LAM = 0.532
K0 = 2 * np.pi / LAM


def _spheres(n, count, rng):
    eps = np.full((n, n, n), 1.33**2, np.complex128)
    idx = np.indices((n, n, n))
    for _ in range(count):
        c = rng.uniform(0.2 * n, 0.8 * n, 3)
        r = rng.uniform(4, 12) * n / 128
        ri = rng.uniform(1.36, 1.40)
        inside = np.sum((idx - c[:, None, None, None]) ** 2, axis=0) <= r * r
        eps[inside] = ri**2
    return eps


def _setup(n, bc, count, seed):
    rng = np.random.default_rng(seed)
    spacing = LAM / 4
    eps0 = _spheres(n, count, rng)
    eps = om.add_absorbing_bc(eps0, spacing, (0, 1, 2), bc * spacing, n_background=1.33)
    J = np.zeros(eps.shape + (3,), np.complex128)
    J[:, :, bc + 1, 0] = 1.0
    return eps, spacing, J


def test_iterate_matches_oracle():
    eps, spacing, J = _setup(32, 4, 3, 0)  # 40^3 with the boundary
    S = om.source_field(J, K0)
    shift, scale = om.get_shift_and_scale(eps)
    m = om.MBS(eps, spacing, K0, shift, scale)
    py, num = m.prec_y(S)
    x_ref, _ = m.iterate(py.copy(), py, 40)
    sample = mbs.Sample(torch.as_tensor(eps.astype(np.complex64)).cuda(), spacing)
    src = mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0)
    res = mbs.solve(sample, src, maxiter=40, tol=0.0, shift_scale=(shift, scale), return_info=True)
    assert res.iterations == 40
    err = rel_l2(res.E.cpu().numpy(), x_ref * num)
    print(f"mbs 40^3 after 40 iterations: {err:.2e}")
    assert err < 1e-5


def test_converged_solution_solves_helmholtz():
    eps, spacing, J = _setup(32, 4, 3, 1)
    S = om.source_field(J, K0)
    sample = mbs.Sample(torch.as_tensor(eps.astype(np.complex64)).cuda(), spacing)
    src = mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0)
    res = mbs.solve(sample, src, maxiter=2000, rtol=1e-6, return_info=True)
    c, r = mbs.get_shift_and_scale(sample.permittivity)
    m = om.MBS(eps, spacing, K0, c, 1 / (1.1j * r))
    pde = m.pde_residual(res.E.cpu().numpy().astype(np.complex128), S)
    print(f"mbs 40^3: {res.iterations} iterations, PDE residual {pde:.2e}")
    assert pde < 1e-4


@pytest.mark.parametrize("seed", [4, 5])
def test_c4_volume_converges(seed):
    """Two of the c4 volumes (128^3 with 20 spheres and a 16-voxel absorbing
    boundary, 160^3), each to a relative residual of 1e-4; the host FP64
    evaluation of the PDE residual confirms the solution (4e-5 measured)."""
    eps, spacing, J = _setup(128, 16, 20, seed)
    S = om.source_field(J, K0)
    sample = mbs.Sample(torch.as_tensor(eps.astype(np.complex64)).cuda(), spacing)
    src = mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0)
    res = mbs.solve(sample, src, maxiter=3000, rtol=1e-4, return_info=True)
    c, r = mbs.get_shift_and_scale(sample.permittivity)
    m = om.MBS(eps, spacing, K0, c, 1 / (1.1j * r))
    pde = m.pde_residual(res.E.cpu().numpy().astype(np.complex128), S)
    print(f"mbs 160^3: {res.iterations} iterations, residual {res.residual:.2e}, PDE {pde:.2e}")
    assert res.iterations < 3000
    assert pde < 2e-4


@pytest.mark.parametrize("tol", [1e-2, 1e-3])
def test_fixed_point_stops_where_the_oracle_does(tol):
    """The device-side stopping test (after every iteration, |r| <= tol,
    jaxopt.FixedPointIteration's rule) stops at the iteration the FP64
    oracle's per-iteration test picks, and the iterate there matches."""
    eps, spacing, J = _setup(24, 4, 2, 3)  # 32^3 with the boundary
    S = om.source_field(J, K0)
    shift, scale = om.get_shift_and_scale(eps)
    m = om.MBS(eps, spacing, K0, shift, scale)
    py, num = m.prec_y(S)
    x, k_ref = py.copy(), 0
    norms = []
    while k_ref < 500:
        x, r = m.iterate(x, py, 1)
        k_ref += 1
        norms.append(np.linalg.norm(r))
        if norms[-1] <= tol:
            break
    sample = mbs.Sample(torch.as_tensor(eps.astype(np.complex64)).cuda(), spacing)
    src = mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0)
    res = mbs.solve(sample, src, maxiter=500, tol=tol, shift_scale=(shift, scale), check_every=8,
                    return_info=True)
    print(f"fixed point tol {tol}: {res.iterations} iterations (oracle {k_ref}), "
          f"residual {res.residual:.4e} (oracle {norms[-1]:.4e}, previous {norms[-2]:.4e})")
    # the two norms straddle tol by far more than float32 rounding here
    assert norms[-1] < tol * (1 - 1e-3) and norms[-2] > tol * (1 + 1e-3)
    assert res.iterations == k_ref
    # r = x_new - x of float32 iterates: cancellation leaves ~1e-3 relative at |r| ~ 1e-3
    assert abs(res.residual - norms[-1]) <= 5e-3 * norms[-1]
    assert rel_l2(res.E.cpu().numpy(), x * num) < 1e-5


@pytest.mark.parametrize("iters", [1, 3, 6])
def test_bicgstab_iterates_match_oracle(iters):
    # solve(use_bicgstab=True) with maxiter k and tol 0 runs exactly k
    # BiCGStab iterations: the iterate matches the FP64 oracle's
    eps, spacing, J = _setup(24, 4, 2, 2)  # 32^3 with the boundary
    S = om.source_field(J, K0)
    shift, scale = om.get_shift_and_scale(eps)
    m = om.MBS(eps, spacing, K0, shift, scale)
    py, num = m.prec_y(S)
    x_ref, k_ref, _ = om.bicgstab(m.prec_forward, py, py, iters, tol=0.0)
    assert k_ref == iters
    sample = mbs.Sample(torch.as_tensor(eps.astype(np.complex64)).cuda(), spacing)
    src = mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0)
    res = mbs.solve(sample, src, maxiter=iters, tol=0.0, shift_scale=(shift, scale), use_bicgstab=True,
                    return_info=True, check_every=2)
    assert res.iterations == iters
    err = rel_l2(res.E.cpu().numpy(), x_ref * num)
    print(f"bicgstab 32^3 after {iters} iterations: {err:.2e}")
    assert err < 1e-5


def test_bicgstab_converged_solution():
    # tol 1e-3 (the reference's default): stops at the oracle's iteration
    # count (+-1 for float32 rounding near the threshold), the field solves
    # the Helmholtz equation and equals the converged fixed point
    eps, spacing, J = _setup(32, 4, 3, 3)
    S = om.source_field(J, K0)
    shift, scale = om.get_shift_and_scale(eps)
    E_ref, k_ref, _ = om.solve_bicgstab(eps, S, spacing, K0, shift, scale, maxiter=500, tol=1e-3)
    sample = mbs.Sample(torch.as_tensor(eps.astype(np.complex64)).cuda(), spacing)
    src = mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0)
    res = mbs.solve(sample, src, maxiter=500, tol=1e-3, shift_scale=(shift, scale), use_bicgstab=True,
                    return_info=True)
    assert abs(res.iterations - k_ref) <= 1
    E = res.E.cpu().numpy().astype(np.complex128)
    assert rel_l2(E, E_ref) < 1e-4
    res5 = mbs.solve(sample, src, maxiter=500, tol=1e-5, shift_scale=(shift, scale), use_bicgstab=True,
                     return_info=True)
    m = om.MBS(eps, spacing, K0, shift, scale)
    E5 = res5.E.cpu().numpy().astype(np.complex128)
    assert m.pde_residual(E5, S) < 1e-4
    E_fp, _, _ = om.solve_bicgstab(eps, S, spacing, K0, shift, scale, maxiter=500, tol=1e-9)
    assert rel_l2(E5, E_fp) < 1e-4
    print(f"bicgstab 40^3: {res.iterations} iterations (oracle {k_ref}) at tol 1e-3, "
          f"{res5.iterations} at 1e-5")


def test_bicgstab_stops_and_keeps_state():
    # extra enqueued steps after the stop change nothing
    eps, spacing, J = _setup(24, 4, 2, 4)
    sample = mbs.Sample(torch.as_tensor(eps.astype(np.complex64)).cuda(), spacing)
    src = mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0)
    sv = mbs.BiCGStabSolver(sample, src, tol=1e-3, maxiter=400)
    while True:
        sv.step(4)
        sv.event.synchronize()
        sv.poll()
        if sv.converged:
            break
    assert sv.stopped and 0 < sv.k <= sv.iterations
    x1 = sv.x.clone()
    sv.stopped = False  # force more steps past the device-side stop
    sv.step(3)
    sv.event.synchronize()
    sv.poll()
    assert sv.stopped and torch.equal(sv.x, x1)


def _aniso_gpu(n=24, bc=4, seed=7):
    from tests.test_oracle_mbs import _aniso

    T, spacing, S = _aniso(n=n, bc=bc, seed=seed)
    J = np.zeros(T.shape[2:] + (3,), np.complex128)
    J[:, :, bc + 1, 0] = 1.0
    return T, spacing, J, om.source_field(J, K0)


@pytest.mark.parametrize("bicg", [False, True])
def test_tensor_iterates_match_oracle(bicg):
    T, spacing, J, S = _aniso_gpu()
    shift, scale = om.get_shift_and_scale_tensor(T)
    m = om.MBS(T, spacing, K0, shift, scale)
    py, num = m.prec_y(S)
    iters = 4 if bicg else 40
    if bicg:
        x_ref, _, _ = om.bicgstab(m.prec_forward, py, py, iters, tol=0.0)
    else:
        x_ref, _ = m.iterate(py.copy(), py, iters)
    sample = mbs.Sample(torch.as_tensor(T.astype(np.complex64)).cuda(), spacing)
    assert sample.tensor
    src = mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0)
    res = mbs.solve(sample, src, maxiter=iters, tol=0.0, shift_scale=(shift, scale), use_bicgstab=bicg,
                    return_info=True)
    err = rel_l2(res.E.cpu().numpy(), x_ref * num)
    print(f"tensor eps 32^3, {'bicgstab' if bicg else 'fixed point'} {iters} iterations: {err:.2e}")
    assert err < 1e-5


def test_identity_tensor_matches_scalar_on_gpu():
    eps, spacing, J = _setup(24, 4, 2, 5)
    T = np.zeros((3, 3) + eps.shape, np.complex64)
    for a in range(3):
        T[a, a] = eps
    shift, scale = om.get_shift_and_scale(eps)
    src = mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0)
    out = []
    for e in (eps.astype(np.complex64), T):
        sample = mbs.Sample(torch.as_tensor(e).cuda(), spacing)
        out.append(mbs.solve(sample, src, maxiter=30, tol=0.0, shift_scale=(shift, scale)).cpu().numpy())
    assert rel_l2(out[1], out[0]) < 1e-6


def test_tensor_converged_solution_solves_helmholtz():
    T, spacing, J, S = _aniso_gpu(n=32, bc=4, seed=9)
    sample = mbs.Sample(torch.as_tensor(T.astype(np.complex64)).cuda(), spacing)
    src = mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0)
    res = mbs.solve(sample, src, maxiter=400, tol=1e-5, use_bicgstab=True, return_info=True)
    shift, scale = mbs.get_shift_and_scale(sample.permittivity)
    m = om.MBS(T, spacing, K0, shift, scale)
    pde = m.pde_residual(res.E.cpu().numpy().astype(np.complex128), S)
    print(f"tensor eps 40^3 bicgstab: {res.iterations} iterations, PDE residual {pde:.2e}")
    assert pde < 1e-4


@pytest.mark.parametrize("case", ["spheres", "random", "single", "collinear", "ring"])
def test_device_enclosing_circle_matches_host(case):
    rng = np.random.default_rng(11)
    if case == "spheres":
        eps, _, _ = _setup(40, 6, 5, 6)
    elif case == "random":
        eps = (rng.standard_normal((30, 40, 50)) + 1j * rng.standard_normal((30, 40, 50))) * 0.1 + 1.8
    elif case == "single":
        eps = np.full((8, 8, 8), 1.33**2 + 0.01j)
    elif case == "collinear":
        eps = (1.7 + 0.1j) + (0.3 - 0.2j) * rng.random((16, 16, 16))
    else:  # values on a circle: every value is a hull vertex
        eps = 1.8 + 0.2j + 0.05 * np.exp(2j * np.pi * rng.random((20, 20, 20)))
    e32 = eps.astype(np.complex64)
    c_h, r_h = mbs.enclosing_circle(e32)
    c_d, s_d = mbs.get_shift_and_scale(torch.as_tensor(e32).cuda())
    r_d = abs(1 / s_d) / 1.1
    if r_h == 0:
        r_h = 1e-6
    assert abs(c_d - c_h) <= 1e-9 * max(1.0, abs(c_h))
    np.testing.assert_allclose(r_d, r_h, rtol=1e-9)
    # every value lies in the circle
    assert np.max(np.abs(e32.astype(np.complex128) - c_d)) <= r_d * (1 + 1e-12) + 1e-15


@pytest.mark.parametrize("n,bc", [(24, 4), (40, 4), (64, 8)])
def test_fused_green_pass_matches_separate_passes(n, bc):
    # the fused x-axis FFT + Green + inverse pass against the three
    # launches (CXM_OPT_MBS_XGREEN 2): the same operator to float32 rounding
    from chromatix_amd import _lib as L

    eps, spacing, J = _setup(n, bc, 2, 12)
    sample = mbs.Sample(torch.as_tensor(eps.astype(np.complex64)).cuda(), spacing)
    src = mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0)
    shift, scale = om.get_shift_and_scale(eps)
    out = []
    try:
        for opt in (0, 2):
            L.check(L.lib().cxm_set_option(L.CXM_OPT_MBS_XGREEN, opt))
            out.append(mbs.solve(sample, src, maxiter=12, tol=0.0, shift_scale=(shift, scale)).cpu().numpy())
    finally:
        L.check(L.lib().cxm_set_option(L.CXM_OPT_MBS_XGREEN, 0))
    assert rel_l2(out[0], out[1]) < 1e-6


@pytest.mark.parametrize("n,bc,tensor", [(24, 4, False), (40, 4, False), (64, 8, False), (120, 20, False),
                                          (24, 4, True)])
def test_fused_z_passes_match_separate_passes(n, bc, tensor):
    # the first z pass fused with V and the last with the update
    # (CXM_OPT_MBS_ZFUSE 0) against separate launches (1): the same
    # iteration to float32 rounding, and the same residual to FP64 order
    from chromatix_amd import _lib as L

    eps, spacing, J = _setup(n, bc, 2, 13)
    if tensor:
        T = np.zeros((3, 3) + eps.shape, np.complex64)
        for a in range(3):
            T[a, a] = eps
        T[0, 1] = T[1, 0] = 0.01
        eps = T
    shift, scale = om.get_shift_and_scale_tensor(eps) if tensor else om.get_shift_and_scale(eps)
    sample = mbs.Sample(torch.as_tensor(eps.astype(np.complex64)).cuda(), spacing)
    src = mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0)
    out, res = [], []
    try:
        for opt in (0, 1):
            L.check(L.lib().cxm_set_option(L.CXM_OPT_MBS_ZFUSE, opt))
            r = mbs.solve(sample, src, maxiter=16, tol=0.0, shift_scale=(shift, scale), return_info=True)
            out.append(r.E.cpu().numpy())
            res.append(float(r.residual))
    finally:
        L.check(L.lib().cxm_set_option(L.CXM_OPT_MBS_ZFUSE, 0))
    err = rel_l2(out[0], out[1])
    print(f"fused z passes {n + 2 * bc}^3 tensor={tensor}: {err:.2e}, residuals {res[0]:.6e} {res[1]:.6e}")
    assert err < 1e-6
    assert abs(res[0] - res[1]) <= 1e-5 * abs(res[1])


@pytest.mark.parametrize("bicg", [False, True])
def test_solve_gradient_matches_oracle(bicg):
    """Autograd through solve (implicit differentiation: the adjoint solve
    plus cxm_mbs_vjp) against the FP64 oracle MBS.vjp, for
    L = Re <C, E(eps, S)>: the permittivity and source gradients."""
    eps, spacing, J = _setup(24, 4, 2, 21)  # 32^3 with the boundary
    S = om.source_field(J, K0)
    shift, scale = om.get_shift_and_scale(eps)
    m = om.MBS(eps, spacing, K0, shift, scale)
    rng = np.random.default_rng(5)
    C = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    eps_bar, S_bar, _ = m.vjp(S, m.solve_exact(S), C)

    eps_t = torch.as_tensor(eps.astype(np.complex64)).cuda().requires_grad_()
    S_t = torch.as_tensor(S.astype(np.complex64)).cuda().requires_grad_()
    src = mbs.Source.__new__(mbs.Source)
    src.k0, src.field = K0, S_t
    kw = dict(use_bicgstab=True, tol=1e-7) if bicg else dict(rtol=1e-7, adjoint_tol=1e-7)
    E = mbs.solve(mbs.Sample(eps_t, spacing), src, maxiter=4000, shift_scale=(shift, scale), **kw)
    Ct = torch.as_tensor(C.astype(np.complex64)).cuda()
    (torch.conj(Ct) * E).real.sum().backward()
    e_eps = rel_l2(eps_t.grad.cpu().numpy(), eps_bar)
    e_src = rel_l2(S_t.grad.cpu().numpy(), S_bar)
    print(f"solve gradient 32^3 (bicgstab={bicg}): eps {e_eps:.2e}, source {e_src:.2e}")
    assert e_eps < 1e-4 and e_src < 1e-4


def test_solve_gradient_flows_to_current_density():
    # J -> Source.field is a torch product: the gradient reaches J through it
    eps, spacing, J = _setup(16, 4, 1, 22)
    shift, scale = om.get_shift_and_scale(eps)
    J_t = torch.as_tensor(J.astype(np.complex64)).cuda().requires_grad_()
    src = mbs.Source(J_t, K0)
    E = mbs.solve(mbs.Sample(torch.as_tensor(eps.astype(np.complex64)).cuda(), spacing), src,
                  maxiter=4000, rtol=1e-6, shift_scale=(shift, scale))
    E.abs().square().sum().backward()
    assert J_t.grad is not None and torch.isfinite(torch.view_as_real(J_t.grad)).all()
    assert float(J_t.grad.abs().max()) > 0


@pytest.mark.parametrize("bicg", [False, True])
def test_solve_batch_matches_solve(bicg):
    """solve_batch (volumes on concurrent streams, two at a time) gives each
    volume's solve() result bit for bit, with the same iteration count."""
    cases = [_setup(n, bc, 3, seed) for n, bc, seed in [(32, 4, 11), (40, 4, 12), (32, 8, 13)]]
    samples = [mbs.Sample(torch.as_tensor(e.astype(np.complex64)).cuda(), sp) for e, sp, _ in cases]
    sources = [mbs.Source(torch.as_tensor(J.astype(np.complex64)).cuda(), K0) for _, _, J in cases]
    kw = dict(tol=1e-4, use_bicgstab=True) if bicg else dict(rtol=1e-4)
    batch = mbs.solve_batch(samples, sources, maxiter=2000, concurrency=2, return_info=True, **kw)
    for s, c, b in zip(samples, sources, batch):
        one = mbs.solve(s, c, maxiter=2000, return_info=True, **kw)
        print(f"solve_batch {tuple(s.shape)} bicgstab={bicg}: {b.iterations} / {one.iterations} iterations")
        assert b.iterations == one.iterations
        assert torch.equal(b.E, one.E)

"""The serving loop (chromatix_amd.serving): every request's result in the
batched, graph-replayed mix equals its result served alone by the plain
one-call-per-operation path; the parameter-table entry point equals the
plain z-list entry point."""
import ctypes

import numpy as np
import pytest
import torch

from chromatix_amd import _lib as L
from chromatix_amd import serving as sv
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _prop(N, Z, rng, t):
    u = torch.as_tensor(np.exp(2j * np.pi * rng.random((N, N))).astype(np.complex64)).cuda()
    ph = torch.as_tensor(rng.uniform(-np.pi, np.pi, (N, N)).astype(np.float32)).cuda()
    return sv.PropRequest(u, ph, rng.uniform(-20, 20, Z), 0.3, 0.532, 1.33, arrival=t)


def _mix(seed=0):
    rng = np.random.default_rng(seed)
    reqs = [_prop(256, 4, rng, 0.001 * i) for i in range(5)] + [_prop(512, 2, rng, 0.002 + 0.001 * i)
                                                                for i in range(3)]
    wl = np.array([0.45, 0.55, 0.65])
    uc = torch.as_tensor(np.exp(2j * np.pi * rng.random((256, 256, 3))).astype(np.complex64)).cuda()
    reqs.append(sv.SasCztRequest(uc, 0.5, wl, 15.0, 1.33, 64, 2.0, arrival=0.003))
    eps = np.full((32, 32, 32), 1.33**2, np.complex64)
    eps[12:20, 12:20, 12:20] = 1.38**2
    J = np.zeros((32, 32, 32, 3), np.complex64)
    J[:, :, 3, 0] = 1
    reqs.append(sv.BornRequest(torch.as_tensor(eps).cuda(), torch.as_tensor(J).cuda(), 0.133,
                               2 * np.pi / 0.532, rtol=1e-4, arrival=0.004))
    reqs.append(sv.BornRequest(torch.as_tensor(eps).cuda(), torch.as_tensor(J).cuda(), 0.133,
                               2 * np.pi / 0.532, rtol=1e-4, arrival=0.005, use_bicgstab=True))
    return reqs


def test_params_table_equals_z_list():
    rng = np.random.default_rng(1)
    N, Z = 128, 3
    u = torch.as_tensor(np.exp(2j * np.pi * rng.random((2, N, N))).astype(np.complex64)).cuda()
    z = rng.uniform(-20, 20, (2, Z))
    outs = []
    for b in range(2):
        d = sv._prop_desc(N, Z, 0.3, 0.532, 1.33, 1)
        zb = L.doubles(z[b])
        d.z = ctypes.cast(zb, ctypes.POINTER(ctypes.c_double))
        o = torch.empty((Z, N, N), dtype=torch.float32, device="cuda")
        ws = torch.empty(L.lib().cxm_prop2d_workspace_size(ctypes.byref(d)), dtype=torch.uint8, device="cuda")
        L.check(L.lib().cxm_prop2d_c64(L.stream_ptr(), u[b].data_ptr(), o.data_ptr(), ws.data_ptr(),
                                       ws.numel(), ctypes.byref(d)))
        outs.append(o)
    d = sv._prop_desc(N, Z, 0.3, 0.532, 1.33, 2)
    rec = L.lib().cxm_prop2d_param_bytes()
    host = (ctypes.c_uint8 * (2 * Z * rec))()
    zz = L.doubles(z.reshape(-1))
    L.check(L.lib().cxm_prop2d_build_params(ctypes.byref(d), ctypes.cast(zz, ctypes.POINTER(ctypes.c_double)),
                                            2 * Z, ctypes.cast(host, ctypes.c_void_p)))
    dev = torch.as_tensor(np.frombuffer(bytes(host), dtype=np.uint8).copy()).cuda()
    o2 = torch.empty((2 * Z, N, N), dtype=torch.float32, device="cuda")
    ws = torch.empty(L.lib().cxm_prop2d_workspace_size(ctypes.byref(d)), dtype=torch.uint8, device="cuda")
    L.check(L.lib().cxm_prop2d_c64_params(L.stream_ptr(), u.data_ptr(), o2.data_ptr(), ws.data_ptr(),
                                          ws.numel(), ctypes.byref(d), dev.data_ptr()))
    assert torch.equal(o2, torch.cat(outs))


@pytest.mark.parametrize("graphs", [True, False])
def test_results_alone_equal_results_in_the_mix(graphs):
    plain = sv.run_plain(_mix(), keep_results=True)
    eng = sv.ServingEngine(max_batch=4, use_graphs=graphs, keep_results=True)
    mix = _mix()
    eng.warm(mix)
    served = eng.run(mix)
    for a, b in zip(plain, served):
        assert b.done is not None
        ra, rb = a.result.cpu().numpy(), b.result.cpu().numpy()
        assert ra.shape == rb.shape
        if isinstance(a, sv.BornRequest):
            assert a.iterations == b.iterations
        err = np.linalg.norm(ra - rb) / np.linalg.norm(ra)
        assert err <= 1e-6, (type(a).__name__, err)
    st = sv.latency_stats(served)
    assert st["n"] == len(served)


# ---- the serving stream at its stated sizes ---------------------------------
C5_WL = np.linspace(0.45, 0.66, 8)


def _c5_mix(seed=3):
    """One 2048^2 x 8-wavelength SAS -> CZT 512^2 request, one 2048^2
    32-plane angular-spectrum intensity request, one 128^3 Born request
    (BiCGStab), four 256^2 propagation requests, and three malformed
    requests (NaN z, a phase mask of the wrong shape, a field with fewer
    wavelength planes than wavelengths)."""
    from oracle import mbs as om

    rng = np.random.default_rng(seed)
    x = (np.arange(2048) - 1024) * 0.5
    aperture = (np.abs(x)[:, None] < 300) & (np.abs(x)[None, :] < 300)
    uc = np.repeat((aperture * np.exp(2j * np.pi * 0.05 * rng.random((2048, 2048))))[..., None], 8, -1)
    sas = sv.SasCztRequest(torch.as_tensor(uc.astype(np.complex64)).cuda(), 0.5, C5_WL, 2000.0, 1.0,
                           512, 2.0, arrival=0.0)
    big = _prop(2048, 32, rng, 0.001)
    small = [_prop(256, 4, rng, 0.002 + 0.0005 * i) for i in range(4)]
    n = 128
    eps = np.full((n, n, n), 1.33**2, np.complex128)
    idx = np.indices((n, n, n))
    for c, r in (((40, 60, 64), 14), ((88, 70, 60), 10)):
        eps[np.sum((idx - np.array(c)[:, None, None, None]) ** 2, axis=0) <= r * r] = 1.38**2
    eps = om.add_absorbing_bc(eps[8:-8, 8:-8, 8:-8], 0.133, (0, 1, 2), 8 * 0.133, n_background=1.33)
    J = np.zeros((n, n, n, 3), np.complex64)
    J[:, :, 12, 0] = 1.0
    born = sv.BornRequest(torch.as_tensor(eps.astype(np.complex64)).cuda(), torch.as_tensor(J).cuda(),
                          0.133, 2 * np.pi / 0.532, rtol=1e-5, arrival=0.003, use_bicgstab=True)
    bad_z = _prop(256, 4, rng, 0.0035)
    bad_z.z = np.array([1.0, np.nan, 2.0, 3.0])
    bad_shape = _prop(256, 4, rng, 0.004)
    bad_shape.phase = bad_shape.phase[:, :128].contiguous()
    bad_wv = sv.SasCztRequest(sas.u[..., :4], 0.5, C5_WL, 2000.0, 1.0, 512, 2.0, arrival=0.0045)
    return [sas, big] + small + [born, bad_z, bad_shape, bad_wv]


def _check_c5_oracle(reqs):
    from oracle import czt as oc
    from oracle import mbs as om
    from oracle import optics as oo
    from oracle import propagation as op

    sas = next(r for r in reqs if isinstance(r, sv.SasCztRequest) and r.error is None)
    big = next(r for r in reqs if isinstance(r, sv.PropRequest) and r.u.shape[-1] == 2048)
    born = next(r for r in reqs if isinstance(r, sv.BornRequest))
    for w in (0, 7):
        v, du = oo.transform_propagate_sas(sas.u[..., w].cpu().numpy(), (0.5, 0.5), C5_WL[w], sas.z, sas.n,
                                           precision="ref32")
        ref, _, _ = oc.czt_kernel_propagate(v, np.ones(v.shape), du, np.asarray(du) / sas.zoom,
                                            (sas.m, sas.m))
        err = rel_l2(sas.result[..., w].cpu().numpy(), ref)
        print(f"C5 SAS -> CZT 512^2, wavelength {C5_WL[w]:.3f}: {err:.2e}")
        assert err < 1e-4
    u = (big.u * torch.exp(1j * big.phase)).cpu().numpy()
    for j in (0, 31):
        ref = np.abs(op.propagate(u, (0.3, 0.3), big.wavelength, big.z[j], big.n, 0, precision="ref32",
                                  fft_precision="fp64")) ** 2
        err = rel_l2(big.result[j].cpu().numpy(), ref)
        print(f"C5 2048^2 plane {j} (z = {big.z[j]:.2f}): {err:.2e}")
        assert err < 1e-5
    eps = born.permittivity.cpu().numpy().astype(np.complex128)
    S = om.source_field(born.current.cpu().numpy(), born.k0)
    c, r = om.get_shift_and_scale(eps)
    m = om.MBS(eps, born.spacing, born.k0, c, r)
    pde = m.pde_residual(born.result.cpu().numpy().astype(np.complex128), S)
    print(f"C5 Born 128^3: {born.iterations} BiCGStab iterations, PDE residual {pde:.2e}")
    assert pde < 1e-4


@pytest.mark.timeout(600)
def test_c5_stream_at_stated_sizes():
    """The C5 stream: every valid request served alone (plain path) and in
    the engine's mix gives the same result, the FP64 oracle agrees with the
    mix (SAS -> CZT at two wavelengths, two of the 32 planes, the Born
    field's PDE residual), the malformed requests are rejected with an
    error while the others complete, and with a 1-byte HBM budget the same
    stream is served one request at a time with the same results."""
    plain = sv.run_plain(_c5_mix(), keep_results=True)
    runs = []
    for budget in (64 << 30, 1):
        eng = sv.ServingEngine(max_batch=4, keep_results=True, hbm_budget=budget)
        mix = _c5_mix()
        eng.warm(mix)
        runs.append(eng.run(mix))
        torch.cuda.synchronize()
    for served in runs:
        for a, b in zip(plain, served):
            assert b.done is not None
            if a.error is not None:
                assert b.error is not None and b.result is None
                continue
            assert b.error is None, b.error
            ra, rb = a.result.cpu().numpy(), b.result.cpu().numpy()
            assert ra.shape == rb.shape
            if isinstance(a, sv.BornRequest):
                assert a.iterations == b.iterations
            err = np.linalg.norm(ra - rb) / np.linalg.norm(ra)
            assert err <= 1e-6, (type(a).__name__, err)
        st = sv.latency_stats(served)
        assert st["rejected"] == 3 and st["n"] == len(served) - 3
        print(f"C5 engine: {st}")
    errors = [r.error for r in plain if r.error is not None]
    assert "z must be" in errors[0] and "phase mask" in errors[1] and "complex64" in errors[2]
    assert next(r for r in runs[0] if isinstance(r, sv.SasCztRequest)).result.shape == (512, 512, 8)
    _check_c5_oracle(runs[0])

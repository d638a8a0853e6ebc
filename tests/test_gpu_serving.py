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

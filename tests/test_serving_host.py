"""Host-side behaviour of the serving loop (chromatix_amd.serving), no GPU:
request validation rejects malformed requests with a message, and the HBM
admission policy queues requests that do not fit instead of failing them,
keeps arrival order and never exceeds the budget except for one request
larger than the whole budget running alone."""
import numpy as np
import pytest
import torch

from chromatix_amd import serving as sv


def _prop(N=16, Z=3, z=None):
    u = torch.ones((N, N), dtype=torch.complex64)
    ph = torch.zeros((N, N), dtype=torch.float32)
    return sv.PropRequest(u, ph, np.linspace(-5, 5, Z) if z is None else z, 0.3, 0.532, 1.33)


def _sas(N=16, wv=3):
    return sv.SasCztRequest(torch.ones((N, N, wv), dtype=torch.complex64), 0.5,
                            np.linspace(0.45, 0.65, wv), 15.0, 1.33, 8, 2.0)


def _born(n=8):
    eps = torch.full((n, n, n), 1.33**2, dtype=torch.complex64)
    J = torch.zeros((n, n, n, 3), dtype=torch.complex64)
    return sv.BornRequest(eps, J, 0.133, 11.8)


def test_valid_requests_pass():
    for r in (_prop(), _sas(), _born()):
        sv.validate(r)


@pytest.mark.parametrize("case", ["nan_z", "inf_z", "empty_z", "phase_shape", "phase_dtype",
                                  "field_not_square", "nan_dx", "neg_wavelength"])
def test_malformed_propagation_requests_are_rejected(case):
    r = _prop()
    if case == "nan_z":
        r.z = np.array([1.0, np.nan, 2.0])
    elif case == "inf_z":
        r.z = np.array([np.inf])
    elif case == "empty_z":
        r.z = np.array([])
    elif case == "phase_shape":
        r.phase = torch.zeros((16, 8), dtype=torch.float32)
    elif case == "phase_dtype":
        r.phase = torch.zeros((16, 16), dtype=torch.float64)
    elif case == "field_not_square":
        r.u = torch.ones((16, 8), dtype=torch.complex64)
    elif case == "nan_dx":
        r.dx = float("nan")
    elif case == "neg_wavelength":
        r.wavelength = -0.5
    with pytest.raises(ValueError):
        sv.validate(r)


def test_malformed_sas_and_born_requests_are_rejected():
    r = _sas()
    r.wavelengths = np.array([0.5, 0.6])  # 3 wavelength planes, 2 wavelengths
    with pytest.raises(ValueError, match="complex64"):
        sv.validate(r)
    r = _sas()
    r.z = float("nan")
    with pytest.raises(ValueError, match="z must be finite"):
        sv.validate(r)
    r = _born()
    r.current = torch.zeros((8, 8, 8, 2), dtype=torch.complex64)
    with pytest.raises(ValueError, match="current must be"):
        sv.validate(r)
    r = _born()
    r.permittivity[1, 2, 3] = complex(float("nan"), 0.0)
    with pytest.raises(ValueError, match="finite"):
        sv.validate(r)


def test_plain_path_rejects_without_touching_the_device():
    bad = [_prop(z=np.array([np.nan])), _sas(wv=2)]
    bad[1].wavelengths = np.array([0.5])
    out = sv.run_plain(bad)
    assert all(r.error is not None and r.result is None and r.done is not None for r in out)
    assert "z must be" in out[0].error


def _simulate(sizes, budget, complete_after=2):
    """The engine's admission loop with the device replaced by a clock: a
    request admitted at tick t completes at t + complete_after.  Returns the
    admission order and the largest in-flight byte count seen with more
    than one request in flight."""
    adm = sv.Admission(budget)
    pending = list(range(len(sizes)))
    inflight, order, peak_multi, t = [], [], 0, 0
    while pending or inflight:
        for i, t_done in list(inflight):
            if t_done <= t:
                inflight.remove((i, t_done))
                adm.release(sizes[i])
        while pending and adm.fits(sizes[pending[0]]):
            i = pending.pop(0)
            adm.admit(sizes[i])
            inflight.append((i, t + complete_after))
            order.append(i)
        if adm.n_active > 1:
            peak_multi = max(peak_multi, adm.in_use)
        t += 1
        assert t < 10000
    assert adm.in_use == 0 and adm.n_active == 0
    return order, peak_multi


def test_tiny_budget_queues_instead_of_failing():
    sizes = [100, 5000, 20, 7, 300]
    order, peak = _simulate(sizes, budget=1)
    assert order == list(range(len(sizes)))  # every request served, in arrival order
    assert peak == 0  # one at a time


def test_budget_is_respected_and_order_kept():
    rng = np.random.default_rng(0)
    sizes = list(rng.integers(1, 60, 40)) + [500] + list(rng.integers(1, 60, 10))
    order, peak = _simulate(sizes, budget=100)
    assert order == list(range(len(sizes)))
    assert peak <= 100  # the 500-byte request ran alone


def test_request_bytes_cover_inputs_and_outputs():
    r = _prop(N=64, Z=4)
    assert sv.ServingEngine._bytes(r) >= 64 * 64 * (8 + 4) + 4 * 64 * 64 * 4
    assert sv.ServingEngine._bytes(_born(16)) >= 16**3 * 8 * 4

"""Request-serving loop for the propagation path: continuous batching of
PSF / propagation requests by shape bucket with one captured HIP graph per
bucket, SAS + CZT zoom and Born requests alongside, admission by an HBM
budget, p50/p99 latency and throughput.

Propagation requests ("a field, a phase mask, Z defocus distances ->
intensity stack") of equal (grid, Z) share a bucket.  Each bucket keeps
device-resident slots: B input planes, a table of B*Z transfer-function
parameter records (cxm_prop2d_c64_params: one record per output plane, so
every request keeps its own distances) and the B*Z output planes.  A batch
is: per request one phase-mask kernel writing u * exp(i phi) into its slot,
then one graph replay (parameter-table copy from pinned host memory + the
three fused propagation passes).  Requests that arrive while batches run
are admitted as slots and memory free up; batches of different buckets run
on different streams and overlap on the device.

SAS + CZT requests (a chromatic field -> scalable angular spectrum
propagation -> per wavelength a CZT zoom onto a smaller output window) run
eagerly on their own stream; Born requests (modified Born series solves)
advance in chunks of iterations on their own streams, their residuals read
back through pinned memory and events, so the loop never waits for an
iteration.  It does wait once per Born request, at admission: the shift of
get_shift_and_scale (cxm_enclosing_circle_c64) returns to the host, which
waits for the permittivity's stream.

Admission keeps the bytes of admitted, unfinished requests within an HBM
budget (a request that does not fit waits, in arrival order; one larger
than the budget runs alone), and malformed requests are rejected with a
message on the request (`validate`) while the others complete.

The plain form for comparison runs every request alone, in arrival order,
one library call per operation, waiting for each to finish.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from .core.field import ChromaticScalarField, ScalarField
from .core.spectrum import Spectrum

__all__ = ["PropRequest", "SasCztRequest", "BornRequest", "ServingEngine", "run_plain",
           "latency_stats", "validate", "Admission"]


@dataclass
class PropRequest:
    """Field u (N, N) complex64 and phase mask (N, N) float32 on the GPU,
    propagated (angular spectrum) to each of the distances z (Z,);
    result: intensity (Z, N, N) float32."""
    u: torch.Tensor
    phase: torch.Tensor
    z: np.ndarray
    dx: float
    wavelength: float
    n: float
    arrival: float = 0.0
    result: torch.Tensor | None = None
    done: float | None = None
    error: str | None = None


@dataclass
class SasCztRequest:
    """Chromatic field (N, N, wv) propagated by z with the scalable angular
    spectrum method, then zoomed by a CZT onto an (m, m) window of spacing
    ``zoom_dx`` per wavelength; result (m, m, wv) complex64."""
    u: torch.Tensor
    dx: float
    wavelengths: np.ndarray
    z: float
    n: float
    m: int
    zoom: float
    arrival: float = 0.0
    result: torch.Tensor | None = None
    done: float | None = None
    error: str | None = None


@dataclass
class BornRequest:
    """Modified Born series solve of a permittivity volume with a source, to
    |prec_y - prec_forward(x)| <= rtol |prec_y|, by the fixed-point iteration
    or (use_bicgstab) BiCGStab."""
    permittivity: torch.Tensor
    current: torch.Tensor
    spacing: float
    k0: float
    rtol: float = 1e-4
    arrival: float = 0.0
    use_bicgstab: bool = False
    result: torch.Tensor | None = None
    done: float | None = None
    iterations: int = 0
    error: str | None = None


def _finite_pos(x) -> bool:
    return bool(np.isfinite(x)) and x > 0


def validate(r) -> None:
    """Raise ValueError for a malformed request: shapes, dtypes, devices and
    the finiteness of its parameters (host checks; a Born request's
    permittivity and source are checked for non-finite values too, since
    the iteration on them would never converge)."""
    if isinstance(r, PropRequest):
        if r.u.ndim != 2 or r.u.shape[0] != r.u.shape[1] or r.u.dtype != torch.complex64:
            raise ValueError(f"field must be (N, N) complex64, got {tuple(r.u.shape)} {r.u.dtype}")
        if tuple(r.phase.shape) != tuple(r.u.shape) or r.phase.dtype != torch.float32:
            raise ValueError(f"phase mask must be {tuple(r.u.shape)} float32, got "
                             f"{tuple(r.phase.shape)} {r.phase.dtype}")
        if r.phase.device != r.u.device:
            raise ValueError("field and phase mask on different devices")
        z = np.asarray(r.z, dtype=np.float64)
        if z.ndim != 1 or z.size == 0 or z.size > 4096 or not np.all(np.isfinite(z)):
            raise ValueError("z must be a non-empty list of finite distances")
        if not (_finite_pos(r.dx) and _finite_pos(r.wavelength) and _finite_pos(r.n)):
            raise ValueError("dx, wavelength and n must be finite and positive")
    elif isinstance(r, SasCztRequest):
        wl = np.asarray(r.wavelengths, dtype=np.float64).reshape(-1)
        if (r.u.ndim != 3 or r.u.shape[0] != r.u.shape[1] or r.u.shape[2] != wl.size
                or r.u.dtype != torch.complex64):
            raise ValueError(f"field must be (N, N, {wl.size}) complex64, got {tuple(r.u.shape)} {r.u.dtype}")
        if not np.all(np.isfinite(wl)) or np.any(wl <= 0):
            raise ValueError("wavelengths must be finite and positive")
        if not np.isfinite(r.z) or r.z == 0:
            raise ValueError("z must be finite and non-zero")
        if not (_finite_pos(r.dx) and _finite_pos(r.n) and _finite_pos(r.zoom)) or int(r.m) < 2:
            raise ValueError("dx, n and zoom must be finite and positive, m >= 2")
    elif isinstance(r, BornRequest):
        eps, J = r.permittivity, r.current
        if eps.ndim != 3 or not eps.is_complex():
            raise ValueError(f"permittivity must be (Nx, Ny, Nz) complex, got {tuple(eps.shape)} {eps.dtype}")
        if tuple(J.shape) != tuple(eps.shape) + (3,):
            raise ValueError(f"current must be {tuple(eps.shape) + (3,)}, got {tuple(J.shape)}")
        if J.device != eps.device:
            raise ValueError("permittivity and current on different devices")
        if not (_finite_pos(r.spacing) and _finite_pos(r.k0) and _finite_pos(r.rtol)):
            raise ValueError("spacing, k0 and rtol must be finite and positive")
        if not (bool(torch.isfinite(torch.view_as_real(eps)).all())
                and bool(torch.isfinite(torch.view_as_real(J.to(torch.complex64))).all())):
            raise ValueError("permittivity and current must be finite")
    else:
        raise ValueError(f"unknown request type {type(r).__name__}")


class Admission:
    """Admission by an HBM budget.  A request is admitted while the bytes
    of the admitted, unfinished requests plus its own stay within the
    budget; one that does not fit waits, and nothing that arrived after it
    overtakes it, until completions free memory; one larger than the whole
    budget runs alone once nothing else is in flight.  A request never
    fails for lack of budget: it waits."""

    def __init__(self, budget: int):
        self.budget = int(budget)
        self.in_use = 0
        self.n_active = 0

    def fits(self, need: int) -> bool:
        return self.n_active == 0 or self.in_use + need <= self.budget

    def admit(self, need: int) -> None:
        self.in_use += need
        self.n_active += 1

    def release(self, need: int) -> None:
        self.in_use -= need
        self.n_active -= 1


_ONES = L.doubles([1.0] * 32)  # placeholder z list for size queries (the table carries the distances)


def _prop_desc(N, Z, dx, wavelength, n, batch):
    d = L.Prop2dDesc()
    d.z = ctypes.cast(_ONES, ctypes.POINTER(ctypes.c_double))
    d.batch, d.height, d.width = batch, N, N
    d.num_z = Z
    d.kind = L.CXM_H_ASM
    d.phase_mode = L.CXM_PHASE_REF32
    d.out_kind = L.CXM_OUT_INTENSITY
    d.num_wv = 1
    d.dy = d.dx = float(dx)
    d.wavelength = float(wavelength)
    d.n = float(n)
    return d


def _phase_into(u, phase, dst, stream_ptr):
    N = u.shape[-1]
    L.check(L.lib().cxm_phase_change_c64(stream_ptr, u.data_ptr(), phase.data_ptr(), dst.data_ptr(),
                                         1, N, N, 1, None))


class _Bucket:
    """Graph instance for (N, Z, B): slots, parameter table, outputs."""

    def __init__(self, N, Z, B, dx, wavelength, n, device, use_graph):
        self.N, self.Z, self.B = N, Z, B
        self.d = _prop_desc(N, Z, dx, wavelength, n, B)
        self.rec = L.lib().cxm_prop2d_param_bytes()
        self.slots = torch.zeros((B, N, N), dtype=torch.complex64, device=device)
        self.out = torch.empty((B * Z, N, N), dtype=torch.float32, device=device)
        self.params_host = torch.empty(B * Z * self.rec, dtype=torch.uint8, pin_memory=True)
        self.params_dev = torch.empty(B * Z * self.rec, dtype=torch.uint8, device=device)
        self.ws = torch.empty(max(1, L.lib().cxm_prop2d_workspace_size(ctypes.byref(self.d))),
                              dtype=torch.uint8, device=device)
        self.stream = torch.cuda.Stream(device=device)
        self.event = torch.cuda.Event()
        self.busy = False
        self.requests: list = []
        self.graph = None
        if use_graph:
            self._capture()

    def _launch(self):
        self.params_dev.copy_(self.params_host, non_blocking=True)
        L.check(L.lib().cxm_prop2d_c64_params(
            L.stream_ptr(self.slots.device), self.slots.data_ptr(), self.out.data_ptr(),
            self.ws.data_ptr(), self.ws.numel(), ctypes.byref(self.d), self.params_dev.data_ptr()))

    def _capture(self):
        # a valid table for the warm-up and capture launches
        self._fill_params([np.zeros(self.Z) + 1.0] * self.B)
        with torch.cuda.stream(self.stream):
            self._launch()  # warm-up outside capture
        self.stream.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            self._launch()
        self.stream.synchronize()

    def _fill_params(self, zlists):
        zs = np.zeros(self.B * self.Z)
        for b, z in enumerate(zlists):
            zs[b * self.Z:(b + 1) * self.Z] = z
        zbuf = L.doubles(zs)
        L.check(L.lib().cxm_prop2d_build_params(ctypes.byref(self.d),
                                                ctypes.cast(zbuf, ctypes.POINTER(ctypes.c_double)),
                                                self.B * self.Z,
                                                ctypes.c_void_p(self.params_host.data_ptr())))

    def start(self, reqs):
        self.requests = list(reqs)
        self._fill_params([r.z for r in reqs] + [np.ones(self.Z)] * (self.B - len(reqs)))
        # the inputs were written on the submitting stream
        self.stream.wait_stream(torch.cuda.current_stream(self.slots.device))
        with torch.cuda.stream(self.stream):
            sp = L.stream_ptr(self.slots.device)
            for b, r in enumerate(reqs):
                _phase_into(r.u, r.phase, self.slots[b], sp)
            if self.graph is not None:
                self.graph.replay()
            else:
                self._launch()
            self.event.record(self.stream)
        self.busy = True

    def poll(self, keep_results: bool, now_fn):
        if not self.busy or not self.event.query():
            return []
        t = now_fn()
        done = []
        for b, r in enumerate(self.requests):
            if keep_results:
                with torch.cuda.stream(self.stream):
                    r.result = self.out[b * self.Z:(b + 1) * self.Z].clone()
            r.done = t
            done.append(r)
        self.busy = False
        self.requests = []
        return done


def _sas_czt(r: SasCztRequest, stream):
    from .functional.czt_propagation import czt_kernel_propagate
    from .functional.optics import transform_propagate_sas

    with torch.cuda.stream(stream):
        f = ChromaticScalarField(r.u, r.dx, 0.0, Spectrum(r.wavelengths))
        s = transform_propagate_sas(f, r.z, r.n)
        N = r.u.shape[0]
        ones = torch.ones((N, N), dtype=torch.complex64, device=r.u.device)
        # per wavelength a zoom onto its own output spacing dx_w / zoom
        return czt_kernel_propagate(s, ones, output_dx=s.dx / r.zoom, output_shape=(r.m, r.m)).u


def _born(r: BornRequest):
    from .experimental import modified_born_series as mbs

    sample = mbs.Sample(r.permittivity, r.spacing)
    src = mbs.Source(r.current, r.k0)
    if r.use_bicgstab:
        res = mbs.solve(sample, src, maxiter=3000, tol=r.rtol, use_bicgstab=True, return_info=True)
    else:
        res = mbs.solve(sample, src, maxiter=3000, rtol=r.rtol, return_info=True)
    return res.E, res.iterations


_BORN_CHUNK = 8  # iterations enqueued between residual checks


class _BornJob:
    """A Born solve advanced by the serving loop without blocking: chunks of
    iterations on its own stream, residual read back through pinned memory."""

    def __init__(self, r: BornRequest, device):
        from .experimental import modified_born_series as mbs

        self.r = r
        self.stream = torch.cuda.Stream(device=device)
        self.stream.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(self.stream):
            sample, src = mbs.Sample(r.permittivity, r.spacing), mbs.Source(r.current, r.k0)
            if r.use_bicgstab:
                self.sv = mbs.BiCGStabSolver(sample, src, tol=r.rtol, maxiter=3000)
            else:
                self.sv = mbs.Solver(sample, src, rtol=r.rtol, maxiter=3000)
            self.sv.step(_BORN_CHUNK)
        self.out_event = None

    def advance(self) -> bool:
        """True when finished (result enqueued and complete)."""
        if self.out_event is not None:
            return self.out_event.query()
        if not self.sv.poll():
            return False
        with torch.cuda.stream(self.stream):
            if self.sv.converged:
                self.r.result = self.sv.result()
                self.r.iterations = getattr(self.sv, "k", self.sv.iterations)
                self.out_event = torch.cuda.Event()
                self.out_event.record(self.stream)
            else:
                self.sv.step(_BORN_CHUNK)
        return False


class ServingEngine:
    """Continuous-batching engine (see the module docstring)."""

    def __init__(self, device=None, hbm_budget: int = 64 << 30, max_batch: int = 8,
                 use_graphs: bool = True, keep_results: bool = False):
        self.device = torch.device(device or "cuda")
        self.hbm_budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.use_graphs = use_graphs
        self.keep_results = keep_results
        self.buckets: dict[tuple, list[_Bucket]] = {}
        self.side_stream = torch.cuda.Stream(device=self.device)
        self.t0 = 0.0

    def _now(self):
        return time.perf_counter() - self.t0

    @staticmethod
    def _bytes(r) -> int:
        if isinstance(r, PropRequest):
            N, Z = r.u.shape[-1], len(r.z)
            return N * N * (8 + 4) + Z * N * N * 4
        if isinstance(r, SasCztRequest):
            N, W = r.u.shape[0], r.u.shape[-1]
            return 8 * N * N * W * 12
        n = int(np.prod(r.permittivity.shape))
        return n * 8 * 12

    def _batch_size(self, N, Z, queued):
        cap = max(1, min(self.max_batch, (2 << 30) // max(1, Z * N * N * 4)))
        b = 1
        while b * 2 <= min(cap, queued):
            b *= 2
        return b

    def _instance(self, r: PropRequest, B):
        N, Z = r.u.shape[-1], len(r.z)
        key = (N, Z, B, r.dx, r.wavelength, r.n)
        insts = self.buckets.setdefault(key, [])
        for inst in insts:
            if not inst.busy:
                return inst
        if len(insts) < 2:
            inst = _Bucket(N, Z, B, r.dx, r.wavelength, r.n, self.device, self.use_graphs)
            insts.append(inst)
            return inst
        return None

    def warm(self, requests):
        """Capture the graphs the request mix will need (outside the timed run)."""
        seen = set()
        for r in requests:
            try:
                validate(r)
            except ValueError:
                continue
            if isinstance(r, PropRequest):
                bmax = self._batch_size(r.u.shape[-1], len(r.z), self.max_batch)
                for B in (1, 2, 4, 8):
                    if B > bmax:
                        break
                    key = (r.u.shape[-1], len(r.z), B, r.dx, r.wavelength, r.n)
                    if key not in seen:
                        seen.add(key)
                        insts = self.buckets.setdefault(key, [])
                        while len(insts) < 2:  # double-buffered per bucket
                            insts.append(_Bucket(r.u.shape[-1], len(r.z), B, r.dx, r.wavelength, r.n,
                                                 self.device, self.use_graphs))
        torch.cuda.synchronize(self.device)

    def run(self, requests):
        """Serve ``requests`` (sorted by .arrival, seconds from start);
        returns them with .done set (and .result with keep_results)."""
        reqs = sorted(requests, key=lambda q: q.arrival)
        pending = list(reqs)
        queue: dict[tuple, list[PropRequest]] = {}
        side: list = []
        born: list[_BornJob] = []
        adm = Admission(self.hbm_budget)
        active = []
        finished = []
        self.t0 = time.perf_counter()

        def reject(r, e):
            r.error, r.result, r.done = f"{type(e).__name__}: {e}", None, self._now()
            finished.append(r)

        while pending or queue or active or side or born:
            now = self._now()
            # admit arrivals within the HBM budget; malformed ones are rejected
            while pending and pending[0].arrival <= now:
                r = pending[0]
                try:
                    validate(r)
                except ValueError as e:
                    pending.pop(0)
                    reject(r, e)
                    continue
                need = self._bytes(r)
                if not adm.fits(need):
                    break
                pending.pop(0)
                if isinstance(r, PropRequest):
                    adm.admit(need)
                    queue.setdefault((r.u.shape[-1], len(r.z), r.dx, r.wavelength, r.n), []).append(r)
                elif isinstance(r, SasCztRequest):
                    self.side_stream.wait_stream(torch.cuda.current_stream(self.device))
                    try:
                        with torch.cuda.stream(self.side_stream):
                            r.result = _sas_czt(r, self.side_stream)
                            ev = torch.cuda.Event()
                            ev.record(self.side_stream)
                    except (ValueError, NotImplementedError, L.CxmError) as e:
                        reject(r, e)
                        continue
                    adm.admit(need)
                    side.append((r, ev))
                else:
                    try:
                        job = _BornJob(r, self.device)
                    except (ValueError, NotImplementedError, L.CxmError) as e:
                        reject(r, e)
                        continue
                    adm.admit(need)
                    born.append(job)
            # schedule batches
            for key in list(queue):
                q = queue[key]
                while q:
                    B = self._batch_size(key[0], key[1], len(q))
                    inst = self._instance(q[0], B)
                    if inst is None:
                        break
                    batch, queue[key] = q[:B], q[B:]
                    q = queue[key]
                    inst.start(batch)
                    active.append(("prop", inst, None))
                if not queue[key]:
                    del queue[key]
            # completions
            still = []
            for kind, obj, t in active:
                done = obj.poll(self.keep_results, self._now)
                if done:
                    for r in done:
                        adm.release(self._bytes(r))
                    finished += done
                else:
                    still.append((kind, obj, t))
            active = still
            keep_b = []
            for job in born:
                if job.advance():
                    job.r.done = self._now()
                    if not self.keep_results:
                        job.r.result = None
                    adm.release(self._bytes(job.r))
                    finished.append(job.r)
                else:
                    keep_b.append(job)
            born = keep_b
            keep = []
            for r, ev in side:
                if ev.query():
                    r.done = self._now()
                    if not self.keep_results:
                        r.result = None
                    adm.release(self._bytes(r))
                    finished.append(r)
                else:
                    keep.append((r, ev))
            side = keep
            if not active and not side and not queue and not born and pending:
                wait = pending[0].arrival - self._now()
                if wait > 0:
                    time.sleep(min(wait, 0.002))
        torch.cuda.synchronize(self.device)
        return reqs


def run_plain(requests, keep_results: bool = False):
    """Each request alone, in arrival order, one call per operation,
    synchronised before the next starts; malformed requests are rejected
    (.error set) as by the engine."""
    from .functional.propagation import asm_propagate
    from .functional.phase_masks import phase_change

    reqs = sorted(requests, key=lambda q: q.arrival)
    t0 = time.perf_counter()
    for r in reqs:
        wait = r.arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        try:
            validate(r)
        except ValueError as e:
            r.error, r.result, r.done = f"{type(e).__name__}: {e}", None, time.perf_counter() - t0
            continue
        if isinstance(r, PropRequest):
            f = ScalarField(r.u, r.dx, 0.0, r.wavelength)
            f = phase_change(f, r.phase)
            out = asm_propagate(f, np.asarray(r.z)[:, None, None] if len(r.z) > 1 else float(r.z[0]),
                                r.n, 0, mode="same")
            res = out.intensity
        elif isinstance(r, SasCztRequest):
            res = _sas_czt(r, torch.cuda.current_stream())
        else:
            res, r.iterations = _born(r)
        torch.cuda.synchronize()
        r.done = time.perf_counter() - t0
        r.result = res if keep_results else None
    return reqs


def latency_stats(reqs) -> dict:
    """p50 / p99 latency (arrival to completion), throughput over the span
    from the first arrival to the last completion, and per request type;
    rejected requests are counted apart and take no part in the rest."""
    ok = [r for r in reqs if r.done is not None and r.error is None]
    lat = np.array([r.done - r.arrival for r in ok])
    span = max(r.done for r in ok) - min(r.arrival for r in ok)
    out = {"n": int(lat.size), "rejected": sum(1 for r in reqs if r.error is not None),
           "p50_ms": float(np.percentile(lat, 50) * 1e3),
           "p99_ms": float(np.percentile(lat, 99) * 1e3), "throughput_rps": float(lat.size / span),
           "span_s": float(span)}
    for cls in (PropRequest, SasCztRequest, BornRequest):
        lt = np.array([r.done - r.arrival for r in ok if isinstance(r, cls)])
        if lt.size:
            out[cls.__name__] = {"n": int(lt.size), "p50_ms": float(np.percentile(lt, 50) * 1e3),
                                 "p99_ms": float(np.percentile(lt, 99) * 1e3)}
    return out

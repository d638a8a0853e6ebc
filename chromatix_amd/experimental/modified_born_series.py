"""Modified Born series solver on libcxm (mirrors
src/chromatix/experimental/modified_born_series/{sample,solver}.py).

The preconditioned fixed-point iteration runs on the GPU (cxm_mbs_*): one
launch sequence per iteration, V multiply, 3D FFT, Green multiply, inverse
3D FFT and the update with an FP64 residual norm.  ``use_bicgstab=True``
runs BiCGStab on the same preconditioned operator (cxm_mbs_bicgstab_*):
two operator applications per iteration, the dot products reduced in FP64
into device scalars, no host round trip inside a chunk of iterations.
Setup (absorbing boundary, the bias/scale of get_shift_and_scale) is host
code.

get_shift_and_scale: the reference minimises max |eps - shift|^2 with a
nonlinear CG to 1e-3 tolerance; here the minimiser is computed exactly, as
the centre of the smallest circle enclosing the permittivity values in the
complex plane (Welzl's algorithm on the distinct values).  Any shift/scale
meeting the accretivity conditions gives the same converged field; the
iteration count depends on it.
"""
from __future__ import annotations

import ctypes
import random

import numpy as np
import scipy.constants as const
import torch

from .. import _lib as L

__all__ = ["Sample", "EmptySample", "Source", "add_absorbing_bc", "get_shift_and_scale", "solve",
           "solve_batch", "MBSResult", "Solver", "BiCGStabSolver"]


class Sample:
    """Permittivity on a uniform grid (sample.py:9-73): (Nx, Ny, Nz), or
    (3, 3, Nx, Ny, Nz) for a permittivity tensor per voxel (the layout the
    reference's solver docstrings give for anisotropic materials)."""

    def __init__(self, permittivity, spacing: float, roi=None):
        self.permittivity = torch.as_tensor(permittivity)
        if self.permittivity.ndim not in (3, 5) or (
                self.permittivity.ndim == 5 and tuple(self.permittivity.shape[:2]) != (3, 3)):
            raise ValueError("permittivity must be (Nx, Ny, Nz) or (3, 3, Nx, Ny, Nz)")
        self.spacing = float(spacing)
        self.ROI = roi if roi is not None else tuple(slice(0, s) for s in self.shape)

    @property
    def tensor(self) -> bool:
        return self.permittivity.ndim == 5

    @property
    def shape(self):
        return tuple(self.permittivity.shape[-3:])

    @property
    def grid(self) -> np.ndarray:
        axes = [np.linspace(0, n - 1, n) - (n // 2 if n != 1 else 0) for n in self.shape]
        return np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1) * self.spacing

    @property
    def k_grid(self) -> np.ndarray:
        axes = [np.fft.fftfreq(n) for n in self.shape]
        return 2 * np.pi * np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1) / self.spacing

    @property
    def extent(self) -> np.ndarray:
        return np.asarray(self.shape) * self.spacing

    def replace(self, **kw) -> "Sample":
        new = Sample(kw.get("permittivity", self.permittivity), kw.get("spacing", self.spacing),
                     kw.get("ROI", self.ROI))
        return new


def EmptySample(shape, spacing: float, device=None) -> Sample:
    return Sample(torch.zeros(tuple(shape), dtype=torch.complex64, device=device), spacing)


class Source:
    """Pre-scaled current density: field = -i / k0 1e-6 c mu0 J (sample.py:76-89)."""

    def __init__(self, current_density, k0: float):
        self.k0 = float(k0)
        J = torch.as_tensor(current_density)
        self.field = (J.to(torch.complex64) * complex(-1j / self.k0 * 1e-6 * const.c * const.mu_0))

    @property
    def current_density(self):
        return self.field / complex(-1j / self.k0 * 1e-6 * const.c * const.mu_0)


def add_absorbing_bc(sample: Sample, axis, thickness: float, max_extinction: float = 0.25,
                     n_background: float = 1.0) -> Sample:
    """Pad by int(thickness / spacing) per side along ``axis`` with a linear
    extinction ramp to ``max_extinction``: eps += (n_bg + i kappa)^2 - n_bg^2
    (sample.py:92-124)."""
    axes = axis if isinstance(axis, tuple) else (axis,)
    w = int(thickness / sample.spacing)
    widths = [(w, w) if i in axes else (0, 0) for i in range(3)]
    ext = np.pad(np.zeros(sample.shape), widths, mode="linear_ramp", end_values=max_extinction)
    eps_bc = (n_background + 1j * ext) ** 2
    eps = sample.permittivity.detach().cpu().numpy().astype(np.complex128)
    roi = tuple(slice(a, a + s) for (a, _), s in zip(widths, sample.shape))
    if sample.tensor:
        # isotropic background n^2 I with the extinction ramp on the diagonal
        perm = np.pad(eps, [(0, 0), (0, 0)] + widths, mode="constant", constant_values=0.0)
        for a in range(3):
            perm[a, a] = np.pad(eps[a, a], widths, mode="constant", constant_values=n_background**2)
            perm[a, a] += eps_bc - n_background**2
    else:
        perm = np.pad(eps, widths, mode="constant", constant_values=n_background**2)
        perm = perm + (eps_bc - n_background**2)
    out = torch.as_tensor(perm.astype(np.complex64), device=sample.permittivity.device)
    return Sample(out, sample.spacing, roi)


def _circle2(a, b):
    c = (a + b) / 2
    return c, abs(a - c)


def _circle3(a, b, c):
    ax, ay, bx, by, cx, cy = a.real, a.imag, b.real, b.imag, c.real, c.imag
    d = 2 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
    if abs(d) < 1e-300:
        # collinear: the widest pair
        best = max(((a, b), (a, c), (b, c)), key=lambda p: abs(p[0] - p[1]))
        return _circle2(*best)
    ux = ((ax * ax + ay * ay) * (by - cy) + (bx * bx + by * by) * (cy - ay) + (cx * cx + cy * cy) * (ay - by)) / d
    uy = ((ax * ax + ay * ay) * (cx - bx) + (bx * bx + by * by) * (ax - cx) + (cx * cx + cy * cy) * (bx - ax)) / d
    o = complex(ux, uy)
    return o, abs(a - o)


def _distinct(values: np.ndarray) -> np.ndarray:
    """Distinct values of a piecewise-constant volume: runs of equal
    neighbours collapse first (cheap), then the sort of np.unique runs on
    what is left."""
    v = np.ascontiguousarray(values).ravel()
    if v.size > 1:
        v = v[np.concatenate(([True], v[1:] != v[:-1]))]
    return np.unique(v)


def enclosing_circle(points) -> tuple[complex, float]:
    """Smallest circle enclosing complex points (Welzl, iterative form)."""
    pts = [complex(p) for p in _distinct(np.asarray(points)).astype(np.complex128)]
    random.Random(0).shuffle(pts)
    eps = 1e-12
    c, r = pts[0], 0.0
    for i, p in enumerate(pts):
        if abs(p - c) <= r * (1 + eps) + eps:
            continue
        c, r = p, 0.0
        for j in range(i):
            q = pts[j]
            if abs(q - c) <= r * (1 + eps) + eps:
                continue
            c, r = _circle2(p, q)
            for k in range(j):
                t = pts[k]
                if abs(t - c) <= r * (1 + eps) + eps:
                    continue
                c, r = _circle3(p, q, t)
    return c, r


def _tensor_shift(eps: np.ndarray) -> tuple[complex, float]:
    """Shift c minimising max over voxels of the spectral norm
    ||eps - c I||_2 and that norm, for (3, 3, ...) permittivities: the centre
    of the circle enclosing every eigenvalue when all the matrices are
    normal (then the norm is the largest |eigenvalue - c|), refined by a
    Nelder-Mead minimisation of the convex maximum otherwise."""
    from scipy.optimize import minimize

    rows = eps.reshape(9, -1).T
    if rows.shape[0] > 1:
        rows = rows[np.concatenate(([True], np.any(rows[1:] != rows[:-1], axis=1)))]
    E = np.unique(rows, axis=0).reshape(-1, 3, 3)
    c0, _ = enclosing_circle(np.linalg.eigvals(E))
    I3 = np.eye(3)

    def norm(c):
        return float(np.max(np.linalg.norm(E - c * I3, ord=2, axis=(1, 2))))

    comm = E @ np.conj(np.swapaxes(E, 1, 2)) - np.conj(np.swapaxes(E, 1, 2)) @ E
    scale = max(1.0, float(np.max(np.abs(E))) ** 2)
    if float(np.max(np.abs(comm))) <= 1e-12 * scale:
        return c0, norm(c0)
    res = minimize(lambda p: norm(complex(p[0], p[1])), np.array([c0.real, c0.imag]),
                   method="Nelder-Mead", options={"xatol": 1e-10, "fatol": 1e-12, "maxiter": 4000})
    c = complex(res.x[0], res.x[1])
    return (c, norm(c)) if norm(c) < norm(c0) else (c0, norm(c0))


def get_shift_and_scale(permittivity) -> tuple[complex, complex]:
    """(shift, scale) with shift minimising max |eps - shift| and
    scale = 1 / (1.1 i sqrt(f)), f = max |eps - shift|^2 (solver.py:10-39);
    for a (3, 3, ...) permittivity tensor |.| is the spectral norm of
    eps - shift I (the norm the reference's note asks for)."""
    t = torch.as_tensor(permittivity).detach()
    if t.ndim != 5 and t.is_cuda:
        # on the device: cxm_enclosing_circle_c64 (candidates pruned on the GPU)
        L.ensure_device(t)
        v = t.to(torch.complex64).resolve_conj().contiguous()
        ws = torch.empty(L.lib().cxm_enclosing_circle_workspace_size(v.numel()), dtype=torch.uint8,
                         device=v.device)
        out = (ctypes.c_double * 3)()
        L.check(L.lib().cxm_enclosing_circle_c64(L.stream_ptr(v.device), v.data_ptr(), v.numel(), out,
                                                 ws.data_ptr(), ws.numel()))
        c, r = complex(out[0], out[1]), float(out[2])
    else:
        eps = t.cpu().numpy()
        if eps.ndim == 5:
            c, r = _tensor_shift(eps.astype(np.complex128))
        else:
            c, r = enclosing_circle(eps)
    if r == 0:
        r = 1e-6  # homogeneous medium: any positive radius keeps the iteration accretive
    return c, 1.0 / (1.1j * r)


class MBSResult:
    def __init__(self, E, iterations, residual):
        self.E, self.iterations, self.residual = E, iterations, residual


def _desc(sample: Sample, k0: float, shift: complex, scale: complex):
    d = L.MbsDesc()
    d.nx, d.ny, d.nz = sample.shape
    d.eps_tensor = 1 if sample.tensor else 0
    d.spacing = sample.spacing
    d.k0 = float(k0)
    d.bias_re, d.bias_im = shift.real, shift.imag
    d.scale_re, d.scale_im = scale.real, scale.imag
    return d


class Solver:
    """Stepwise solve: the work is enqueued in chunks of iterations and the
    stopping test runs on the device after every iteration
    (cxm_mbs_iterate_stop, jaxopt.FixedPointIteration's rule: stop once the
    update norm is at most the tolerance), so iterations enqueued past the
    stop cost nothing.  The status of each chunk comes back through pinned
    memory and an event, so a caller can interleave other work without
    blocking (the serving loop does) or keep the next chunk queued while it
    waits for the last one (``_run`` does)."""

    _RING = 2  # chunks whose status read-back may be in flight at once

    def __init__(self, sample: Sample, source: Source, initial_E=None, *, shift_scale=None,
                 tol: float = 1e-3, rtol: float | None = None, maxiter: int = 1000):
        eps = sample.permittivity.to(torch.complex64).resolve_conj().contiguous()
        L.ensure_device(eps)
        self.eps, self.dev, self.shape = eps, eps.device, sample.shape
        S = torch.as_tensor(source.field, device=self.dev).to(torch.complex64).resolve_conj().contiguous()
        if tuple(S.shape) != sample.shape + (3,):
            raise ValueError(f"source field must have shape {sample.shape + (3,)}")
        shift, scale = shift_scale if shift_scale is not None else get_shift_and_scale(eps)
        self.d = _desc(sample, source.k0, complex(shift), complex(scale))
        self.ws = torch.empty(L.lib().cxm_mbs_workspace_size(ctypes.byref(self.d)), dtype=torch.uint8,
                              device=self.dev)
        self.py = torch.empty((3,) + sample.shape, dtype=torch.complex64, device=self.dev)
        self.num = torch.empty(2, dtype=torch.float64, device=self.dev)
        s = L.stream_ptr(self.dev)
        L.check(L.lib().cxm_mbs_precondition(s, eps.data_ptr(), S.data_ptr(), self.py.data_ptr(),
                                             self.num.data_ptr(), self.ws.data_ptr(), self.ws.numel(),
                                             ctypes.byref(self.d)))
        if initial_E is None:
            self.x = self.py.clone()
        else:
            E0 = torch.as_tensor(initial_E, device=self.dev).to(torch.complex64)
            self.x = (E0.permute(3, 0, 1, 2) / self.num[0].to(torch.float32)).resolve_conj().contiguous()
        self.rtol, self.tol = rtol, float(tol)
        self.maxiter = int(maxiter)
        # device status: [|r|^2 of the last iteration, iterations, stopped, |prec_y|^2]
        self.status_dev = torch.zeros(4, dtype=torch.float64, device=self.dev)
        if rtol is not None:
            pws = torch.empty(max(1, L.lib().cxm_plane_sumsq_workspace_size(1)), dtype=torch.uint8,
                              device=self.dev)
            L.check(L.lib().cxm_plane_sumsq_c64(s, self.py.data_ptr(), self.status_dev[3:].data_ptr(), 1,
                                                self.py.numel(), 1, None, pws.data_ptr(), pws.numel()))
            self._pws = pws
        self._ring = [(torch.zeros(4, dtype=torch.float64, pin_memory=True), torch.cuda.Event())
                      for _ in range(self._RING)]
        self._inflight: list[int] = []
        self._next = 0
        self.event = self._ring[0][1]
        self.enqueued = 0
        self.iterations, self.residual = 0, float("inf")
        self.stopped = self.maxiter <= 0

    @property
    def pending(self) -> bool:
        return bool(self._inflight)

    def step(self, k: int) -> None:
        """Enqueue k iterations (fewer at maxiter) and the status read-back;
        a no-op once the solve is known to have stopped, or while the
        read-back ring is full."""
        k = min(int(k), self.maxiter - self.enqueued)
        if k <= 0 or self.stopped or len(self._inflight) >= self._RING:
            return
        L.check(L.lib().cxm_mbs_iterate_stop(
            L.stream_ptr(self.dev), self.eps.data_ptr(), self.py.data_ptr(), self.x.data_ptr(), k,
            self.maxiter, self.tol, -1.0 if self.rtol is None else float(self.rtol),
            self.status_dev.data_ptr(), self.ws.data_ptr(), self.ws.numel(), ctypes.byref(self.d)))
        host, ev = self._ring[self._next]
        host.copy_(self.status_dev, non_blocking=True)
        ev.record()
        self.event = ev
        self._inflight.append(self._next)
        self._next = (self._next + 1) % self._RING
        self.enqueued += k

    def _take(self, slot: int) -> None:
        r2, it, stop, _ = self._ring[slot][0].tolist()
        self.residual = float(np.sqrt(r2)) if it > 0 else float("inf")
        self.iterations = int(it)
        self.stopped = self.stopped or bool(stop)

    def poll(self) -> bool:
        """True once every enqueued chunk's status has arrived (non-blocking)."""
        while self._inflight and self._ring[self._inflight[0]][1].query():
            self._take(self._inflight.pop(0))
        return not self._inflight

    def wait(self) -> None:
        """Block until the oldest chunk in flight has finished."""
        if self._inflight:
            self._ring[self._inflight[0]][1].synchronize()
            self._take(self._inflight.pop(0))

    @property
    def converged(self) -> bool:
        return self.stopped or self.iterations >= self.maxiter

    def result(self) -> torch.Tensor:
        E = torch.empty(self.shape + (3,), dtype=torch.complex64, device=self.dev)
        L.check(L.lib().cxm_mbs_output(L.stream_ptr(self.dev), self.x.data_ptr(), self.num.data_ptr(),
                                       E.data_ptr(), ctypes.byref(self.d)))
        return E


class BiCGStabSolver(Solver):
    """Stepwise BiCGStab on the preconditioned system (solve(...,
    use_bicgstab=True), solver.py:141-143): the iteration of
    jax.scipy.sparse.linalg.bicgstab without a preconditioner, its scalars
    kept on the device; it stops once |r| <= max(tol |prec_y|, atol) or on
    breakdown, and the steps enqueued after that change nothing."""

    def __init__(self, sample: Sample, source: Source, initial_E=None, *, shift_scale=None,
                 tol: float = 1e-3, atol: float = 0.0, maxiter: int = 1000):
        super().__init__(sample, source, initial_E, shift_scale=shift_scale, tol=tol,
                         maxiter=maxiter)
        lib = L.lib()
        nbytes = lib.cxm_mbs_bicgstab_state_size(ctypes.byref(self.d))
        self.state = torch.empty(nbytes, dtype=torch.uint8, device=self.dev)
        # [|r|^2, k, stopped, |prec_y|^2]
        self.status_dev = torch.zeros(4, dtype=torch.float64, device=self.dev)
        self.status_host = torch.zeros(4, dtype=torch.float64, pin_memory=True)
        L.check(lib.cxm_mbs_bicgstab_init(L.stream_ptr(self.dev), self.eps.data_ptr(), self.py.data_ptr(),
                                          self.x.data_ptr(), self.state.data_ptr(), nbytes, float(tol),
                                          float(atol), self.status_dev.data_ptr(), self.ws.data_ptr(),
                                          self.ws.numel(), ctypes.byref(self.d)))
        self.k, self.stopped = 0, False
        self._readback()

    @property
    def pending(self) -> bool:
        return self._bpending

    def _readback(self):
        self.status_host.copy_(self.status_dev, non_blocking=True)
        self.event.record()
        self._bpending = True

    def step(self, k: int) -> None:
        k = min(int(k), self.maxiter - self.iterations)
        if k <= 0 or self.stopped:
            return
        L.check(L.lib().cxm_mbs_bicgstab_iterate(L.stream_ptr(self.dev), self.eps.data_ptr(),
                                                 self.x.data_ptr(), self.state.data_ptr(),
                                                 self.state.numel(), k, self.status_dev.data_ptr(),
                                                 self.ws.data_ptr(), self.ws.numel(),
                                                 ctypes.byref(self.d)))
        self.iterations += k
        self._readback()

    def poll(self) -> bool:
        if self._bpending and self.event.query():
            rs, k, stopped, _ = self.status_host.tolist()
            self.residual = float(np.sqrt(rs))
            self.k, self.stopped = int(k), bool(stopped)
            self._bpending = False
        return not self._bpending

    def wait(self) -> None:
        self.event.synchronize()
        self.poll()

    @property
    def converged(self) -> bool:
        return self.stopped or self.iterations >= self.maxiter


def _run(sample: Sample, source: Source, initial_E, maxiter, tol, use_bicgstab, rtol, check_every,
         shift_scale):
    """The forward solve; returns the stepped solver and E."""
    if use_bicgstab:
        if rtol is not None:
            raise ValueError("rtol applies to the fixed-point iteration; BiCGStab's tol is relative")
        sv = BiCGStabSolver(sample, source, initial_E, shift_scale=shift_scale, tol=tol,
                            maxiter=maxiter)
        sv.event.synchronize()
        sv.poll()
        while not sv.converged:
            sv.step(check_every)
            sv.event.synchronize()
            sv.poll()
        return sv, sv.result()
    sv = Solver(sample, source, initial_E, shift_scale=shift_scale, tol=tol, rtol=rtol,
                maxiter=maxiter)
    sv.step(check_every)
    while not sv.converged:
        sv.step(check_every)  # the next chunk stays queued while the host waits
        sv.wait()
    return sv, sv.result()


def solve_batch(samples, sources, maxiter: int = 1000, tol: float = 1e-3, use_bicgstab: bool = False,
                *, rtol: float | None = None, check_every: int = 8, concurrency: int = 2,
                return_info: bool = False) -> list:
    """Independent solves of (sample, source) pairs -- what ``jax.vmap`` of
    solve (solver.py:101-146) gives in the reference -- run ``concurrency``
    at a time, each on a stream of its own, so that one volume's setup and
    the gaps between its launches overlap the other volumes' iterations.
    Each result is that of solve() for its pair: the same kernels in the same
    order on its stream.  Forward only (no gradients).  Returns the list of
    E (or MBSResult with ``return_info``), in input order, ready on the
    caller's current stream."""
    samples, sources = list(samples), list(sources)
    if len(samples) != len(sources):
        raise ValueError("solve_batch needs one source per sample")
    if use_bicgstab and rtol is not None:
        raise ValueError("rtol applies to the fixed-point iteration; BiCGStab's tol is relative")
    if concurrency < 1:
        raise ValueError("concurrency must be >= 1")
    if torch.is_grad_enabled() and any(
            s.permittivity.requires_grad or torch.as_tensor(c.field).requires_grad
            for s, c in zip(samples, sources)):
        raise NotImplementedError("solve_batch: no gradients; use solve() per volume")
    out: list = [None] * len(samples)
    if not samples:
        return out
    dev = torch.as_tensor(samples[0].permittivity).device
    L.ensure_device(torch.as_tensor(samples[0].permittivity))
    caller = torch.cuda.current_stream(dev)
    streams = [torch.cuda.Stream(device=dev) for _ in range(min(concurrency, len(samples)))]
    for st in streams:
        st.wait_stream(caller)  # the inputs may still be in flight on the caller's stream
    todo = list(range(len(samples)))
    active: dict = {}

    def start(slot):
        i = todo.pop(0)
        with torch.cuda.stream(streams[slot]):
            if use_bicgstab:
                sv = BiCGStabSolver(samples[i], sources[i], tol=tol, maxiter=maxiter)
            else:
                sv = Solver(samples[i], sources[i], tol=tol, rtol=rtol, maxiter=maxiter)
            # two chunks queued: the stream stays busy while the host reads
            # the first chunk's status
            sv.step(check_every)
            sv.step(check_every)
        active[slot] = (i, sv)

    for slot in range(len(streams)):
        start(slot)
    while active:
        for slot in list(active):
            i, sv = active[slot]
            with torch.cuda.stream(streams[slot]):
                sv.wait()
                if not sv.converged:
                    sv.step(check_every)
                    continue
                E = sv.result()
                E.record_stream(caller)
            its = sv.k if use_bicgstab else sv.iterations
            out[i] = MBSResult(E, its, sv.residual) if return_info else E
            del active[slot]
            if todo:
                start(slot)
    for st in streams:
        caller.wait_stream(st)
    return out


def _adjoint_desc(d):
    """The desc of the conjugated problem: bias and scale conjugated (kb
    follows), the operators of A^H = -V^H (G^H V^H + I)."""
    a = L.MbsDesc.from_buffer_copy(d)
    a.bias_im, a.scale_im = -d.bias_im, -d.scale_im
    return a


def _adjoint_solve(eps, d, Ebar, *, tol, rtol, maxiter, use_bicgstab, check_every):
    """lam (3, Nx, Ny, Nz planar) solving A^H lam = Ebar, A the solve's
    prec_forward: the same iteration as the forward solve on conj(eps) and
    the conjugated desc, with the right-hand side normalised to max |.| = 1
    as the forward normalises prec_y, and the forward's stopping rule."""
    dev = eps.device
    lib = L.lib()
    da = _adjoint_desc(d)
    eps_c = torch.conj_physical(eps).contiguous()
    b = Ebar.permute(3, 0, 1, 2).contiguous()
    norm = float(b.abs().max())
    if norm == 0.0:
        return torch.zeros_like(b)
    if not np.isfinite(norm):
        raise ValueError("the gradient of the solve's output is not finite")
    py = b / norm
    x = py.clone()
    s = L.stream_ptr(dev)
    ws = torch.empty(lib.cxm_mbs_workspace_size(ctypes.byref(da)), dtype=torch.uint8, device=dev)
    it = 0
    if use_bicgstab:
        nbytes = lib.cxm_mbs_bicgstab_state_size(ctypes.byref(da))
        state = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        status = torch.zeros(4, dtype=torch.float64, device=dev)
        L.check(lib.cxm_mbs_bicgstab_init(s, eps_c.data_ptr(), py.data_ptr(), x.data_ptr(),
                                          state.data_ptr(), nbytes, float(tol), 0.0, status.data_ptr(),
                                          ws.data_ptr(), ws.numel(), ctypes.byref(da)))
        while it < maxiter and not bool(status[2].item()):
            k = min(check_every, maxiter - it)
            L.check(lib.cxm_mbs_bicgstab_iterate(s, eps_c.data_ptr(), x.data_ptr(), state.data_ptr(),
                                                 nbytes, k, status.data_ptr(), ws.data_ptr(), ws.numel(),
                                                 ctypes.byref(da)))
            it += k
    else:
        # the forward's stopping rule, tested on the device after every iteration
        status = torch.zeros(4, dtype=torch.float64, device=dev)
        if rtol is not None:
            pws = torch.empty(max(1, lib.cxm_plane_sumsq_workspace_size(1)), dtype=torch.uint8, device=dev)
            L.check(lib.cxm_plane_sumsq_c64(s, py.data_ptr(), status[3:].data_ptr(), 1, py.numel(), 1, None,
                                            pws.data_ptr(), pws.numel()))
        while it < maxiter:
            k = min(check_every, maxiter - it)
            L.check(lib.cxm_mbs_iterate_stop(s, eps_c.data_ptr(), py.data_ptr(), x.data_ptr(), k, maxiter,
                                             float(tol), -1.0 if rtol is None else float(rtol),
                                             status.data_ptr(), ws.data_ptr(), ws.numel(), ctypes.byref(da)))
            it += k
            if bool(status[2].item()):
                break
    return x * norm


class _BornSolveFn(torch.autograd.Function):
    """E = solve(eps, S) with the implicit reverse-mode rule: the adjoint
    system A^H lam = Ebar solved by the forward's own iteration, then
    cxm_mbs_vjp for the permittivity and source gradients.  This is the
    gradient jaxopt's implicit differentiation of the fixed point
    (solver.py:136-142) and jax's custom_linear_solve rule of bicgstab
    (solver.py:143-144) both produce; the shift and scale of
    get_shift_and_scale are constants of it (the converged E does not depend
    on them)."""

    @staticmethod
    def forward(ctx, eps, S, cfg):
        sample = Sample(eps, cfg["spacing"], cfg["roi"])
        src = Source.__new__(Source)
        src.k0, src.field = cfg["k0"], S
        sv, E = _run(sample, src, cfg["initial_E"], cfg["maxiter"], cfg["tol"], cfg["use_bicgstab"],
                     cfg["rtol"], cfg["check_every"], cfg["shift_scale"])
        Sc = S.to(device=sv.dev, dtype=torch.complex64).resolve_conj().contiguous()
        ctx.save_for_backward(sv.eps, Sc, E)
        ctx.d, ctx.cfg, cfg["solver"] = sv.d, cfg, sv
        ctx.src_device = S.device
        return E

    @staticmethod
    def backward(ctx, gE):
        eps, S, E = ctx.saved_tensors
        cfg = ctx.cfg
        ge = gs = None
        if gE is None or not (ctx.needs_input_grad[0] or ctx.needs_input_grad[1]):
            return None, None, None
        gE = gE.to(torch.complex64).resolve_conj().contiguous()
        tol = cfg["adjoint_tol"] if cfg["adjoint_tol"] is not None else cfg["tol"]
        rtol = None if cfg["use_bicgstab"] else cfg["rtol"]
        if cfg["adjoint_tol"] is not None and rtol is not None:
            rtol = cfg["adjoint_tol"]
        lam = _adjoint_solve(eps, ctx.d, gE, tol=tol, rtol=rtol, maxiter=cfg["maxiter"],
                             use_bicgstab=cfg["use_bicgstab"], check_every=cfg["check_every"])
        lib = L.lib()
        if ctx.needs_input_grad[0]:
            ge = torch.empty_like(eps)
        if ctx.needs_input_grad[1]:
            gs = torch.empty_like(S)
        ws = torch.empty(lib.cxm_mbs_vjp_workspace_size(ctypes.byref(ctx.d)), dtype=torch.uint8,
                         device=eps.device)
        L.check(lib.cxm_mbs_vjp(L.stream_ptr(eps.device), eps.data_ptr(), S.data_ptr(), E.data_ptr(),
                                lam.data_ptr(), ge.data_ptr() if ge is not None else None,
                                gs.data_ptr() if gs is not None else None, ws.data_ptr(), ws.numel(),
                                ctypes.byref(ctx.d)))
        if gs is not None and gs.device != ctx.src_device:
            gs = gs.to(ctx.src_device)
        return ge, gs, None


def solve(sample: Sample, source: Source, initial_E=None, maxiter: int = 1000, tol: float = 1e-3,
          implicit_diff: bool = True, use_bicgstab: bool = False, *, rtol: float | None = None,
          check_every: int = 8, shift_scale=None, return_info: bool = False,
          adjoint_tol: float | None = None):
    """Electric field E (Nx, Ny, Nz, 3) solving the scattering problem
    (solver.py:101-146).  Fixed-point iteration (default): stops after the
    first iteration whose update of the normalised iterate has l2 norm at
    most ``tol`` (jaxopt's FixedPointIteration criterion, tested on the
    device after every iteration) or, with ``rtol``, at most rtol |prec_y|.
    ``use_bicgstab``: BiCGStab until |r| <= tol |prec_y|
    (jax.scipy.sparse.linalg.bicgstab's criterion) or breakdown.  The host
    enqueues ``check_every`` iterations at a time and reads the status of
    each chunk back while the next one runs.

    Differentiable in the permittivity and the source field (scalar
    permittivity) by implicit differentiation: the backward pass solves the
    adjoint system with the same iteration and stopping rule (tolerance
    ``adjoint_tol`` when given) -- the gradient jaxopt's implicit_diff=True
    and the bicgstab rule of jax give.  ``initial_E`` gets no gradient (as
    under implicit differentiation); implicit_diff=False, the reference's
    differentiation through the unrolled iteration, is refused."""
    eps = sample.permittivity
    S = torch.as_tensor(source.field)
    if torch.is_grad_enabled() and (eps.requires_grad or S.requires_grad):
        if not implicit_diff:
            raise NotImplementedError(
                "solve: differentiation through the unrolled iteration (implicit_diff=False) is "
                "not implemented; use implicit_diff=True")
        if sample.tensor:
            raise NotImplementedError("solve: gradients for a permittivity tensor are not implemented")
        cfg = dict(spacing=sample.spacing, roi=sample.ROI, k0=source.k0, initial_E=initial_E,
                   maxiter=maxiter, tol=tol, use_bicgstab=use_bicgstab, rtol=rtol,
                   check_every=check_every, shift_scale=shift_scale, adjoint_tol=adjoint_tol)
        E = _BornSolveFn.apply(eps.to(torch.complex64), S, cfg)
        sv = cfg["solver"]
    else:
        sv, E = _run(sample, source, initial_E, maxiter, tol, use_bicgstab, rtol, check_every,
                     shift_scale)
    if return_info:
        return MBSResult(E, sv.k if use_bicgstab else sv.iterations, sv.residual)
    return E

"""Oracle: modified Born series solve (test infrastructure, FP64).

Follows (paths relative to the chromatix repository root):
  Sample.k_grid, add_absorbing_bc, Source  src/chromatix/experimental/modified_born_series/sample.py:9-124
  get_shift_and_scale, precondition, solve src/chromatix/experimental/modified_born_series/solver.py:10-146
  solve(..., use_bicgstab=True)            solver.py:141-143, which calls
      jax.scipy.sparse.linalg.bicgstab (JAX >= 0.4.1, a dependency absent
      here): ``bicgstab`` restates its published iteration (van der Vorst's
      BiCGStab, no preconditioner M, stop on |r|^2 <= max(tol^2 |b|^2, atol^2)
      or breakdown) and is pinned against scipy.sparse.linalg.bicgstab's
      iterates and numpy.linalg.solve on dense systems.
  tensorial permittivity (3, 3, Nx, Ny, Nz): listed as future work in
      modified_born_series/README.md; here V(x) = (s eps - kb I) x per voxel
      and the shift minimises the largest spectral norm of eps - shift I,
      pinned by the PDE residual with the tensor and by eps = e I reducing to
      the scalar problem (tests/test_oracle_mbs.py).
Fields are (Nx, Ny, Nz, 3) complex128.  get_shift_and_scale minimises
max |eps - s|^2 directly (scipy Nelder-Mead from the reference's start
point [1.6, 0.25]); the solution is pinned by the residual of the PDE it
solves, (|K|^2 P_T - eps) E = -S (tests/test_oracle_mbs.py).
"""
from __future__ import annotations

import numpy as np
import scipy.constants as const
from scipy.optimize import minimize

ax3 = (0, 1, 2)


def k_grid(shape, spacing):
    axes = [np.fft.fftfreq(n) for n in shape]
    return 2 * np.pi * np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1) / spacing


def source_field(J, k0):
    return -1j / k0 * 1e-6 * const.c * const.mu_0 * np.asarray(J, np.complex128)


def add_absorbing_bc(eps, spacing, axis, thickness, max_extinction=0.25, n_background=1.0):
    axes = axis if isinstance(axis, tuple) else (axis,)
    w = int(thickness / spacing)
    widths = [(w, w) if i in axes else (0, 0) for i in range(3)]
    ext = np.pad(np.zeros(np.shape(eps)), widths, mode="linear_ramp", end_values=max_extinction)
    perm = np.pad(np.asarray(eps, np.complex128), widths, mode="constant",
                  constant_values=n_background**2)
    return perm + ((n_background + 1j * ext) ** 2 - n_background**2)


def get_shift_and_scale(eps):
    vals = np.unique(np.asarray(eps, np.complex128).ravel())

    def loss(p):
        return np.max(np.abs(vals - (p[0] + 1j * p[1]))) ** 2

    res = minimize(loss, np.array([1.6, 0.25]), method="Nelder-Mead",
                   options={"xatol": 1e-12, "fatol": 1e-14, "maxiter": 20000})
    shift = res.x[0] + 1j * res.x[1]
    return shift, 1 / (1.1j * np.sqrt(loss(res.x)))


def get_shift_and_scale_tensor(eps):
    """Tensor permittivity: shift minimising max ||eps - shift I||_2 (direct
    Nelder-Mead over the distinct matrices from the mean diagonal)."""
    E = np.unique(np.asarray(eps, np.complex128).reshape(9, -1).T, axis=0).reshape(-1, 3, 3)

    def loss(p):
        return np.max(np.linalg.norm(E - (p[0] + 1j * p[1]) * np.eye(3), ord=2, axis=(1, 2)))

    c0 = np.mean(np.trace(E, axis1=1, axis2=2)) / 3
    res = minimize(loss, np.array([c0.real, c0.imag]), method="Nelder-Mead",
                   options={"xatol": 1e-12, "fatol": 1e-14, "maxiter": 20000})
    return res.x[0] + 1j * res.x[1], 1 / (1.1j * loss(res.x))


def add_absorbing_bc_tensor(eps, spacing, axis, thickness, max_extinction=0.25, n_background=1.0):
    """add_absorbing_bc for (3, 3, ...) permittivities: background n^2 I,
    extinction ramp on the diagonal."""
    eps = np.asarray(eps, np.complex128)
    diag = [add_absorbing_bc(eps[a, a], spacing, axis, thickness, max_extinction, n_background)
            for a in range(3)]
    out = np.zeros((3, 3) + diag[0].shape, np.complex128)
    axes = axis if isinstance(axis, tuple) else (axis,)
    w = int(thickness / spacing)
    widths = [(w, w) if i in axes else (0, 0) for i in range(3)]
    for a in range(3):
        for b in range(3):
            out[a, b] = diag[a] if a == b else np.pad(eps[a, b], widths)
    return out


class MBS:
    def __init__(self, eps, spacing, k0, shift, scale):
        self.eps = np.asarray(eps, np.complex128)
        self.tensor = self.eps.ndim == 5
        shape = self.eps.shape[-3:]
        self.K = k_grid(shape, spacing) / k0
        self.K2 = np.sum(self.K**2, axis=-1, keepdims=True)
        self.scale, self.kb = scale, shift * scale + 1

    def eps_mul(self, E):
        """eps E per voxel (scalar or 3x3)."""
        if self.tensor:
            return np.einsum("abxyz,xyzb->xyza", self.eps, E)
        return self.eps[..., None] * E

    def V(self, x):
        return self.scale * self.eps_mul(x) - self.kb * x

    def G(self, y):
        yf = np.fft.fftn(y, axes=ax3)
        proj = np.sum(self.K * yf, axis=-1, keepdims=True)
        with np.errstate(invalid="ignore", divide="ignore"):
            yl = np.where(self.K2 != 0, self.K / self.K2 * proj, 0.0)
        yt = yf - yl
        return np.fft.ifftn(yt / (self.kb - self.scale * self.K2) + yl / self.kb, axes=ax3)

    def prec_y(self, S):
        p = -self.V(self.G(S * self.scale))
        num = np.max(np.abs(p))
        return p / num, num

    def prec_forward(self, x):
        return -self.V(self.G(self.V(x)) + x)

    def iterate(self, x, py, n):
        r = None
        for _ in range(n):
            r = py - self.prec_forward(x)
            x = x + r
        return x, r

    def adjoint(self) -> "MBS":
        """The operator set of the conjugated problem: eps -> eps^H per voxel,
        scale -> conj(scale), kb -> conj(kb).  Its G is the adjoint of this
        G (the Fourier multipliers conjugate, the projections are real
        symmetric) and its prec_forward = -V^H (G^H V^H + I) is the adjoint
        of -V (G V + I): the operator the implicit differentiation of solve
        (jaxopt FixedPointIteration, implicit_diff=True, solver.py:136-142)
        solves with."""
        o = MBS.__new__(MBS)
        o.eps = np.conj(np.swapaxes(self.eps, 0, 1)) if self.tensor else np.conj(self.eps)
        o.tensor, o.K, o.K2 = self.tensor, self.K, self.K2
        o.scale, o.kb = np.conj(self.scale), np.conj(self.kb)
        return o

    def vjp(self, S, E, Ebar, tol=1e-12, maxiter=20000):
        """Reverse-mode rule of E = solve(eps, S) (scalar eps) for the
        cotangent Ebar (dL/dRe E + i dL/dIm E).  With A = prec_forward and
        b = -V G(scale S) (E = A^-1 b; the num_scale normalisation cancels):
          A^H lam = Ebar  (the same iteration on the adjoint operators),
          mu = G^H(V^H lam),  w = G(V E) + E - G(scale S),
          eps_bar = conj(scale sum_c (conj(lam_c) w_c + conj(mu_c) E_c)),
          S_bar = -conj(scale) mu.
        Returns (eps_bar, S_bar, lam)."""
        if self.tensor:
            raise NotImplementedError("vjp: scalar permittivity")
        adj = self.adjoint()
        lam = np.asarray(Ebar, np.complex128).copy()
        nb = np.linalg.norm(Ebar)
        for _ in range(maxiter):
            r = Ebar - adj.prec_forward(lam)
            lam = lam + r
            if np.linalg.norm(r) <= tol * nb:
                break
        mu = adj.G(adj.V(lam))
        w = self.G(self.V(E)) + E - self.G(S * self.scale)
        eps_bar = np.conj(self.scale * np.sum(np.conj(lam) * w + np.conj(mu) * E, axis=-1))
        S_bar = -np.conj(self.scale) * mu
        return eps_bar, S_bar, lam

    def solve_exact(self, S, tol=1e-13, maxiter=20000):
        """E = A^-1 b by the fixed-point iteration run to `tol` (FP64)."""
        b = -self.V(self.G(S * self.scale))
        x = b.copy()
        nb = np.linalg.norm(b)
        for _ in range(maxiter):
            r = b - self.prec_forward(x)
            x = x + r
            if np.linalg.norm(r) <= tol * nb:
                break
        return x

    def pde_residual(self, E, S):
        """|(|K|^2 P_T - eps) E + S| / |S|, the equation the fixed point solves."""
        Ef = np.fft.fftn(E, axes=ax3)
        proj = np.sum(self.K * Ef, axis=-1, keepdims=True)
        with np.errstate(invalid="ignore", divide="ignore"):
            El = np.where(self.K2 != 0, self.K / self.K2 * proj, 0.0)
        curl = np.fft.ifftn(self.K2 * (Ef - El), axes=ax3)
        return np.linalg.norm(curl - self.eps_mul(E) + S) / np.linalg.norm(S)


def bicgstab(A, b, x0, maxiter, tol=1e-5, atol=0.0):
    """BiCGStab of jax.scipy.sparse.linalg.bicgstab in FP64, M = identity.
    Returns (x, k, |r|^2): k iterations run, -10 on rho breakdown, -11 on
    alpha / omega breakdown."""
    def vdot(a, c):
        return np.vdot(a.ravel(), c.ravel())

    b = np.asarray(b, np.complex128)
    x = np.asarray(x0, np.complex128).copy()
    atol2 = max(tol**2 * vdot(b, b).real, atol**2)
    r = b - A(x)
    rh, p, q = r.copy(), r.copy(), r.copy()
    rho = alpha = omega = 1.0 + 0j
    k = 0
    rs = vdot(r, r).real
    while rs > atol2 and k < maxiter and k >= 0:
        rho_n = vdot(rh, r)
        beta = rho_n / rho * alpha / omega
        p = r + beta * (p - omega * q)
        q = A(p)
        alpha_n = rho_n / vdot(rh, q)
        s = r - alpha_n * q
        early = vdot(s, s).real < atol2
        t = A(s)
        with np.errstate(invalid="ignore", divide="ignore"):
            omega_n = vdot(t, s) / vdot(t, t)
        if early:
            x = x + alpha_n * p
            r = s
        else:
            x = x + alpha_n * p + omega_n * s
            r = s - omega_n * t
        k = -11 if (omega_n == 0 or alpha_n == 0) else k + 1
        k = -10 if rho_n == 0 else k
        rho, alpha, omega = rho_n, alpha_n, omega_n
        rs = vdot(r, r).real
    return x, k, rs


def solve_bicgstab(eps, S, spacing, k0, shift, scale, maxiter=1000, tol=1e-3):
    """solve(..., use_bicgstab=True) (solver.py:101-146): E = x num_scale,
    x from BiCGStab on prec_forward(x) = prec_y from x0 = prec_y."""
    m = MBS(eps, spacing, k0, shift, scale)
    py, num = m.prec_y(np.asarray(S, np.complex128))
    x, k, rs = bicgstab(m.prec_forward, py, py, maxiter, tol)
    return x * num, k, rs

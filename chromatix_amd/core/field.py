"""Field containers mirroring src/chromatix/core/field.py.

A ``Field`` holds the complex field ``u`` as a complex64 torch tensor on a
HIP device, and its sampling metadata (``dx``, ``origin``, ``spectrum``) as
float64 numpy arrays on the host, so that parameter arithmetic never forces a
device synchronisation.  Shapes and axis conventions are the reference's:

* ``ScalarField``: ``u`` is ``(... H W)``, ``dx``/``origin`` are ``(2,)``
  (field.py:616-665);
* ``ChromaticScalarField``: ``u`` is ``(... H W wv)``, ``dx``/``origin`` are
  ``(wv, 2)`` (field.py:668-740).

Vector fields are not part of the FFT-propagation path and are not provided.
"""
from __future__ import annotations

from typing import Any

import numpy as np
import torch

from .spectrum import MonoSpectrum, Spectrum

__all__ = [
    "Field",
    "ScalarField",
    "ChromaticScalarField",
    "grid",
    "freq_grid",
    "pad",
    "crop",
    "shift_grid",
    "shift_field",
]


def grid(shape: tuple[int, int], spacing: np.ndarray) -> np.ndarray:
    """Sample coordinates ``(i - N/2) * dx`` (field.py:23-33); shape
    ``(... y x 2)`` for spacing ``(... 2)``."""
    ny, nx = shape
    gy, gx = np.meshgrid(np.arange(ny) - ny / 2, np.arange(nx) - nx / 2, indexing="ij")
    g = np.stack([gy, gx], axis=-1)
    sp = np.asarray(spacing, dtype=np.float64)
    return sp[..., None, None, :] * g


def freq_grid(shape: tuple[int, int], spacing: np.ndarray) -> np.ndarray:
    """Centred frequency grid ``fftshift(fftfreq(N)) / dx`` (field.py:36-46)."""
    ny, nx = shape
    fy, fx = np.meshgrid(
        np.fft.fftshift(np.fft.fftfreq(ny)), np.fft.fftshift(np.fft.fftfreq(nx)), indexing="ij"
    )
    g = np.stack([fy, fx], axis=-1)
    sp = np.asarray(spacing, dtype=np.float64)
    return (1.0 / sp)[..., None, None, :] * g


def _broadcast_dx(dx, num_wv: int) -> np.ndarray:
    """(wv, 2) spacing from a scalar, (wv,) or (wv, 2) input
    (utils/shapes.py::_broadcast_dx_to_grid)."""
    dx = np.atleast_1d(np.asarray(dx, dtype=np.float64))
    if dx.shape == (1,):
        out = np.repeat(dx[None, :], 2, axis=1).repeat(num_wv, axis=0)
    elif dx.ndim == 1:
        if dx.shape[0] != num_wv:
            if num_wv == 1 and dx.shape[0] == 2:
                out = dx[None, :]
            else:
                raise ValueError("Number of wavelengths does not match")
        else:
            out = np.repeat(dx[:, None], 2, axis=1)
    elif dx.ndim == 2 and dx.shape[1] == 2:
        if dx.shape[0] != num_wv:
            raise ValueError("Number of wavelengths does not match")
        out = dx
    else:
        raise ValueError(f"dx must be scalar or have shape (2,), (C,), or (C, 2); got {dx.shape}.")
    if np.any(out < 0):
        raise ValueError(f"dx must be larger than 0, got {out}.")
    return out.astype(np.float64)


def _to_complex_tensor(u: Any) -> torch.Tensor:
    if isinstance(u, torch.Tensor):
        return u.to(torch.complex64)
    return torch.as_tensor(np.asarray(u), dtype=torch.complex64)


class Field:
    """Common behaviour of the scalar field classes (field.py:49-531)."""

    u: torch.Tensor
    dx: np.ndarray
    origin: np.ndarray
    spectrum: Spectrum
    # axis indices, negative (set by subclasses)
    y: int
    x: int

    # ---- shape -----------------------------------------------------------
    @property
    def shape(self) -> tuple[int, ...]:
        return tuple(self.u.shape)

    @property
    def spatial_shape(self) -> tuple[int, int]:
        return (self.u.shape[self.y], self.u.shape[self.x])

    @property
    def spatial_dims(self) -> tuple[int, int]:
        return (self.y, self.x)

    @property
    def ndim(self) -> int:
        return self.u.ndim

    @property
    def batch_dims(self) -> tuple[int, ...]:
        return tuple(range(self.ndim + self.spatial_dims[0]))

    @property
    def batch_shape(self) -> tuple[int, ...]:
        return tuple(self.u.shape[: self.ndim + self.y])

    @property
    def num_batch_dims(self) -> int:
        return len(self.batch_dims)

    # ---- sampling ---------------------------------------------------------
    @property
    def extent(self) -> np.ndarray:
        return self.dx * np.asarray(self.spatial_shape, dtype=np.float64)

    @property
    def df(self) -> np.ndarray:
        return 1.0 / self.extent

    @property
    def k_grid(self) -> np.ndarray:
        return 2 * np.pi * self.f_grid

    @property
    def spatial_limits(self) -> tuple[tuple[float, float], tuple[float, float]]:
        """(y_min, y_max), (x_min, x_max) of ``grid`` (field.py:221-229).  The
        coordinates are monotone in the index, so the extremes are those of
        the first and last sample of each axis (and wavelength), computed
        with the same float64 operations as ``grid`` instead of the whole
        grid."""
        sp = np.asarray(self.dx, dtype=np.float64)
        org = np.broadcast_to(np.asarray(self.origin, dtype=np.float64), sp.shape)
        out = []
        for d, n in enumerate(self.spatial_shape):
            ends = np.array([0.0, float(n - 1)]) - n / 2
            v = sp[..., d:d + 1] * ends + org[..., d:d + 1]
            out.append((float(v.min()), float(v.max())))
        return out[0], out[1]

    @property
    def wavelength(self) -> np.ndarray:
        return self.spectrum.wavelength

    @property
    def central_wavelength(self) -> float:
        return self.spectrum.central_wavelength

    @property
    def k(self) -> np.ndarray:
        return 2 * np.pi / self.wavelength

    @property
    def surface_area(self) -> np.ndarray:
        return self.extent

    @property
    def phase(self) -> torch.Tensor:
        return torch.angle(self.u)

    @property
    def amplitude(self) -> torch.Tensor:
        return torch.abs(self.u)

    @property
    def conj(self) -> "Field":
        return self.replace(u=torch.conj(self.u).resolve_conj())

    # ---- construction -----------------------------------------------------
    def replace(self, **kwargs) -> "Field":
        new = object.__new__(type(self))
        new.__dict__.update(self.__dict__)
        for k, v in kwargs.items():
            if k == "u":
                v = _to_complex_tensor(v) if not isinstance(v, torch.Tensor) else v
            elif k in ("dx", "origin"):
                v = np.asarray(v, dtype=np.float64)
            setattr(new, k, v)
        return new

    @classmethod
    def build(cls, u, dx, spectrum) -> "Field":
        """Type dispatch of field.py:364-384 (scalar fields only)."""
        spectrum = Spectrum.build(spectrum)
        u = _to_complex_tensor(u)
        mono = isinstance(spectrum, MonoSpectrum)
        if u.ndim == 2 and mono:
            return ScalarField(u, dx, 0.0, spectrum)
        if u.ndim == 3 and not mono:
            return ChromaticScalarField(u, dx, 0.0, spectrum)
        raise ValueError(
            f"unsupported field: u.ndim={u.ndim} with {'mono' if mono else 'chromatic'} "
            "spectrum (vector fields are not on the propagation path)"
        )

    @classmethod
    def empty(cls, shape, dx, spectrum, scalar: bool = True, device=None) -> "Field":
        return cls.zeros(shape, dx, spectrum, scalar, device)

    @classmethod
    def zeros(cls, shape, dx, spectrum, scalar: bool = True, device=None) -> "Field":
        if not scalar:
            raise ValueError("vector fields are not on the propagation path")
        spectrum = Spectrum.build(spectrum)
        if isinstance(spectrum, MonoSpectrum):
            u = torch.zeros(tuple(shape), dtype=torch.complex64, device=device)
        else:
            u = torch.zeros((*shape, spectrum.size), dtype=torch.complex64, device=device)
        return Field.build(u, dx, spectrum)

    @classmethod
    def empty_like(cls, field: "Field", shape=None, dx=None, origin=None, spectrum=None) -> "Field":
        """field.py:316-362: same attributes, new (uninitialised) u."""
        if dx is None:
            dx = field.dx
        else:
            dx = np.asarray(dx, dtype=np.float64)
            if dx.ndim != field.dx.ndim:
                raise ValueError("New spacing must have same number of dimensions as old spacing")
        new_shape = list(field.shape)
        if shape is not None:
            if len(shape) != 2:
                raise ValueError("shape must be (height, width)")
            new_shape[field.y] = shape[0]
            new_shape[field.x] = shape[1]
        if spectrum is None:
            spectrum = field.spectrum
        u = torch.empty(tuple(new_shape), dtype=torch.complex64, device=field.u.device)
        return field.replace(u=u, dx=dx, origin=field.origin if origin is None else origin,
                             spectrum=spectrum)

    def expand_dims(self) -> "Field":
        return self.replace(u=self.u.unsqueeze(0))

    def unsqueeze(self) -> "Field":
        return self.expand_dims()

    def squeeze(self) -> "Field":
        u = self.u
        for d in reversed(self.batch_dims):
            if u.shape[d] == 1:
                u = u.squeeze(d)
        return self.replace(u=u)

    # ---- arithmetic (field.py:445-531) -----------------------------------
    def _other(self, other):
        if isinstance(other, Field):
            return other.u
        if isinstance(other, np.ndarray):
            return torch.as_tensor(other, device=self.u.device)
        return other

    def __add__(self, o): return self.replace(u=self.u + self._other(o))
    def __radd__(self, o): return self + o
    def __sub__(self, o): return self.replace(u=self.u - self._other(o))
    def __rsub__(self, o): return self.replace(u=self._other(o) - self.u)
    def __mul__(self, o): return self.replace(u=(self.u * self._other(o)).to(torch.complex64))
    def __rmul__(self, o): return self * o
    def __truediv__(self, o): return self.replace(u=(self.u / self._other(o)).to(torch.complex64))
    def __rtruediv__(self, o): return self.replace(u=(self._other(o) / self.u).to(torch.complex64))
    def __neg__(self): return self.replace(u=-self.u)

    # ---- derived quantities, computed by HIP kernels ------------------------
    @property
    def intensity(self) -> torch.Tensor:
        from ..functional.elementwise import intensity

        return intensity(self)

    @property
    def power(self) -> torch.Tensor:
        from ..functional.elementwise import power

        return power(self)

    def __repr__(self) -> str:
        return (f"{type(self).__name__}(shape={self.shape}, dx={self.dx.tolist()}, "
                f"spectrum={self.spectrum})")


class ScalarField(Field):
    """Monochromatic scalar field, ``u`` of shape ``(... H W)``."""

    y, x = -2, -1

    def __init__(self, u, dx, origin, spectrum):
        spectrum = Spectrum.build(spectrum)
        self.u = _to_complex_tensor(u)
        self.dx = _broadcast_dx(dx, 1)[0]
        self.origin = _broadcast_dx(origin, 1)[0] if np.ndim(origin) else np.full(2, float(origin))
        self.spectrum = spectrum
        if self.u.ndim < 2:
            raise ValueError(f"Expected at least 2-dimensional field, got shape {tuple(self.u.shape)}.")

    @property
    def grid(self) -> np.ndarray:
        return grid(self.spatial_shape, self.dx) + self.origin

    @property
    def f_grid(self) -> np.ndarray:
        return freq_grid(self.spatial_shape, self.dx)

    @property
    def central_dx(self) -> float:
        return float(self.dx[0])

    @property
    def central_rectangular_dx(self) -> np.ndarray:
        return self.dx

    @property
    def broadcasted_wavelength(self) -> np.ndarray:
        return self.spectrum.wavelength.reshape(())


class ChromaticScalarField(Field):
    """Scalar field over a spectrum, ``u`` of shape ``(... H W wv)``."""

    y, x = -3, -2
    wv = -1

    def __init__(self, u, dx, origin, spectrum):
        spectrum = Spectrum.build(spectrum)
        self.u = _to_complex_tensor(u)
        self.dx = _broadcast_dx(dx, spectrum.size)
        if np.ndim(origin) == 0:
            self.origin = np.full((spectrum.size, 2), float(origin))
        else:
            self.origin = _broadcast_dx(origin, spectrum.size)
        self.spectrum = spectrum
        if self.u.ndim < 3:
            raise ValueError(f"Expected at least 3-dimensional field, got shape {tuple(self.u.shape)}.")
        if self.u.shape[-1] != spectrum.size:
            raise ValueError("Expected last dimension of u to be same as wavelengths.")

    @property
    def grid(self) -> np.ndarray:
        g = grid(self.spatial_shape, self.dx)  # (wv, y, x, 2)
        return np.moveaxis(g, 0, 2) + self.origin

    @property
    def f_grid(self) -> np.ndarray:
        return np.moveaxis(freq_grid(self.spatial_shape, self.dx), 0, 2)

    @property
    def central_dx(self) -> float:
        return float(self.dx[0, 0])

    @property
    def central_rectangular_dx(self) -> np.ndarray:
        return self.dx[0]

    @property
    def broadcasted_wavelength(self) -> np.ndarray:
        return self.spectrum.wavelength


# ---- free functions (field.py:534-613) --------------------------------------
def _pair(v) -> tuple[int, int]:
    if isinstance(v, (int, np.integer)):
        return int(v), int(v)
    a, b = v
    return int(a), int(b)


def pad(field: Field, pad_width, cval: float = 0) -> Field:
    """Symmetric constant padding of the spatial axes (field.py:534-549)."""
    py, px = _pair(pad_width)
    u = field.u
    if isinstance(field, ChromaticScalarField):
        u = u.movedim(-1, -3)
    u = torch.nn.functional.pad(u, (px, px, py, py), mode="constant", value=0.0)
    if cval != 0:
        mask = torch.ones(u.shape[-2:], dtype=torch.bool, device=u.device)
        mask[py: u.shape[-2] - py, px: u.shape[-1] - px] = False
        u = torch.where(mask, torch.as_tensor(complex(cval), dtype=u.dtype, device=u.device), u)
    if isinstance(field, ChromaticScalarField):
        u = u.movedim(-3, -1)
    return field.replace(u=u.resolve_conj().contiguous())


def crop(field: Field, crop_width) -> Field:
    """Symmetric crop of the spatial axes (field.py:552-570)."""
    cy, cx = _pair(crop_width)
    sl = [slice(None)] * field.ndim
    ny, nx = field.spatial_shape
    sl[field.ndim + field.y] = slice(cy, ny - cy)
    sl[field.ndim + field.x] = slice(cx, nx - cx)
    return field.replace(u=field.u[tuple(sl)].resolve_conj().contiguous())


def shift_grid(field: Field, shift_yx) -> Field:
    """Set the grid origin to ``shift_yx`` (field.py:573-585)."""
    s = np.atleast_1d(np.asarray(shift_yx, dtype=np.float64))
    if s.size < 2:
        s = np.array([s.item(), s.item()])
    if s.size != 2:
        raise ValueError("Shift must be an array of size (2,) in the order (y x)")
    return field.replace(origin=np.zeros_like(field.origin) + s)


def shift_field(field: Field, shift_by) -> Field:
    """Integer-pixel shift with zero fill (field.py:588-613)."""
    sy, sx = _pair(shift_by)
    u = field.u
    dims = (field.ndim + field.y, field.ndim + field.x)
    for d, s in zip(dims, (sy, sx)):
        n = u.shape[d]
        if s == 0:
            continue
        if s > 0:
            kept = u.narrow(d, s, n - s)
            u = torch.cat([kept, torch.zeros_like(u.narrow(d, 0, s))], dim=d)
        else:
            kept = u.narrow(d, 0, n + s)
            u = torch.cat([torch.zeros_like(u.narrow(d, 0, -s)), kept], dim=d)
    return field.replace(u=u)


"""Parameter structs and optics helpers of the forward-solver API.

Mirrors the reference's chained get_default_parameters/update_struct config
system (Codes/sub_code/BASIC_OPTICAL_PARAMETER.m:1-14, Codes/FORWARD_SOLVER.m:9-21,
Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:32-45, Codes/sub_code/update_struct.m:1-15).
"""
from __future__ import annotations

import copy
import os
import warnings

import numpy as np
import scipy.fft as sfft


def update_struct(base, update, warn=True):
    """Merge ``update`` into a copy of ``base``; unknown keys warn and are still set
    (Codes/sub_code/update_struct.m:1-15)."""
    out = copy.deepcopy(base)
    for k, v in (update or {}).items():
        if k not in base and warn:
            warnings.warn(f'Unexpected field : "{k}" was set ?! Verify the field name of your parameter structures.')
        out[k] = copy.deepcopy(v)
    return out


def basic_optical_parameter(init=None):
    """Codes/sub_code/BASIC_OPTICAL_PARAMETER.m:1-14."""
    p = {'size': [100, 100, 100], 'wavelength': 0.532, 'NA': 1.2, 'RI_bg': 1.336,
         'resolution': [0.1, 0.1, 0.1], 'vector_simulation': True, 'use_abbe_sine': True}
    return update_struct(p, init) if init is not None else p


def round_half_away(x):
    """MATLAB round()."""
    x = np.asarray(x, dtype=np.float64)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def ri_to_potential(RI, wavelength, n_m):
    """Vt = k^2 (RI^2 / n_m^2 - 1), k = 2 pi n_m / lambda (Codes/sub_code/RI2potential.m:1-4)."""
    k = 2.0 * np.pi * n_m / wavelength
    return k * k * (np.asarray(RI) ** 2 / (n_m * n_m) - 1.0)


def fourier_grid(size, resolution):
    """Centred spatial-frequency coordinates (cycles/um) per axis
    (Codes/sub_code/DERIVE_OPTICAL_TOOL.m:34-47)."""
    return [(np.arange(n) - n // 2) / (float(r) * n) for n, r in zip(size, resolution)]


def pupil(size_xy, resolution, wavelength, NA, n_m):
    """NA circle and cos(theta) on a centred 2D frequency grid
    (Codes/sub_code/DERIVE_OPTICAL_TOOL.m:51-65)."""
    kx, ky = fourier_grid(size_xy, resolution[:2])
    kxy = np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2)
    k0_nm = n_m / wavelength
    na_circle = kxy < NA / wavelength
    k3 = np.sqrt(np.maximum(k0_nm ** 2 - kxy ** 2, 0.0))
    return na_circle, k3 / k0_nm, k3


_FFT_WORKERS = max(1, min(8, os.cpu_count() or 1))


def _centered(a, fn):
    """fftshift(fn(ifftshift(a))) over the first two axes (result types as numpy.fft:
    single precision stays single), on several threads when the array holds several 2D
    transforms."""
    a = np.asarray(a)
    ax = (0, 1)
    workers = _FFT_WORKERS if a.size >= 4 * a.shape[0] * a.shape[1] else 1
    return np.fft.fftshift(fn(np.fft.ifftshift(a, axes=ax), axes=ax, workers=workers), axes=ax)


def centered_fft2(a):
    """fftshift(fft2(ifftshift(a))) over the first two axes."""
    return _centered(a, sfft.fft2)


def centered_ifft2(a):
    """fftshift(ifft2(ifftshift(a))) over the first two axes."""
    return _centered(a, sfft.ifft2)


def polarisation_basis(size_xy, resolution, wavelength, n_m):
    """Per transverse frequency of a centred 2D grid: the radial and perpendicular
    unit vectors in the pupil, the unit tangent of the Ewald sphere in the plane of
    incidence, and the mask of propagating frequencies (Codes/FORWARD_SOLVER.m:131-180).

    With K = k_xy / k_m, rho = |K|, K3 = sqrt(1 - rho^2) and phi the azimuth:
    radial = (cos phi, sin phi), perpendicular = (sin phi, -cos phi) and tangent =
    (K3 cos phi, K3 sin phi, -rho); at rho = 0 the azimuth is taken as 0.  The
    reference builds its K_2 axis with size(1) samples (:141), so only square field
    grids work there; other grids are refused here."""
    n0, n1 = int(size_xy[0]), int(size_xy[1])
    if n0 != n1:
        raise ValueError('vector fields need a square field grid (FORWARD_SOLVER.m:141 builds K_2 with size(1) '
                         'samples)')
    kx, ky = fourier_grid((n0, n1), resolution[:2])
    k_m = n_m / wavelength
    K1 = kx[:, None] / k_m
    K2 = ky[None, :] / k_m
    rho = np.sqrt(K1 ** 2 + K2 ** 2)
    K3 = np.sqrt(np.maximum(1.0 - rho ** 2, 0.0))
    inside = K3 > 0
    safe = np.where(rho > 0, rho, 1.0)
    cphi = np.where(rho > 0, K1 / safe, 1.0)
    sphi = np.where(rho > 0, K2 / safe, 0.0)
    radial = np.stack([cphi, sphi], axis=2)
    perp = np.stack([sphi, -cphi], axis=2)
    tangent = np.stack([K3 * cphi, K3 * sphi, -rho], axis=2) * inside[:, :, None]
    return radial, perp, tangent, inside

"""Optics grids, potential conversion and centred-FFT helpers (FP64 restatement).

Test infrastructure only (see oracle/__init__.py).

Array convention: MATLAB index order (x, y, z, ...) maps to NumPy axes 0, 1, 2, ...
"""
from __future__ import annotations

import copy
import warnings

import numpy as np
import scipy.fft as sfft


def basic_optical_parameter(init=None):
    """Codes/sub_code/BASIC_OPTICAL_PARAMETER.m:1-14."""
    p = dict(size=[100, 100, 100], wavelength=0.532, NA=1.2, RI_bg=1.336,
             resolution=[0.1, 0.1, 0.1], vector_simulation=True, use_abbe_sine=True)
    if init is not None:
        p = update_struct(p, init)
    return p


def update_struct(base, update, warn=True):
    """Codes/sub_code/update_struct.m:1-15 (unknown keys warn and are still set).

    ``warn=False`` mirrors the reference's warning('off','all') around callers
    such as DERIVE_OPTICAL_TOOL (FORWARD_SOLVER.m:25-27)."""
    out = copy.deepcopy(base)
    for k, v in update.items():
        if k not in base and warn:
            warnings.warn(f'Unexpected field : "{k}" was set')
        out[k] = copy.deepcopy(v)
    return out


def centered_coor(n):
    """(1:n) - (floor(n/2)+1) as 0-based array: DERIVE_OPTICAL_TOOL.m:22-24,42-44."""
    return np.arange(n, dtype=np.float64) - (n // 2)


def derive_optical_tool(params):
    """Codes/sub_code/DERIVE_OPTICAL_TOOL.m:1-81 (FP64; MATLAB stores these as single).

    Returns a dict with image/fourier coordinates broadcastable over (x, y, z).
    """
    p = update_struct(basic_optical_parameter(), params, warn=False)
    size = [int(s) for s in p['size']]
    res = [float(r) for r in p['resolution']]
    u = {}
    img = [centered_coor(size[j]) * res[j] for j in range(3)]          # :22-27
    fres = [1.0 / (res[j] * size[j]) for j in range(3)]               # :34-36
    four = [centered_coor(size[j]) * fres[j] for j in range(3)]        # :42-47
    shp = [(-1, 1, 1), (1, -1, 1), (1, 1, -1)]
    u['image_coor'] = [img[j].reshape(shp[j]) for j in range(3)]
    u['fourier_coor'] = [four[j].reshape(shp[j]) for j in range(3)]
    u['fourier_res'] = fres
    u['res'] = res
    u['size'] = size
    u['coorxy'] = np.sqrt(u['fourier_coor'][0] ** 2 + u['fourier_coor'][1] ** 2)  # :51-53
    u['lambda'] = p['wavelength']
    u['k0'] = 1.0 / p['wavelength']                                   # :56
    u['k0_nm'] = u['k0'] * p['RI_bg']                                 # :57
    u['nm'] = p['RI_bg']
    u['kmax'] = p['NA'] / p['wavelength']                             # :59
    u['NA_circle'] = u['coorxy'] < u['kmax']                          # :60
    k3 = u['k0_nm'] ** 2 - u['coorxy'] ** 2                           # :61
    k3[k3 < 0] = 0
    u['k3'] = np.sqrt(k3)
    u['dV'] = res[0] * res[1] * res[2]                                # :62
    u['refocusing_kernel'] = 1j * 2 * np.pi * u['k3']                  # :64
    u['cos_theta'] = u['k3'] / u['k0_nm']                             # :65
    return u


def ri2potential(RI, wavelength, base_RI):
    """Codes/sub_code/RI2potential.m:1-4: k^2 (RI^2/n_m^2 - 1), k = 2 pi n_m / lambda."""
    k = 2 * np.pi * base_RI / wavelength
    return k ** 2 * (np.asarray(RI) ** 2 / base_RI ** 2 - 1)


def potential2ri(pot, wavelength, base_RI):
    """Codes/sub_code/potential2RI.m:1-4."""
    k = 2 * np.pi * base_RI / wavelength
    return base_RI * np.sqrt(1 + np.asarray(pot) / k ** 2)


def fft_flip(a, flip_bool, use_shift):
    """Codes/sub_code/fft_flip.m:1-11: flip each flagged axis, then circshift by 1
    when the axis length is even or use_shift is false."""
    for ax, f in enumerate(flip_bool):
        if f:
            a = np.flip(a, axis=ax)
            if a.shape[ax] % 2 == 0 or not use_shift:
                a = np.roll(a, 1, axis=ax)
    return a


# MATLAB fftshift/ifftshift with no dim argument act on every axis; on axes that
# are not transformed the two shifts cancel, so only the transformed axes matter.
def cfft2(a, workers=None):
    """fftshift(fft2(ifftshift(a))) over axes (0, 1)."""
    ax = (0, 1)
    return sfft.fftshift(sfft.fft2(sfft.ifftshift(a, axes=ax), axes=ax, workers=workers), axes=ax)


def cifft2(a, workers=None):
    """fftshift(ifft2(ifftshift(a))) over axes (0, 1)."""
    ax = (0, 1)
    return sfft.fftshift(sfft.ifft2(sfft.ifftshift(a, axes=ax), axes=ax, workers=workers), axes=ax)


def cifftn3(a, workers=None):
    """fftshift(ifftn(ifftshift(a))) over axes (0, 1, 2)."""
    ax = (0, 1, 2)
    return sfft.fftshift(sfft.ifftn(sfft.ifftshift(a, axes=ax), axes=ax, workers=workers), axes=ax)


def cfft1(a, axis, inverse=False, workers=None):
    """fftshift(fft(ifftshift(a), [], axis)) (or ifft) along one axis."""
    f = sfft.ifft if inverse else sfft.fft
    return sfft.fftshift(f(sfft.ifftshift(a, axes=axis), axis=axis, workers=workers), axes=axis)


def polarisation_utility(ZP, n_m, lam, dx):
    """Codes/FORWARD_SOLVER.m:131-180 (FP64).  Returns Radial_2D (x, y, 2), Perp_2D
    (x, y, 2), ewald_TanVec (x, y, 3) and K_mask (x, y).

    The reference builds K_2 over -floor(ZP(2)/2):ZP(1)-floor(ZP(2)/2)-1 (:141), i.e.
    with ZP(1) samples, so a non-square field grid fails inside MATLAB; the same
    grids are refused here."""
    ZP = [int(ZP[0]), int(ZP[1])]
    if ZP[0] != ZP[1]:
        raise ValueError('polarisation_utility: the reference supports square field grids only (FORWARD_SOLVER.m:141)')
    k0 = 1.0 / lam                                                          # :135
    k = 2 * np.pi * n_m * k0                                                # :136
    kres = [1.0 / dx[0] / ZP[0], 1.0 / dx[1] / ZP[1]]                       # :138
    K1 = (2 * np.pi * kres[0] / k * (np.arange(ZP[0]) - ZP[0] // 2)).reshape(-1, 1)   # :140,143
    K2 = (2 * np.pi * kres[1] / k * (np.arange(ZP[0]) - ZP[1] // 2)).reshape(1, -1)   # :141,144
    K3 = np.real(np.sqrt((1 - (K1 ** 2 + K2 ** 2)).astype(np.complex128)))  # :146
    K_mask = ~(K3 == 0)                                                     # :147
    norm_rad = np.sqrt(K1 ** 2 + K2 ** 2)                                   # :151
    with np.errstate(divide='ignore', invalid='ignore'):
        t1 = K1 / norm_rad
        t2 = K2 / norm_rad
    t1 = np.where(norm_rad == 0, 1.0, t1)                                   # :155-156
    t2 = np.where(norm_rad == 0, 0.0, t2)
    Radial_2D = np.stack([t1, t2], axis=2)                                  # :158-159
    Perp_2D = np.stack([Radial_2D[:, :, 1], -Radial_2D[:, :, 0]], axis=2)   # :163-164
    norm_sph = np.sqrt(K1 ** 2 + K2 ** 2 + K3 ** 2)                         # :167
    Radial_3D = np.stack([np.broadcast_to(K1 / norm_sph, K3.shape), np.broadcast_to(K2 / norm_sph, K3.shape),
                          K3 / norm_sph], axis=2)                           # :168-170
    proj = np.sum(Radial_3D[:, :, :2] * Radial_2D, axis=2)                  # :172
    T = np.zeros(K3.shape + (3,))
    T[:, :, :2] = Radial_2D                                                 # :175
    T = T - proj[:, :, None] * Radial_3D                                    # :176
    nrm = np.sqrt(np.sum(T ** 2, axis=2))                                   # :177
    nrm[~K_mask] = 1                                                        # :178
    T = T / nrm[:, :, None]                                                 # :179
    return Radial_2D, Perp_2D, T, K_mask

"""Refocusing and transmission/reflection kernels of the 2D input/output
projections (host set-up, FP64).

Computed once per solver as FORWARD_SOLVER_CONVERGENT_BORN's constructor does
(Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:62-118):
  * cyclic xy boundary: the analytic angular-spectrum forms (:65-72);
  * otherwise the axially truncated free-space Green's function of
    Codes/sub_code/truncated_green_plus.m:1-110, built with the quasi-discrete
    Hankel transform of Codes/sub_code/hankel_transform/discrete_hankel_croped_fourier.m
    on the zeros of J0 (Codes/sub_code/hankel_transform/JnRoots.m).
The arrays handed to the GPU are (conv_x, conv_y, RI_z), complex64, column
major, centred (kx, ky) order with z in real space.
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sfft
import scipy.special as sp

from .params import fourier_grid, pupil, update_struct, basic_optical_parameter


def bessel_j0_zeros(count):
    """Positive zeros of J0 by Newton's method from McMahon's first-order
    estimate (k - 1/4) pi, as JnRoots.m iterates (clipped steps)."""
    r = np.pi * (np.arange(1, count + 1, dtype=np.float64) - 0.25)
    for _ in range(60):
        step = np.clip(sp.j0(r) / (-sp.j1(r)), -0.05, 0.05)
        r -= step
        if np.max(np.abs(step)) < 1e-15 * r[-1]:
            break
    return r


class HankelQD:
    """Quasi-discrete Hankel transform of order 0 on [0, R] with `samples`
    points, output cropped to the first `pos` frequencies
    (discrete_hankel_croped_fourier.m:19-64)."""

    def __init__(self, R, samples, cropping):
        roots = bessel_j0_zeros(samples + 1)
        C = roots[-1]
        c = roots[:-1]
        self.r = R / C * c
        k = c / R
        self.pos = int(np.argmin(np.abs(k - k.max() * cropping))) + 1
        J1 = np.abs(sp.j1(c))
        self.K = 2 * np.pi * R / C * J1
        self.Rv = J1 / R
        w = np.sqrt(2 / C) / J1
        P = self.pos
        # I[i, j] = w_i w_j J0(c_i c_j / C) for i < pos; the backward matrix is its transpose
        self.I = (w[:P, None] * w[None, :]) * sp.j0(np.outer(c[:P] / C, c))
        self.k = k[:P]

    def forward(self, vals):
        return (self.I @ (vals / self.Rv[:, None])) * self.K[:self.pos, None]

    def backward(self, vals):
        return (self.I.T @ (vals / self.K[:self.pos, None])) * self.Rv[:, None]


def _bracket(value, grid):
    """Last index with grid[idx] < value (0 when none) and its upper neighbour,
    with the blend weight (value - g0) / (g1 - g0); the kernel mixes
    green[idx] * w + green[idx + 1] * (1 - w) as truncated_green_plus.m:89 does."""
    L = grid.size
    n1 = np.clip(np.searchsorted(grid, value, side='left') - 1, 0, L - 1)
    n2 = n1.copy()
    n2[grid[n1] < value] += 1
    n2 = np.minimum(n2, L - 1)
    with np.errstate(divide='ignore', invalid='ignore'):
        w = (value - grid[n1]) / (grid[n2] - grid[n1])
    w[n2 == n1] = 1.0
    return n1, n2, w


def truncated_green(params, refocus=False, chunk=16):
    """Axially truncated Green's function (or refocusing kernel) in centred
    (kx, ky, kz) order on the grid params['size'] (truncated_green_plus.m:1-110)."""
    req = basic_optical_parameter()
    req.update(oversize_z=10, oversample_xy=20000, oversample_xy_fourier=2)
    p = update_struct(req, params, warn=False)
    out = np.array(p['size'], dtype=np.float64)
    res = np.array(p['resolution'], dtype=np.float64)
    creation = out * np.array([np.sqrt(2) * 1.1, np.sqrt(2) * 1.1, 1.0]) + p['oversize_z']
    max_size = max(creation[0] * res[0], creation[1] * res[1])
    creation[0:2] = np.ceil(max_size / res[0:2])
    cs = [int(round(v)) for v in creation]
    out = out.astype(int)
    fxg, fyg, fzg = fourier_grid(cs, res)
    coorxy = np.sqrt(fxg[:, None] ** 2 + fyg[None, :] ** 2)
    k0 = p['RI_bg'] / p['wavelength']
    kmax = p['NA'] / p['wavelength']
    samples = p['oversample_xy'] * p['oversample_xy_fourier']
    diag = np.hypot(cs[0], cs[1])
    if samples < diag * 5:
        raise ValueError('grid too large for the Hankel sampling (oversample_xy)')
    hk = HankelQD(1.0, samples, diag * p['oversample_xy_fourier'] / (2 * samples))
    H = hk.r / hk.r.max() * coorxy.max() * p['oversample_xy_fourier']
    Hk3 = np.sqrt(np.maximum(k0 * k0 - H * H, 0.0))
    Hc = H < kmax
    nz = cs[2]
    fac = (nz - p['oversize_z'] / 2) / nz
    dkz = 1.0 / (res[2] * nz)
    g = np.sinc((Hk3[:, None] - fzg[None, :]) / dkz * fac) * fac
    if refocus:
        g = g * Hc[:, None]
    else:
        g = g * Hc[:, None] / (Hk3[:, None] + (~Hc)[:, None]) / (1j * 4 * np.pi)
    img = hk.forward(g.astype(np.complex128))
    nk = hk.k / hk.k.max()
    mask = (nk < 0.9) + (nk < 1) * (nk > 0.9) * (1 - nk) * 10
    g = hk.backward(img * mask[:, None])
    n1, n2, w = _bracket(coorxy, H)
    scale = nz / (res[0] * res[1])
    sx, sy = cs[0] // 2 - out[0] // 2, cs[1] // 2 - out[1] // 2
    res3 = np.empty((out[0], out[1], nz), dtype=np.complex128)
    ax = (0, 1)
    for z0 in range(0, nz, chunk):
        zz = slice(z0, min(nz, z0 + chunk))
        gs = (g[n1, zz] * w[..., None] + g[n2, zz] * (1 - w[..., None])) * scale
        gs = sfft.fftshift(sfft.ifft2(sfft.ifftshift(gs, axes=ax), axes=ax), axes=ax)
        gs = gs[sx:sx + out[0], sy:sy + out[1]]
        res3[:, :, zz] = sfft.fftshift(sfft.fft2(sfft.ifftshift(gs, axes=ax), axes=ax), axes=ax)
    res3 = sfft.fftshift(sfft.ifft(sfft.ifftshift(res3, axes=2), axis=2), axes=2)
    zs = nz // 2 - out[2] // 2
    res3 = res3[:, :, zs:zs + out[2]]
    return sfft.fftshift(sfft.fft(sfft.ifftshift(res3, axes=2), axis=2), axes=2)


def _c_ifftn3(a):
    ax = (0, 1, 2)
    return sfft.fftshift(sfft.ifftn(sfft.ifftshift(a, axes=ax), axes=ax), axes=ax)


def _c_fft2(a):
    ax = (0, 1)
    return sfft.fftshift(sfft.fft2(sfft.ifftshift(a, axes=ax), axes=ax), axes=ax)


def _c_ifft2(a):
    ax = (0, 1)
    return sfft.fftshift(sfft.ifft2(sfft.ifftshift(a, axes=ax), axes=ax), axes=ax)


def build_io_kernels(params, expected_ri_size, cyclic_xy):
    """(refocus, kernel_trans, kernel_ref) as complex128 (conv_x, conv_y, RI_z) arrays
    (Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:62-118)."""
    p = params
    res = np.array(p['resolution'], float)
    if cyclic_xy:
        na, cos_t, k3 = pupil(p['size'][:2], res, p['wavelength'], p['NA'], p['RI_bg'])
        zc = (np.arange(p['size'][2]) - p['size'][2] // 2) * res[2]
        ref = np.exp(1j * 2 * np.pi * k3[:, :, None] * zc[None, None, :]) * na[:, :, None]       # :66-68
        fsg = ref / (1j * 4 * np.pi) * (na / (k3 + (~na)))[:, :, None] / (res[0] * res[1])       # :69-71
        fsg = _c_ifft2(fsg)                                                                       # :72
    else:
        rc = p['RI_center']
        ptg = dict(p)
        ptg['size'] = [p['size'][0] + expected_ri_size[0] + rc[0], p['size'][1] + expected_ri_size[1] + rc[1],
                       p['size'][2]]

        def crop_shift(a):                                                                        # :82-85
            a = a[-min(0, rc[0]):a.shape[0] - max(0, rc[0]), -min(0, rc[1]):a.shape[1] - max(0, rc[1])]
            return np.roll(a, (-rc[0], -rc[1]), axis=(0, 1))
        ref = _c_fft2(_c_ifftn3(crop_shift(truncated_green(ptg, refocus=True)))) * (res[0] * res[1])  # :80-92
        fsg = _c_ifftn3(crop_shift(truncated_green(ptg, refocus=False)))                          # :97-107
    ktrans = _c_fft2(np.conj(fsg))                                                                # :114
    kref = _c_fft2(fsg)                                                                           # :115
    return ref, ktrans, kref

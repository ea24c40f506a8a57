"""Bessel zeros, quasi-discrete Hankel transform and the axially truncated Green's
kernels (FP64 restatement).  Test infrastructure only (see oracle/__init__.py).

MATLAB computes all of this in single precision; the oracle computes it in FP64,
so parity with the reference is pinned to the formula, not to MATLAB's bits.
"""
from __future__ import annotations

import numpy as np
import scipy.special as sp

from .optics import derive_optical_tool, basic_optical_parameter, update_struct, cfft2, cifft2, cfft1


def jn_roots(n, N, max_iter=100):
    """Codes/sub_code/hankel_transform/JnRoots.m:10-36.

    First N positive roots of J_n (for n > 0 the root at 0 is prepended and the
    list cut to N entries, as the reference does).  The reference uses a
    Gauss-Newton step with a finite-difference derivative; here the analytic
    derivative J_n' = (J_{n-1} - J_{n+1})/2 is used, which converges to the same
    roots (to FP64 accuracy) from the same McMahon-type starting guess.
    """
    r = n * np.pi / 2 - np.pi / 4 + np.pi * np.arange(1, N + 1, dtype=np.float64)   # :12
    for _ in range(max_iter):                                                       # :15-21
        f = sp.jv(n, r)
        d = 0.5 * (sp.jv(n - 1, r) - sp.jv(n + 1, r))
        step = f / d
        step = np.clip(np.nan_to_num(step), -0.05, 0.05)                            # :34-35
        r = r - step
        if np.max(np.abs(step)) < 1e-15 * np.max(r):
            break
    if n:                                                                           # :22-25
        r = np.concatenate([[0.0], r[r > 0]])[:N]
    return r


class DiscreteHankelCropedFourier:
    """Codes/sub_code/hankel_transform/discrete_hankel_croped_fourier.m:1-66."""

    def __init__(self, max_r, samples, order, cropping):
        n = order
        R = float(max_r)
        roots = jn_roots(n, samples + 1)                                   # :21
        C = roots[-1]                                                      # :27
        c = roots[:-1]                                                     # :28
        self.r = R / C * c                                                 # :29
        k = c / R                                                          # :30
        self.pos = int(np.argmin(np.abs(k - k.max() * cropping))) + 1     # :32 (count, 1-based)
        I = np.abs(sp.jv(1 + n, c))                                        # :35
        self.K = 2 * np.pi * R / C * I                                     # :36
        self.Rv = I / R                                                    # :37
        Iw = np.sqrt(2 / C) / I                                            # :38
        bess = sp.jv(n, np.outer(c[:self.pos] / C, c))                     # :39 (pos x samples)
        self.I2 = Iw[:, None] * Iw[None, :self.pos] * bess.T               # :40-41 (samples x pos)
        self.I = Iw[:self.pos, None] * Iw[None, :] * bess                  # :42 (pos x samples)
        self.k = k[:self.pos]                                              # :44

    def forward(self, vals):
        """:52-57  I * (vals ./ R) .* K(1:pos)."""
        return (self.I @ (vals / self.Rv[:, None])) * self.K[:self.pos, None]

    def backward(self, vals):
        """:58-63  I2 * (vals ./ K(1:pos)) .* R."""
        return (self.I2 @ (vals / self.K[:self.pos, None])) * self.Rv[:, None]


def bin_search_nearest(value, searched_pos):
    """Codes/sub_code/truncated_green_plus.m:111-129, restated literally (1-based
    indices internally, returned 0-based).  Note that the reference blends
    green(nearest)*blend + green(nearest2)*(1-blend), i.e. the interpolation
    weights are swapped relative to a textbook lerp; truncated_green_plus below
    keeps that as the reference computes it."""
    sp_ = np.asarray(searched_pos).ravel()
    L = sp_.size
    nc = np.ones(value.shape, dtype=np.int64)
    step_up = L
    while step_up > 1:
        step_up = -(-step_up // 2)                 # ceil(step_up/2)
        upper = nc + step_up
        upper[upper > L] = L
        change = sp_[upper - 1] < value
        nc = np.where(change, upper, nc)
    nc2 = nc.copy()
    below = sp_[nc - 1] < value
    nc2[below] += 1
    nc2[nc2 > L] = L
    with np.errstate(divide='ignore', invalid='ignore'):
        blend = (value - sp_[nc - 1]) / (sp_[nc2 - 1] - sp_[nc - 1])
    blend[nc2 == nc] = 1
    return nc - 1, nc2 - 1, blend


def truncated_green_plus(params, refocus_version=False):
    """Codes/sub_code/truncated_green_plus.m:1-110 (FP64).

    Returns the (size_x, size_y, size_z) kernel in centred (kx, ky, kz) order.
    """
    req = basic_optical_parameter()
    req.update(use_GPU=True, oversize_z=10, oversample_xy=20000,
               oversample_xy_fourier=2, simultanous_process=30)             # :8-13
    p = update_struct(req, params, warn=False)                              # :16 (warnings off, :79)
    out_px = np.array(p['size'], dtype=np.float64)                          # :19
    res = np.array(p['resolution'], dtype=np.float64)
    creation = out_px * np.array([np.sqrt(2) * 1.1, np.sqrt(2) * 1.1, 1.0]) + p['oversize_z']  # :21
    creation_size = creation * res                                          # :22
    max_size = max(creation_size[0], creation_size[1])                      # :23
    creation[0:2] = np.ceil(max_size / res[0:2])                            # :24
    p['size'] = [int(round(v)) for v in creation]                           # :25
    u = derive_optical_tool(p)                                              # :28
    out_px = out_px.astype(int)

    rng = 1                                                                 # :32
    samples = p['oversample_xy'] * p['oversample_xy_fourier']               # :33
    diag = np.sqrt(p['size'][0] ** 2 + p['size'][1] ** 2)
    if samples < diag * 5:                                                  # :34-36
        raise ValueError('the code might not be valid for such big array')
    hank = DiscreteHankelCropedFourier(rng, samples, 0, diag * p['oversample_xy_fourier'] / (2 * samples))  # :38

    H = hank.r / hank.r.max()                                               # :41-42
    H = H * u['coorxy'].max() * p['oversample_xy_fourier']                  # :43
    Hk3 = u['k0_nm'] ** 2 - H ** 2                                          # :44
    Hk3[Hk3 < 0] = 0
    Hk3 = np.sqrt(Hk3)
    Hcircle = H < u['kmax']                                                 # :45

    nz = u['size'][2]
    fac = (nz - p['oversize_z'] / 2) / nz
    coor3 = u['fourier_coor'][2].ravel()
    green = np.sinc((Hk3[:, None] - coor3[None, :]) / u['fourier_res'][2] * fac) * fac   # :47-49
    if not refocus_version:                                                  # :50-55
        green = green * Hcircle[:, None] / (Hk3[:, None] + (~Hcircle)[:, None])
        green = green / (1j * 4 * np.pi)
    else:
        green = green * Hcircle[:, None]
    green = green.astype(np.complex128)

    img = hank.forward(green)                                               # :58
    nk = hank.k / hank.k.max()                                              # :59
    mask = (nk < 0.9) + (nk < 1) * (nk > 0.9) * (1 - nk) * 10               # :62
    img = img * mask[:, None]                                               # :65
    green = hank.backward(img)                                              # :66

    coorxy = u['coorxy'][:, :, 0]
    n1, n2, blend = bin_search_nearest(coorxy, H)                           # :70
    out = np.zeros((out_px[0], out_px[1], nz), dtype=np.complex128)         # :77
    sx, sy = coorxy.shape
    st = [sx // 2 - out_px[0] // 2, sy // 2 - out_px[1] // 2]               # :94 (0-based)
    scale = nz / (u['res'][0] * u['res'][1])                                # :90
    for z in range(nz):                                                     # :82-102 (slab loop, per slice)
        gs = green[n1, z] * blend + green[n2, z] * (1 - blend)              # :89
        gs = gs * scale
        gs = cifft2(gs)                                                     # :93
        gs = gs[st[0]:st[0] + out_px[0], st[1]:st[1] + out_px[1]]           # :96-99
        out[:, :, z] = cfft2(gs)                                            # :100
    out = cfft1(out, axis=2, inverse=True)                                  # :103
    zs = nz // 2 - out_px[2] // 2                                           # :105-107 (0-based)
    out = out[:, :, zs:zs + out_px[2]]
    out = cfft1(out, axis=2)                                                # :109
    return out

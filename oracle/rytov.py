"""Rytov reconstruction and phase unwrapping (FP64/FP32 restatement).

Test infrastructure only (see oracle/__init__.py).

Follows Codes/BACKWARD_SOLVER_RYTOV.m:22-127, Codes/sub_code/vector2scalarfield.m:1-14 and
the residue / raster-unwrap kernels of Codes/sub_code/goldstein_GPU
(get_residue_cuda.cu:35-66, unwrapp_phase_goldstein_cuda.cu:60-113).  The branch-cut
placement (goldstein_branch_cut.m) is restated in oracle/goldstein.py, which
``rytov_solve`` uses by default; ``unwrap_residue_free`` is the closed case with no
residues (zero step maps).
"""
from __future__ import annotations

import numpy as np

from .optics import derive_optical_tool, cfft2, potential2ri
from .cbs import matlab_round

PI32 = np.float32(3.141592654)


def vector2scalarfield(input_field, output_field):
    """Codes/sub_code/vector2scalarfield.m:1-14: project two polarisations on the dominant one."""
    inp = np.asarray(input_field)
    out = np.asarray(output_field)
    if inp.shape[2] == 2:
        dom = (np.mean(np.abs(inp), axis=(0, 1), keepdims=True)[:, :, :1]
               > np.mean(np.abs(inp[:, :, 1:2]), axis=(0, 1), keepdims=True))                 # :5
        dom = np.concatenate([dom, ~dom], axis=2)                                               # :6
        ph = np.angle(np.mean(inp * dom, axis=2, keepdims=True))
        main = np.sum(inp * np.exp(-1j * ph), axis=(0, 1), keepdims=True)                      # :7
        main = main / np.sqrt(np.sum(np.abs(main) ** 2, axis=2, keepdims=True))                # :8
    else:
        main = 1
    return np.sum(inp * main, axis=2), np.sum(out * main, axis=2)                             # :11-12


def _mod32(a, b):
    """fmodf(fmodf(a, b) + b, b) in single precision (get_residue_cuda.cu:30-32)."""
    a = np.asarray(a, np.float32)
    b = np.float32(b)
    return np.fmod(np.fmod(a, b) + b, b).astype(np.float32)


def residues(wrapped):
    """get_residue_cuda.cu:35-66: residue of each 2x2 loop, (n1-1) x (n2-1) x n3 (FP32)."""
    p = np.asarray(wrapped, np.float32)
    v11, v21 = p[:-1, :-1], p[1:, :-1]
    v12, v22 = p[:-1, 1:], p[1:, 1:]
    tp = np.float32(2.0) * PI32
    d1_1 = _mod32(v11 - v12 + PI32, tp)
    d1_2 = _mod32(v21 - v22 + PI32, tp)
    d2_1 = _mod32(v11 - v21 + PI32, tp)
    d2_2 = _mod32(v12 - v22 + PI32, tp)
    return np.round((d1_1 + d2_2 - d1_2 - d2_1) / tp).astype(np.float32)


def _fma_step(x, step, tp):
    """float32(x - step * tp) with one rounding: fmaf(-step, tp, x)."""
    return np.float32(np.float64(np.float32(x)) - np.float64(step) * np.float64(tp))


def raster_unwrap(wrapped, step1, step2):
    """unwrapp_phase_goldstein_cuda.cu:60-113: integrate the wrapped differences down the
    first column (dim 2, minus 2 pi step2) and then along every row (dim 1, minus 2 pi
    step1), in single precision.  wrapped: (n1, n2, n3); step1: (n1, n2, n3); step2: (n2, n3).
    The update last + d - step * 2 pi is one fused multiply-add, as nvcc contracts the
    reference's expression by default: the product is not rounded on its own (formed in
    float64, where it and the difference are exact for phases below 2^20 rad, then rounded
    once to float32)."""
    p = np.array(wrapped, np.float32, copy=True)
    n1, n2, n3 = p.shape
    tp = np.float32(2.0) * PI32
    for z in range(n3):                                                      # unwrapp_kernel_1
        last = p[0, 0, z]
        for i in range(1, n2):
            cur = p[0, i, z]
            d = _mod32(cur - last + PI32, tp) - PI32
            cur = _fma_step(last + d, step2[i, z], tp)
            p[0, i, z] = cur
            last = cur
    for z in range(n3):                                                      # unwrapp_kernel_2
        for j in range(n2):
            last = p[0, j, z]
            for i in range(1, n1):
                cur = p[i, j, z]
                d = _mod32(cur - last + PI32, tp) - PI32
                cur = _fma_step(last + d, step1[i, j, z], tp)
                p[i, j, z] = cur
                last = cur
    return p


def unwrap_residue_free(wrapped):
    """unwrapp2_gpu.m:1-5 for phase maps without residues (zero step maps)."""
    w = np.asarray(wrapped, np.float32)
    if np.any(residues(w) != 0):
        raise ValueError('phase map holds residues: use oracle.goldstein.unwrap')
    n1, n2, n3 = w.shape
    return raster_unwrap(w, np.zeros((n1, n2, n3), np.int32), np.zeros((n2, n3), np.int32))


def rytov_solve(params, input_field, output_field, unwrap=None):
    """Codes/BACKWARD_SOLVER_RYTOV.m:22-127 (FP64 except the FP32 unwrap).

    input_field, output_field: (x, y, pol, A).  Returns (RI, ORytov)."""
    p = params
    u = derive_optical_tool(p)
    inp = np.asarray(input_field)
    out = np.asarray(output_field)
    if inp.ndim != 4:                                                         # :27-29
        raise ValueError('You need to provide the field with 4 dimenssion')
    if inp.shape[0] != inp.shape[1]:
        raise ValueError('Please input a square field')
    if inp.shape != out.shape:
        raise ValueError('Please input field and bg of same size')
    if p['resolution'][0] != p['resolution'][1]:
        raise ValueError('x/y input resolution must be isotropic')
    if p['size'][0] != p['size'][1]:
        raise ValueError('x/y output size must be isotropic')
    if p['size'][0] != inp.shape[0] or p['size'][1] != inp.shape[1]:
        raise ValueError('declare size in the parameter must be the same as the field size')
    if unwrap is None:
        from .goldstein import unwrap
    bg, sp = vector2scalarfield(inp, out)                                     # :47
    ratio = sp / bg
    ret_phase = unwrap(np.angle(ratio).astype(np.float32)).astype(np.float64)   # :48-49
    ret_amp = np.abs(ratio)                                                   # :53
    A = ret_phase.shape[2]
    n1, n2, n3 = [int(s) for s in p['size']]
    res = u['fourier_res']
    f_d1 = np.zeros(A)
    f_d2 = np.zeros(A)
    for ii in range(A):                                                       # :60-70
        F = cfft2(bg[:, :, ii])
        m = np.abs(F)
        # MATLAB max over complex values compares magnitudes; find() returns column-major order
        idx = np.flatnonzero((m == m.max()).ravel(order='F'))[0]
        i1, i2 = idx % n1, idx // n1
        if ii == 0:
            i1_0, i2_0 = i1, i2
        f_d1[ii] = i1 - i1_0                                                  # f_dy (dim 1)
        f_d2[ii] = i2 - i2_0                                                  # f_dx (dim 2)
    k0_2 = res[1] * f_d2                                                      # :72 (k0_x, dim 2)
    k0_1 = res[0] * f_d1                                                      # :73 (k0_y, dim 1)
    k0_3 = np.real(np.sqrt((u['k0_nm'] ** 2 - k0_2 ** 2 - k0_1 ** 2).astype(complex)))   # :74
    na = u['NA_circle'][:, :, 0]
    k3 = u['k3'][:, :, 0]
    c1 = np.broadcast_to(u['fourier_coor'][0][:, :, 0], (n1, n2))
    c2 = np.broadcast_to(u['fourier_coor'][1][:, :, 0], (n1, n2))
    OR = np.zeros((n1, n2, n3), complex)
    CNT = np.zeros((n1, n2, n3))
    dxy = p['resolution'][0] * p['resolution'][1]
    for kk in range(A):                                                       # :79-104
        FR = np.log(ret_amp[:, :, kk]) + 1j * ret_phase[:, :, kk]
        Us = cfft2(FR) * dxy
        Us = np.roll(Us, (int(np.round(k0_1[kk] / res[0])), int(np.round(k0_2[kk] / res[1]))), axis=(0, 1))
        Us = Us * na
        K2 = c2 - k0_2[kk]
        K1 = c1 - k0_1[kk]
        K3 = k3 - k0_3[kk]
        Up = (k3 / 1j) * Us                                                   # :88
        sel = ((k3 > 0) & na
               & (K2 > res[1] * (-(n2 // 2))) & (K1 > res[0] * (-(n1 // 2))) & (K3 > res[2] * (-(n3 // 2)))
               & (K2 < res[1] * (n2 // 2 - 1)) & (K1 < res[0] * (n1 // 2 - 1)) & (K3 < res[2] * (n3 // 2 - 1)))
        # find() order is column-major; a duplicate 3D index keeps the value written last (:97-99)
        sel_f = np.flatnonzero(sel.ravel(order='F'))
        Up = Up.ravel(order='F')[sel_f]
        i1 = np.round(K1.ravel(order='F')[sel_f] / res[0] + n1 // 2).astype(int)   # 0-based
        i2 = np.round(K2.ravel(order='F')[sel_f] / res[1] + n2 // 2).astype(int)
        i3 = matlab_round(K3.ravel(order='F')[sel_f] / res[2] + n3 // 2 + 1).astype(int) - 1
        temp_or = OR.copy()
        temp_cnt = CNT.copy()
        for j in range(len(sel_f)):
            OR[i1[j], i2[j], i3[j]] = temp_or[i1[j], i2[j], i3[j]] + Up[j]
            CNT[i1[j], i2[j], i3[j]] = temp_cnt[i1[j], i2[j], i3[j]] + (Up[j] != 0)
    nz = CNT > 0
    OR[nz] = OR[nz] / CNT[nz] / float(np.prod(p['resolution']))             # :105
    from scipy import fft as sfft
    rec = sfft.fftshift(sfft.ifftn(sfft.ifftshift(OR)))                       # :106
    RI = potential2ri(rec * 4 * np.pi, p['wavelength'], p['RI_bg'])           # :107
    if p.get('use_non_negativity', False):                                    # :109-125
        emask = u['coorxy'][:, :, 0] < (2 * p['NA'] / p['wavelength'])
        rec = sfft.fftshift(sfft.ifftn(sfft.ifftshift(OR)))
        for _ in range(int(p.get('non_negativity_iteration', 100))):
            neg = np.real(rec) < 0
            rec[neg] = 1j * np.imag(rec[neg])
            new = sfft.fftshift(sfft.fftn(sfft.ifftshift(rec)))
            new = emask[:, :, None] * new * (np.abs(OR) == 0) + OR
            rec = sfft.fftshift(sfft.ifftn(sfft.ifftshift(new)))
        neg = np.real(rec) < 0
        rec[neg] = 1j * np.imag(rec[neg])
        RI = potential2ri(rec * 4 * np.pi, p['wavelength'], p['RI_bg'])
    return RI, OR

"""Isotropic 3D total-variation proximal step (FISTA on the dual), host FP64.

Follows Codes/sub_code/TV_FISTA_inner_v2.m:1-123 for 3D input: periodic forward
differences, the projection of the potential onto its bounds, dual projection
onto the unit ball and FISTA momentum on the dual variable.
"""
from __future__ import annotations

import numpy as np


def _grad(x, out):
    """Periodic forward differences x - circshift(x, -1) along each axis (:114-123)."""
    for ax in range(3):
        np.subtract(x, np.roll(x, -1, axis=ax), out=out[ax])
    return out


def _div(P, out):
    """Adjoint of _grad: P0 + P1 + P2 - circshift(P_k, +1, axis k) (:99-113)."""
    np.add(P[0], P[1], out=out)
    out += P[2]
    for ax in range(3):
        out -= np.roll(P[ax], 1, axis=ax)
    return out


def project_bounds(x, Vmin, Vmax, Vimag_max, non_neg, dirichlet):
    """Bounds of the potential (:57-75), in place on a copy."""
    x = np.array(x, copy=True)
    if non_neg:
        if np.iscomplexobj(x):
            re = x.real
            x.real = np.clip(re, Vmin, Vmax)
        else:
            np.clip(x, Vmin, Vmax, out=x)
    if np.iscomplexobj(x):
        im = x.imag
        im[im < 0] = 0
        if Vimag_max > 0:
            np.minimum(im, Vimag_max, out=im)
        x.imag = im
    if dirichlet:
        for ax in range(3):
            idx = [slice(None)] * 3
            idx[ax] = 0
            x[tuple(idx)] = 0
            idx[ax] = -1
            x[tuple(idx)] = 0
    return x


def tv_fista_inner(in_mat, lam, Vmin, Vmax, Vimag_max, non_neg, dirichlet, inner_itt):
    """TV_FISTA_inner_v2 for a 3D array (returns the projected TV proximal point)."""
    in_mat = np.asarray(in_mat)
    if Vimag_max == 0:
        in_mat = np.real(in_mat)
    in_mat = in_mat.astype(np.complex128 if np.iscomplexobj(in_mat) else np.float64)
    P = np.zeros((3,) + in_mat.shape, dtype=in_mat.dtype)
    R = P.copy()
    D = np.empty((3,) + in_mat.shape, dtype=in_mat.dtype)
    div = np.empty_like(in_mat)
    t = 1.0
    c = 1.0 / (12.0 * lam)
    for _ in range(int(inner_itt)):
        P_old = P
        y = project_bounds(in_mat - lam * _div(R, div), Vmin, Vmax, Vimag_max, non_neg, dirichlet)
        _grad(y, D)
        P = R + c * D
        nrm = np.sqrt(np.maximum(np.abs(P[0]) ** 2 + np.abs(P[1]) ** 2 + np.abs(P[2]) ** 2, 1.0))
        P /= nrm
        t_new = (1 + np.sqrt(1 + 4 * t * t)) / 2
        R = P + ((t - 1) / t_new) * (P - P_old)
        t = t_new
    return project_bounds(in_mat - lam * _div(P, div), Vmin, Vmax, Vimag_max, non_neg, dirichlet)


def tv_fista_inner_device(solver, in_mat, lam, Vmin, Vmax, Vimag_max, non_neg, dirichlet, inner_itt):
    """The same proximal step on the GPU of `solver` (a FORWARD_SOLVER_CONVERGENT_BORN), in
    single precision as the reference computes it: real when Vimag_max == 0 (cbs_tv_prox),
    complex otherwise (cbs_tv_prox_complex; NaN, the value the defaults nmax = Inf and
    kappamax = 0 give, keeps the imaginary part with no upper cap)."""
    import ctypes as C
    from ._lib import check, lib, CBS_MEM_HOST
    dims = (C.c_uint32 * 3)(*np.shape(in_mat))
    if Vimag_max == 0:
        x = np.asfortranarray(np.real(in_mat).astype(np.float32))
        out = np.empty_like(x, order='F')
        check(lib.cbs_tv_prox(solver._h, x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), dims,
                              float(lam), float(Vmin), float(Vmax), int(bool(non_neg)), int(bool(dirichlet)),
                              int(inner_itt), CBS_MEM_HOST))
        return out.astype(np.float64)
    x = np.asfortranarray(np.asarray(in_mat).astype(np.complex64))
    out = np.empty_like(x, order='F')
    check(lib.cbs_tv_prox_complex(solver._h, x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), dims,
                                  float(lam), float(Vmin), float(Vmax), float(Vimag_max), int(bool(non_neg)),
                                  int(bool(dirichlet)), int(inner_itt), CBS_MEM_HOST))
    return out.astype(np.complex128)

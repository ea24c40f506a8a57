"""BACKWARD_SOLVER_RYTOV, MI355X backend: the Rytov reconstruction that
BACKWARD_SOLVER_MULTI starts from (Codes/BACKWARD_SOLVER_RYTOV.m:22-127), with the
phase unwrapping on the GPU.

Phase unwrapping (Codes/sub_code/goldstein_GPU/unwrapp2_gpu.m) is split as the
reference splits it:
- the residue map, the dipole removal, the nearest-residue lists and the rastering of
  branches into 2 pi step maps run on the device;
- the branch search (which residues a cut joins) runs on the host, as
  get_branches_goldstein_v2 runs on the CPU there, with the reference's ordering and
  search rules, so the step maps are the reference's;
- the raster integration with the step maps runs on the device (cbs_phase_integrate).

The k-space mapping follows the MATLAB code, including two behaviours:
- the illumination peak is measured relative to the first angle;
- a 3D frequency hit twice within one angle keeps the last value written.
"""
from __future__ import annotations

import ctypes as C
import numpy as np

from ._lib import check, lib, CBS_MEM_HOST
from .params import update_struct, basic_optical_parameter, fourier_grid, pupil, centered_fft2

_TWO_PI32 = np.float32(2.0) * np.float32(3.141592654)


def get_default_parameters(init=None):
    """BACKWARD_SOLVER_RYTOV.get_default_parameters (:6-17) on top of the optical defaults."""
    p = basic_optical_parameter()
    p.update(use_non_negativity=False, non_negativity_iteration=100, use_GPU=True)
    return update_struct(p, init) if init is not None else p


def phase_residues(wrapped, device=0):
    """Residues of the 2x2 loops of (n1, n2, n3) phase maps on the GPU."""
    w = np.asfortranarray(np.asarray(wrapped, np.float32))
    if w.ndim == 2:
        w = w[:, :, None]
    n1, n2, n3 = w.shape
    out = np.empty((n1 - 1, n2 - 1, n3), np.float32, order='F')
    dims = (C.c_uint32 * 3)(n1, n2, n3)
    check(lib.cbs_phase_residues(int(device), w.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), dims,
                                 CBS_MEM_HOST))
    return out


def phase_branch_cuts(residues, device=0):
    """Goldstein step maps from a residue map (n1 - 1, n2 - 1, n3): dipole removal and the
    reference's branch search (goldstein_branch_cut.m:1-87).  Returns (step1 (n1, n2, n3),
    step2 (n2, n3)) int32, column-major."""
    r = np.asfortranarray(np.asarray(residues, np.float32))
    if r.ndim == 2:
        r = r[:, :, None]
    m1, m2, n3 = r.shape
    step1 = np.empty((m1 + 1, m2 + 1, n3), np.int32, order='F')
    step2 = np.empty((m2 + 1, n3), np.int32, order='F')
    dims = (C.c_uint32 * 3)(m1 + 1, m2 + 1, n3)
    check(lib.cbs_phase_branch_cuts(int(device), r.ctypes.data_as(C.c_void_p), step1.ctypes.data_as(C.c_void_p),
                                    step2.ctypes.data_as(C.c_void_p), dims, CBS_MEM_HOST))
    return step1, step2


def unwrap_phase(wrapped, device=0):
    """unwrapp2_gpu.m:1-5: (n1, n2[, n3]) wrapped phases -> unwrapped phases (float32):
    residues, Goldstein branch cuts and the raster integration (cbs_phase_unwrap)."""
    w = np.asfortranarray(np.asarray(wrapped, np.float32))
    squeeze = w.ndim == 2
    if squeeze:
        w = w[:, :, None]
    n1, n2, n3 = w.shape
    out = np.empty_like(w, order='F')
    dims = (C.c_uint32 * 3)(n1, n2, n3)
    check(lib.cbs_phase_unwrap(int(device), w.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), dims,
                               CBS_MEM_HOST))
    return out[:, :, 0] if squeeze else out


def vector2scalarfield(input_field, output_field):
    """Codes/sub_code/vector2scalarfield.m:1-14.  For two polarisations: per angle, the
    polarisation with the larger mean magnitude is dominant; the polarisation vector is
    the field summed over the image after removing the dominant component's phase,
    normalised; both fields are then summed over polarisations weighted by that vector
    (as the reference does, without conjugation)."""
    a = np.asarray(input_field)
    b = np.asarray(output_field)
    if a.shape[2] != 2:
        return a.sum(axis=2), b.sum(axis=2)
    mag = np.mean(np.abs(a), axis=(0, 1))                          # (2, A)
    dominant = np.where((mag[0] > mag[1])[None, None, :], a[:, :, 0, :], a[:, :, 1, :])
    pol = np.sum(a * np.exp(-1j * np.angle(dominant))[:, :, None, :], axis=(0, 1))   # (2, A)
    pol = pol / np.sqrt(np.sum(np.abs(pol) ** 2, axis=0, keepdims=True))
    return np.sum(a * pol[None, None], axis=2), np.sum(b * pol[None, None], axis=2)


class BACKWARD_SOLVER_RYTOV:
    get_default_parameters = staticmethod(get_default_parameters)

    def __init__(self, params, device=0):
        self.parameters = dict(params)
        self.device = int(device)

    def solve(self, input_field, output_field):
        """:22-127.  input_field, output_field: (x, y, pol, A).  Returns (RI, ORytov)."""
        p = self.parameters
        a = np.asarray(input_field)
        b = np.asarray(output_field)
        if a.ndim != 4:
            raise ValueError('You need to provide the field with 4 dimenssion : dim1 x dim2 x polarisation x '
                             'illuminationnumber')
        if a.shape[0] != a.shape[1]:
            raise ValueError('Please input a square field')
        if a.shape != b.shape:
            raise ValueError('Please input field and bg of same size')
        if p['resolution'][0] != p['resolution'][1]:
            raise ValueError('x/y input resolution must be isotropic')
        if p['size'][0] != p['size'][1]:
            raise ValueError('x/y output size must be isotropic')
        if p['size'][0] != a.shape[0] or p['size'][1] != a.shape[1]:
            raise ValueError('declare size in the parameter must be the same as the field size')
        bg, sp = vector2scalarfield(a, b)
        ratio = sp / bg
        phase = unwrap_phase(np.angle(ratio), self.device).astype(np.float64)
        amp = np.abs(ratio)
        n1, n2, n3 = [int(v) for v in p['size']]
        A = phase.shape[2]
        res = [1.0 / (p['resolution'][j] * [n1, n2, n3][j]) for j in range(3)]
        # illumination peaks relative to the first angle (:60-70)
        spec_bg = centered_fft2(bg)
        flat = np.abs(spec_bg).reshape(n1 * n2, A, order='F')
        peak = np.argmax(flat, axis=0)                 # first maximum in column-major order
        d1 = (peak % n1) - (peak[0] % n1)
        d2 = (peak // n1) - (peak[0] // n1)
        k0_nm = p['RI_bg'] / p['wavelength']
        k01 = res[0] * d1
        k02 = res[1] * d2
        k03 = np.sqrt(np.maximum(k0_nm ** 2 - k01 ** 2 - k02 ** 2, 0.0))
        na, _, k3 = pupil((n1, n2), p['resolution'], p['wavelength'], p['NA'], p['RI_bg'])
        f1, f2 = fourier_grid((n1, n2), p['resolution'][:2])
        F1 = np.broadcast_to(f1[:, None], (n1, n2))
        F2 = np.broadcast_to(f2[None, :], (n1, n2))
        dxy = p['resolution'][0] * p['resolution'][1]
        OR = np.zeros(n1 * n2 * n3, complex)
        CNT = np.zeros(n1 * n2 * n3)
        lo = [-(n // 2) for n in (n1, n2, n3)]
        hi = [n // 2 - 1 for n in (n1, n2, n3)]
        for kk in range(A):
            spec = centered_fft2(np.log(amp[:, :, kk]) + 1j * phase[:, :, kk]) * dxy
            spec = np.roll(spec, (int(d1[kk]), int(d2[kk])), axis=(0, 1)) * na
            K1 = F1 - k01[kk]
            K2 = F2 - k02[kk]
            K3 = k3 - k03[kk]
            vals = (k3 / 1j) * spec                                                       # :88
            ok = ((k3 > 0) & na & (K1 > res[0] * lo[0]) & (K2 > res[1] * lo[1]) & (K3 > res[2] * lo[2])
                  & (K1 < res[0] * hi[0]) & (K2 < res[1] * hi[1]) & (K3 < res[2] * hi[2]))
            sel = np.flatnonzero(ok.ravel(order='F'))
            i1 = np.rint(K1.ravel(order='F')[sel] / res[0]).astype(np.int64) + n1 // 2
            i2 = np.rint(K2.ravel(order='F')[sel] / res[1]).astype(np.int64) + n2 // 2
            x3 = K3.ravel(order='F')[sel] / res[2] + n3 // 2 + 1
            i3 = np.floor(x3 + 0.5).astype(np.int64) - 1                  # MATLAB round (x3 > 0)
            lin = i1 + n1 * (i2 + n2 * i3)
            v = vals.ravel(order='F')[sel]
            # last write wins for a repeated target: keep the final occurrence of each index
            rev_unique, rev_first = np.unique(lin[::-1], return_index=True)
            keep = len(lin) - 1 - rev_first
            OR[rev_unique] += v[keep]
            CNT[rev_unique] += (v[keep] != 0)
        OR = OR.reshape((n1, n2, n3), order='F')
        CNT = CNT.reshape((n1, n2, n3), order='F')
        hit = CNT > 0
        OR[hit] = OR[hit] / CNT[hit] / float(np.prod(p['resolution']))
        rec = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(OR)))
        kk2 = (2 * np.pi * p['RI_bg'] / p['wavelength']) ** 2
        to_ri = lambda pot: p['RI_bg'] * np.sqrt(1 + pot * 4 * np.pi / kk2)           # potential2RI(4 pi rec)
        RI = to_ri(rec)
        if p.get('use_non_negativity', False):                                          # :109-125
            kxy = np.sqrt(F1 ** 2 + F2 ** 2)
            emask = (kxy < 2 * p['NA'] / p['wavelength'])[:, :, None]
            empty = np.abs(OR) == 0
            for _ in range(int(p.get('non_negativity_iteration', 100))):
                rec = np.where(np.real(rec) < 0, 1j * np.imag(rec), rec)
                new = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(rec)))
                new = emask * new * empty + OR
                rec = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(new)))
            rec = np.where(np.real(rec) < 0, 1j * np.imag(rec), rec)
            RI = to_ri(rec)
        return RI, OR

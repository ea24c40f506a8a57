"""GPU: the centred pupil filter of the field grid (cbs_field_pupil) and the scalar
solve that applies it on the device (cbs_solve_field, which solve() calls for
pole_num 1).

cbs_field_pupil is checked against NumPy in FP64,
fftshift(ifft2(ifftshift(P * fftshift(fft2(ifftshift(in + add)))))), on field grids of
even and odd, 5-smooth, prime and mixed lengths (compiled line FFTs, run-time Stockham
and Bluestein transforms), for host and device memory, with and without the added
field and in place: <= 2e-6 relative L2 (complex64 transforms of <= 512 points a side).
The solve is checked against the FP64 oracle with the Abbe-sine factors on and off, on
an even field grid and on one of prime lengths."""
import ctypes as C

import numpy as np
import pytest

from helpers import params, rel, random_field

pytestmark = pytest.mark.gpu


def _np_pupil(x, P, add=None):
    a = x + (add if add is not None else 0)
    ax = (0, 1)
    F = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(a, axes=ax), axes=ax), axes=ax)
    F = F * P[:, :, None]
    return np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(F, axes=ax), axes=ax), axes=ax)


def _solver(fx, fy):
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    p = params((fx, fy, 8), iterations=2)
    s = FORWARD_SOLVER_CONVERGENT_BORN(p)
    s.set_RI(np.full((fx, fy, 8), 1.34, np.float32))
    return s


@pytest.mark.parametrize('fx,fy', [(96, 96), (75, 45), (83, 71), (16, 20), (49, 98), (512, 512)])
def test_field_pupil_matches_numpy(fx, fy):
    from inverse_solver_amd import _lib
    from inverse_solver_amd import device as dv
    lib, check = _lib.lib, _lib.check
    rng = np.random.default_rng(fx * 1000 + fy)
    nf = 3
    x = random_field((fx, fy, nf), fx + fy).astype(np.complex64)
    b = random_field((fx, fy, nf), fx * fy).astype(np.complex64)
    P = rng.random((fx, fy)).astype(np.float32)
    P[rng.random((fx, fy)) < 0.3] = 0
    s = _solver(fx, fy)
    h = s._h
    xf, bf, Pf = (np.asfortranarray(v) for v in (x, b, P))
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    ref = _np_pupil(x.astype(np.complex128), P.astype(np.float64))
    ref_add = _np_pupil(x.astype(np.complex128), P.astype(np.float64), b.astype(np.complex128))
    # host memory, without and with the added field
    out = np.empty_like(xf, order='F')
    check(lib.cbs_field_pupil(h, ptr(xf), None, nf, ptr(Pf), ptr(out), _lib.CBS_MEM_HOST))
    e0 = rel(out, ref)
    check(lib.cbs_field_pupil(h, ptr(xf), ptr(bf), nf, ptr(Pf), ptr(out), _lib.CBS_MEM_HOST))
    e1 = rel(out, ref_add)
    # device memory, in place
    d = dv.DeviceArray.from_host(xf.T)
    db = dv.DeviceArray.from_host(bf.T)
    check(lib.cbs_field_pupil(h, C.c_void_p(d.ptr), C.c_void_p(db.ptr), nf, ptr(Pf), C.c_void_p(d.ptr),
                              _lib.CBS_MEM_DEVICE))
    e2 = rel(d.to_host().T, ref_add)
    s.close()
    print((fx, fy), 'rel err host %.1e, host+add %.1e, device in place %.1e' % (e0, e1, e2), flush=True)
    assert max(e0, e1, e2) < 2e-6


@pytest.mark.parametrize('abbe', [True, False])
@pytest.mark.parametrize('ri', [(16, 16, 16), (17, 19, 12)])
def test_scalar_solve_with_device_pupils_matches_oracle(abbe, ri):
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd.phantoms import bead_ri, circle_illumination
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    p = params(ri, iterations=12, use_abbe_sine=abbe)
    RI = bead_ri(ri, 5, 1.46, 1.336)
    u = circle_illumination(ri[:2], 3, 1.2, 0.532, [0.1] * 3)
    o = Oracle(p)
    o.set_RI(RI)
    t_ref, r_ref, f_ref = o.solve(u)
    s = FORWARD_SOLVER_CONVERGENT_BORN(p)
    s.set_RI(RI)
    t, r, f = s.solve(u)
    s.close()
    et, er, ef = rel(t, t_ref), rel(r, r_ref), rel(f, f_ref)
    print(ri, 'abbe', abbe, 'trans %.1e ref %.1e 3d %.1e' % (et, er, ef), flush=True)
    assert t.shape == t_ref.shape and r.shape == r_ref.shape and f.shape == f_ref.shape
    assert et < 1e-5 and ef < 1e-5 and er < 1e-4

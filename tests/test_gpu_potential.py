"""GPU: set_RI's RI -> V conversion on the device (cbs_potential_max,
cbs_set_potential_ri) against the host formula it replaces: Vt = k^2 (RI^2/n_m^2 - 1)
(Codes/sub_code/RI2potential.m:1-4), eps = 1.01 max|Vt| (:205), V = pad(Vt) - i eps
(:212-215, :238-241), and the adjoint potential from conj(flip_z(RI))
(Codes/BACKWARD_SOLVER_MULTI.m:48).

For a real RI the device evaluates the same FP64 operations in the same order, so
max|Vt|, V and the solves that follow are identical bit for bit.  For a complex RI
the modulus may differ in the last bits (hypot), so max|Vt| agrees to a few ulps and
the solves to 1e-6."""
import ctypes as C

import numpy as np
import pytest

from helpers import params, sphere_ri, random_field, rel

pytestmark = pytest.mark.gpu


def _host_V(s, RI, flip_conj):
    from inverse_solver_amd.params import ri_to_potential
    p = s.parameters
    RI = np.asarray(RI)
    if flip_conj:
        RI = np.conj(np.flip(RI, axis=2))
    Vt = ri_to_potential(RI.astype(np.complex128), p['wavelength'], p['RI_bg'])
    vmax = float(np.max(np.abs(Vt)))
    b = s.boundary_thickness_pixel
    Vpad = np.pad(Vt, [(int(b[0]),) * 2, (int(b[1]),) * 2, (int(b[2]),) * 2])
    return np.asfortranarray((Vpad - 1j * (vmax * 1.01)).astype(np.complex64)), vmax


def _set_host(s, RI, flip_conj):
    from inverse_solver_amd._lib import lib, check, CBS_MEM_HOST
    V, vmax = _host_V(s, RI, flip_conj)
    prm = s._potential_params(float(vmax * 1.01))
    check(lib.cbs_set_potential(s._h, C.byref(prm), V.ctypes.data_as(C.c_void_p), CBS_MEM_HOST))
    s._host_V = V
    return vmax


def _device_vmax(s, RI):
    from inverse_solver_amd._lib import lib, check, CBS_MEM_HOST
    cplx = bool(np.iscomplexobj(RI) and np.any(np.imag(RI) != 0))
    arr = np.asfortranarray(RI, dtype=np.complex128) if cplx else np.asfortranarray(np.real(RI), dtype=np.float64)
    out = C.c_double()
    check(lib.cbs_potential_max(s._h, C.c_void_p(arr.ctypes.data), int(cplx), float(s.parameters['wavelength']),
                                float(s.parameters['RI_bg']), CBS_MEM_HOST, C.byref(out)))
    return out.value


@pytest.mark.parametrize('cplx,flip', [(False, False), (False, True), (True, False), (True, True)])
def test_device_potential_matches_host_formula(cplx, flip):
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    ri = (16, 16, 12)
    p = params(ri, iterations=7)
    RI = sphere_ri(ri, 0.5, n_s=1.41, center=(7, 8, 5))           # off centre in z: the flip matters
    if cplx:
        RI = RI + 0.01j * sphere_ri(ri, 0.3, n_s=1.0, n_m=0.0, center=(7, 8, 4))
    fields = np.stack([random_field(ri, 30 + a) for a in range(2)], axis=-1)
    a = FORWARD_SOLVER_CONVERGENT_BORN(p)
    vh = _set_host(a, RI, flip)
    ua = a.solve_incident(fields)
    b = FORWARD_SOLVER_CONVERGENT_BORN(p)
    b.set_RI(RI, flip_conj=flip)
    ub = b.solve_incident(fields)
    vd = _device_vmax(b, RI)
    print(cplx, flip, 'vmax host %.17g device %.17g' % (vh, vd), 'rel %.2e' % rel(ub, ua), flush=True)
    assert a.Bornmax == b.Bornmax and a.ROI == b.ROI
    if cplx:
        assert abs(vd - vh) <= 1e-15 * vh
        assert rel(ub, ua) < 1e-6
    else:
        assert vd == vh and a.eps_imag == b.eps_imag
        np.testing.assert_array_equal(ub, ua)
    a.close()
    b.close()


def test_device_potential_from_device_buffer_and_errors():
    """RI already in device memory (the backward solver's resident copy), a max|Vt|
    passed in, and the argument checks of the two entry points."""
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd import device as dv
    from inverse_solver_amd._lib import lib, CbsError, CBS_MEM_DEVICE, check
    ri = (16, 16, 12)
    p = params(ri, iterations=5)
    RI = sphere_ri(ri, 0.5, n_s=1.40)
    fields = random_field(ri, 3)
    a = FORWARD_SOLVER_CONVERGENT_BORN(p)
    a.set_RI(RI)
    ua = a.solve_incident(fields)
    buf = dv.DeviceArray((RI.size,), np.float64)
    buf.upload(np.asfortranarray(RI))
    b = FORWARD_SOLVER_CONVERGENT_BORN(p)
    vmax = b.set_RI_device(buf.ptr, False)
    np.testing.assert_array_equal(b.solve_incident(fields), ua)
    b.set_RI_device(buf.ptr, False, vmax=vmax)
    np.testing.assert_array_equal(b.solve_incident(fields), ua)
    out = C.c_double()
    with pytest.raises(CbsError):
        check(lib.cbs_potential_max(b._h, C.c_void_p(buf.ptr), 2, 0.532, 1.336, CBS_MEM_DEVICE, C.byref(out)))
    with pytest.raises(CbsError):
        check(lib.cbs_potential_max(b._h, C.c_void_p(buf.ptr), 0, -1.0, 1.336, CBS_MEM_DEVICE, C.byref(out)))
    with pytest.raises(CbsError):
        check(lib.cbs_potential_max(b._h, None, 0, 0.532, 1.336, CBS_MEM_DEVICE, C.byref(out)))
    prm = b._potential_params(b.eps_imag)
    with pytest.raises(CbsError):
        check(lib.cbs_set_potential_ri(b._h, C.byref(prm), C.c_void_p(buf.ptr), 0, 2, 0.532, 1.336, CBS_MEM_DEVICE))
    with pytest.raises(ValueError):                     # homogeneous RI: eps = 0
        b.set_RI(np.full(ri, 1.336))
    bad = RI.copy()
    bad[3, 4, 5] = np.nan
    with pytest.raises(ValueError, match='non-finite'):
        b.set_RI(bad)
    bad[3, 4, 5] = np.inf
    with pytest.raises(ValueError, match='non-finite'):
        b.set_RI(bad)
    buf.free()
    a.close()
    b.close()


@pytest.mark.parametrize('ri,bt,flip', [((64, 64, 64), 16, False), ((128, 128, 128), 16, True)])
def test_real_ri_float_potential_on_the_fused_kernels(ri, bt, flip):
    """A real RI's potential is read by the fused real-space pass as its real part alone
    (4 bytes a voxel, Im V = -eps rebuilt in registers), a caller's V as complex64: on the
    96^3 (C1) and 160^3 (headline) grids, in storage and rhs-first block order, both give
    the same fields bit for bit."""
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd.phantoms import bead_ri, incident_plane_waves
    p = params(ri, bt_px_target=bt, iterations=5, return_reflection=False)
    RI = bead_ri(ri, ri[0] // 3, 1.46, 1.336)
    u = incident_plane_waves(list(ri), 3, 1.2, 0.532, 1.336, [0.1, 0.1, 0.1]).astype(np.complex64)
    for order in (4, 0):
        a = FORWARD_SOLVER_CONVERGENT_BORN(p, options=dict(block_order=order))
        _set_host(a, RI, flip)                      # complex V from the caller
        assert a.plan()['form'] == 'fused'
        ua = a.solve_incident(u)
        a.close()
        b = FORWARD_SOLVER_CONVERGENT_BORN(p, options=dict(block_order=order))
        b.set_RI(RI, flip_conj=flip)                # real RI: the float form
        assert b.plan()['form'] == 'fused' and b.simulation_size == a.simulation_size
        ub = b.solve_incident(u)
        b.close()
        print(ri, 'flip', flip, 'order', order, 'sim', b.simulation_size, 'max |diff|', float(np.max(np.abs(ub - ua))),
              flush=True)
        np.testing.assert_array_equal(ub, ua)

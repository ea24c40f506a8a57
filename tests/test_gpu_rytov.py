"""GPU parity of the phase unwrapping kernels and the Rytov reconstruction.

* cbs_phase_residues equals the oracle's residue map exactly (same single-precision
  arithmetic) on random maps (many residues), vortices and smooth maps, odd sizes.
* cbs_phase_integrate equals the oracle's raster unwrap exactly, with zero and with
  random step maps.
* cbs_phase_branch_cuts equals the oracle's Goldstein step maps exactly (dipole removal,
  nearest lists, branch search, raster) on random maps (thousands of residues), a noisy
  vortex field, and a map with a vortex pair, a dipole and vortices cut to both borders;
  unwrap_phase (cbs_phase_unwrap) equals the oracle's unwrap exactly on the same maps;
  both entry points give the same on device buffers (the unwrap also in place).
* BACKWARD_SOLVER_RYTOV equals the oracle's restatement to 1e-6, on measured fields and on
  fields with phase noise that leaves residues.
"""
import numpy as np
import pytest

from helpers import params, sphere_ri, rel
from oracle import rytov as R

pytestmark = pytest.mark.gpu


def _maps():
    rng = np.random.default_rng(3)
    out = [rng.uniform(-np.pi, np.pi, (13, 11, 3)).astype(np.float32)]
    x, y = np.meshgrid(np.arange(17), np.arange(15), indexing='ij')
    v = np.stack([np.arctan2(y - 6.4, x - 8.2), np.arctan2(y - 3.7, x - 5.1) - np.arctan2(y - 9.3, x - 11.6),
                  np.angle(np.exp(1j * (0.8 * x + 0.5 * y + 0.03 * x * y)))], axis=2)
    out.append(np.angle(np.exp(1j * v)).astype(np.float32))
    return out


def test_gpu_residues_match_oracle_exactly():
    from inverse_solver_amd.rytov import phase_residues
    for m in _maps():
        got = phase_residues(m)
        ref = R.residues(m)
        print('residues', m.shape, int(np.abs(ref).sum()))
        np.testing.assert_array_equal(got, ref)


def test_gpu_raster_integration_matches_oracle_exactly():
    import ctypes as C
    from inverse_solver_amd._lib import check, lib, CBS_MEM_HOST
    rng = np.random.default_rng(4)
    for m in _maps():
        n1, n2, n3 = m.shape
        for steps in (0, 1):
            s1 = (rng.integers(-1, 2, (n1, n2, n3)) * steps).astype(np.int32)
            s2 = (rng.integers(-1, 2, (n2, n3)) * steps).astype(np.int32)
            w = np.asfortranarray(m)
            out = np.empty_like(w, order='F')
            dims = (C.c_uint32 * 3)(n1, n2, n3)
            check(lib.cbs_phase_integrate(0, w.ctypes.data_as(C.c_void_p),
                                          np.asfortranarray(s1).ctypes.data_as(C.c_void_p),
                                          np.asfortranarray(s2).ctypes.data_as(C.c_void_p),
                                          out.ctypes.data_as(C.c_void_p), dims, CBS_MEM_HOST))
            np.testing.assert_array_equal(out, R.raster_unwrap(m, s1, s2))


def _residue_maps():
    rng = np.random.default_rng(7)
    out = [rng.uniform(-np.pi, np.pi, s).astype(np.float32) for s in ((40, 40, 2), (33, 29, 1), (9, 9, 2), (128, 128, 3))]
    x, y = np.meshgrid(np.arange(100), np.arange(100), indexing='ij')
    p = x + 1j * y
    c = rng.uniform(5, 95, (30, 2))
    phi = 0.005 * x * y + sum(np.angle(p - complex(*cc)) * s for cc, s in zip(c, rng.choice([-1, 1], 30)))
    out.append(np.angle(np.exp(1j * (phi + 0.6 * rng.standard_normal(phi.shape))))[:, :, None].astype(np.float32))
    # a far vortex pair, a dipole, and a vortex cut to each border
    phi = np.angle((p - complex(19.5, 16.5)) / (p - complex(19.5, 22.5))) + np.angle(p - complex(3.5, 60.5)) \
        - np.angle(-1j * (p - complex(60.5, 3.5)))
    phi[50, 50], phi[51, 50] = -1.75, 1.75
    out.append(np.angle(np.exp(1j * phi))[:, :, None].astype(np.float32))
    return out


def test_gpu_branch_cuts_match_oracle_exactly():
    from oracle import goldstein as G
    from inverse_solver_amd.rytov import phase_branch_cuts
    for m in _residue_maps() + _maps():
        s1_ref, s2_ref, info = G.step_maps(m)
        s1, s2 = phase_branch_cuts(R.residues(m))
        print('branch cuts', m.shape, 'residues', int(np.abs(info['residues']).sum()),
              'branches', info['branches'].shape[1], flush=True)
        np.testing.assert_array_equal(s1, s1_ref)
        np.testing.assert_array_equal(s2, s2_ref)


def test_gpu_unwrap_matches_oracle_exactly():
    from oracle import goldstein as G
    from inverse_solver_amd.rytov import unwrap_phase
    for m in _residue_maps() + _maps():
        u = unwrap_phase(m)
        np.testing.assert_array_equal(u, G.unwrap(m))
        k = (u - m) / (2 * np.pi)
        np.testing.assert_allclose(k, np.round(k), atol=1e-4)              # congruent modulo 2 pi
    m = _maps()[1]
    # a residue-free map unwraps exactly as the oracle's raster integration; 2D input
    np.testing.assert_array_equal(unwrap_phase(m[:, :, 2:3]), R.unwrap_residue_free(m[:, :, 2:3]))
    np.testing.assert_array_equal(unwrap_phase(m[:, :, 0]), G.unwrap(m[:, :, :1])[:, :, 0])
    # no residues at all: a constant map picks up the single-precision wrap of a zero
    # difference (fmod(pi + 2 pi, 2 pi) - pi = -2.4e-7 per step), as the reference's does
    z = np.zeros((6, 5, 2), np.float32)
    np.testing.assert_array_equal(unwrap_phase(z), G.unwrap(z))
    assert np.max(np.abs(unwrap_phase(z))) < 1e-5


def test_gpu_unwrap_and_cuts_on_device_buffers():
    # CBS_MEM_DEVICE: column-major device arrays (a row-major (n3, n2, n1) buffer), in place
    # for the unwrap, and the residue map left unchanged by the branch cuts
    import ctypes as C
    from oracle import goldstein as G
    from inverse_solver_amd._lib import check, lib, CBS_MEM_DEVICE
    from inverse_solver_amd.device import DeviceArray
    m = _residue_maps()[0]
    n1, n2, n3 = m.shape
    dims = (C.c_uint32 * 3)(n1, n2, n3)
    ref = G.unwrap(m)
    dw = DeviceArray.from_host(np.ascontiguousarray(m.T))
    dout = DeviceArray((n3, n2, n1), np.float32)
    check(lib.cbs_phase_unwrap(0, C.c_void_p(dw.ptr), C.c_void_p(dout.ptr), dims, CBS_MEM_DEVICE))
    np.testing.assert_array_equal(dout.to_host().T, ref)
    check(lib.cbs_phase_unwrap(0, C.c_void_p(dw.ptr), C.c_void_p(dw.ptr), dims, CBS_MEM_DEVICE))
    np.testing.assert_array_equal(dw.to_host().T, ref)
    res = R.residues(m)
    dres = DeviceArray.from_host(np.ascontiguousarray(res.T))
    d1 = DeviceArray((n3, n2, n1), np.int32)
    d2 = DeviceArray((n3, n2), np.int32)
    check(lib.cbs_phase_branch_cuts(0, C.c_void_p(dres.ptr), C.c_void_p(d1.ptr), C.c_void_p(d2.ptr), dims,
                                    CBS_MEM_DEVICE))
    s1, s2, _ = G.step_maps(m)
    np.testing.assert_array_equal(d1.to_host().T, s1)
    np.testing.assert_array_equal(d2.to_host().T, s2)
    np.testing.assert_array_equal(dres.to_host().T, res)
    for d in (dw, dout, dres, d1, d2):
        d.free()


def test_gpu_branch_cuts_reject_bad_sizes():
    import ctypes as C
    from inverse_solver_amd._lib import lib, CBS_MEM_HOST
    r = np.zeros((1, 4, 1), np.float32, order='F')
    s1 = np.zeros((2, 5, 1), np.int32, order='F')
    s2 = np.zeros((5, 1), np.int32, order='F')
    dims = (C.c_uint32 * 3)(1, 5, 1)
    assert lib.cbs_phase_branch_cuts(0, r.ctypes.data_as(C.c_void_p), s1.ctypes.data_as(C.c_void_p),
                                     s2.ctypes.data_as(C.c_void_p), dims, CBS_MEM_HOST) != 0


def _measured(ri, n_s, radius, A, iters=20):
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd.phantoms import circle_illumination
    p = params(ri, iterations=iters, return_reflection=False, return_3D=False)
    s = FORWARD_SOLVER_CONVERGENT_BORN(p)
    s.set_RI(sphere_ri(ri, radius, n_s=n_s))
    u = circle_illumination(ri[:2], A, 0.9, 0.532, [0.1] * 3)
    t, _, _ = s.solve(u)
    s.close()
    return p, u.astype(np.complex128), t.astype(np.complex128)


@pytest.mark.parametrize('non_neg', [False, True])
def test_gpu_rytov_matches_oracle(non_neg):
    from inverse_solver_amd.rytov import BACKWARD_SOLVER_RYTOV, get_default_parameters
    ri = (24, 24, 24)
    p, u, t = _measured(ri, 1.35, 0.6, 7)
    rp = get_default_parameters(dict(size=list(p['size']), wavelength=0.532, NA=1.2, RI_bg=1.336,
                                     resolution=[0.1] * 3, use_non_negativity=non_neg, non_negativity_iteration=20))
    RI, OR = BACKWARD_SOLVER_RYTOV(rp).solve(u, t)
    RI_ref, OR_ref = R.rytov_solve(rp, u, t)
    print('Rytov vs oracle: RI contrast %.2e  ORytov %.2e' % (rel(RI - 1.336, RI_ref - 1.336), rel(OR, OR_ref)))
    assert rel(OR, OR_ref) < 1e-6
    assert rel(RI - 1.336, RI_ref - 1.336) < 1e-6


def test_gpu_rytov_with_residues_matches_oracle():
    from oracle import goldstein as G
    from inverse_solver_amd.rytov import BACKWARD_SOLVER_RYTOV, get_default_parameters
    ri = (24, 24, 24)
    p, u, t = _measured(ri, 1.35, 0.6, 5)
    t = t * np.exp(1j * 1.4 * np.random.default_rng(11).standard_normal(t.shape))      # phase noise
    n_res = int(np.abs(R.residues(np.angle(t[:, :, 0] / u[:, :, 0]).astype(np.float32))).sum())
    assert n_res > 20
    rp = get_default_parameters(dict(size=list(p['size']), wavelength=0.532, NA=1.2, RI_bg=1.336,
                                     resolution=[0.1] * 3))
    RI, OR = BACKWARD_SOLVER_RYTOV(rp).solve(u, t)
    RI_ref, OR_ref = R.rytov_solve(rp, u, t, unwrap=G.unwrap)
    print('Rytov with %d residues vs oracle: RI contrast %.2e  ORytov %.2e'
          % (n_res, rel(RI - 1.336, RI_ref - 1.336), rel(OR, OR_ref)))
    assert rel(OR, OR_ref) < 1e-6
    assert rel(RI - 1.336, RI_ref - 1.336) < 1e-6


def test_gpu_multi_starts_from_rytov():
    from inverse_solver_amd.rytov import BACKWARD_SOLVER_RYTOV, get_default_parameters as rdef
    from inverse_solver_amd.backward_solver import BACKWARD_SOLVER_MULTI, get_default_parameters
    ri = (16, 16, 17)
    p, u, t = _measured(ri, 1.36, 0.5, 5, iters=12)
    fp = dict(p, return_reflection=True, return_3D=True)
    b = BACKWARD_SOLVER_MULTI(get_default_parameters(dict(forward_solver_parameters=fp, inner_itt=10, itter_max=2,
                                                          verbose=False)))
    init = BACKWARD_SOLVER_RYTOV(rdef(dict(size=list(p['size']), resolution=[0.1] * 3)))
    RI = b.solve(u, t, init_solver=init)
    assert RI.shape == ri and np.all(np.isfinite(RI))
    assert b.err_list[1] < b.err_list[0]
    b.close()

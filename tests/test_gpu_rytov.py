"""GPU parity of the phase unwrapping kernels and the Rytov reconstruction.

* cbs_phase_residues equals the oracle's residue map exactly (same single-precision
  arithmetic) on random maps (many residues), vortices and smooth maps, odd sizes.
* cbs_phase_integrate equals the oracle's raster unwrap exactly, with zero and with
  random step maps.
* unwrap_phase on maps with residues (the branch-cut placement is this library's own):
  the result is congruent to the input modulo 2 pi, and every neighbour difference
  between pixels that are not on a cut is the wrapped difference.
* BACKWARD_SOLVER_RYTOV equals the oracle's restatement (residue-free inputs) to 1e-6.
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


def test_gpu_unwrap_with_residues_is_consistent():
    from inverse_solver_amd.rytov import unwrap_phase, _goldstein_cut_mask
    m = _maps()[1]
    u = unwrap_phase(m)
    k = (u - m) / (2 * np.pi)
    np.testing.assert_allclose(k, np.round(k), atol=1e-4)              # congruent modulo 2 pi
    res = R.residues(m)
    for z in range(m.shape[2]):
        cut = _goldstein_cut_mask(res[:, :, z])
        free = ~cut
        for ax in (0, 1):
            du = np.diff(u[:, :, z], axis=ax)
            dw = np.angle(np.exp(1j * np.diff(m[:, :, z].astype(np.float64), axis=ax)))
            both = free[1:, :] & free[:-1, :] if ax == 0 else free[:, 1:] & free[:, :-1]
            assert np.max(np.abs((du - dw)[both])) < 1e-3, (z, ax)
    # a residue-free map unwraps exactly as the oracle's raster integration
    np.testing.assert_array_equal(unwrap_phase(m[:, :, 2:3]), R.unwrap_residue_free(m[:, :, 2:3]))


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

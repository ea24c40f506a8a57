"""GPU parity of the full 2D-in / 2D-out solve (FORWARD_SOLVER_CONVERGENT_BORN.solve)
against the FP64 oracle: transmitted, reflected and 3D fields.

Tolerances (relative L2, complex64 GPU vs complex128 oracle): 1e-5 for the 3D
and transmitted fields; 1e-4 for the reflected field, which is ~100x weaker
than the transmitted one so its relative error carries the transmitted
field's absolute FP32 rounding."""
import numpy as np
import pytest

from helpers import params, rel

pytestmark = pytest.mark.gpu


def _setup(ri, cyclic=False, iters=16, A=3):
    from inverse_solver_amd.phantoms import bead_ri, circle_illumination
    p = params(ri, iterations=iters)
    if cyclic:
        p['boundary_thickness'] = [0, 0, p['boundary_thickness'][2]]
    RI = bead_ri(ri, 8, 1.46, 1.336)
    u = circle_illumination(ri[:2], A, 1.2, 0.532, [0.1] * 3)
    return p, RI, u


@pytest.mark.parametrize('cyclic,opts', [(False, {}), (False, dict(max_batch=2)), (True, {}),
                                         (False, dict(form=1)), (False, dict(form=2)),
                                         (False, dict(form=2, max_batch=2))])
def test_solve_matches_oracle(cyclic, opts):
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    ri = (16, 16, 16)
    p, RI, u = _setup(ri, cyclic)
    o = Oracle(p)
    o.set_RI(RI)
    t_ref, r_ref, f_ref = o.solve(u)
    s = FORWARD_SOLVER_CONVERGENT_BORN(p, options=opts)
    s.set_RI(RI)
    t, r, f = s.solve(u)
    et, er, ef = rel(t, t_ref), rel(r, r_ref), rel(f, f_ref)
    print('cyclic', cyclic, opts, s.plan(), 'trans %.2e ref %.2e 3d %.2e' % (et, er, ef))
    assert t.shape == t_ref.shape and r.shape == r_ref.shape and f.shape == f_ref.shape
    if opts.get('form') == 2:
        assert s.plan()['form'] == 'rows'
    assert ef < 1e-5
    assert et < 1e-5
    assert er < 1e-4
    s.close()


def test_return_flags_and_repeat_solve():
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    ri = (16, 16, 16)
    p, RI, u = _setup(ri, A=2)
    p['return_reflection'] = False
    p['return_3D'] = False
    s = FORWARD_SOLVER_CONVERGENT_BORN(p)
    s.set_RI(RI)
    t1, r1, f1 = s.solve(u)
    assert r1 is None and f1 is None and t1.shape == (16, 16, 1, 2)
    t2, _, _ = s.solve(u)       # even Bornmax: the shift state returns to its start
    np.testing.assert_allclose(t1, t2, rtol=0, atol=1e-6 * np.abs(t1).max())
    s.close()


@pytest.mark.parametrize('rc', [[1, 0], [0, 2], [0, -2]])
def test_ri_center_refused_like_the_reference(rc):
    """Every 2D solve of the reference stops when RI_center is not [0 0]: crop_field2RI
    (:173-179, reached from :554) indexes outside the field grid for a positive offset, and
    the refocusing product (:550) meets a smaller cropped kernel for a negative one."""
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd._lib import CbsError
    ri = (16, 16, 16)
    p, RI, u = _setup(ri, A=1)
    p['RI_center'] = rc
    s = FORWARD_SOLVER_CONVERGENT_BORN(p)
    s.set_RI(RI)
    with pytest.raises((CbsError, ValueError)):
        s.solve(u)
    s.close()


@pytest.mark.parametrize('ri', [(16, 16, 16), (16, 16, 17)])
def test_octant_kernels_give_the_full_kernel_solve(ri):
    """CudaBorn action 0 hands over first octants (FORWARD_SOLVER_CONVERGENT_BORN_CUDA.m:28-38);
    cbs_set_io_kernels_octant expands them (the symmetry is pinned on the oracle's kernels in
    test_io_kernels.py).  The 2D solve then equals the one from the whole kernels."""
    import ctypes as C
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd._lib import lib, check, CBS_MEM_HOST
    p, RI, u = _setup(ri, A=2)
    full = FORWARD_SOLVER_CONVERGENT_BORN(p)
    full.set_RI(RI)
    t0, r0, f0 = full.solve(u)
    octs = FORWARD_SOLVER_CONVERGENT_BORN(p)
    octs.set_RI(RI)
    octs._ensure_io()                               # builds the whole kernels on the host
    R, T = octs._io_arrays[0], octs._io_arrays[1]
    cx, cy, rz = R.shape
    ro = np.asfortranarray(R[:cx // 2 + 1, :cy // 2 + 1, :rz // 2 + 1])
    to = np.asfortranarray(T[:cx // 2 + 1, :cy // 2 + 1, :rz // 2 + 1])
    check(lib.cbs_set_io_kernels_octant(octs._h, ro.ctypes.data_as(C.c_void_p), to.ctypes.data_as(C.c_void_p),
                                        CBS_MEM_HOST))
    t1, r1, f1 = octs.solve(u)
    e = (rel(t1, t0), rel(r1, r0), rel(f1, f0))
    print(ri, 'octant vs whole kernels: trans %.2e ref %.2e 3d %.2e' % e, flush=True)
    assert e[0] < 1e-6 and e[2] < 1e-6 and e[1] < 1e-5
    full.close()
    octs.close()

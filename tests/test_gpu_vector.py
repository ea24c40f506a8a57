"""GPU parity of the vector (dyadic Green's function) CBS path against the FP64 oracle.

The oracle's vector branch is pinned in tests/test_oracle_vector.py (dense dyadic
brute force, the scalar limit on layered media, the polarisation transforms).
Tolerances as for the scalar path: relative L2 error <= 1e-5 for the Born loop, the
3D fields and the transmitted field; 1e-4 for the (weaker) reflected field.
"""
import numpy as np
import pytest

from helpers import params, sphere_ri, random_field, rel
from oracle.cbs import ForwardSolverConvergentBorn as Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _gpu():
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    return FORWARD_SOLVER_CONVERGENT_BORN


CASES = [
    # (ri_size, iterations, options)
    ((16, 16, 16), 12, dict(form=0)),
    ((16, 16, 16), 12, dict(form=1)),
    ((16, 16, 16), 12, dict(form=0, split_a=6)),
    ((16, 12, 10), 10, dict(form=0)),
    ((17, 17, 19), 8, dict(form=0)),
    ((16, 16, 16), 9, dict(form=2)),
]


@pytest.mark.parametrize('ri_size,iters,opts', CASES)
def test_vector_born_loop_matches_oracle(ri_size, iters, opts):
    p = params(ri_size, iterations=iters, vector_simulation=True)
    RI = sphere_ri(ri_size, 0.5, n_s=1.42)
    fields = [np.stack([random_field(ri_size, 3 * a + j) for j in range(3)], axis=-1) for a in range(2)]
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    ref = np.stack([o.born_loop(f) for f in fields], axis=-1)           # (x, y, z, 3, A)
    s = _gpu()(p, options=opts)
    s.set_RI(RI)
    got = s.solve_incident(np.stack(fields, axis=-1))
    err = rel(got, ref)
    print(ri_size, opts, s.plan(), 'vector rel err', err, flush=True)
    if opts.get('form') == 2:
        assert s.plan()['form'] == 'rows'
    assert got.shape == ref.shape
    assert err < TOL
    s.close()


def test_vector_odd_bornmax_alternation_and_batch():
    ri = (16, 16, 16)
    p = params(ri, iterations=5, vector_simulation=True)
    RI = sphere_ri(ri, 0.6, n_s=1.40)
    fields = [np.stack([random_field(ri, 40 + 3 * a + j) for j in range(3)], axis=-1) for a in range(3)]
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    ref = np.stack([o.born_loop(f) for f in fields], axis=-1)
    s = _gpu()(p, options=dict(max_batch=2))
    s.set_RI(RI)
    got = s.solve_incident(np.stack(fields, axis=-1))
    for a in range(3):
        assert rel(got[..., a], ref[..., a]) < TOL, a
    s.close()


def _circle_pol(ri, n_angles, pol_seed):
    from inverse_solver_amd.phantoms import circle_illumination
    u = circle_illumination(ri[:2], n_angles, 1.2, 0.532, [0.1] * 3)[:, :, 0, :]
    rng = np.random.default_rng(pol_seed)
    w = rng.standard_normal((2, n_angles)) + 1j * rng.standard_normal((2, n_angles))
    return u[:, :, None, :] * w[None, None, :, :]                       # (x, y, 2, A)


@pytest.mark.parametrize('ri,opts', [((16, 16, 17), {}), ((16, 16, 16), dict(form=2))])
def test_vector_solve_2d_matches_oracle(ri, opts):
    p = params(ri, iterations=12, vector_simulation=True)
    RI = sphere_ri(ri, 0.5, n_s=1.41)
    u = _circle_pol(ri, 3, 1)
    o = Oracle(p)
    o.set_RI(RI)
    t_ref, r_ref, f_ref = o.solve(u)
    s = _gpu()(p, options=opts)
    s.set_RI(RI)
    t, r, f = s.solve(u)
    if opts.get('form') == 2:
        assert s.plan()['form'] == 'rows'
    print('vector 2D: trans %.2e  ref %.2e  3D %.2e' % (rel(t, t_ref), rel(r, r_ref), rel(f, f_ref)), flush=True)
    assert t.shape == t_ref.shape == (16, 16, 2, 3)
    assert f.shape == f_ref.shape == ri + (3, 3)
    assert rel(t, t_ref) < TOL
    assert rel(f, f_ref) < TOL
    assert rel(r, r_ref) < 1e-4
    s.close()


def test_vector_gradient_matches_oracle():
    from inverse_solver_amd.backward_solver import BACKWARD_SOLVER_MULTI, get_default_parameters
    from oracle.backward import get_gradiant_static
    ri = (16, 16, 17)
    p = params(ri, iterations=10, vector_simulation=True)
    RI_true = sphere_ri(ri, 0.5, n_s=1.39)
    u = _circle_pol(ri, 3, 2)
    o = Oracle(p)
    o.set_RI(RI_true)
    meas, _, _ = o.solve(u)
    RI0 = sphere_ri(ri, 0.4, n_s=1.37)
    b = BACKWARD_SOLVER_MULTI(get_default_parameters(dict(forward_solver_parameters=p, verbose=False)))
    g, err = b.get_gradiant_static(RI0, u, meas)
    g_ref, err_ref = get_gradiant_static(RI0, u.astype(np.complex128), meas, o)
    print('vector gradient rel err %.2e  misfit %.2e' % (rel(g, g_ref), abs(err - err_ref) / err_ref), flush=True)
    assert rel(g, g_ref) < 1e-4
    assert abs(err - err_ref) < 1e-4 * err_ref
    b.close()


def test_vector_component_sequential_kspace_equals_all_component_kspace():
    # 160^3 simulation grid (128^3 RI + 16-px boundary), a = 40, b = 4: the k-space pass with
    # all three components of an 8-wide x tile in LDS (pass_kspace<4, 160, 8, 1024, 3>) and
    # the one taking the components through LDS in turn with a 16-wide tile
    # (pass_kspace_seq3<4, 160, 16, 1024>) apply the same operations to every value
    ri = (128, 128, 128)
    p = params(ri, iterations=6, vector_simulation=True, boundary_thickness=[8, 8, 8])
    RI = sphere_ri(ri, 0.5, n_s=1.40)
    fields = np.stack([np.stack([random_field(ri, 70 + 3 * a + j) for j in range(3)], axis=-1) for a in range(2)],
                      axis=-1)
    # (8, 1024): the all-component kernel; (16, 1024) and (8, 512): the sequential one
    out = {}
    for t, nt in ((8, 1024), (16, 1024), (8, 512)):
        s = _gpu()(p, options=dict(form=0, split_a=40, tile_t=t, threads_kspace=nt))
        s.set_RI(RI)
        out[t, nt] = s.solve_incident(fields)
        print('tile', t, nt, s.plan(), flush=True)
        assert s.plan()['tile_t'] == t and s.plan()['a'] == 40 and s.plan()['b'] == 4
        s.close()
    np.testing.assert_array_equal(out[16, 1024], out[8, 1024])
    np.testing.assert_array_equal(out[8, 512], out[8, 1024])

"""GPU: grids whose line lengths have no compiled kernel, against the FP64 oracle.

The reference plans its FFTs on whatever simulation size it gets
(Codes/CUDA_CODE/execution.cu:11-34, execution2.cu:11-35; MATLAB fftn), with no
rounding of the size (FORWARD_SOLVER_CONVERGENT_BORN.m:120).  Such grids run the
per-axis form: compiled line FFTs for every 5-smooth length up to 1296 (but 1000, 1125,
1200, 1215, 1250, 1280), run-time-planned Stockham stages of radix 2, 3, 4, 5, 7, 8, 9,
11, 13 and 16 for the other lengths that factor into those, and Bluestein's chirp-z
transform for lengths with a larger prime factor (on a compiled convolution length up
to 1296, so lines up to 648; on run-time Stockham stages above that).  Checked at
reduced iteration counts (<= 1e-5 relative L2 against the oracle):
  * Born loops on grids with factors 7, 11, 13, with prime lengths (97, 83, 19 and the
    661 of the run-time Bluestein), with lengths 2 x a prime, and with 5-smooth lengths
    that have no fused kernel (100, 120, 75);
  * the reference demo's forward geometry (SIMULATION_BEAD.m:11-19: 161 x 161 x 61 at
    0.532 / 4 / 1.2 um, default boundary_thickness 6 -> 183 x 183 x 83, convolution grid
    322 x 322), the whole 2D-in/2D-out solve;
  * its backward geometry (boundary_thickness [2 2 4], SIMULATION_BEAD.m:48 ->
    169 x 169 x 75)."""
import numpy as np
import pytest

from helpers import params, sphere_ri, random_field, plane_wave, rel
from oracle.cbs import ForwardSolverConvergentBorn as Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5
DEMO_RES = 0.532 / 4 / 1.2


def _loop_case(ri, bt_px, iters, nrhs=2, **kw):
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    p = params(ri, bt_px_target=bt_px, iterations=iters, **kw)
    res = p['resolution'][0]
    RI = sphere_ri(ri, 0.3 * min(ri) * res, n_s=1.42, dx=res)
    fields = [plane_wave(ri, 0.15, -0.1, dx=res)] + [random_field(ri, 7 + a) for a in range(nrhs - 1)]
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    refs = [o.born_loop(f) for f in fields]
    s = FORWARD_SOLVER_CONVERGENT_BORN(p)
    s.set_RI(RI)
    assert s.simulation_size == list(o.V.shape)
    got = s.solve_incident(np.stack(fields, axis=-1))
    plan = s.plan()
    errs = [rel(got[..., a], refs[a]) for a in range(nrhs)]
    print(ri, '-> sim', s.simulation_size, plan['form'], 'rel err', errs, flush=True)
    s.close()
    return s.simulation_size, plan, errs


@pytest.mark.parametrize('ri,bt_px', [
    ((34, 36, 44), 4),       # sim 42 x 44 x 52: factors 7, 11, 13 (Stockham)
    ((81, 60, 41), 8),       # sim 97 x 76 x 57: the prime 97, 4 x 19, 3 x 19 (Bluestein)
    ((30, 22, 20), 3),       # sim 36 x 28 x 26: 4 x 7, 2 x 13 with a compiled x length
    ((100, 120, 56), 8),     # sim 116 x 136 x 72: 4 x 29, 8 x 17 (Bluestein), 72 (compiled)
    ((84, 104, 59), 8),      # sim 100 x 120 x 75: compiled line FFTs on every axis
    ((651, 9, 9), 5),        # sim 661 x 19 x 19: 661 (Bluestein on run-time stages), 19 (compiled, L = 48)
])
def test_born_loop_any_grid(ri, bt_px):
    sim, plan, errs = _loop_case(ri, bt_px, iters=6)
    assert plan['form'] == 'plain'
    assert max(errs) < TOL


def test_reference_demo_backward_geometry():
    # SIMULATION_BEAD.m: 161 x 161 x 61 at 0.1108 um, boundary_thickness [2 2 4] -> 169 x 169 x 75
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    ri = (161, 161, 61)
    p = params(ri, iterations=4, boundary_thickness=[2, 2, 4], resolution=[DEMO_RES] * 3)
    RI = sphere_ri(ri, 2.5, n_s=1.4609, dx=DEMO_RES)
    u = plane_wave(ri, 0.1, 0.05, dx=DEMO_RES)
    o = Oracle(p, io_kernels=False, workers=16)
    o.set_RI(RI)
    ref = o.born_loop(u)
    s = FORWARD_SOLVER_CONVERGENT_BORN(p)
    s.set_RI(RI)
    assert s.simulation_size == [169, 169, 75]
    err = rel(s.solve_incident(u), ref)
    print('169 x 169 x 75', s.plan()['form'], 'rel err', err, flush=True)
    s.close()
    assert err < TOL


def test_reference_demo_forward_geometry_2d_solve():
    # SIMULATION_BEAD.m: 161 x 161 x 61 at 0.1108 um, default boundary_thickness 6 ->
    # 183 x 183 x 83, convolution grid 322 x 322; bead of 5 um diameter, n = 1.4609;
    # one normal-incidence plane wave; transmitted, reflected and 3D fields
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd.phantoms import circle_illumination
    ri = (161, 161, 61)
    p = params(ri, iterations=4, boundary_thickness=[6, 6, 6], resolution=[DEMO_RES] * 3)
    RI = sphere_ri(ri, 2.5, n_s=1.4609, dx=DEMO_RES)
    u = circle_illumination(ri[:2], 1, 1.2, 0.532, [DEMO_RES] * 3)
    s = FORWARD_SOLVER_CONVERGENT_BORN(p)
    s.set_RI(RI)
    assert s.simulation_size == [183, 183, 83] and s.conv_size == [322, 322]
    t, r, f = s.solve(u)
    s.close()
    o = Oracle(p, workers=16)
    o.set_RI(RI)
    tr, rr, fr = o.solve(u)
    e = (rel(t, tr), rel(f, fr), rel(r, rr))
    print('183 x 183 x 83: trans %.2e 3D %.2e ref %.2e' % e, flush=True)
    assert e[0] < TOL and e[1] < TOL
    assert e[2] < 1e-4      # the reflected field is ~100x weaker and carries the transmitted field's rounding

"""GPU: the FP64 convergence test (cbs_set_convergence) against the oracle's.

The reference always runs Bornmax iterations; the option stops a right-hand side
after the first iteration n >= 3 with |f_n| <= tol |F_1 + F_2| (norms summed in FP64).
Checked here:
  * the stop iteration of every rhs equals the one the oracle's FP64 norms give, for a
    stop on an even iteration (its increment was deferred, and the next launch adds it
    to Field) and on an odd one, in the fused, four-pass and plain forms;
  * the fields equal the oracle's series truncated there (<= 1e-5), with the persistent
    flip parity of an odd Bornmax carried across the rhs as without the test;
  * a tolerance that never triggers gives the fields of the fixed-Bornmax solve bit for
    bit, eager and replayed from a captured graph;
  * the same stops on a 23^3 grid, whose per-axis form runs the fused Bluestein kernels
    (x inverse + update + x forward, z forward + Green's multiply + z inverse).
The tolerances are set halfway (in log scale) between the oracle's ratio at the stop
and the smallest ratio before it, so FP32 rounding cannot move the decision."""
import numpy as np
import pytest

from helpers import params, sphere_ri, random_field, plane_wave, rel
from oracle.cbs import ForwardSolverConvergentBorn as Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5
RI_SIZE = (16, 16, 16)
M = 31


def _setup(ri=RI_SIZE, radius=0.5, **kw):
    p = params(ri, iterations=M, **kw)
    RI = sphere_ri(ri, radius, n_s=1.40)
    fields = [plane_wave(ri, 0.2, -0.1), plane_wave(ri, -0.1, 0.25), random_field(ri, 5)]
    return p, RI, fields


def _ratios(p, RI, fields):
    """Per rhs, sqrt(|f_n|^2 / |F_1 + F_2|^2) for n = 1..M from the oracle (flip state
    running on from rhs to rhs as in the batched solve)."""
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    out = []
    for f in fields:
        inc, st = [], {}
        o.born_loop(f, increments=inc, tol=0.0, stats=st)   # tol 0: the test only stops on exact zeros
        assert st['iterations'] == M
        out.append(np.sqrt(np.array(inc) / st['fref']))
    return np.array(out)


def _tol_for(r, even):
    """A tolerance whose stop (first n >= 3 with ratio <= tol) is unambiguous and on an
    even or odd n: geometric mean of the ratio at n and the smallest ratio before it."""
    for n in range(6, len(r) + 1):
        if (n % 2 == 0) != even:
            continue
        before = np.min(r[2:n - 1])
        if r[n - 1] < before / 1.2:
            return float(np.sqrt(r[n - 1] * before)), n
    raise AssertionError('no clean stop found')


@pytest.mark.parametrize('form', [0, 2, 1, 3])
@pytest.mark.parametrize('even', [True, False])
def test_stop_iteration_and_fields_match_oracle(form, even):
    _stop_case(_setup(), form, even)


@pytest.mark.parametrize('even', [True, False])
def test_stop_on_a_bluestein_grid(even):
    # 9^3 RI + 7 px = 23^3: every axis a prime length, the per-axis form on Bluestein
    # transforms of compiled length 48, x and z in the fused kernels
    plan = _stop_case(_setup((9, 9, 9), radius=0.4, bt_px_target=7), None, even)
    assert plan['form'] == 'plain'


def _stop_case(setup, form, even):
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    p, RI, fields = setup
    ratios = _ratios(p, RI, fields)
    tol, n0 = _tol_for(ratios[0], even)
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    refs, stops = [], []
    for f in fields:
        st = {}
        refs.append(o.born_loop(f, tol=tol, stats=st))
        stops.append(st['iterations'])
    assert stops[0] == n0 and min(stops) < M
    for a in range(len(fields)):     # no ratio of any rhs sits within FP32 noise of tol
        assert np.min(np.abs(ratios[a, 2:] / tol - 1.0)) > 1e-3, a
    s = FORWARD_SOLVER_CONVERGENT_BORN(p, options=dict(form=form) if form is not None else None)
    s.set_RI(RI)
    s.set_convergence(tol)
    got = s.solve_incident(np.stack(fields, axis=-1))
    its = s.iterations(len(fields))
    print('form', form, s.plan()['form'], 'tol %.3e' % tol, 'stops', list(its), 'oracle', stops, flush=True)
    assert list(its) == stops
    for a in range(len(fields)):
        assert rel(got[..., a], refs[a]) < TOL, a
    # the norms after the stop are zero (those launches skipped the rhs)
    nrm = s.norms(len(fields))
    for a in range(len(fields)):
        assert np.all(nrm[a, stops[a]:] == 0) and np.all(nrm[a, :stops[a]] > 0)
    plan = s.plan()
    s.close()
    return plan


def test_tolerance_that_never_triggers_is_bit_identical():
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    p, RI, fields = _setup()
    u = np.stack(fields, axis=-1)
    base = FORWARD_SOLVER_CONVERGENT_BORN(p)
    base.set_RI(RI)
    # 3 rhs x 31 iterations flip the persistent shift state: consecutive calls differ
    refs = [base.solve_incident(u) for rep in range(3)]
    base.close()
    assert rel(refs[1], refs[0]) > 1e-3
    for opts in (dict(), dict(use_graph=1)):
        s = FORWARD_SOLVER_CONVERGENT_BORN(p, options=opts)
        s.set_RI(RI)
        s.set_convergence(1e-30)
        for rep in range(3):
            got = s.solve_incident(u)
            np.testing.assert_array_equal(got, refs[rep])
            assert list(s.iterations(3)) == [M] * 3
        s.close()


def test_converged_graph_replays_equal_eager():
    """With a stop inside the loop, a captured graph replays to the eager result."""
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    p, RI, fields = _setup()
    u = np.stack(fields[:2], axis=-1)
    ratios = _ratios(p, RI, fields[:1])
    tol, _ = _tol_for(ratios[0], True)
    e = FORWARD_SOLVER_CONVERGENT_BORN(p)
    e.set_RI(RI)
    e.set_convergence(tol)
    eager = e.solve_incident(u)
    g = FORWARD_SOLVER_CONVERGENT_BORN(p, options=dict(use_graph=1))
    g.set_RI(RI)
    g.set_convergence(tol)
    for rep in range(3):
        np.testing.assert_array_equal(g.solve_incident(u), eager)
        e.solve_incident(u[..., :1])
    e.close()
    g.close()

"""GPU parity at the full sizes of the measured configurations, through properties
that do not depend on the size (the FP64 oracle cannot follow every case there):

* 160^3 simulation grid (128^3 RI), 49 right-hand sides in one batch, as the default
  bench line runs it: each rhs solved in the batch equals the same rhs solved alone
  bit for bit; the solve is linear in the incident field; one rhs against the FP64
  oracle (a few iterations: the oracle finishes in seconds).
* 576^2 x 320 simulation grid (512^2 x 256 RI), four-pass form: linearity and batch
  invariance over two rhs.
Tolerances: 1e-5 relative L2 (complex64 against exact linearity / FP64)."""
import numpy as np
import pytest

from helpers import params, rel

pytestmark = pytest.mark.gpu


def _c2():
    from inverse_solver_amd.phantoms import spheres, incident_plane_waves
    ri = (128, 128, 128)
    RI, _ = spheres(ri, 8, (0.5, 2.0), (1.36, 1.48), seed=2)
    u = incident_plane_waves(list(ri), 49, 1.2, 0.532, 1.336, [0.1, 0.1, 0.1])
    p = params(ri, bt_px_target=16, iterations=6, return_reflection=False)
    return p, RI, u


def test_160cubed_49rhs_batch_invariance_linearity_and_oracle():
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    p, RI, u = _c2()
    s = FORWARD_SOLVER_CONVERGENT_BORN(p)
    s.set_RI(RI)
    assert s.simulation_size == [160, 160, 160]
    print('plan', s.plan(), flush=True)
    batch = s.solve_incident(u)                                    # (x, y, z, 49)
    for k in (0, 17, 48):
        alone = s.solve_incident(u[..., k])
        np.testing.assert_array_equal(alone, batch[..., k])
    a, b = 0.7 - 0.2j, -0.4 + 0.9j
    mix = s.solve_incident(a * u[..., 3] + b * u[..., 30])
    lin = rel(mix, a * batch[..., 3] + b * batch[..., 30])
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    ref = o.born_loop(u[..., 5].astype(np.complex128))
    err = rel(batch[..., 5], ref)
    print('160^3 x 49: linearity %.2e, oracle %.2e' % (lin, err), flush=True)
    assert lin < 1e-5
    assert err < 1e-5
    s.close()


@pytest.mark.parametrize('order', [4, 14, 15])
def test_160cubed_49rhs_block_orders_equal(order):
    # at this size the state streams are nontemporal, so order bit 3 (nontemporal k-space
    # loads) takes effect; the numbering of the blocks changes no value
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    p, RI, u = _c2()
    p['iterations_number'] = 3
    outs = []
    for opts in (dict(block_order=0), dict(block_order=order)):
        s = FORWARD_SOLVER_CONVERGENT_BORN(p, options=opts)
        s.set_RI(RI)
        outs.append(s.solve_incident(u))
        assert s.plan()['block_order'] == opts['block_order']
        s.close()
    np.testing.assert_array_equal(outs[1], outs[0])


def test_576sq_x320_fourpass_linearity_and_batch():
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd.phantoms import spheres, incident_plane_waves
    ri = (512, 512, 256)
    RI, _ = spheres(ri, 6, (1.0, 3.0), (1.36, 1.44), seed=5)
    u = incident_plane_waves(list(ri), 2, 1.2, 0.532, 1.336, [0.1, 0.1, 0.1])
    p = params(ri, bt_px_target=32, iterations=4, return_reflection=False)
    s = FORWARD_SOLVER_CONVERGENT_BORN(p)
    s.set_RI(RI)
    assert s.simulation_size == [576, 576, 320]
    assert s.plan()['form'] == 'rows'
    batch = s.solve_incident(u)
    alone = s.solve_incident(u[..., 1])
    np.testing.assert_array_equal(alone, batch[..., 1])
    a, b = 1.3 + 0.1j, 0.2 - 0.8j
    mix = s.solve_incident(a * u[..., 0] + b * u[..., 1])
    lin = rel(mix, a * batch[..., 0] + b * batch[..., 1])
    print('576^2 x 320 four-pass: linearity %.2e' % lin, flush=True)
    assert lin < 1e-5
    s.close()

"""GPU: a seeded sweep of small random grids against the FP64 oracle.

The planner sends a grid to the fused, four-pass or per-axis form, and the per-axis form
transforms each axis with a compiled line FFT, a Bluestein transform on a compiled
convolution length or run-time Stockham stages, with the x and z passes fused around the
update and the Green's multiply where it can.  Fourteen RI sizes drawn from [8, 34) per
axis with 2-7 boundary pixels (prime, 7-, 11-, 17- and 19-smooth and 5-smooth simulation
lengths) each run in the planner's form and in the unfused per-axis form, two rhs (a
plane wave and a random field) with an odd iteration count, so the persistent flip
parity is carried across the batch: <= 1e-5 relative L2 against the oracle."""
import numpy as np
import pytest

from helpers import params, sphere_ri, random_field, plane_wave, rel
from oracle.cbs import ForwardSolverConvergentBorn as Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _cases(n=14, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        ri = tuple(int(v) for v in rng.integers(8, 34, 3))
        out.append((ri, int(rng.integers(2, 8))))
    return out


@pytest.mark.parametrize('ri,bt_px', _cases())
def test_random_grid_matches_oracle(ri, bt_px):
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    p = params(ri, bt_px_target=bt_px, iterations=5)
    res = p['resolution'][0]
    RI = sphere_ri(ri, 0.3 * min(ri) * res, n_s=1.42, dx=res)
    fields = [plane_wave(ri, 0.1, -0.05, dx=res), random_field(ri, sum(ri))]
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    refs = [o.born_loop(f) for f in fields]
    for opts in (None, dict(form=3)):
        s = FORWARD_SOLVER_CONVERGENT_BORN(p, options=opts)
        s.set_RI(RI)
        assert s.simulation_size == list(o.V.shape)
        got = s.solve_incident(np.stack(fields, axis=-1))
        form = s.plan()['form']
        s.close()
        errs = [rel(got[..., a], refs[a]) for a in range(len(fields))]
        print(ri, bt_px, '-> sim', list(o.V.shape), form, 'rel err', ['%.1e' % e for e in errs], flush=True)
        assert max(errs) < TOL, (form, errs)

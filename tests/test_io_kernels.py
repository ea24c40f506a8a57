"""The product's host construction of the refocusing / transmission kernels
(inverse_solver_amd/io_kernels.py) equals the oracle's line-by-line restatement
of truncated_green_plus and the constructor (oracle/hankel.py, oracle/cbs.py)."""
import numpy as np
import pytest

from helpers import params, rel


@pytest.mark.parametrize('cyclic,rc', [(False, [0, 0]), (True, [0, 0]), (False, [2, -1]), (False, [-3, 1])])
def test_io_kernels_match_oracle(cyclic, rc):
    from inverse_solver_amd.io_kernels import build_io_kernels
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    ri = (12, 10, 8)
    p = params(ri, bt_px_target=2, iterations=4, RI_center=rc)
    if cyclic:
        p['boundary_thickness'] = [0, 0, 1.0]
    o = Oracle(p)
    ref, kt, kr = build_io_kernels(o.parameters, o.expected_RI_size, o.cyclic_boundary_xy)
    assert rel(ref, o.refocusing_util) < 1e-10
    assert rel(kt, o.kernel_trans) < 1e-10
    assert rel(kr, o.kernel_ref) < 1e-10


def test_bessel_zeros_product_vs_reference_table():
    import os
    from inverse_solver_amd.io_kernels import bessel_j0_zeros
    c = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'dht_bessel_zeros.npz'))['c']
    np.testing.assert_allclose(bessel_j0_zeros(c.shape[1]), c[0], rtol=0, atol=1e-11)


def test_oracle_refuses_ri_center_in_solve():
    """crop_field2RI (:173-179) crops parameters.size from an array of that size, so a
    nonzero RI_center indexes outside it and the reference's 2D solve stops (:554)."""
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    from helpers import sphere_ri
    from inverse_solver_amd.phantoms import circle_illumination
    ri = (12, 12, 8)
    for rc in ([1, 0], [0, 2], [0, -2]):
        o = Oracle(params(ri, bt_px_target=2, iterations=2, RI_center=rc))
        o.set_RI(sphere_ri(ri, 0.3, n_s=1.4))
        # a positive offset stops in the crop (:554), a negative one already in the
        # refocusing product (:550), whose cropped kernel is smaller than the grid
        with pytest.raises((IndexError, ValueError)):
            o.solve(circle_illumination(ri[:2], 1, 1.2, 0.532, [0.1] * 3))
